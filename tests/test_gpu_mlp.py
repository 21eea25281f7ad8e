"""MLP surrogate on the GPU.

- forward values and parameter gradients of the GPU kernels (bf16 MFMA,
  fp32 accumulation) against the FP64 oracle that emulates the same bf16
  rounding points (oracle/mlp_oracle.py, itself pinned by finite
  differences on the CPU);
- training on the GPU: what the trained network gains over the same
  network untrained, measured against the FP64 tables;
- the surrogate pair pass (engine_flag_mlp): same pairs and exclusions as
  the table pass, excluded pairs cancel exactly, Newton's third law,
  bitwise repeatable, force error against the table pass bounded.
"""
import numpy as np
import pytest

import mdcore as md
import mlp_oracle as mo
import surrogate as sg
import systems

pytestmark = pytest.mark.gpu

BLOCKS = ["W1", "B1", "E1", "W2", "B2", "EMB", "W3", "B3", "N"]


@pytest.fixture(scope="module")
def mix3():
    sysm = systems.mixture(3000, seed=4, nspecies=3)
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    dom = sg.pair_domains(eng, sysm)
    eng.mlp_init(*dom, seed=3)
    yield sysm, eng, dom
    eng.close()


def test_forward_matches_oracle(mix3):
    sysm, eng, dom = mix3
    npair = len(dom[0])
    theta = eng.mlp_params().astype(np.float64)
    sc = np.stack(dom, 1)
    rng = np.random.default_rng(1)
    p = rng.integers(0, npair, 4096)
    r = (dom[0][p] + (dom[1][p] - dom[0][p]) * rng.random(4096)).astype(np.float32).astype(np.float64)
    E, F = eng.mlp_eval(r, p)
    o = mo.forward(theta, npair, sc, r, p, rounding=True)
    dy0 = np.arcsinh(E / sc[p, 2]) - o["y0"]
    dy1 = np.arcsinh(F / sc[p, 3]) - o["y1"]
    print(f"forward: rms dy0 {np.sqrt(np.mean(dy0 ** 2)):.2e} max {np.abs(dy0).max():.2e}; "
          f"rms dy1 {np.sqrt(np.mean(dy1 ** 2)):.2e} max {np.abs(dy1).max():.2e}", flush=True)
    # fp32 vs fp64 before the bf16 rounding points flips a rounding now and then
    assert np.sqrt(np.mean(dy0 ** 2)) < 1e-4 and np.sqrt(np.mean(dy1 ** 2)) < 1e-4
    assert np.abs(dy0).max() < 5e-3 and np.abs(dy1).max() < 5e-3


def test_gradient_matches_oracle(mix3):
    sysm, eng, dom = mix3
    npair = len(dom[0])
    theta = eng.mlp_params().astype(np.float64)
    sc = np.stack(dom, 1)
    rng = np.random.default_rng(2)
    n = 4096
    p = np.repeat(rng.integers(0, npair, n // 16), 16)
    r = (dom[0][p] + (dom[1][p] - dom[0][p]) * rng.random(n)).astype(np.float32).astype(np.float64)
    pots, _, _ = sg.pair_potentials(eng, sysm)
    E = np.zeros(n)
    F = np.zeros(n)
    for k in range(npair):
        m = p == k
        E[m], F[m] = pots[k].eval_r(r[m])
    E = E.astype(np.float32).astype(np.float64)
    F = F.astype(np.float32).astype(np.float64)
    g, loss = eng.mlp_grad(r, p, E, F, 0.5, 2.0)
    lo, go = mo.loss_and_grad(theta, npair, sc, r, p, E, F, 0.5, 2.0, rounding=True)
    assert loss == pytest.approx(lo, rel=1e-5)
    off = mo.offsets(npair)
    for a, b in zip(BLOCKS[:-1], BLOCKS[1:]):
        ref = go[off[a]:off[b]]
        err = np.linalg.norm(g[off[a]:off[b]] - ref) / np.linalg.norm(ref)
        print(f"gradient {a}: rel L2 {err:.2e}", flush=True)
        assert err < 1e-4, a


def test_grad_rejects_mixed_tiles(mix3):
    _, eng, dom = mix3
    r = np.full(16, 0.5)
    p = np.zeros(16, np.int32)
    p[3] = 1
    with pytest.raises(md.MdcoreError):
        eng.mlp_grad(r, p, np.zeros(16), np.zeros(16))
    md.library().errs_clear()


@pytest.fixture(scope="module")
def water_trained():
    sysm = systems.water(1000, 3.1, (10, 10, 10), seed=2)
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    eng.shuffle()
    eng.nonbond()
    f_table = eng.forces()
    e_table = eng.gpu_results()
    dom = sg.pair_domains(eng, sysm)
    S = sg.make_samples(eng, sysm, 1 << 21, dom[0], dom[1], configs=[sysm.x], seed=5)
    eng.mlp_init(*dom, seed=7)
    eng.setflags(md.engine_flag_mlp, 0)
    eng.nonbond()
    f_untrained = eng.forces()
    eng.setflags(0, md.engine_flag_mlp)
    hist = eng.mlp_train(*S, steps=20000, batch_tiles=1024, lr0=1e-2, lr1=1e-5, seed=11)
    yield {"sys": sysm, "eng": eng, "dom": dom, "f_table": f_table, "e_table": e_table,
           "f_untrained": f_untrained, "hist": hist}
    eng.close()


def rel_rmse(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_training_gain_over_untrained_network(water_trained):
    """Same network, same kernels: training on the GPU cuts the pair-pass
    force error against the FP64-built tables by orders of magnitude."""
    w = water_trained
    eng = w["eng"]
    eng.setflags(md.engine_flag_mlp, 0)
    eng.nonbond()
    f_trained = eng.forces()
    eng.setflags(0, md.engine_flag_mlp)
    e0 = rel_rmse(w["f_untrained"], w["f_table"])
    e1 = rel_rmse(f_trained, w["f_table"])
    print(f"force rel RMSE vs tables: untrained {e0:.3e}, trained {e1:.3e}; loss {w['hist'][:20].mean():.3e} -> "
          f"{w['hist'][-50:].mean():.3e}", flush=True)
    assert w["hist"][-50:].mean() < 1e-3 * w["hist"][:20].mean()
    assert e1 < 0.03
    assert e1 < e0 / 30


def test_mlp_pass_counts_exclusions_newton_repeatable(water_trained):
    w = water_trained
    eng = w["eng"]
    eng.setflags(md.engine_flag_mlp, 0)
    eng.nonbond()
    f1 = eng.forces()
    r1 = eng.gpu_results()
    eng.nonbond()
    f2 = eng.forces()
    eng.setflags(0, md.engine_flag_mlp)
    assert np.array_equal(f1, f2)
    assert r1["pairs"] == w["e_table"]["pairs"]
    assert r1["exclusions"] == w["e_table"]["exclusions"]
    assert np.abs(f1.sum(0)).max() <= 1e-6 * np.abs(f1).sum(0).max()


def test_isolated_excluded_pair_cancels_through_the_network():
    """One water molecule alone: every pair is excluded.  With the network
    evaluating them (r >= r_lo), the pair pass and the exclusion pass see
    bitwise the same pair inputs, so they cancel to rounding of the sums."""
    sysm = systems.water(1, 3.0, (1, 1, 1), seed=1, jitter=0.0)
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    dom = sg.pair_domains(eng, sysm)
    ti, tj = eng.mlp_pairs()
    assert np.all(dom[0][(ti == 2) | (tj == 2)] <= 0.1)   # O-H (0.1 nm) and H-H go through the network
    eng.mlp_init(*dom, seed=5)
    eng.setflags(md.engine_flag_mlp, 0)
    eng.shuffle()
    eng.nonbond()
    r = eng.gpu_results()
    f = eng.forces()
    eng.close()
    assert r["pairs"] == 3 and r["exclusions"] == 3
    assert abs(r["epot"]) > 0.0
    assert abs(r["epot"] + r["epot_exclusion"]) <= 1e-6 * abs(r["epot"])
    assert np.abs(f).max() <= 1e-3


def test_mlp_flag_requires_init():
    sysm = systems.argon(6, jitter=0.05)
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    eng.setflags(md.engine_flag_mlp, 0)
    with pytest.raises(md.MdcoreError):
        eng.nonbond()
    md.library().errs_clear()
    eng.close()


def test_mlp_half_and_full_shell_agree(water_trained):
    """The surrogate pass on the half shell (the default: each pair's network
    evaluation once, opposite forces folded per atom) and on the full shell
    (engine_flag_fullshell: each pair at both cells) count the same pairs and
    agree to FP32 summation order."""
    w = water_trained
    eng = w["eng"]
    out = {}
    for label, flags in (("half", 0), ("full", md.engine_flag_fullshell)):
        eng.setflags(md.engine_flag_mlp | flags, md.engine_flag_fullshell if flags == 0 else 0)
        eng.nonbond()
        out[label] = (eng.forces().copy(), eng.gpu_results())
    eng.setflags(0, md.engine_flag_mlp | md.engine_flag_fullshell)
    (fh, rh), (ff, rf) = out["half"], out["full"]
    assert rh["pairs"] == rf["pairs"] and rh["exclusions"] == rf["exclusions"]
    assert rel_rmse(fh, ff) <= 1e-5
    assert rh["epot"] == pytest.approx(rf["epot"], rel=1e-6)
    assert rh["virial"] == pytest.approx(rf["virial"], rel=1e-6)
