"""MLP surrogate on the GPU.

- forward values, r-derivatives and parameter gradients of the GPU kernels
  (bf16 MFMA, fp32 accumulation) against the FP64 oracle that emulates the
  same bf16 rounding points (oracle/mlp_oracle.py, itself pinned by finite
  differences on the CPU);
- the switch and the energy shift of the device evaluation;
- training on the GPU: the trained network's force error against the FP64
  tables, and what it gains over the method without it: the configured
  plain-cutoff table drifts in NVE, the network switched off to zero at the
  cutoff conserves its energy;
- the surrogate pair pass (engine_flag_mlp): same pairs and exclusions as
  the table pass, excluded pairs cancel exactly, Newton's third law,
  bitwise repeatable, force error against the table pass bounded, and
  forces and energy equal to the network evaluated pair by pair over a
  brute force of the pairs.
"""
import numpy as np
import pytest

import mdcore as md
import mlp_oracle as mo
import surrogate as sg
import systems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix3():
    sysm = systems.mixture(3000, seed=4, nspecies=3)
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    dom = sg.pair_domains(eng, sysm)
    eng.mlp_init(*dom, seed=3)
    yield sysm, eng, dom
    eng.close()


def table_energy(eng, sysm, r, p):
    pots, _, _ = sg.pair_potentials(eng, sysm)
    E = np.zeros(len(r))
    F = np.zeros(len(r))
    for k in range(len(pots)):
        m = p == k
        if m.any():
            E[m], F[m] = pots[k].eval_r(r[m])
    return E, F


def domain_samples(dom, n, seed):
    rng = np.random.default_rng(seed)
    p = rng.integers(0, len(dom[0]), n)
    r = (dom[0][p] + (dom[1][p] - dom[0][p]) * rng.random(n)).astype(np.float32).astype(np.float64)
    return r, p


def test_forward_matches_oracle(mix3):
    """y and yd of the device against the oracle, read through E and dE/dr
    without a switch: y = asinh((E + shift) / sE), yd = dE/dr / (sE cosh y);
    the shift makes the unswitched energy at r_hi the table's."""
    sysm, eng, dom = mix3
    npair = len(dom[0])
    theta = eng.mlp_params().astype(np.float64)
    sc = np.stack(dom, 1)
    assert np.all(sc[:, 3] >= sc[:, 1])
    r, p = domain_samples(dom, 4096, 1)
    E, F = eng.mlp_eval(r, p)
    et_hi, _ = table_energy(eng, sysm, dom[1], np.arange(npair))
    sh = mo.shifts(theta, npair, sc, et_hi, rounding=True)
    o = mo.forward(theta, npair, sc, r, p, rounding=True, shift=sh)
    y = np.arcsinh((E + sh[p]) / sc[p, 2])
    yd = F / (sc[p, 2] * np.cosh(y))
    dy = y - o["y"]
    dyd = (yd - o["yd"]) / np.sqrt(np.mean(o["yd"] ** 2))
    print(f"forward: rms dy {np.sqrt(np.mean(dy ** 2)):.2e} max {np.abs(dy).max():.2e}; "
          f"rel rms dyd {np.sqrt(np.mean(dyd ** 2)):.2e} max {np.abs(dyd).max():.2e}", flush=True)
    # fp32 vs fp64 before the bf16 rounding points flips a rounding now and then
    assert np.sqrt(np.mean(dy ** 2)) < 1e-4 and np.abs(dy).max() < 5e-3
    assert np.sqrt(np.mean(dyd ** 2)) < 1e-3 and np.abs(dyd).max() < 2e-2
    # the energy at r_hi is the table's
    Eh, _ = eng.mlp_eval(dom[1], np.arange(npair))
    assert np.allclose(Eh, et_hi, rtol=1e-5, atol=1e-5 * sc[:, 2].max())


def test_switch_and_shift_on_device(mix3):
    """Same weights (same seed) with and without a switch from r_hi - 0.3 nm:
    E_s = S E_u and dE_s/dr = S' E_u + S dE_u/dr, S from the reference's
    switch; both are zero at r_hi."""
    sysm, eng, dom = mix3
    npair = len(dom[0])
    r, p = domain_samples(dom, 4096, 5)
    eng.mlp_init(*dom, seed=3)
    Eu, Fu = eng.mlp_eval(r, p)
    rsw = dom[1] - 0.3
    try:
        eng.mlp_init(dom[0], dom[1], dom[2], rsw, seed=3)
        Es, Fs = eng.mlp_eval(r, p)
        Eh, Fh = eng.mlp_eval(dom[1], np.arange(npair))
    finally:
        eng.mlp_init(*dom, seed=3)
    S, dS = mo.switch(r, rsw[p], dom[1][p])
    assert (S < 1).sum() > 100 and (S == 1).sum() > 100
    scale = np.abs(Eu) + 1e-3
    assert np.all(np.abs(Es - S * Eu) <= 1e-5 * scale)
    fsc = np.abs(Fu) + np.abs(dS * Eu) + 1e-3
    assert np.all(np.abs(Fs - (dS * Eu + S * Fu)) <= 1e-4 * fsc)
    assert np.all(Eh == 0.0) and np.all(Fh == 0.0)


def test_gradient_matches_oracle(mix3):
    sysm, eng, dom = mix3
    npair = len(dom[0])
    theta = eng.mlp_params().astype(np.float64)
    sc = np.stack(dom, 1)
    rng = np.random.default_rng(2)
    n = 4096
    p = np.repeat(rng.integers(0, npair, n // 16), 16)
    r = (dom[0][p] + (dom[1][p] - dom[0][p]) * rng.random(n)).astype(np.float32).astype(np.float64)
    E, F = table_energy(eng, sysm, r, p)
    E = E.astype(np.float32).astype(np.float64)
    F = F.astype(np.float32).astype(np.float64)
    g, loss = eng.mlp_grad(r, p, E, F, 0.5, 2.0)
    lo, go = mo.loss_and_grad(theta, npair, sc, r, p, E, F, 0.5, 2.0, rounding=True)
    assert loss == pytest.approx(lo, rel=1e-5)
    off = mo.offsets(npair)
    for a, b in zip(mo.BLOCKS[:-1], mo.BLOCKS[1:]):
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


def test_init_rejects_switch_below_domain(mix3):
    _, eng, dom = mix3
    with pytest.raises(md.MdcoreError):
        eng.mlp_init(dom[0], dom[1], dom[2], dom[0] * 0.5, seed=3)
    md.library().errs_clear()
    eng.mlp_init(*dom, seed=3)


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
    hist = eng.mlp_train(*S, steps=20000, batch_tiles=1024, lr0=1e-2, lr1=1e-5, seed=11)
    yield {"sys": sysm, "eng": eng, "dom": dom, "f_table": f_table, "e_table": e_table, "hist": hist}
    eng.close()


def rel_rmse(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_trained_network_force_error(water_trained):
    """Training on the GPU brings the pair-pass forces within 3 % (relative
    RMS) of the FP64-built tables' pass.  (The box's total energy is a small
    residual of large pair energies of both signs, so it is only printed.)"""
    w = water_trained
    eng = w["eng"]
    eng.setflags(md.engine_flag_mlp, 0)
    eng.nonbond()
    f_trained = eng.forces()
    r = eng.gpu_results()
    eng.setflags(0, md.engine_flag_mlp)
    e1 = rel_rmse(f_trained, w["f_table"])
    tot = r["epot"] + r["epot_exclusion"]
    tot_t = w["e_table"]["epot"] + w["e_table"]["epot_exclusion"]
    print(f"force rel RMSE vs tables {e1:.3e}; energy {tot:.6e} vs {tot_t:.6e}; loss "
          f"{w['hist'][:20].mean():.3e} -> {w['hist'][-50:].mean():.3e}", flush=True)
    assert w["hist"][-50:].mean() < 1e-3 * w["hist"][:20].mean()
    assert e1 < 0.03


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


def test_pass_equals_the_network_pair_by_pair(mix3):
    """The surrogate pass against the network evaluated pair by pair by the
    evaluation kernel (engine_mlp_eval) over a host brute force of the pairs
    within the cutoff, in the projected window (which takes every such pair):
    per-atom forces and the energy agree to the rounding of r.  The pass's
    dense lists here run to several 64-pair chunks with partial tiles, so a
    pair given another pair's inputs or results would show at once."""
    sysm, _, dom = mix3
    eng = systems.build_engine(sysm, md.engine_flag_gpu | md.engine_flag_projwindow)
    eng.mlp_init(*dom, seed=3)
    eng.setflags(md.engine_flag_mlp, 0)
    eng.shuffle()
    eng.nonbond()
    f = eng.forces()
    res = eng.gpu_results()
    x = np.asarray(sysm.x, np.float64).reshape(-1, 3)
    box = np.asarray(sysm.dim, np.float64)
    _, ti, tj = sg.pair_potentials(eng, sysm)
    mt = int(max(ti.max(), tj.max(), np.max(sysm.types))) + 1
    pid = -np.ones((mt, mt), np.int64)
    for k, (a, b) in enumerate(zip(ti, tj)):
        pid[a, b] = pid[b, a] = k
    I, J = np.triu_indices(len(x), 1)
    d = x[I] - x[J]
    d -= box * np.round(d / box)
    r2 = np.einsum("ij,ij->i", d, d)
    m = r2 <= sysm.cutoff ** 2
    I, J, d, r = I[m], J[m], d[m], np.sqrt(r2[m])
    p = pid[sysm.types[I], sysm.types[J]]
    assert np.all(p >= 0) and len(r) == res["pairs"]
    E, F = table_energy(eng, sysm, r, p)
    net = r >= dom[0][p]
    En, Fn = eng.mlp_eval(r[net], p[net])
    eng.close()
    E[net], F[net] = En, Fn
    assert net.mean() > 0.9
    fh = np.zeros_like(x)
    g = (F / r)[:, None] * d
    np.add.at(fh, I, -g)
    np.add.at(fh, J, g)
    err = rel_rmse(f, fh)
    print(f"{len(r)} pairs, {net.sum()} through the network: force rel RMSE {err:.2e}, energy "
          f"{res['epot']:.9e} vs {E.sum():.9e}", flush=True)
    # measured 1.6e-6 and 1.6e-7
    assert err <= 2e-5
    assert abs(res["epot"] - E.sum()) <= 1e-5 * np.abs(E).sum()


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


def test_gain_over_plain_table_energy_drift():
    """What the trained network gains over the method without it.  The
    argon system's configured table (LJ 12-6 cut at 1.0 nm without a shift)
    drifts in NVE: every pair crossing the cutoff jumps in energy.  The
    network trained on that same table and switched off to zero between
    0.9 nm and the cutoff has force = -dE/dr exactly and zero energy and
    force at the cutoff, and conserves its energy over the same 20 ps (32,000 atoms from the fcc
    start at 100 K, dt 2 fs)."""
    sysm = systems.argon(20)
    t = sg.nve_drift(sysm, equil=1000, steps=10000)
    n = sg.nve_drift(sysm, mlp=True, switch=0.9, equil=1000, steps=10000)
    print(f"drift per ns, relative: plain table {t['drift_rel_per_ns']:.2e} "
          f"(halves {t['drift_rel_per_ns_halves'][0]:.2e}, {t['drift_rel_per_ns_halves'][1]:.2e}); "
          f"switched network {n['drift_rel_per_ns']:.2e} "
          f"(halves {n['drift_rel_per_ns_halves'][0]:.2e}, {n['drift_rel_per_ns_halves'][1]:.2e})", flush=True)
    assert abs(n["drift_rel_per_ns"]) < 1e-4
    assert max(abs(v) for v in n["drift_rel_per_ns_halves"]) < 2e-4
    assert abs(t["drift_rel_per_ns"]) > 5 * abs(n["drift_rel_per_ns"])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_switched_network_drift_over_seeds(seed):
    """The switched network's energy conservation does not hang on one fit:
    per seed its own Maxwell start, training samples, initial weights and
    batch order; 20 ps of 32,000 argon atoms at dt 2 fs drift below 1e-4 per
    ns overall and in each half."""
    sysm = systems.argon(20, seed=6178 + seed)
    n = sg.nve_drift(sysm, mlp=True, switch=0.9, equil=1000, steps=10000, seed=100 + seed,
                     sample_seed=200 + seed, train_seed=300 + seed)
    print(f"seed {seed}: switched network drift {n['drift_rel_per_ns']:.2e}/ns "
          f"(halves {n['drift_rel_per_ns_halves'][0]:.2e}, {n['drift_rel_per_ns_halves'][1]:.2e})", flush=True)
    assert abs(n["drift_rel_per_ns"]) < 1e-4
    assert max(abs(v) for v in n["drift_rel_per_ns_halves"]) < 1e-4
