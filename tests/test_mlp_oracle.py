"""The FP64 reference of the MLP surrogate (oracle/mlp_oracle.py), pinned
against what does not depend on it:

- bf16 rounding: torch's float32 -> bfloat16 conversion (round to nearest,
  ties to even) on random values and on hand-made ties;
- the forward-mode derivative: yd against central finite differences of y
  in r, and dE/dr (switch and shift included) against those of E;
- the switch: the reference's own compiled switched LJ (potential.c:69-100
  through potential_create_LJ126_switch_f/_dfdr in oracle/_ref) and its
  end conditions S(a) = 1, S(b) = 0, S'(a) = S'(b) = 0;
- gradients: central finite differences of the loss in float64 (rounding
  off, so the loss is smooth), for every parameter block, with weights on
  both the energy and the force term, so the second-derivative terms of
  the force loss are pinned too;
- the network output: a hand-computed 1-unit special case (zero weights
  everywhere but one path), values and derivative;
- the bf16 pair hi + lo that carries layer 2's inputs: within 2^-16 of the
  FP32 value.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import mlp_oracle as mo
import pyoracle as po

NP = 3


def random_theta(seed, npair=NP):
    rng = np.random.default_rng(seed)
    o = mo.offsets(npair)
    t = rng.normal(0.0, 0.3, o["N"])
    return t


def random_samples(seed, n, npair=NP):
    rng = np.random.default_rng(seed)
    sc = np.array([[0.2, 1.0, 1.0, 1.0], [0.15, 1.2, 2.0, 0.9], [0.3, 0.9, 0.5, 0.9]])[:npair]
    p = rng.integers(0, npair, n)
    r = sc[p, 0] + (sc[p, 1] - sc[p, 0]) * rng.random(n)
    E = rng.normal(0, 5, n)
    F = rng.normal(0, 50, n)
    return sc, r, p, E, F


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e4, 1000), rng.normal(0, 1e-4, 1000)]).astype(np.float32)
    # exact ties: 1 + 2^-8 (-> 1, even), 1 + 3 * 2^-8 (-> 1 + 2^-6, even), and their negatives
    ties = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), -(1 + 3 * 2 ** -8)], np.float32)
    x = np.concatenate([x, ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), ref)
    assert list(mo.bf16(ties)) == [1.0, 1 + 2 ** -6, -1.0, -(1 + 2 ** -6)]


def test_single_path_closed_form():
    """One unit on each layer, with layer 2's input as a bf16 pair hi + lo:
    y  = w3 silu(w2 split(silu(z1)) + b2) + b3,   z1 = w1 rn + b1,
    yd = w3 silu'(z2) w2 split(silu'(z1) w1 drn)."""
    npair = 1
    P = {k: np.zeros_like(v) for k, v in mo.unpack(np.zeros(mo.offsets(npair)["N"]), npair).items()}
    P["W1"][5, 0] = 0.7
    P["b1"][5] = 0.1
    P["W2"][9, 5] = 1.25
    P["b2"][9] = -0.2
    P["W3"][0, 9] = 2.0
    P["b3"][0] = 0.3
    theta = mo.pack(P, npair)
    sc = np.array([[0.2, 1.0, 3.0, 1.0]])
    r = np.array([0.45])
    rn = (2 * 0.45 - 0.2 - 1.0) / 0.8
    drn = 2.0 / 0.8

    def sig(z):
        return 1 / (1 + np.exp(-z))

    def silu(z):
        return z * sig(z)

    def dsilu(z):
        return sig(z) + z * sig(z) * (1 - sig(z))

    def tbf(v):
        return float(torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).float())

    def split(v):
        hi = tbf(v)
        return hi + tbf(float(np.float32(v) - np.float32(hi)))

    z1 = 0.7 * rn + 0.1
    h1 = split(silu(z1))
    g1 = split(dsilu(z1) * 0.7 * drn)
    z2 = 1.25 * h1 - 0.2
    y = 2.0 * silu(z2) + 0.3
    yd = 2.0 * dsilu(z2) * 1.25 * g1
    f = mo.forward(theta, npair, sc, r, [0], rounding=True)
    assert f["y"][0] == pytest.approx(y, rel=1e-12)
    assert f["yd"][0] == pytest.approx(yd, rel=1e-12)
    assert f["E"][0] == pytest.approx(3.0 * np.sinh(y), rel=1e-12)
    assert f["dEdr"][0] == pytest.approx(3.0 * np.cosh(y) * yd, rel=1e-12)


@pytest.mark.parametrize("seed", [1, 2])
def test_forward_mode_matches_finite_differences(seed):
    """yd = dy/dr and dE/dr (switch and shift included) are the r-derivatives
    of y and E."""
    theta = random_theta(seed)
    sc, r, p, _, _ = random_samples(seed + 10, 64)
    sc = sc.copy()
    sc[:, 3] = np.minimum(sc[:, 3], sc[:, 1] - 0.15)   # every pair switched over its last 0.15 nm
    r = np.clip(r, sc[p, 0] + 1e-3, sc[p, 1] - 1e-3)
    shift = np.random.default_rng(seed).normal(0, 0.5, NP)
    h = 1e-6
    f = mo.forward(theta, NP, sc, r, p, rounding=False, shift=shift)
    fp = mo.forward(theta, NP, sc, r + h, p, rounding=False, shift=shift)
    fm = mo.forward(theta, NP, sc, r - h, p, rounding=False, shift=shift)
    fd_y = (fp["y"] - fm["y"]) / (2 * h)
    fd_e = (fp["E"] - fm["E"]) / (2 * h)
    assert np.allclose(f["yd"], fd_y, rtol=1e-6, atol=1e-7 * np.abs(fd_y).max())
    assert np.allclose(f["dEdr"], fd_e, rtol=1e-6, atol=1e-7 * np.abs(fd_e).max())
    # the switch is live somewhere in the sample, and so is the shift
    S, _ = mo.switch(r, sc[p, 3], sc[p, 1])
    assert (S < 1.0).any() and (S > 0.0).any()


def test_switch_end_conditions():
    a, b = 0.9, 1.2
    r = np.array([0.5, a, b, 1.5])
    S, dS = mo.switch(r, a, b)
    assert list(S) == [1.0, 1.0, 0.0, 0.0]
    assert list(dS) == [0.0, 0.0, 0.0, 0.0]
    rr = np.linspace(a, b, 1001)
    S, dS = mo.switch(rr, a, b)
    assert np.all(np.diff(S) <= 0.0)
    h = 1e-7
    fd = (mo.switch(rr[1:-1] + h, a, b)[0] - mo.switch(rr[1:-1] - h, a, b)[0]) / (2 * h)
    assert np.allclose(dS[1:-1], fd, atol=1e-6)
    assert abs(dS[0]) < 1e-12 and abs(dS[-1]) < 1e-12
    # a >= b: no switch
    S, dS = mo.switch(rr, b, b)
    assert np.all(S == 1.0) and np.all(dS == 0.0)


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_switch_matches_reference_build():
    """The reference's switched LJ (potential.c:855-867) divided by its plain
    LJ gives its switch S(r); its derivative gives S'(r)."""
    lib = po.ref()
    A, B, s, rc = 9.677e-6, 6.209e-3, 0.9, 1.0
    for name, v in (("A", A), ("B", B), ("s", s), ("cutoff", rc)):
        C.c_double.in_dll(lib, f"potential_create_LJ126_switch_{name}").value = v
    f = lib.potential_create_LJ126_switch_f
    df = lib.potential_create_LJ126_switch_dfdr
    lj = lib.potential_LJ126
    ljp = lib.potential_LJ126_p
    for fn in (f, df):
        fn.restype = C.c_double
        fn.argtypes = [C.c_double]
    for fn in (lj, ljp):
        fn.restype = C.c_double
        fn.argtypes = [C.c_double, C.c_double, C.c_double]
    r = np.linspace(0.85, 1.0, 301)
    S, dS = mo.switch(r, s, rc)
    for k, rk in enumerate(r):
        e = lj(rk, A, B)
        assert f(rk) == pytest.approx(e * S[k], rel=1e-12, abs=1e-18)
        assert df(rk) == pytest.approx(ljp(rk, A, B) * S[k] + e * dS[k], rel=1e-10, abs=1e-15)


def test_shift_and_switch_at_rhi():
    """Unswitched, the energy at r_hi is the table's; switched, it is zero
    with zero derivative."""
    theta = random_theta(5)
    sc, _, _, _, _ = random_samples(6, 4)
    et = np.array([-0.3, 0.01, 0.2])
    sh = mo.shifts(theta, NP, sc, et)
    f = mo.forward(theta, NP, sc, sc[:, 1], np.arange(NP), rounding=True, shift=sh)
    plain = sc[:, 3] >= sc[:, 1]
    assert np.allclose(f["E"][plain], et[plain], rtol=1e-12, atol=1e-14)
    assert np.all(f["E"][~plain] == 0.0) and np.all(f["dEdr"][~plain] == 0.0)
    assert plain.any() and (~plain).any()


@pytest.mark.parametrize("block", mo.BLOCKS[:-1])
def test_gradient_matches_finite_differences(block):
    theta = random_theta(1)
    sc, r, p, E, F = random_samples(2, 48)
    loss, g = mo.loss_and_grad(theta, NP, sc, r, p, E, F, 0.7, 1.3, rounding=False)
    o = mo.offsets(NP)
    lo, hi = o[block], o[mo.BLOCKS[mo.BLOCKS.index(block) + 1]]
    rng = np.random.default_rng(3)
    idx = rng.choice(np.arange(lo, hi), min(12, hi - lo), replace=False)
    h = 1e-6
    for k in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += h
        tm[k] -= h
        lp, _ = mo.loss_and_grad(tp, NP, sc, r, p, E, F, 0.7, 1.3, rounding=False)
        lm, _ = mo.loss_and_grad(tm, NP, sc, r, p, E, F, 0.7, 1.3, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9 * max(1.0, loss)), (block, k)


@pytest.mark.parametrize("block", ["W1", "W2", "B2", "W3"])
def test_force_term_gradient_alone(block):
    """wE = 0: only the derivative output yd enters the loss, so the whole
    gradient comes through the tangent path and silu''."""
    theta = random_theta(7)
    sc, r, p, E, F = random_samples(8, 32)
    loss, g = mo.loss_and_grad(theta, NP, sc, r, p, E, F, 0.0, 1.0, rounding=False)
    o = mo.offsets(NP)
    lo, hi = o[block], o[mo.BLOCKS[mo.BLOCKS.index(block) + 1]]
    idx = np.random.default_rng(9).choice(np.arange(lo, hi), min(8, hi - lo), replace=False)
    h = 1e-6
    for k in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (mo.loss_and_grad(tp, NP, sc, r, p, E, F, 0.0, 1.0, rounding=False)[0] -
              mo.loss_and_grad(tm, NP, sc, r, p, E, F, 0.0, 1.0, rounding=False)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9 * max(1.0, loss)), (block, k)
    assert np.abs(g[o["B3"]:o["N"]]).max() == 0.0   # b3 does not move yd


def test_rounding_changes_little():
    theta = random_theta(4)
    sc, r, p, E, F = random_samples(5, 200)
    a = mo.forward(theta, NP, sc, r, p, rounding=True)
    b = mo.forward(theta, NP, sc, r, p, rounding=False)
    assert np.abs(a["y"] - b["y"]).max() < 0.05
    assert np.abs(a["y"] - b["y"]).max() > 0.0
    assert np.abs(a["yd"] - b["yd"]).max() < 0.05 * np.abs(b["yd"]).max()


def test_bf16_pair_split_precision():
    rng = np.random.default_rng(11)
    v = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 30, 2000)]).astype(np.float32).astype(np.float64)
    hi = mo.bf16(v)
    lo = mo.bf16(np.float32(v) - np.float32(hi))
    err = np.abs(hi + lo - v)
    assert np.all(err <= 2.0 ** -16 * np.abs(v))
    assert np.all(np.abs(hi - v) <= 2.0 ** -8 * np.abs(v))
    assert (np.abs(hi - v) > 2.0 ** -12 * np.abs(v)).mean() > 0.5   # a single bf16 is far coarser
