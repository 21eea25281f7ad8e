"""The FP64 reference of the MLP surrogate (oracle/mlp_oracle.py), pinned
against what does not depend on it:

- bf16 rounding: torch's float32 -> bfloat16 conversion (round to nearest,
  ties to even) on random values and on hand-made ties;
- gradients: central finite differences of the loss in float64 (rounding
  off, so the loss is smooth), for every parameter block;
- the network output: a hand-computed 1-unit special case (zero weights
  everywhere but one path gives y0 = w3 * silu(w2 * bf16(silu(z1))) + b3).
"""
import numpy as np
import pytest
import torch

import mlp_oracle as mo

NP = 3


def random_theta(seed, npair=NP):
    rng = np.random.default_rng(seed)
    o = mo.offsets(npair)
    t = rng.normal(0.0, 0.3, o["N"])
    return t


def random_samples(seed, n, npair=NP):
    rng = np.random.default_rng(seed)
    sc = np.array([[0.2, 1.0, 1.0, 10.0], [0.15, 1.2, 2.0, 5.0], [0.3, 0.9, 0.5, 20.0]])[:npair]
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
    """One unit on each layer: y0 = w3 silu(w2 bf16(silu(w1 rn + b1)) + b2) + b3."""
    npair = 1
    P = {k: np.zeros_like(v) for k, v in mo.unpack(np.zeros(mo.offsets(npair)["N"]), npair).items()}
    P["W1"][5, 0] = 0.7
    P["b1"][5] = 0.1
    P["W2"][9, 5] = 1.25
    P["b2"][9] = -0.2
    P["W3"][0, 0, 9] = 2.0
    P["b3"][0, 0] = 0.3
    theta = mo.pack(P, npair)
    sc = np.array([[0.2, 1.0, 3.0, 1.0]])
    r = np.array([0.45])
    rn = (2 * 0.45 - 0.2 - 1.0) / 0.8

    def silu(z):
        return z / (1 + np.exp(-z))

    h1 = float(torch.tensor(silu(0.7 * rn + 0.1), dtype=torch.float32).to(torch.bfloat16).float())
    y0 = 2.0 * silu(1.25 * h1 - 0.2) + 0.3
    f = mo.forward(theta, npair, sc, r, [0], rounding=True)
    assert f["y0"][0] == pytest.approx(y0, rel=1e-12)
    assert f["E"][0] == pytest.approx(3.0 * np.sinh(y0), rel=1e-12)
    # every other unit contributes silu(0) = 0, so y1 = b3[0, 1] = 0
    assert f["y1"][0] == 0.0


@pytest.mark.parametrize("block", ["W1", "B1", "E1", "W2", "B2", "EMB", "W3", "B3"])
def test_gradient_matches_finite_differences(block):
    theta = random_theta(1)
    sc, r, p, E, F = random_samples(2, 48)
    loss, g = mo.loss_and_grad(theta, NP, sc, r, p, E, F, 0.7, 1.3, rounding=False)
    o = mo.offsets(NP)
    names = ["W1", "B1", "E1", "W2", "B2", "EMB", "W3", "B3", "N"]
    lo, hi = o[block], o[names[names.index(block) + 1]]
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


def test_rounding_changes_little():
    theta = random_theta(4)
    sc, r, p, E, F = random_samples(5, 200)
    a = mo.forward(theta, NP, sc, r, p, rounding=True)
    b = mo.forward(theta, NP, sc, r, p, rounding=False)
    assert np.abs(a["y0"] - b["y0"]).max() < 0.05
    assert np.abs(a["y0"] - b["y0"]).max() > 0.0
