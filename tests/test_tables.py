"""Table build and evaluation (host library and oracle), pinned three ways:

1. bit-exact against the reference's own potential.c, compiled unmodified
   into oracle/_ref (skipped where that build is absent);
2. bit-exact against tests/golden/tables.npz, written from that build by
   tests/golden/make_golden.py (always runs);
3. against what the mathematics fixes: a quadratic is reproduced exactly,
   the table meets its tolerance against the analytic function, values and
   slopes are continuous across interval edges, and the left extension is
   the quadratic Taylor expansion at a.
"""
import math
from pathlib import Path

import numpy as np
import pytest

import mdcore as md
import pyoracle as po

GOLDEN = Path(__file__).with_name("golden") / "tables.npz"
KE = 138.935458

BUILDERS = {
    "LJ126": md.LJ126,
    "LJ126_switch": md.LJ126_switch,
    "LJ126_Coulomb": md.LJ126_Coulomb,
    "LJ126_Ewald": md.LJ126_Ewald,
    "Ewald": md.Ewald,
    "Coulomb": md.Coulomb,
    "harmonic": md.harmonic,
    "LJ126_RF": md.LJ126_RF,
}


def golden_cases():
    g = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in g.files})
    return g, names


G, NAMES = golden_cases()


@pytest.mark.parametrize("name", NAMES)
def test_host_table_matches_golden(name):
    kind = str(G[f"{name}/kind"])
    args = tuple(G[f"{name}/args"])
    d = BUILDERS[kind](*args).data()
    assert d["n"] == int(G[f"{name}/n"])
    assert np.array_equal(d["alpha"], G[f"{name}/alpha"])
    assert np.array_equal(d["c"], G[f"{name}/c"])


@pytest.mark.parametrize("name", NAMES)
def test_host_and_oracle_eval_match_golden(name):
    kind = str(G[f"{name}/kind"])
    args = tuple(G[f"{name}/args"])
    p = BUILDERS[kind](*args)
    d = p.data()
    for r2, (e_ref, f_ref) in zip(G[f"{name}/r2"], G[f"{name}/ef"]):
        e, f = p.eval(r2)
        assert (e, f) == (e_ref, f_ref)
        eo, fo = po.evaluate(d, r2)
        assert (eo, fo) == (e_ref, f_ref)


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("seed", range(6))
def test_host_table_bitexact_vs_reference_build(seed):
    """Random parameters, same answer as the compiled reference."""
    rng = np.random.default_rng(seed)
    r = po.ref()
    sig = rng.uniform(0.25, 0.40)
    eps = rng.uniform(0.1, 1.0)
    q = rng.uniform(-0.6, 0.6)
    A, B = 4 * eps * sig ** 12, 4 * eps * sig ** 6
    cases = [
        (md.LJ126(0.8 * sig, 1.2, A, B, 1e-3), r.potential_create_LJ126(0.8 * sig, 1.2, A, B, 1e-3)),
        (md.LJ126_RF(0.12, 1.2, A, B, q, KE, 0.0, 1e-3),
         r.ref_potential_create_LJ126_RF(0.12, 1.2, A, B, q, KE, 0.0, 1e-3)),
        (md.LJ126_Coulomb(0.1, 1.0, A, B, q, 1e-3), r.potential_create_LJ126_Coulomb(0.1, 1.0, A, B, q, 1e-3)),
    ]
    for mine, theirs in cases:
        d = mine.data()
        t = po.ref_table(theirs)
        assert d["n"] == t["n"]
        assert np.array_equal(d["alpha"], t["alpha"])
        assert np.array_equal(d["c"], t["c"])
        for r2 in rng.uniform(d["a"] ** 2, d["b"] ** 2, 50):
            assert mine.eval(r2) == po.ref_eval(t, r2)


def test_quadratic_is_reproduced_exactly():
    """A degree-2 function lies in the quintic space: the table is exact."""
    K, r0 = 1234.5, 0.1
    p = md.harmonic(0.05, 0.2, K, r0, 1e-8)
    for r in np.linspace(0.05, 0.2, 101):
        e, f = p.eval(r * r)
        assert e == pytest.approx(K * (r - r0) ** 2, rel=1e-12, abs=1e-9)
        assert f == pytest.approx(2 * K * (r - r0) / r, rel=1e-9, abs=1e-7)


def lj(r, A, B):
    return A / r ** 12 - B / r ** 6, (-12 * A / r ** 13 + 6 * B / r ** 7)


def test_lj_table_meets_tolerance():
    A, B = 4 * 0.996 * 0.3405 ** 12, 4 * 0.996 * 0.3405 ** 6
    p = md.LJ126(0.275, 1.0, A, B, 1e-3)
    d = p.data()
    x = np.linspace(0.275, 1.0, 4001)
    # per-interval relative error bound of potential_getcoeffs: |err| <= tol * max(|f|, 1)
    for r in x:
        e, _ = p.eval(r * r)
        ex, _ = lj(r, A, B)
        assert abs(e - ex) <= 1e-3 * max(abs(ex), 1.0) * 2
    assert 10 <= d["n"] <= 64


def test_continuity_across_interval_edges():
    A, B = 2.6378e-6, 2.6192e-3
    p = md.LJ126_RF(0.05, 1.0, A, B, 0.71842576, KE, 0.0, 1e-3)
    d = p.data()
    n, c = d["n"], d["c"].reshape(-1, 8)
    for k in range(1, n):
        edge = c[k, 0] + 1.0 / c[k, 1]  # right end of interval k
        eps = 1e-9 * edge
        el, fl = p.eval((edge - eps) ** 2)
        er, fr = p.eval((edge + eps) ** 2)
        assert abs(el - er) <= 1e-6 * max(abs(el), 1.0)
        assert abs(fl - fr) <= 1e-4 * max(abs(fl), 1.0)


def test_left_extension_is_quadratic_taylor():
    """Below a the table is e(a) + e'(a)(r-a) + e''(a)(r-a)^2/2 of interval 1."""
    A, B = 4 * 0.996 * 0.3405 ** 12, 4 * 0.996 * 0.3405 ** 6
    p = md.LJ126(0.275, 1.0, A, B, 1e-3)
    a = 0.275
    h = 1e-5
    e0, _ = p.eval(a * a)
    e1, _ = p.eval((a + h) ** 2)
    e2, _ = p.eval((a + 2 * h) ** 2)
    d1 = (e1 - e0) / h
    for r in (0.2, 0.25, 0.27):
        e, _ = p.eval(r * r)
        taylor1 = e0 + d1 * (r - a)
        # the quadratic term is O((r-a)^2 e''); compare to first order with a generous bound
        assert abs(e - taylor1) <= 0.6 * abs(e2 - 2 * e1 + e0) / h ** 2 * (r - a) ** 2 + 1e-6 * abs(e0) + 1e-3


def test_ivalsmax_is_reported():
    """A tolerance no table can meet fails with potential_err_ivalsmax, not a crash."""
    with pytest.raises(md.MdcoreError):
        md.LJ126(0.05, 3.0, 1.0, 1.0, 1e-14)
    md.library().errs_clear()
