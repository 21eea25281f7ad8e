"""Pseudo-Verlet sort reuse (engine_flag_verlet_pseudo) through engine_step
on the GPU: engine_verlet_update (engine.c:193-315) decides the rebuilds,
the pass sorts only on a rebuild (runner.c:773) and otherwise reuses the
sort lists with each pair window widened by both cells' maxdx
(runner_dopair.c:428), and engine_advance only kicks between rebuilds.

Each pass until the first rebuild after the start is checked against the
FP64 oracle given the same cell-relative positions, the same per-cell maxdx
and the sort lists of the start: pair counts bit-exact, forces to 1e-5 and
energies to 1e-6 relative (the FP32 tables).  The oracle's reuse mode is
pinned against brute force in test_oracle.py.
"""
import dataclasses

import numpy as np
import pytest

import mdcore as md
import pyoracle as po
import systems

pytestmark = pytest.mark.gpu


def argon_skin():
    """2,048 argon atoms at 300 K in 4^3 cells of 1.13 nm (cell option
    1.05 nm): skin min(h) - rc = 0.13 nm."""
    sysm = systems.argon(8, jitter=0.05, T=300.0)
    return dataclasses.replace(sysm, L=np.full(3, 1.05))


def test_sort_reuse_passes_match_oracle():
    sysm = argon_skin()
    dt, nmax = 0.004, 60
    eng = systems.build_engine(sysm, md.engine_flag_gpu | md.engine_flag_verlet_pseudo)
    eng.setdt(dt)
    o = po.from_system(sysm)
    nc = o.ncells
    reused = 0
    for step in range(nmax):
        xl, cell = eng.local()
        counts0, ids0 = eng.cells(nc)
        eng.step()
        maxdx, cell_maxdx, _ = eng.maxdx(nc)
        counts1, ids1 = eng.cells(nc)
        rebuilt = step == 0 or maxdx == 0.0
        if step > 0 and rebuilt:
            break
        # no migration between rebuilds: the layout is the start's
        assert np.array_equal(counts0, counts1) and np.array_equal(ids0, ids1)
        o.set_verlet(xl, cell_maxdx, keep_sort=step > 0)
        r = o.nonbond(po.OR_WINDOW_REF)
        en = eng.energies()
        f = eng.forces()
        assert en["pairs"] == r["pairs"], step
        scale = np.abs(r["f"]).max()
        assert np.abs(f - r["f"]).max() <= 1e-5 * scale, step
        assert en["epot"] == pytest.approx(r["epot"] + r["epot_exclusion"], rel=1e-6), step
        reused += step > 0
    print(f"{reused} passes reused the start's sort lists; last maxdx {maxdx:.4f} nm", flush=True)
    assert reused >= 5
    assert step < nmax - 1, "no rebuild within the run"
    eng.close()


def test_rebuild_follows_the_skin_criterion():
    """Rebuilds happen exactly when 2 maxdx > skin or maxdx > rc/4, and the
    per-cell maxdx are the largest displacements since the last rebuild
    (displacements taken here from absolute positions, minimum image)."""
    sysm = argon_skin()
    dt = 0.004
    eng = systems.build_engine(sysm, md.engine_flag_gpu | md.engine_flag_verlet_pseudo)
    eng.setdt(dt)
    nc = int(np.prod(np.floor(sysm.dim / sysm.L)))
    h = sysm.dim / np.floor(sysm.dim / sysm.L)
    skin = h.min() - sysm.cutoff
    ref = eng.positions()
    nrebuild = 0
    for step in range(120):
        x = eng.positions()
        _, cell = eng.local()
        eng.step()
        maxdx, cell_maxdx, _ = eng.maxdx(nc)
        if step == 0:
            continue
        dv = x - ref
        dv -= sysm.dim * np.round(dv / sysm.dim)
        disp = np.sqrt((dv ** 2).sum(1))
        expect = disp.max()
        if 2.0 * expect > skin or expect > 0.25 * sysm.cutoff:
            assert maxdx == 0.0, step
            nrebuild += 1
            ref = x
            continue
        assert maxdx == pytest.approx(expect, rel=1e-9), step
        cm = np.zeros(nc)
        np.maximum.at(cm, cell, disp)
        np.testing.assert_allclose(cell_maxdx, cm, rtol=1e-9, atol=1e-15)
    print(f"{nrebuild} rebuilds in 120 steps", flush=True)
    assert nrebuild >= 2
    eng.close()
