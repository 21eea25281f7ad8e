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


def mixture_moving():
    sysm = systems.mixture(4000, seed=11)
    return dataclasses.replace(sysm, v=md.gen_maxwell(sysm.masses[sysm.types], systems.KB * 300.0, 5))


# (system, dt, steps): argon rebuilds every ~20 steps; the mixture's 0.6 nm
# cells under a 1.2 nm cutoff have a negative skin and rebuild every step
RESIDENT = {"argon_skin": (argon_skin, 0.004, 60), "mixture_4k": (mixture_moving, 0.0005, 12)}


def kinetic(sysm, v):
    return 0.5 * float(np.sum(sysm.masses[sysm.types][:, None] * v * v))


def same_state(eh, ed, nc):
    ch, ih = eh.cells(nc)
    cd, idd = ed.cells(nc)
    assert np.array_equal(ch, cd) and np.array_equal(ih, idd)
    assert np.array_equal(eh.positions(), ed.positions())
    assert np.array_equal(eh.velocities(), ed.velocities())
    assert np.array_equal(eh.forces(), ed.forces())
    assert eh.energies() == ed.energies()
    mh, cmh, rbh = eh.maxdx(nc)
    md_, cmd, rbd = ed.maxdx(nc)
    assert mh == md_ and rbh == rbd
    assert np.array_equal(cmh, cmd)


@pytest.mark.parametrize("name", sorted(RESIDENT))
def test_resident_verlet_steps_match_host_steps(name):
    make, dt, n = RESIDENT[name]
    sysm = make()
    print(f"case {name}: N={sysm.N} dt={dt} steps={n}", flush=True)
    flags = md.engine_flag_gpu | md.engine_flag_verlet_pseudo
    eh = systems.build_engine(sysm, flags)
    ed = systems.build_engine(sysm, flags)
    nc = int(np.prod(np.floor(sysm.dim / sysm.L)))
    for e in (eh, ed):
        e.setdt(dt)
    epot, ekin, rebuilds = [], [], 0
    for _ in range(n):
        eh.step()
        epot.append(eh.energies()["epot"])
        ekin.append(kinetic(sysm, eh.velocities()))
        rebuilds += eh.maxdx(nc)[0] == 0.0
    ep, ek = ed.gpu_steps(n)
    assert np.array_equal(np.array(epot), ep)
    np.testing.assert_allclose(ek, np.array(ekin), rtol=1e-12)
    same_state(eh, ed, nc)
    print(f"{name}: {rebuilds} rebuild steps in {n}", flush=True)
    assert 2 <= rebuilds
    eh.close()
    ed.close()


def test_resident_verlet_interleaves_with_host_steps():
    sysm = argon_skin()
    flags = md.engine_flag_gpu | md.engine_flag_verlet_pseudo
    eh = systems.build_engine(sysm, flags)
    ed = systems.build_engine(sysm, flags)
    nc = int(np.prod(np.floor(sysm.dim / sysm.L)))
    for e in (eh, ed):
        e.setdt(0.004)
    for _ in range(45):
        eh.step()
    ed.gpu_steps(17)
    for _ in range(6):
        ed.step()
    ed.gpu_steps(22)
    same_state(eh, ed, nc)
    eh.close()
    ed.close()
