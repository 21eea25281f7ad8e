"""Rigid water on the GPU: engine_step with the nonbonded pass on the device
and the constraints solved after each advance (engine.c:2196-2232; SHAKE
with engine_flag_shake, P-SHAKE otherwise), and engine_gpu_steps solving
them on the device (k_rigid, one thread per molecule).

- host-driven: every O-H and H-H distance stays within tol_rigid of its
  constraint (|d^2 - r^2| < 1e-6 nm^2) along the trajectory;
- device-resident steps equal host-driven ones bit for bit: positions,
  velocities, per-cell order, forces (P-SHAKE adds its constraint forces),
  energies, the P-SHAKE coupling matrices through their effect on later
  steps, and with sort reuse each cell's maxdx.
"""
import dataclasses

import numpy as np
import pytest

import mdcore as md
import systems

pytestmark = pytest.mark.gpu


def water(rigid=True, seed=5, T=300.0):
    sysm = systems.water(1000, 3.1, (10, 10, 10), seed=seed, rigid=rigid)
    return dataclasses.replace(sysm, v=md.gen_maxwell(sysm.masses[sysm.types], systems.KB * T, seed + 4))


def residues(eng, sysm):
    x = eng.positions()
    out = []
    for i, j, d in sysm.rigid:
        dv = x[i] - x[j]
        dv -= sysm.dim * np.round(dv / sysm.dim)
        out.append(abs(d * d - dv @ dv))
    return np.array(out)


@pytest.mark.parametrize("flags", [0, md.engine_flag_shake], ids=["pshake", "shake"])
def test_host_driven_constraints_hold(flags):
    sysm = water()
    eng = systems.build_engine(sysm, md.engine_flag_gpu | flags)
    eng.setdt(0.002)
    assert eng.nrigids() == 1000
    worst = 0.0
    for _ in range(40):
        eng.step()
        worst = max(worst, residues(eng, sysm).max())
    print(f"largest |d^2 - r^2| over 40 steps: {worst:.2e} nm^2", flush=True)
    # the solvers stop at max |d^2 - r^2| < tol_rigid = 1e-6; absolute
    # positions add rounding of order 1e-17 here
    assert worst < 1e-6 * (1 + 1e-6)
    assert np.isfinite(eng.energies()["epot"])
    eng.close()


def kinetic(sysm, v):
    return 0.5 * float(np.sum(sysm.masses[sysm.types][:, None] * v * v))


CASES = {"pshake": 0, "shake": md.engine_flag_shake, "pshake_verlet": md.engine_flag_verlet_pseudo}


@pytest.mark.parametrize("name", sorted(CASES))
def test_resident_rigid_steps_match_host_steps(name):
    sysm = water()
    flags = md.engine_flag_gpu | CASES[name]
    n = 30  # P-SHAKE refines its coupling on steps 1..19
    print(f"case {name}: {sysm.N} atoms, {n} steps", flush=True)
    eh = systems.build_engine(sysm, flags)
    ed = systems.build_engine(sysm, flags)
    nc = int(np.prod(np.floor(sysm.dim / sysm.L)))
    for e in (eh, ed):
        e.setdt(0.002)
    epot, ekin = [], []
    for _ in range(n):
        eh.step()
        epot.append(eh.energies()["epot"])
        ekin.append(kinetic(sysm, eh.velocities()))
    ep, ek = ed.gpu_steps(n)
    assert np.array_equal(np.array(epot), ep)
    np.testing.assert_allclose(ek, np.array(ekin), rtol=1e-12)
    ch, ih = eh.cells(nc)
    cd, idd = ed.cells(nc)
    assert np.array_equal(ch, cd) and np.array_equal(ih, idd)
    assert np.array_equal(eh.positions(), ed.positions())
    assert np.array_equal(eh.velocities(), ed.velocities())
    assert np.array_equal(eh.forces(), ed.forces())
    assert eh.energies() == ed.energies()
    mh, cmh, _ = eh.maxdx(nc)
    md_, cmd, _ = ed.maxdx(nc)
    assert mh == md_ and np.array_equal(cmh, cmd)
    # both continue identically (the coupling matrices came back too)
    eh.step()
    ed.step()
    assert np.array_equal(eh.positions(), ed.positions())
    assert residues(ed, sysm).max() < 1e-6
    eh.close()
    ed.close()
