"""engine_gpu_steps: whole MD steps with the particles resident on the device
(shuffle, nonbonded pass, kick-drift, migration), against the same number of
host-driven engine_step calls (engine.c:2070-2240, engine_advance 1945-2052,
space_shuffle_local space.c:307-364).

The device migration keeps the host's per-cell storage order (swap-with-last
departures, arrivals by source cell then departure order), so the sorted
lists, the pair sets and the whole trajectory are the host's bit for bit:
positions, velocities, the cell contents in storage order, the last pass's
forces and every step's pass energy are compared with ==.  The kinetic
energy is summed per cell on the device and in numpy on the host, so it is
compared to 1e-12 relative.
"""
import dataclasses

import numpy as np
import pytest

import mdcore as md
import surrogate as sg
import systems

pytestmark = pytest.mark.gpu


def with_velocities(sysm, T, seed):
    m = sysm.masses[sysm.types]
    return dataclasses.replace(sysm, v=md.gen_maxwell(m, systems.KB * T, seed))


def argon_on_faces():
    """2,048 argon atoms at 600 K, the lattice shifted so that a plane of
    sites lies 0.0125 nm inside each cell face: many cell crossings."""
    sysm = systems.argon(8, jitter=0.05, T=600.0)
    return dataclasses.replace(sysm, x=np.mod(sysm.x + 0.125, sysm.dim))


CASES = {
    "argon_hot": (argon_on_faces, 0.002, 30),
    # 4,000-atom 20-species LJ+RF mixture, random packing (close contacts)
    "mixture_4k": (lambda: with_velocities(systems.mixture(4000, seed=11), 300.0, 5), 0.0005, 20),
    # 1,000 water molecules with intramolecular exclusions
    "water_1k": (lambda: with_velocities(systems.water(1000, 3.1, (10, 10, 10), seed=5), 300.0, 9), 0.0005, 20),
}


def ncells(sysm):
    return int(np.prod(np.floor(sysm.dim / sysm.L).astype(int)))


def kinetic(sysm, v):
    return 0.5 * float(np.sum(sysm.masses[sysm.types][:, None] * v * v))


def host_steps(eng, sysm, n):
    epot, ekin = [], []
    for _ in range(n):
        eng.step()
        epot.append(eng.energies()["epot"])
        ekin.append(kinetic(sysm, eng.velocities()))
    return np.array(epot), np.array(ekin)


def assert_same_state(eh, ed, sysm):
    nc = ncells(sysm)
    ch, ih = eh.cells(nc)
    cd, idd = ed.cells(nc)
    assert np.array_equal(ch, cd), "cell counts differ"
    assert np.array_equal(ih, idd), "per-cell storage order differs"
    assert np.array_equal(eh.positions(), ed.positions())
    assert np.array_equal(eh.velocities(), ed.velocities())
    assert np.array_equal(eh.forces(), ed.forces())
    assert eh.energies() == ed.energies()


@pytest.mark.parametrize("name", sorted(CASES))
def test_resident_steps_match_host_steps(name):
    make, dt, n = CASES[name]
    sysm = make()
    print(f"case {name}: N={sysm.N} dt={dt} steps={n}", flush=True)
    eh = systems.build_engine(sysm, md.engine_flag_gpu)
    ed = systems.build_engine(sysm, md.engine_flag_gpu)
    nc = ncells(sysm)
    _, ids0 = eh.cells(nc)
    counts0, _ = eh.cells(nc)
    for e in (eh, ed):
        e.setdt(dt)
    epot_h, ekin_h = host_steps(eh, sysm, n)
    epot_d, ekin_d = ed.gpu_steps(n)
    assert np.array_equal(epot_h, epot_d)
    np.testing.assert_allclose(ekin_d, ekin_h, rtol=1e-12)
    assert_same_state(eh, ed, sysm)
    counts1, ids1 = ed.cells(nc)
    cell0 = np.repeat(np.arange(nc), counts0)[np.argsort(ids0)]
    cell1 = np.repeat(np.arange(nc), counts1)[np.argsort(ids1)]
    moved = int(np.count_nonzero(cell0 != cell1))
    print(f"{name}: {moved} of {sysm.N} particles changed cell", flush=True)
    assert moved > 0
    eh.close()
    ed.close()


def test_resident_steps_match_host_steps_surrogate():
    """The surrogate pass (half-shell network kernel) inside resident steps:
    a briefly trained network copied between two engines; device steps equal
    host-driven steps bit for bit."""
    make, dt, n = CASES["water_1k"]
    sysm = make()
    eh = systems.build_engine(sysm, md.engine_flag_gpu)
    ed = systems.build_engine(sysm, md.engine_flag_gpu)
    dom = sg.pair_domains(eh, sysm)
    S = sg.make_samples(eh, sysm, 1 << 16, dom[0], dom[1], seed=3)
    eh.mlp_init(*dom, seed=4)
    eh.mlp_train(*S, steps=300, batch_tiles=64, seed=5)
    ed.mlp_init(*dom, seed=4)
    ed.mlp_params(eh.mlp_params())
    for e in (eh, ed):
        e.setdt(dt)
        e.setflags(md.engine_flag_mlp, 0)
    epot_h, _ = host_steps(eh, sysm, 10)
    epot_d, _ = ed.gpu_steps(10)
    assert np.all(np.isfinite(epot_h))
    assert np.array_equal(epot_h, epot_d)
    assert_same_state(eh, ed, sysm)
    eh.close()
    ed.close()


def test_resident_steps_interleave_with_host_steps():
    """Device steps, host steps and device steps again: the state goes back and
    forth through the host cells without losing a bit."""
    make, dt, _ = CASES["argon_hot"]
    sysm = make()
    eh = systems.build_engine(sysm, md.engine_flag_gpu)
    ed = systems.build_engine(sysm, md.engine_flag_gpu)
    for e in (eh, ed):
        e.setdt(dt)
    host_steps(eh, sysm, 12)
    ed.gpu_steps(4)
    host_steps(ed, sysm, 3)
    ed.gpu_steps(5)
    assert_same_state(eh, ed, sysm)
    eh.close()
    ed.close()


def test_resident_steps_zero_and_negative():
    make, dt, _ = CASES["argon_hot"]
    sysm = make()
    ed = systems.build_engine(sysm, md.engine_flag_gpu)
    ed.setdt(dt)
    x0 = ed.positions()
    epot, ekin = ed.gpu_steps(0)
    assert len(epot) == 0 and len(ekin) == 0
    assert np.array_equal(ed.positions(), x0)
    with pytest.raises(RuntimeError):
        ed.gpu_steps(-1)
    ed.close()
