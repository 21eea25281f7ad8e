"""engine_step end to end on the GPU -- shuffle, space_prepare, the nonbonded
pass, kick-drift and cell migration (engine.c:1945-2052) -- against the same
integrator driven by FP64 oracle forces on the host.

Argon, 2,048 atoms at 120 K, dt = 2 fs, 25 steps: atoms cross cell faces, so
migration and the re-sort of changed cells are exercised.  Tolerances: the
GPU forces carry ~1e-6 relative error, which moves positions by well under
1e-6 nm over 25 steps of this (short, non-chaotic) horizon.
"""
import numpy as np
import pytest

import mdcore as md
import pyoracle as po
import systems

pytestmark = pytest.mark.gpu


def oracle_pass(sysm, x):
    o = po.Oracle(sysm.origin, sysm.dim, sysm.L, sysm.cutoff, np.mod(x, sysm.dim), sysm.types,
                  systems.tables(sysm), sysm.excl)
    r = o.nonbond(po.OR_WINDOW_REF)
    o.close()
    return r


def test_trajectory_matches_fp64_integration():
    sysm = systems.argon(8, jitter=0.05, T=120.0)
    dt, steps = 0.002, 25
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    eng.setdt(dt)
    m = sysm.masses[sysm.types][:, None]
    x = sysm.x.copy()
    v = sysm.v.copy()
    x0 = np.mod(x, sysm.dim)
    for s in range(steps):
        eng.step()
        r = oracle_pass(sysm, x)
        v += r["f"] * (dt / m)
        x += dt * v
        en = eng.energies()
        assert en["pairs"] == r["pairs"], s
        assert en["epot"] == pytest.approx(r["epot"], rel=1e-6), s
    xg = eng.positions()
    d = np.mod(xg - x + 0.5 * sysm.dim, sysm.dim) - 0.5 * sysm.dim
    moved = np.mod(x - x0 + 0.5 * sysm.dim, sysm.dim) - 0.5 * sysm.dim
    print(f"max |x_gpu - x_fp64| after {steps} steps: {np.abs(d).max():.3e} nm "
          f"(max displacement {np.abs(moved).max():.3f} nm)", flush=True)
    assert np.abs(moved).max() > 0.01          # atoms actually moved
    assert np.abs(d).max() < 1e-6
    eng.close()

