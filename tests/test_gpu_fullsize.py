"""The benchmark systems at full size on the GPU against the FP64 oracle.

Bit-exact: pair counts (reference window), exclusion counts, and every
sorted list of every cell (all 46,656 cells at 1M atoms, in one bulk copy).
Floating point: force relative RMSE <= 1e-5 and energy relative error
<= 1e-6 against FP64.  The oracle runs on up to 16
host threads (a few seconds per system).
"""
import os

import numpy as np
import pytest

import mdcore as md
import pyoracle as po
import systems

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
FULL = {
    "mixture_1m": lambda: systems.mixture(1048576),
    "water_100k": lambda: systems.water(),
    "argon_32k": lambda: systems.argon(20, jitter=0.05),
}


@pytest.mark.parametrize("name", sorted(FULL))
def test_full_size_parity(name):
    sysm = FULL[name]()
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    eng.shuffle()
    eng.nonbond()
    g = eng.gpu_results()
    f = eng.forces()
    eng.sync(md.engine_gpu_sync_sort_down)
    orc = po.from_system(sysm)
    o = orc.nonbond(po.OR_WINDOW_REF, THREADS)
    print(f"{name}: pairs {g['pairs']} (oracle {o['pairs']}), exclusions {g['exclusions']}", flush=True)
    assert g["pairs"] == o["pairs"]
    assert g["exclusions"] == o["exclusions"]
    err = float(np.sqrt(np.sum((f - o["f"]) ** 2) / np.sum(o["f"] ** 2)))
    print(f"{name}: force rel RMSE {err:.3e}, epot {g['epot']:.9e} vs {o['epot']:.9e}", flush=True)
    assert err <= 1e-5
    assert abs(g["epot"] - o["epot"]) <= 1e-6 * abs(o["epot"])
    assert abs(g["epot_exclusion"] - o["epot_exclusion"]) <= 1e-6 * max(1.0, abs(o["epot_exclusion"]))
    # sorted lists: every flagged direction of every cell, in one bulk copy
    orc.sort()
    counts, ids = eng.cells(orc.ncells)
    oids = np.concatenate([orc.cell(cid)[0] for cid in range(orc.ncells)])
    assert np.array_equal(ids, oids)
    flags = np.array([orc.cell(cid)[1] for cid in range(orc.ncells)], np.int64)
    got, want = eng.sortlists(), orc.sortlists()
    assert len(got) == len(want) == 13 * len(ids)
    # word k of cell c's block lies in direction (k - 13*start_c) // count_c
    cell = np.repeat(np.arange(orc.ncells), 13 * counts)
    start = np.concatenate([[0], np.cumsum(13 * counts.astype(np.int64))])[:-1]
    sid = (np.arange(len(got)) - start[cell]) // np.maximum(counts[cell], 1)
    flagged = ((flags[cell] >> sid) & 1).astype(bool)
    bad = np.nonzero(flagged & (got != want))[0]
    print(f"{name}: {int(flagged.sum())} sorted words in {orc.ncells} cells compared", flush=True)
    assert len(bad) == 0, f"first mismatch: cell {cell[bad[0]]} sid {sid[bad[0]]}"
    eng.close()
    orc.close()


@pytest.mark.parametrize("name", sorted(FULL))
def test_full_size_modes_agree(name):
    """At full size, the default pass and the passes with the other pair-kernel
    shell, without the owner split and without paired visits count the same
    pairs and exclusions and agree to FP32 summation order."""
    sysm = FULL[name]()
    res = {}
    for label, flags in (("default", 0), ("full", md.engine_flag_fullshell), ("half", md.engine_flag_halfshell),
                         ("nosplit", md.engine_flag_nosplit),
                         ("unpaired", md.engine_flag_fullshell | md.engine_flag_unpaired)):
        eng = systems.build_engine(sysm, md.engine_flag_gpu | flags)
        eng.shuffle()
        eng.nonbond()
        res[label] = (eng.forces().copy(), eng.gpu_results())
        eng.close()
    f0, g0 = res["default"]
    for label in ("full", "half", "nosplit", "unpaired"):
        f, g = res[label]
        err = float(np.sqrt(np.sum((f - f0) ** 2) / np.sum(f0 ** 2)))
        print(f"{name} {label}: pairs {g['pairs']}, rel diff to default {err:.2e}", flush=True)
        assert g["pairs"] == g0["pairs"] and g["exclusions"] == g0["exclusions"]
        assert g["pairs_refwindow"] == g0["pairs_refwindow"]
        assert err <= 2e-6
        assert abs(g["epot"] - g0["epot"]) <= 1e-6 * abs(g0["epot"])
