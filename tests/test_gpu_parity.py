"""GPU path vs the FP64 oracle on the same inputs.

Bit-exact: cell contents, every sorted list, pair counts (both windows,
including pair tasks that the reference window's unsigned break skips
whole), exclusion counts.  Floating point: per-atom force relative RMSE <= 1e-5 and
energy / virial relative error <= 1e-6 against the FP64 oracle (FP32 pair
arithmetic, two-sum accumulation, FP64 reductions).  Every system runs with
both pair-kernel shells (full: each pair at both cells, the neighbour cells
at opposite offsets visited together, or one at a time; half: each pair task
once, opposite forces folded per atom), with each owner cell's neighbours
split over several work items (the default for these small systems) and
with one item per cell.
"""
import numpy as np
import pytest

import mdcore as md
import pyoracle as po
import systems
from cases import face_pairs
from gpuutil import engine_pass, rel_rmse

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-6

SYSTEMS = {
    # span 1, 24 atoms/cell
    "argon_n8": lambda: systems.argon(8, jitter=0.05),
    # span 2 (0.6 nm cells, rc 1.2): the reference window drops in-cutoff pairs
    "mixture_4k": lambda: systems.mixture(4000, seed=11),
    # ~100 atoms/cell (two 50-atom chunks per wave), exclusions, three tables
    "water_1k": lambda: systems.water(1000, 3.1, (10, 10, 10), seed=5),
    # ~330 atoms/cell: six chunks per wave and the sort path for cells too
    # large for the LDS sort
    "dense_bigcells": lambda: dense_big_cells(),
    # more types than 6 bits hold: the packed record moves to 8 type bits
    # (<= 256 particles per cell), then to 10 (<= 64 per cell)
    "many_types_100": lambda: many_types(100),
    "many_types_600": lambda: many_types(600),
    # simple cubic lattice, 64 atoms per cell: every direction's keys repeat
    # many times, so nearly every sort list goes through the tie sort
    "lattice_ties": lambda: lattice_ties(),
    # perfect fcc crystal, 32 atoms per cell: the cell's one wave sorts
    # nearly all 13 lists with repeated keys together
    "fcc_lattice": lambda: fcc_ties(),
    # the same with a 2 pm jitter: a few lists per cell keep repeated keys
    "fcc_nearlattice": lambda: fcc_ties(jitter=0.002),
    # 8 to 27 atoms per cell on a cubic lattice: lists of up to 15 atoms are
    # a single selection-sort leaf, the longer ones partition first
    "lattice_small_cells": lambda: lattice_ties(n=10, a=0.45),
    # 27 to 64 atoms per cell: lists above 32 take whole-wave partitions, their
    # parts and the shorter lists half-wave ones
    "lattice_mixed_cells": lambda: lattice_ties(n=13, a=0.33),
    # pairs planted across cell faces: the reference window's unsigned break
    # wraps and skips whole pair tasks that hold in-range pairs
    "face_pairs": lambda: face_pairs(),
}


def many_types(ntypes, N=4000, nbase=5, seed=13):
    """The 4k mixture's geometry with ntypes particle types; type t has the
    parameters of base species t % nbase, so only 15 distinct tables."""
    base = systems.mixture(N, seed=seed, nspecies=nbase)
    types = md.gen_types(N, ntypes, 1, seed + 1)
    pot_of = {}
    for ti, tj, p in base.pots:
        pot_of[(ti - 1, tj - 1)] = p
        pot_of[(tj - 1, ti - 1)] = p
    pots = [(i, j, pot_of[((i - 1) % nbase, (j - 1) % nbase)]) for i in range(1, ntypes + 1)
            for j in range(i, ntypes + 1)]
    masses = np.concatenate([[0.0], [base.masses[1 + (t % nbase)] for t in range(ntypes)]])
    return systems.System("many_types", base.origin, base.dim, base.L, base.cutoff, base.x, types, masses, pots)


def lattice_ties(n=16, a=0.25, seed=17):
    """Two species at random on the sites, so that the net forces do not
    cancel by symmetry."""
    g = (np.arange(n) + 0.5) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    L = n * a
    types = md.gen_types(len(x), 2, 1, seed)
    pots = [(1, 1, md.LJ126(0.05, 1.0, 1e-6, 1e-3, 1e-3)), (1, 2, md.LJ126(0.05, 1.0, 2e-6, 1.5e-3, 1e-3)),
            (2, 2, md.LJ126(0.05, 1.0, 4e-6, 2.5e-3, 1e-3))]
    return systems.System("lattice", np.zeros(3), np.full(3, L), np.full(3, 1.0), 1.0, x, types,
                          np.array([0.0, 1.0, 2.0]), pots)


def fcc_ties(n=8, a=0.565, jitter=0.0, seed=19):
    """fcc crystal (argon's lattice constant), two species at random on the
    sites so that the net forces do not cancel by symmetry."""
    x = md.gen_fcc(n, n, n, a, a / 4)
    L = n * a
    if jitter > 0:
        x = np.mod(x + md.gen_uniform(x.size, -0.5 * jitter, 0.5 * jitter, seed + 1).reshape(x.shape), L)
    types = md.gen_types(len(x), 2, 1, seed)
    # species 2 larger and deeper than argon: the pair forces on an atom
    # cancel to about a fifth of their magnitude, not to a twenty-fifth
    pots = [(1, 1, md.LJ126(0.275, 1.0, 9.7e-6, 6.2e-3, 1e-3)), (1, 2, md.LJ126(0.275, 1.0, 3e-5, 1.2e-2, 1e-3)),
            (2, 2, md.LJ126(0.275, 1.0, 9e-5, 2.4e-2, 1e-3))]
    return systems.System("fcc", np.zeros(3), np.full(3, L), np.full(3, 1.0), 1.0, x, types,
                          np.array([0.0, 39.9, 39.9]), pots)


def dense_big_cells(L=3.9, density=150.0, seed=3):
    N = int(density * L ** 3)
    x = md.gen_uniform(3 * N, 0.0, L, seed).reshape(-1, 3)
    pot = md.LJ126(0.05, 1.2, 1e-6, 1e-3, 1e-3)
    return systems.System("dense", np.zeros(3), np.full(3, L), np.full(3, 1.3), 1.2, x, np.ones(N, np.int32),
                          np.array([0.0, 1.0]), [(1, 1, pot)])


SHELLS = {"full": md.engine_flag_fullshell, "half": md.engine_flag_halfshell,
          "fullunpaired": md.engine_flag_fullshell | md.engine_flag_unpaired}
SPLITS = {"split": 0, "nosplit": md.engine_flag_nosplit}
MODES = {f"{sh}-{sp}": SHELLS[sh] | SPLITS[sp] for sh in sorted(SHELLS) for sp in sorted(SPLITS)}


@pytest.fixture(scope="module", params=[(s, m) for s in sorted(SYSTEMS) for m in MODES],
                ids=lambda p: f"{p[0]}-{p[1]}")
def case(request):
    name, mode = request.param
    sys = SYSTEMS[name]()
    orc = po.from_system(sys)
    eng = engine_pass(sys, md.engine_flag_gpu | MODES[mode])
    yield f"{name}-{mode}", sys, orc, eng
    eng.close()
    orc.close()


def test_cells_and_sortlists_bitexact(case):
    name, sys, orc, eng = case
    eng.sync(md.engine_gpu_sync_sort_down)
    orc.sort()
    counts, ids = eng.cells(orc.ncells)
    pos = 0
    nlists = 0
    for cid in range(orc.ncells):
        oids, flags = orc.cell(cid)
        n = counts[cid]
        assert n == len(oids), f"cell {cid} count"
        assert np.array_equal(ids[pos:pos + n], oids), f"cell {cid} order"
        pos += n
        for sid in range(13):
            if n == 0 or not (flags >> sid) & 1:
                continue
            got = eng.sortlist(cid, sid, n)
            want = orc.sortlist(cid, sid, n)
            assert np.array_equal(got, want), f"{name}: cell {cid} sid {sid}"
            nlists += 1
    assert nlists > 0


def test_counts_reference_window(case):
    name, sys, orc, eng = case
    o = orc.nonbond(po.OR_WINDOW_REF)
    g = eng.gpu_results()
    assert g["pairs"] == o["pairs"]
    assert g["pairs_refwindow"] == o["pairs"]
    assert g["exclusions"] == o["exclusions"]


def test_forces_energy_vs_fp64(case):
    name, sys, orc, eng = case
    o = orc.nonbond(po.OR_WINDOW_REF)
    en = eng.energies()
    f = eng.forces()
    err = rel_rmse(f, o["f"])
    assert err <= FORCE_RTOL, f"{name}: force rel RMSE {err:.3e}"
    assert abs(en["epot_nonbond"] - o["epot"]) <= ENERGY_RTOL * abs(o["epot"])
    assert abs(en["epot_exclusion"] - o["epot_exclusion"]) <= ENERGY_RTOL * max(1.0, abs(o["epot_exclusion"]))
    assert abs(en["virial"] - o["virial"]) <= ENERGY_RTOL * abs(o["virial"])
    # Newton's third law: both cells of a pair see opposite pair forces (to
    # FP32 rounding in the full shell, exactly in the half shell).  On the
    # mixed lattice the net forces are small against the pair forces they
    # sum, so the same rounding is a larger share of them.
    newton = 3e-6 if name.startswith("lattice_mixed_cells") else 1e-6
    assert np.abs(f.sum(0)).max() <= newton * np.abs(f).sum(0).max()


def test_projected_window_finds_every_pair(case):
    name, sys, orc, _ = case
    mode = MODES[name.split("-", 1)[1]]
    eng = engine_pass(sys, md.engine_flag_gpu | md.engine_flag_projwindow | mode)
    try:
        g = eng.gpu_results()
        p = orc.nonbond(po.OR_WINDOW_PROJ)
        r = orc.nonbond(po.OR_WINDOW_REF)
        b = orc.bruteforce()
        # The pair pass evaluates excluded pairs and the exclusion pass takes
        # them out again; the brute force never counts them.
        assert g["pairs"] == p["pairs"] == b["pairs"] + p["exclusions"]
        assert g["pairs_refwindow"] == r["pairs"]
        err = rel_rmse(eng.forces(), b["f"] if sys.excl is None else p["f"])
        assert err <= FORCE_RTOL
    finally:
        eng.close()


def test_repeatable(case):
    """Two passes over the same input give bitwise-identical forces/energy."""
    name, sys, orc, eng = case
    f1 = eng.forces().copy()
    e1 = eng.gpu_results()
    eng.nonbond()
    f2 = eng.forces()
    e2 = eng.gpu_results()
    assert np.array_equal(f1, f2), f"{name}: forces differ between passes"
    assert e1 == e2


@pytest.mark.parametrize("name,expect", [("mixture_4k", "full"), ("water_1k", "half"), ("dense_bigcells", "half")])
def test_shell_choice(name, expect):
    """Without a shell flag the engine takes the half shell for cells of
    engine_gpu_half_mean atoms or more on average when they are fewer than
    the device's compute units (all three systems have fewer): its forces
    are bitwise those of the forced shell (the two shells add in different
    orders)."""
    sys = SYSTEMS[name]()
    ncells = np.prod(np.floor(sys.dim / sys.L))
    mean = len(sys.x) / ncells
    assert ncells < 256
    assert (mean >= md.engine_gpu_half_mean) == (expect == "half")
    out = {}
    for label, flags in (("auto", 0), ("full", md.engine_flag_fullshell), ("half", md.engine_flag_halfshell)):
        eng = engine_pass(sys, md.engine_flag_gpu | flags)
        out[label] = eng.forces().copy()
        eng.close()
    other = "half" if expect == "full" else "full"
    assert np.array_equal(out["auto"], out[expect])
    assert not np.array_equal(out["auto"], out[other])


def test_unsorted_flag_evaluates_every_pair():
    """engine_flag_unsorted (the reference's unsorted runner, runner.c:801-804)
    evaluates every pair within the cutoff: on the span-2 mixture, where the
    reference's sorted window misses some, the pass equals the brute force."""
    sys = SYSTEMS["mixture_4k"]()
    orc = po.from_system(sys)
    eng = engine_pass(sys, md.engine_flag_gpu | md.engine_flag_unsorted)
    try:
        g = eng.gpu_results()
        b = orc.bruteforce()
        r = orc.nonbond(po.OR_WINDOW_REF)
        assert g["pairs"] == b["pairs"] > r["pairs"]
        assert rel_rmse(eng.forces(), b["f"]) <= FORCE_RTOL
        assert abs(g["epot"] - b["epot"]) <= ENERGY_RTOL * abs(b["epot"])
    finally:
        eng.close()
        orc.close()
