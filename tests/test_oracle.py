"""The FP64 oracle, pinned against what does not depend on it:

- direction tables: bit-exact vs the reference's compiled cell.c, and the
  geometry they encode (unit axis of each flipped offset);
- sort: hand-traced tie orders of the reference quicksort/selection sort,
  plus sortedness and permutation on random keys;
- sort keys: range and spacing that the constants bias/dscale fix;
- task list: identical to the host library's, and the counts of the
  1M-atom geometry (62 pair tasks per cell);
- pair loops: all-pairs brute force (the projected window finds every pair,
  the reference window misses some on span-2 cells), Newton's third law,
  translation invariance, threaded == serial, exact exclusion cancellation;
- sort reuse (pseudo-Verlet): with stale sort lists the widened projected
  window still finds every pair brute force finds, and the unwidened one
  does not.
"""
import ctypes as C

import numpy as np
import pytest

import mdcore as md
import pyoracle as po
import systems

# ---------------------------------------------------------------- directions


def offset_geometry():
    """For pattern p = 9(ox+1)+3(oy+1)+(oz+1): flip iff the offset is
    lexicographically negative; the axis is the unit vector of the
    flipped (canonical) offset."""
    out = []
    for p in range(27):
        o = np.array([p // 9 - 1, (p // 3) % 3 - 1, p % 3 - 1], float)
        out.append((p, o, tuple(o) < (0, 0, 0)))
    return out


def test_direction_tables_encode_the_geometry():
    lib = md.library()
    sid = np.frombuffer((C.c_char * 27).in_dll(lib, "cell_sortlistID"), dtype=np.int8)
    axes = np.frombuffer((C.c_double * 39).in_dll(lib, "cell_shift"), dtype=np.float64).reshape(13, 3)
    flip = np.frombuffer((C.c_char * 27).in_dll(lib, "cell_flip"), dtype=np.int8)
    seen = set()
    for p, o, neg in offset_geometry():
        assert bool(flip[p]) == neg, p
        if p == 13:
            continue
        canon = -o if neg else o
        assert np.allclose(axes[sid[p]], canon / np.linalg.norm(canon), atol=1e-15, rtol=0)
        seen.add(int(sid[p]))
    assert seen == set(range(13))
    oracle_axes = np.ctypeslib.as_array((C.c_double * 39).in_dll(po.lib(), "or_axis"))
    assert np.array_equal(oracle_axes, axes.reshape(-1))


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_direction_tables_bitexact_vs_reference():
    sid, shift, flip = po.ref_cell_tables()
    lib = md.library()
    assert np.array_equal(np.frombuffer((C.c_char * 27).in_dll(lib, "cell_sortlistID"), np.int8), sid)
    assert np.array_equal(np.frombuffer((C.c_double * 39).in_dll(lib, "cell_shift"), np.float64), shift)
    assert np.array_equal(np.frombuffer((C.c_char * 27).in_dll(lib, "cell_flip"), np.int8), flip)


# ---------------------------------------------------------------- sort


def words(keys):
    return np.array([(i << 16) | k for i, k in enumerate(keys)], np.uint32)


def test_selection_sort_tie_order_hand_traced():
    # N < 16: selection sort with a strict '>' keeps the first maximum and
    # swaps it forward.  [5,7,5,7,3] -> indices 1,3,2,0,4 (traced by hand).
    out = po.sort_desc(words([5, 7, 5, 7, 3]))
    assert list(out >> 16) == [1, 3, 2, 0, 4]
    assert list(out & 0xFFFF) == [7, 7, 5, 5, 3]


@pytest.mark.parametrize("n", [16, 17, 31])
def test_quicksort_equal_keys_reverse_hand_traced(n):
    # All keys equal: the partition swaps (0,n-1),(1,n-2),... until the
    # pointers cross, then both halves go to selection sort, which never
    # swaps equal keys.  Result: the indices reversed.
    out = po.sort_desc(words([9] * n))
    assert list(out >> 16) == list(range(n - 1, -1, -1))


@pytest.mark.parametrize("seed", range(5))
def test_sort_descending_permutation(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    keys = rng.integers(0, 40 if seed % 2 else 32768, n)
    out = po.sort_desc(words(keys))
    k = out & 0xFFFF
    assert np.all(k[:-1] >= k[1:])
    assert sorted(out >> 16) == list(range(n))
    assert np.array_equal(keys[out >> 16], k)


def test_key_range_and_spacing():
    """bias = |h| and dscale = 32767/(2|h|) map x.n in [-|h|, |h|] onto [0, 32767]."""
    h = np.array([0.60793, 0.60793, 0.60793])
    bias = np.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])
    dscale = 32767.0 / (2 * bias)
    axes = np.ctypeslib.as_array((C.c_double * 39).in_dll(po.lib(), "or_axis")).reshape(13, 3)
    rng = np.random.default_rng(3)
    for sid in range(13):
        n = axes[sid]
        corner_lo = np.where(n < 0, h * (1 - 1e-12), 0.0)   # minimises x.n in the cell
        corner_hi = np.where(n > 0, h * (1 - 1e-12), 0.0)   # maximises x.n
        klo = po.key(corner_lo, n, dscale, bias)
        khi = po.key(corner_hi, n, dscale, bias)
        assert 0 <= klo <= khi <= 32767
        # the full cell spans |h| * sum|n_k| of the axis: keys spread accordingly
        span = dscale * float(np.dot(np.abs(n), h))
        assert abs((khi - klo) - span) <= 1.0 + 1e-9
        # a displacement d along the axis moves the key by dscale*d (+-1 for truncation)
        x = rng.uniform(0.1, 0.3, 3)
        d = 0.05
        assert abs(int(po.key(x + d * n, n, dscale, bias)) - int(po.key(x, n, dscale, bias)) - dscale * d) <= 1.0
    # the cell's own corner (0,0,0) projects onto the middle of the key range
    assert po.key(np.zeros(3), axes[12], dscale, bias) == 16383


# ---------------------------------------------------------------- tasks


def test_task_list_matches_host_library():
    for sysm in (systems.argon(6), systems.mixture(3000, seed=2)):
        eng = systems.build_engine(sysm, md.engine_flag_gpu, start=False)
        orc = po.from_system(sysm)
        mine = eng.tasks()
        theirs = orc.tasks()
        # host list = pairs/selves in the reference order, then one sort task per cell
        nps = len(theirs)
        assert np.array_equal(mine[:nps, :3], theirs[:, :3])
        assert np.array_equal(mine[:nps, 3][mine[:nps, 0] == 2], theirs[:, 3][theirs[:, 0] == 2])
        assert np.all(mine[nps:, 0] == 3) and len(mine) - nps == orc.ncells
        eng.close()


def test_task_counts_of_the_1m_geometry():
    box = (1048576 / 100.0) ** (1 / 3)
    e = md.Engine([0, 0, 0], [box] * 3, [0.6] * 3, 1.2, 1, md.engine_flag_gpu)
    t = e.tasks()
    assert (t[:, 0] == 2).sum() == 36 ** 3 * 62
    assert (t[:, 0] == 1).sum() == 36 ** 3
    e.close()


# ---------------------------------------------------------------- pair loops


@pytest.fixture(scope="module")
def small_systems():
    return {
        "argon": systems.argon(6, jitter=0.05),
        "mixture": systems.mixture(3000, seed=4),
        "water": systems.water(1000, 3.1, (10, 10, 10), seed=9),
    }


@pytest.mark.parametrize("name", ["argon", "mixture", "water"])
def test_windows_against_bruteforce(small_systems, name):
    sysm = small_systems[name]
    o = po.from_system(sysm)
    ref = o.nonbond(po.OR_WINDOW_REF)
    prj = o.nonbond(po.OR_WINDOW_PROJ)
    bf = o.bruteforce()
    # excluded pairs are evaluated by the pair pass and removed again
    assert prj["pairs"] == bf["pairs"] + prj["exclusions"]
    assert ref["pairs"] <= prj["pairs"]
    if name == "mixture":   # span 2: the |shift| window drops in-cutoff pairs
        assert ref["pairs"] < prj["pairs"]
    else:                   # span 1, cubic cells: the two windows agree
        assert ref["pairs"] == prj["pairs"]
    scale = np.abs(bf["f"]).max()
    assert np.abs(prj["f"] - bf["f"]).max() <= 1e-9 * scale
    assert prj["epot"] + prj["epot_exclusion"] == pytest.approx(bf["epot"], rel=1e-10, abs=1e-8)


@pytest.mark.parametrize("name", ["argon", "mixture", "water"])
def test_newton_third_law(small_systems, name):
    r = po.from_system(small_systems[name]).nonbond(po.OR_WINDOW_REF)
    assert np.abs(r["f"].sum(0)).max() <= 1e-10 * np.abs(r["f"]).sum(0).max()


def test_threads_match_serial(small_systems):
    o = po.from_system(small_systems["mixture"])
    a = o.nonbond(po.OR_WINDOW_REF, 1)
    b = o.nonbond(po.OR_WINDOW_REF, 4)
    assert a["pairs"] == b["pairs"]
    assert np.abs(a["f"] - b["f"]).max() <= 1e-11 * np.abs(a["f"]).max()
    assert a["epot"] == pytest.approx(b["epot"], rel=1e-13)


def test_translation_invariance(small_systems):
    sysm = small_systems["argon"]
    o1 = po.from_system(sysm)
    r1 = o1.nonbond(po.OR_WINDOW_PROJ)
    x2 = np.mod(sysm.x + np.array([0.137, 0.071, 0.213]), sysm.dim)
    o2 = po.Oracle(sysm.origin, sysm.dim, sysm.L, sysm.cutoff, x2, sysm.types, systems.tables(sysm), sysm.excl)
    r2 = o2.nonbond(po.OR_WINDOW_PROJ)
    assert abs(r1["pairs"] - r2["pairs"]) <= 2
    assert r1["epot"] == pytest.approx(r2["epot"], rel=1e-9)
    assert np.abs(r1["f"] - r2["f"]).max() <= 1e-7 * np.abs(r1["f"]).max()


def test_exclusion_cancels_the_pair():
    """An isolated excluded pair: pair pass + exclusion pass = 0."""
    sysm = systems.water(1, 3.0, (1, 1, 1), seed=1, jitter=0.0)
    o = po.from_system(sysm)
    r = o.nonbond(po.OR_WINDOW_REF)
    assert r["pairs"] == 3 and r["exclusions"] == 3
    assert abs(r["epot"] + r["epot_exclusion"]) <= 1e-12 * abs(r["epot"])
    assert np.abs(r["f"]).max() <= 1e-12 * 1e4


def test_reference_window_miss_fraction_matches_probe():
    """The |shift| window of runner_dopair.c:451-452 drops in-cutoff pairs on
    span-2 offsets.  At the 1M-mixture geometry (0.6 nm cells, rc 1.2 nm,
    100 atoms/nm^3, uniform random positions) the compiled FP64 reference
    runner was measured to miss 2.186 % of the pairs the exact window finds;
    the restatement must land on the same figure (sampling noise ~0.02 %)."""
    L = 6.0
    N = int(100 * L ** 3)
    x = md.gen_uniform(3 * N, 0.0, L, 1).reshape(-1, 3)
    pot = md.LJ126(0.05, 1.2, 1e-6, 1e-3, 1e-3)
    sysm = systems.System("uniform", np.zeros(3), np.full(3, L), np.full(3, 0.6), 1.2, x, np.ones(N, np.int32),
                          np.array([0.0, 1.0]), [(1, 1, pot)])
    o = po.from_system(sysm)
    ref = o.nonbond(po.OR_WINDOW_REF, 4)["pairs"]
    prj = o.nonbond(po.OR_WINDOW_PROJ, 4)["pairs"]
    miss = 1.0 - ref / prj
    assert 0.0210 <= miss <= 0.0228, miss


def test_sort_reuse_window_widening():
    """Pseudo-Verlet sort reuse (runner.c:773: no re-sort between rebuilds;
    runner_dopair.c:428: the window widened by both cells' maxdx).  After the
    particles move up to 0.06 nm (within half the 0.13 nm skin) without a
    re-sort, the projected window still finds every in-cutoff pair when
    widened by each cell's largest displacement, and misses some without the
    widening; a fresh sort needs none.  Brute force on the moved positions is
    the check."""
    sysm = systems.argon(6, jitter=0.05)  # 864 atoms, 3^3 cells of 1.13 nm
    o = po.from_system(sysm)
    o.nonbond(po.OR_WINDOW_PROJ)  # sorts at the start positions
    nc, cdim = o.ncells, 3
    h = sysm.dim / cdim
    cell = np.empty(sysm.N, np.int64)
    for cid in range(nc):
        ids, _ = o.cell(cid)
        cell[ids] = cid
    loc = np.stack([cell // (cdim * cdim), (cell // cdim) % cdim, cell % cdim], 1)
    xrel0 = np.mod(sysm.x, sysm.dim) - loc * h
    rng = np.random.default_rng(3)
    d = rng.normal(size=(sysm.N, 3))
    d *= (0.06 * rng.uniform(0.0, 1.0, sysm.N) / np.linalg.norm(d, axis=1))[:, None]
    maxdx = np.zeros(nc)
    np.maximum.at(maxdx, cell, np.linalg.norm(d, axis=1))
    o.set_verlet(xrel0 + d, maxdx, keep_sort=True)
    reuse = o.nonbond(po.OR_WINDOW_PROJ)
    bf = o.bruteforce()
    assert reuse["pairs"] == bf["pairs"]
    assert np.abs(reuse["f"] - bf["f"]).max() <= 1e-9 * np.abs(bf["f"]).max()
    o.set_verlet(None, np.zeros(nc), keep_sort=True)
    narrow = o.nonbond(po.OR_WINDOW_PROJ)
    assert narrow["pairs"] < bf["pairs"]
    o.set_verlet(None, None, keep_sort=False)
    fresh = o.nonbond(po.OR_WINDOW_PROJ)
    assert fresh["pairs"] == bf["pairs"]
    print(f"pairs: brute force {bf['pairs']}, reused sort widened {reuse['pairs']}, "
          f"not widened {narrow['pairs']}")
