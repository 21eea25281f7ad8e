"""The FP64 oracle, pinned against what does not depend on it:

- direction tables: bit-exact vs the reference's compiled cell.c, and the
  geometry they encode (unit axis of each flipped offset);
- sort: tie orders of the reference quicksort/selection sort traced by hand
  (all-equal lists and three mixed repeated-key lists of 16, 20 and 24
  words, one with a nested partition), each planted mistake in the oracle's
  quicksort failing at least one trace; sortedness and permutation on
  random keys;
- sort keys: range and spacing that the constants bias/dscale fix;
- task list: identical to the host library's, and the counts of the
  1M-atom geometry (62 pair tasks per cell);
- per-task window counts: the sorted window's visited set does not depend
  on the order of equal keys, so every task's count has a closed form over
  the keys, the shift and the cutoff, checked task by task on span-2
  geometries in both window modes and with widened windows;
- pair loops: all-pairs brute force (the projected window finds every pair,
  the reference window misses some on span-2 cells), Newton's third law,
  translation invariance, threaded == serial, exact exclusion cancellation;
- sort reuse (pseudo-Verlet): with stale sort lists the widened projected
  window still finds every pair brute force finds, and the unwidened one
  does not.
"""
import ctypes as C
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import mdcore as md
import pyoracle as po
import systems
from cases import face_pairs

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


# Mixed repeated keys, traced by hand through runner.c:110-174 (pivot =
# key at (lo+hi)/2; scan i up while key > pivot, j down while key < pivot;
# swap when i < j, then step both; push [lo,j] and [i,hi]; selection sort
# keeping the first maximum when hi-lo < 15).  Positions are p0..; entries
# are (index, key).
#
# A, n=16: pivot key 5 (index 7).  Swaps p0/p15, p1/p13, p2/p11, p4/p9,
#   p5/p7; i=j=6 holds a 5 and is stepped over: i=7, j=5.  Left p0..p5 =
#   (15,6)(13,8)(11,5)(3,8)(9,7)(7,5); p6 = (6,5); right p7..p15 =
#   (5,1)(8,2)(4,5)(10,3)(2,3)(12,0)(1,5)(14,3)(0,3).  Both sides go to
#   selection sort: left -> 13 3 9 15 11 7; right -> 4 1 10 2 14 0 8 5 12.
# B, n=20: pivot key 2 (index 9).  Swaps p0/p19, p1/p18, p3/p17, p4/p15,
#   p5/p14, p7/p13, p8/p11, p9/p10 -> i=10, j=9.  Left p0..p9 =
#   19 18 2 17 15 14 6 13 11 10, right p10..p19 = 9 8 12 7 5 4 16 3 1 0;
#   selection sorts give 18 2 14 6 10 19 17 13 11 15 and
#   9 7 4 3 0 12 16 8 1 5.
# C, n=24: pivot key 2 (index 11).  Swaps p11/p23, p17/p22; i=j=21 holds
#   a 2: i=22, j=20.  [22,23] = (17,1)(11,2) -> 11 17.  [0,20] has 21
#   words: pivot key 6 at p10 (index 10); swaps p1/p20, p4/p18, p5/p17,
#   p9/p16, p10/p15 -> i=14, j=13.  [14,20] = 14 10 9 5 4 19 1 ->
#   10 4 9 19 14 5 1; [0,13] = 0 20 2 3 18 22 6 7 8 16 15 23 12 13 ->
#   0 3 8 15 6 20 13 7 2 23 12 18 22 16.
TRACED = {
    "A": ([3, 5, 3, 8, 5, 1, 5, 5, 2, 7, 3, 5, 0, 8, 3, 6],
          [13, 3, 9, 15, 11, 7, 6, 4, 1, 10, 2, 14, 0, 8, 5, 12]),
    "B": ([2, 1, 3, 2, 2, 1, 3, 2, 1, 2, 3, 2, 1, 2, 3, 2, 1, 2, 3, 2],
          [18, 2, 14, 6, 10, 19, 17, 13, 11, 15, 9, 7, 4, 3, 0, 12, 16, 8, 1, 5]),
    "C": ([9, 4, 7, 9, 6, 4, 8, 7, 9, 5, 6, 2, 7, 8, 4, 9, 6, 1, 7, 5, 8, 2, 6, 7],
          [0, 3, 8, 15, 6, 20, 13, 7, 2, 23, 12, 18, 22, 16, 10, 4, 9, 19, 14, 5, 1, 21, 11, 17]),
}


@pytest.mark.parametrize("case", sorted(TRACED))
def test_quicksort_mixed_keys_hand_traced(case):
    keys, order = TRACED[case]
    out = po.sort_desc(words(keys))
    assert list(out >> 16) == order
    assert list(out & 0xFFFF) == [keys[i] for i in order]


PLANTS = ["OR_PLANT_PIVOT_UP", "OR_PLANT_CUT16", "OR_PLANT_LAST_MAX", "OR_PLANT_GE_SCANS"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("plant", PLANTS)
def test_planted_sort_mistakes_fail_the_traces(plant, tmp_path):
    """The oracle compiled with one planted mistake (pivot rounded up, the
    selection-sort threshold at 16, the last maximum kept, scans that pass
    keys equal to the pivot) gives another order for at least one traced
    list: the traces pin the tie order, not just sortedness."""
    so = tmp_path / f"lib_{plant}.so"
    src = Path(po.HERE) / "oracle.c"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", f"-D{plant}", "-o", str(so), str(src),
                    "-lm", "-lpthread"], check=True)
    lib = C.CDLL(str(so))
    lib.or_sort_desc.argtypes = [C.POINTER(C.c_uint), C.c_int]
    wrong = 0
    for keys, order in TRACED.values():
        a = words(keys).copy()
        lib.or_sort_desc(a.ctypes.data_as(C.POINTER(C.c_uint)), len(a))
        wrong += list(a >> 16) != order
    assert wrong >= 1, plant


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


def window_counts_closed_form(o, sysm, mode, maxdx=None):
    """Pairs every task of the oracle's task list evaluates, from a formula
    that does not depend on the order of equal keys.

    Pair task (ci, cj, sid) (runner_dopair.c:397-537): i walks ci's list in
    descending key order, j walks cj's in ascending order while
    key_j + dnshift - key_i < dmaxdist, in unsigned arithmetic for the
    reference window (dnshift = dscale*|shift|) and signed for the projected
    one (dscale*shift.n).  The break on i (line 462) removes only atoms for
    which no j qualifies, except when it wraps: with key_i,max > key_j,min +
    dnshift the unsigned difference is huge at the first i and the task is
    empty.  The visited set is therefore {(i, j): key_j + dnshift - key_i <
    dmaxdist} (empty on a wrap), of which the pairs with a potential and
    r^2 <= rc^2 are evaluated, r^2 from dx = (x_i - shift) - x_j
    (fptype.h:275-278).  Self task (753-855): every j < i in storage order.
    Keys: (unsigned)(dscale*(bias + x.n)) (runner_dosort.c:148-149), from
    the cell-relative FP64 positions."""
    dim = np.asarray(sysm.dim, float)
    L = np.asarray(sysm.L, float)
    cdim = np.floor(dim / L).astype(int)
    h = dim / cdim
    origin = np.asarray(sysm.origin, float)
    bias = math.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])
    dscale = 32767.0 / (2 * bias)
    assert dscale == o.dscale()
    rc = float(sysm.cutoff)
    rc2 = rc * rc
    axes = np.ctypeslib.as_array((C.c_double * 39).in_dll(po.lib(), "or_axis")).reshape(13, 3)
    tabs = systems.tables(sysm)
    mt = tabs["max_type"]
    haspot = (tabs["pind"] >= 0).reshape(mt, mt)
    if maxdx is None:
        maxdx = np.zeros(o.ncells)
    cells = []
    for cid in range(o.ncells):
        ids, _ = o.cell(cid)
        cells.append((o.cell_x(cid, len(ids)), sysm.types[ids]))
    loc = np.stack(np.unravel_index(np.arange(o.ncells), tuple(cdim)), 1).astype(float)
    corg = origin + loc * h

    def keys(x, sid):
        n = axes[sid]
        return (dscale * (((bias + x[:, 0] * n[0]) + x[:, 1] * n[1]) + x[:, 2] * n[2])).astype(np.int64)

    out = []
    wraps = 0
    wrapped_pairs = 0
    for typ, ci, cj, sid in o.tasks():
        xi, ti = cells[ci]
        if typ == 1:
            d = xi[:, None, :] - xi[None, :, :]
            r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
            ok = np.tril(np.ones((len(xi), len(xi)), bool), -1) & (r2 <= rc2) & haspot[ti[:, None], ti[None, :]]
            out.append(int(ok.sum()))
            continue
        xj, tj = cells[cj]
        if len(xi) == 0 or len(xj) == 0:
            out.append(0)
            continue
        sh = corg[cj] - corg[ci]
        sh = np.where(sh * 2 > dim, sh - dim, np.where(sh * 2 < -dim, sh + dim, sh))
        dmaxdist = int(2 + dscale * ((rc + maxdx[ci]) + maxdx[cj]))
        if mode == po.OR_WINDOW_REF:
            dnshift = int(dscale * math.sqrt(sh[0] * sh[0] + sh[1] * sh[1] + sh[2] * sh[2]))
        else:
            n = axes[sid]
            dnshift = int(dscale * ((sh[0] * n[0] + sh[1] * n[1]) + sh[2] * n[2]))
        ki, kj = keys(xi, sid), keys(xj, sid)
        d = (xi - sh)[:, None, :] - xj[None, :, :]
        r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        ok = (kj[None, :] + dnshift - ki[:, None] < dmaxdist) & (r2 <= rc2) & haspot[ti[:, None], tj[None, :]]
        if mode == po.OR_WINDOW_REF and ki.max() > kj.min() + dnshift:
            wraps += 1
            wrapped_pairs += int(ok.sum())
            out.append(0)
        else:
            out.append(int(ok.sum()))
    return np.array(out, np.int64), wraps, wrapped_pairs


def uniform_system(dim, L, cutoff, density, seed):
    dim = np.asarray(dim, float)
    N = int(density * np.prod(dim))
    x = md.gen_uniform(3 * N, 0.0, 1.0, seed).reshape(-1, 3) * dim
    pot = md.LJ126(0.05, cutoff, 1e-6, 1e-3, 1e-3)
    return systems.System("uniform", np.zeros(3), dim, np.asarray(L, float), cutoff, x, np.ones(N, np.int32),
                          np.array([0.0, 1.0]), [(1, 1, pot)])


WINDOW_GEOMETRIES = {
    # the 1M mixture's geometry: 0.6 nm cells, rc 1.2 nm (span 2 on every axis)
    "mixture_span2": lambda: systems.mixture(5000, seed=4),
    # span (2,1,1): offsets such as (2,1,0), where |shift| and shift.n differ most
    "aniso_span211": lambda: uniform_system((4.4, 5.25, 5.25), (0.55, 1.05, 1.05), 1.0, 60.0, 3),
    # pairs planted across faces, edges and corners: the unsigned break wraps
    "face_pairs": lambda: face_pairs(),
}


@pytest.mark.parametrize("mode", [po.OR_WINDOW_REF, po.OR_WINDOW_PROJ], ids=["ref", "proj"])
@pytest.mark.parametrize("geom", sorted(WINDOW_GEOMETRIES))
def test_window_counts_per_task_closed_form(geom, mode):
    sysm = WINDOW_GEOMETRIES[geom]()
    o = po.from_system(sysm)
    got = o.nonbond(mode)
    counts = o.task_counts()
    want, wraps, wrapped = window_counts_closed_form(o, sysm, mode)
    assert counts.sum() == got["pairs"]
    assert np.array_equal(counts, want)
    if mode == po.OR_WINDOW_REF:
        if geom == "face_pairs":
            # the wrap-around rule is exercised: tasks it empties hold pairs
            # a signed window would have evaluated
            assert wraps > 0 and wrapped > 0
        prj = po.from_system(sysm).nonbond(po.OR_WINDOW_PROJ)["pairs"]
        print(f"{geom}: reference window misses {1 - got['pairs'] / prj:.4%} of the in-cutoff pairs; "
              f"{wraps} wrapped tasks holding {wrapped} pairs")
    o.close()


def test_window_counts_widened_closed_form():
    """Per-cell maxdx widens dmaxdist = 2 + dscale*(rc + maxdx_i + maxdx_j)
    (runner_dopair.c:428) with a fresh sort: every task's count still
    follows the closed form."""
    sysm = WINDOW_GEOMETRIES["mixture_span2"]()
    o = po.from_system(sysm)
    maxdx = np.random.default_rng(5).uniform(0.0, 0.06, o.ncells)
    o.set_verlet(None, maxdx, keep_sort=False)
    for mode in (po.OR_WINDOW_REF, po.OR_WINDOW_PROJ):
        o.nonbond(mode)
        want, _, _ = window_counts_closed_form(o, sysm, mode, maxdx)
        assert np.array_equal(o.task_counts(), want), mode
    o.close()


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


def test_bulk_sortlists_match_per_cell_lists():
    """or_sortlists (the oracle's every-cell copy used by the full-size GPU
    comparison) holds each cell's flagged lists at sid*count in its block,
    cells in id order, and the host library's engine_getsortlists has the
    same size."""
    sysm = systems.mixture(3000, seed=2)
    o = po.from_system(sysm)
    o.sort()
    bulk = o.sortlists()
    off = 0
    for cid in range(o.ncells):
        ids, flags = o.cell(cid)
        n = len(ids)
        for sid in range(13):
            if n and (flags >> sid) & 1:
                assert np.array_equal(bulk[off + sid * n:off + (sid + 1) * n], o.sortlist(cid, sid, n))
        off += 13 * n
    assert off == len(bulk) == 13 * sysm.N
    eng = systems.build_engine(sysm, md.engine_flag_gpu, start=False)
    assert len(eng.sortlists()) == 13 * sysm.N
    eng.close()
    o.close()


def test_advance_is_the_kick_drift():
    """or_advance: v += dt f/m, x += dt v per particle with the last pass's
    forces (engine.c:1945-2052), the same on 1 and 3 threads; at dt = 0
    nothing moves."""
    sysm = systems.mixture(3000, seed=2)
    out = []
    for threads in (1, 3):
        o = po.from_system(sysm)
        r = o.nonbond(po.OR_WINDOW_REF)
        v0 = np.random.default_rng(1).normal(size=(sysm.N, 3))
        im = 1.0 / np.where(sysm.masses > 0, sysm.masses, 1.0)
        o.set_motion(v0, im)
        x0 = np.concatenate([o.cell_x(c, len(o.cell(c)[0])) for c in range(o.ncells)])
        ids = np.concatenate([o.cell(c)[0] for c in range(o.ncells)])
        assert o.advance(0.0, threads) == 0
        assert np.array_equal(np.concatenate([o.cell_x(c, len(o.cell(c)[0])) for c in range(o.ncells)]), x0)
        dt = 1e-5
        o.advance(dt, threads)
        x1 = np.concatenate([o.cell_x(c, len(o.cell(c)[0])) for c in range(o.ncells)])
        v1 = v0[ids] + r["f"][ids] * (dt * im[sysm.types[ids]])[:, None]
        assert np.allclose(x1, x0 + dt * v1, rtol=0, atol=1e-15)
        out.append(x1)
        o.close()
    assert np.array_equal(out[0], out[1])
