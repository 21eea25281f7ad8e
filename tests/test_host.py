"""Host library: the C-ABI exports, the reference's error behaviour, and the
host data model (cell contents after load + shuffle) against the oracle."""
import re
from pathlib import Path

import numpy as np
import pytest

import mdcore as md
import pyoracle as po
import systems

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    """Names of the functions declared in include/*.h (prototypes ending in ');')."""
    names = set()
    decl = re.compile(r"^(?:extern\s+)?(?:const\s+)?(?:struct\s+\w+|unsigned\s+\w+|\w+)\s*\**\s*([A-Za-z_]\w*)\s*\(")
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
        text = re.sub(r"__attribute__\s*\(\(.*?\)\)", "", text)
        for stmt in re.findall(r"[^;{}]*?\)\s*;", text):
            m = decl.match(" ".join(stmt.split()))
            if m:
                names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = md.library()
    names = header_functions()
    assert "engine_nonbond_gpu" in names and "engine_gpu_load" in names and "engine_init" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_engine_init_rejects_bad_input():
    lib = md.library()
    e = lib.engine_alloc()
    o = np.zeros(3)
    d = np.full(3, 5.0)
    L = np.full(3, 1.0)
    assert lib.engine_init(e, None, md.ptr(d), md.ptr(L), 1.0, 7, 2, md.engine_flag_gpu) == -1  # engine_err_null
    assert lib.engine_init(e, md.ptr(o), md.ptr(d), md.ptr(L), 1.0, 7, 0, md.engine_flag_gpu) == -6  # range
    assert "engine_init" in md.errors()
    lib.errs_clear()
    lib.engine_free(e)


def test_load_outside_box_and_start_without_gpu_flag():
    e = md.Engine([0, 0, 0], [3.0] * 3, [1.0] * 3, 1.0, 1, md.engine_flag_none)
    e.addtype(1.0)
    with pytest.raises(md.MdcoreError) as ex:
        e.load(np.array([[3.5, 0.1, 0.1]]), np.array([0]))
    assert "space_addpart" in str(ex.value)
    e.load(np.array([[0.5, 0.5, 0.5]]), np.array([0]))
    with pytest.raises(md.MdcoreError) as ex:
        e.start()
    assert "(-22)" in str(ex.value)  # engine_err_nogpu: no CPU fallback
    with pytest.raises(md.MdcoreError) as ex:
        e.gpu_steps(3)                   # device steps need engine_gpu_load
    assert "(-22)" in str(ex.value)
    e.close()


def test_addpot_and_type_ranges():
    e = md.Engine([0, 0, 0], [3.0] * 3, [1.0] * 3, 1.0, 2, md.engine_flag_gpu)
    assert e.addtype(1.0) == 1 and e.addtype(2.0) == 2   # type 0 is the reserved null type
    with pytest.raises(md.MdcoreError):
        e.addtype(3.0)                                     # max_type reached
    p = md.LJ126(0.275, 1.0, 9.677e-6, 6.209e-3, 1e-3)
    with pytest.raises(md.MdcoreError):
        e.addpot(p, 1, 3)
    e.addpot(p, 1, 2)
    e.close()


@pytest.mark.parametrize("make", [lambda: systems.argon(6, jitter=0.08), lambda: systems.mixture(2500, seed=6),
                                  lambda: systems.water(1000, 3.1, (10, 10, 10), seed=2)])
def test_host_cells_match_oracle(make):
    """engine_load + engine_shuffle place particles in the same cells, in the
    same order, as the oracle's restatement of space_addpart/shuffle."""
    sysm = make()
    eng = systems.build_engine(sysm, md.engine_flag_gpu, start=False)
    eng.shuffle()
    orc = po.from_system(sysm)
    counts, ids = eng.cells(orc.ncells)
    pos = 0
    for cid in range(orc.ncells):
        oids, _ = orc.cell(cid)
        assert np.array_equal(ids[pos:pos + counts[cid]], oids)
        pos += counts[cid]
    assert pos == sysm.N
    x = eng.positions()
    assert np.abs(np.mod(x - sysm.x + 0.5 * sysm.dim, sysm.dim) - 0.5 * sysm.dim).max() < 1e-12
    eng.close()


def test_exclusions_are_symmetric_and_deduplicated():
    e = md.Engine([0, 0, 0], [3.0] * 3, [1.0] * 3, 1.0, 1, md.engine_flag_gpu)
    for i, j in [(3, 1), (1, 3), (2, 2), (5, 4), (4, 5), (1, 3)]:
        e.exclusion_add(i, j)
    e.exclusion_shrink()
    e.close()


def test_setdt_rejects_negative():
    e = md.Engine([0, 0, 0], [3.0] * 3, [1.0] * 3, 1.0, 1, md.engine_flag_gpu)
    e.setdt(0.002)
    with pytest.raises(md.MdcoreError):
        e.setdt(-1.0)
    md.library().errs_clear()
    e.close()


def test_potential_eval_r_matches_scalar_eval():
    p = md.LJ126_RF(0.12, 1.2, 2e-6, 2e-3, 0.3, 138.935458, 0.0, 1e-3)
    r = np.linspace(0.1, 1.2, 257)
    e, dedr = p.eval_r(r)
    for k in range(0, 257, 16):
        es, fs = p.eval(r[k] * r[k])
        assert e[k] == es
        assert dedr[k] == fs * r[k]


def test_sample_pairs_finds_every_pair_once_per_side():
    """All in-cutoff ordered pairs of a small periodic box, against numpy."""
    rng = np.random.default_rng(5)
    L, N, rc = 3.3, 300, 1.0
    x = rng.random((N, 3)) * L
    t = rng.integers(1, 4, N).astype(np.int32)
    r, ti, tj = md.sample_pairs(x, t, [L] * 3, rc, 10 ** 6, 7)
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    want = np.sort(np.sqrt(r2[r2 <= rc * rc]))
    assert len(r) == len(want)
    assert np.allclose(np.sort(r), want, rtol=0, atol=1e-12)
    # types follow the pairs: counts of (ti, tj) combinations agree
    ii, jj = np.nonzero(r2 <= rc * rc)
    got = np.bincount(ti * 4 + tj, minlength=16)
    ref = np.bincount(t[ii] * 4 + t[jj], minlength=16)
    assert np.array_equal(got, ref)


def test_sample_pairs_rejects_small_box():
    with pytest.raises(md.MdcoreError):
        md.sample_pairs(np.zeros((4, 3)), np.ones(4, np.int32), [2.0] * 3, 1.0, 10, 1)


def test_set_velocities_round_trip():
    """engine_setvelocities writes p->v by particle id; engine_getvelocities
    reads the same values back, after a shuffle has reordered the cells."""
    sysm = systems.argon(4, jitter=0.05)
    eng = systems.build_engine(sysm, md.engine_flag_gpu, start=False)
    eng.shuffle()
    v = md.gen_uniform(3 * sysm.N, -1.0, 1.0, 5).reshape(-1, 3)
    eng.set_velocities(v)
    assert np.array_equal(eng.velocities(), v)
    with pytest.raises(ValueError):
        eng.set_velocities(v[:-1])
    eng.close()


def _serial_advance(cells, xl, v, dt, cdim, h):
    """The reference's advance on the host cells (engine.c:1945-2052), one
    serial loop: per cell, each particle is kicked (no forces here) and
    drifted, and one that left the cell moves at once to its neighbour's
    incoming list, its slot taken by the cell's last particle; then each
    cell appends its incoming particles in arrival order."""
    cells = [list(c) for c in cells]
    incoming = [[] for _ in cells]
    xl = xl.copy()
    for cid, parts in enumerate(cells):
        loc = np.array(np.unravel_index(cid, cdim))
        pid = 0
        while pid < len(parts):
            q = parts[pid]
            xl[q] = xl[q] + dt * v[q]
            delta = (xl[q] >= h).astype(int) - (xl[q] < 0.0).astype(int)
            if delta.any():
                xl[q] = xl[q] - delta * h
                dest = np.ravel_multi_index(tuple((loc + delta) % cdim), cdim)
                incoming[dest].append(q)
                last = parts.pop()
                if pid < len(parts):
                    parts[pid] = last  # kicked and drifted when the loop reaches it
            else:
                pid += 1
    return [c + inc for c, inc in zip(cells, incoming)], xl


def test_parallel_advance_keeps_the_serial_order():
    """engine_advance kick-drifts every particle in parallel and moves only
    the particles that left their cell, cell by cell in order: the cells'
    contents and order, and every position, equal one serial pass of the
    reference's loop, with many particles crossing faces (a step of 0.25 of
    the cell width per unit velocity)."""
    import ctypes as C
    sysm = systems.argon(6, jitter=0.05)
    eng = systems.build_engine(sysm, md.engine_flag_gpu, start=False)
    eng.shuffle()
    N = sysm.N
    v = md.gen_uniform(3 * N, -1.0, 1.0, 11).reshape(-1, 3)
    eng.set_velocities(v)
    cdim = tuple(int(n) for n in np.floor(sysm.dim / sysm.L))
    h = sysm.dim / np.array(cdim)
    ncells = int(np.prod(cdim))
    counts, ids = eng.cells(ncells)
    cells = np.split(ids, np.cumsum(counts)[:-1])

    def local():
        x = np.zeros(3 * N)
        cell = np.zeros(N, dtype=np.int32)
        assert eng.lib.engine_getlocal(eng.e, x.ctypes.data_as(C.POINTER(C.c_double)),
                                       cell.ctypes.data_as(C.POINTER(C.c_int)), N) == N
        return x.reshape(-1, 3)

    dt = 0.25 * float(h.min())
    want_cells, want_x = _serial_advance(cells, local(), eng.velocities(), dt, cdim, h)
    eng.setdt(dt)
    eng.advance()
    counts2, ids2 = eng.cells(ncells)
    got = np.split(ids2, np.cumsum(counts2)[:-1])
    moved = sum(len(a) != len(b) or not np.array_equal(a, b) for a, b in zip(cells, got))
    assert moved > ncells // 2
    for cid in range(ncells):
        assert np.array_equal(got[cid], np.array(want_cells[cid], dtype=ids.dtype)), f"cell {cid}"
    assert np.array_equal(local(), want_x)
    eng.close()
