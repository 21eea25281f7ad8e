"""ORACLE — test infrastructure only.

ctypes access to
  liboracle.so      the FP64 CPU restatement of the nonbonded path (oracle.c);
  _ref/libmdref.so  the reference's own table code (potential.c, errs.c,
                    cell.c compiled unmodified, see Makefile), when built.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF = HERE / "_ref" / "libmdref.so"

OR_WINDOW_REF = 0
OR_WINDOW_PROJ = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)
_llp = C.POINTER(C.c_longlong)

_lib = None
_ref = None


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


class _Tables(C.Structure):
    _fields_ = [("max_type", C.c_int), ("ntab", C.c_int), ("pind", _ip), ("n", _ip), ("alpha", _dp),
                ("c", _dp), ("off", _ip)]


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run make -C {HERE}")
        l = C.CDLL(str(LIB))
        sig = {
            "or_create": (C.c_void_p, [_dp, _dp, _dp, C.c_double, C.c_int, _dp, _ip, C.POINTER(_Tables), C.c_int,
                                       _ip]),
            "or_free": (None, [C.c_void_p]),
            "or_sort": (C.c_int, [C.c_void_p]),
            "or_nonbond": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _llp]),
            "or_bruteforce": (C.c_int, [C.c_void_p, _dp, _dp, _llp]),
            "or_ncells": (C.c_int, [C.c_void_p]),
            "or_cell": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _up]),
            "or_sortlist": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _up]),
            "or_ntasks": (C.c_int, [C.c_void_p, _ip, _ip]),
            "or_task": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip]),
            "or_dscale": (C.c_double, [C.c_void_p]),
            "or_task_counts": (C.c_int, [C.c_void_p, _llp]),
            "or_cell_x": (C.c_int, [C.c_void_p, C.c_int, _dp]),
            "or_sortlists": (C.c_longlong, [C.c_void_p, _up]),
            "or_set_motion": (C.c_int, [C.c_void_p, _dp, _dp]),
            "or_advance": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
            "or_set_verlet": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
            "or_sort_desc": (None, [_up, C.c_int]),
            "or_key": (C.c_uint, [_dp, _dp, C.c_double, C.c_double]),
            "or_eval": (None, [C.c_int, _dp, _dp, C.c_double, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def ref_available():
    return REF.exists()


def ref():
    """The reference's compiled table code (raises if _ref was not built)."""
    global _ref
    if _ref is None:
        if not REF.exists():
            raise RuntimeError(f"{REF} missing: run make -C {HERE} ref where the reference tree exists")
        r = C.CDLL(str(REF))
        sig = {
            "potential_create_LJ126": (C.c_void_p, [C.c_double] * 5),
            "potential_create_LJ126_switch": (C.c_void_p, [C.c_double] * 6),
            "potential_create_LJ126_Coulomb": (C.c_void_p, [C.c_double] * 6),
            "potential_create_LJ126_Ewald": (C.c_void_p, [C.c_double] * 7),
            "potential_create_Ewald": (C.c_void_p, [C.c_double] * 5),
            "potential_create_Coulomb": (C.c_void_p, [C.c_double] * 4),
            "potential_create_harmonic": (C.c_void_p, [C.c_double] * 5),
            "ref_potential_create_LJ126_RF": (C.c_void_p, [C.c_double] * 8),
            "ref_potential_data": (C.c_int, [C.c_void_p, _ip, _dp, _dp, _dp, C.c_int]),
            "ref_potential_eval": (None, [C.c_void_p, C.c_double, _dp, _dp]),
            "potential_getalpha": (C.c_double, [C.CFUNCTYPE(C.c_double, C.c_double), C.c_double, C.c_double]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(r, name)
            fn.restype = res
            fn.argtypes = args
        _ref = r
    return _ref


def ref_table(handle):
    """(n, a, b, alpha[4], c) of a reference-built table."""
    r = ref()
    if not handle:
        raise RuntimeError("reference table construction failed")
    n = C.c_int(0)
    ab = np.zeros(2)
    alpha = np.zeros(4)
    c = np.zeros(8 * 400)
    rc = r.ref_potential_data(C.c_void_p(handle), C.byref(n), _p(ab), _p(alpha), _p(c), len(c))
    if rc < 0:
        raise RuntimeError("ref_potential_data failed")
    return {"n": n.value, "a": ab[0], "b": ab[1], "alpha": alpha, "c": c[: (n.value + 1) * 8].copy(),
            "handle": C.c_void_p(handle)}


def ref_eval(table, r2):
    e = C.c_double()
    f = C.c_double()
    ref().ref_potential_eval(table["handle"], float(r2), C.byref(e), C.byref(f))
    return e.value, f.value


def ref_cell_tables():
    """cell_sortlistID[27], cell_shift[39], cell_flip[27] of the reference's cell.c."""
    r = ref()
    sid = np.frombuffer((C.c_char * 27).in_dll(r, "cell_sortlistID"), dtype=np.int8).copy()
    shift = np.frombuffer((C.c_double * 39).in_dll(r, "cell_shift"), dtype=np.float64).copy()
    flip = np.frombuffer((C.c_char * 27).in_dll(r, "cell_flip"), dtype=np.int8).copy()
    return sid, shift, flip


# ---------------------------------------------------------------- system oracle

class Oracle:
    """or_create over the same inputs as the engine (tables passed as data)."""

    def __init__(self, origin, dim, L, cutoff, x, types, tables, excl=None):
        l = lib()
        self._keep = []
        o = np.ascontiguousarray(origin, np.float64)
        d = np.ascontiguousarray(dim, np.float64)
        LL = np.ascontiguousarray(L, np.float64)
        xx = np.ascontiguousarray(x, np.float64).reshape(-1)
        tt = np.ascontiguousarray(types, np.int32)
        pind = np.ascontiguousarray(tables["pind"], np.int32)
        n = np.ascontiguousarray(tables["n"], np.int32)
        alpha = np.ascontiguousarray(tables["alpha"], np.float64)
        c = np.ascontiguousarray(tables["c"], np.float64)
        off = np.ascontiguousarray(tables["off"], np.int32)
        self._keep += [o, d, LL, xx, tt, pind, n, alpha, c, off]
        T = _Tables(tables["max_type"], len(n), _p(pind, C.c_int), _p(n, C.c_int), _p(alpha), _p(c),
                    _p(off, C.c_int))
        ex = None
        nex = 0
        if excl is not None and len(excl):
            ex = np.ascontiguousarray(excl, np.int32).reshape(-1)
            nex = len(ex) // 2
            self._keep.append(ex)
        self.N = len(tt)
        self.h = l.or_create(_p(o), _p(d), _p(LL), float(cutoff), self.N, _p(xx), _p(tt, C.c_int), C.byref(T), nex,
                             _p(ex, C.c_int))
        if not self.h:
            raise RuntimeError("or_create failed (particle outside the box?)")
        self.h = C.c_void_p(self.h)
        self.ncells = l.or_ncells(self.h)

    def nonbond(self, mode=OR_WINDOW_REF, nthreads=1):
        f = np.zeros(3 * self.N)
        en = np.zeros(3)
        cnt = np.zeros(2, np.int64)
        lib().or_nonbond(self.h, mode, nthreads, _p(f), _p(en), _p(cnt, C.c_longlong))
        return {"f": f.reshape(-1, 3), "epot": en[0], "epot_exclusion": en[1], "virial": en[2],
                "pairs": int(cnt[0]), "exclusions": int(cnt[1])}

    def bruteforce(self):
        f = np.zeros(3 * self.N)
        en = np.zeros(3)
        cnt = np.zeros(1, np.int64)
        if lib().or_bruteforce(self.h, _p(f), _p(en), _p(cnt, C.c_longlong)) < 0:
            raise RuntimeError("brute force needs at least 2*span+1 cells per axis")
        return {"f": f.reshape(-1, 3), "epot": en[0], "virial": en[2], "pairs": int(cnt[0])}

    def set_verlet(self, xrel=None, cell_maxdx=None, keep_sort=True):
        """Pseudo-Verlet state: cell-relative positions by id (cells unchanged),
        per-cell maxdx, and whether nonbond() keeps the last sort lists."""
        xr = None if xrel is None else np.ascontiguousarray(xrel, dtype=np.float64).reshape(-1)
        md = None if cell_maxdx is None else np.ascontiguousarray(cell_maxdx, dtype=np.float64)
        lib().or_set_verlet(self.h, _p(xr), _p(md), 1 if keep_sort else 0)

    def sort(self):
        lib().or_sort(self.h)

    def cell(self, cid):
        cnt = C.c_int()
        flags = C.c_uint()
        lib().or_cell(self.h, cid, C.byref(cnt), None, C.byref(flags))
        ids = np.zeros(max(cnt.value, 1), np.int32)
        lib().or_cell(self.h, cid, None, _p(ids, C.c_int), None)
        return ids[: cnt.value], flags.value

    def sortlist(self, cid, sid, count):
        out = np.zeros(max(count, 1), np.uint32)
        lib().or_sortlist(self.h, cid, sid, _p(out, C.c_uint))
        return out[:count]

    def tasks(self):
        npair = C.c_int()
        nself = C.c_int()
        nt = lib().or_ntasks(self.h, C.byref(npair), C.byref(nself))
        out = np.zeros((nt, 4), np.int32)
        a, b, c_, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        for k in range(nt):
            lib().or_task(self.h, k, C.byref(a), C.byref(b), C.byref(c_), C.byref(d))
            out[k] = (a.value, b.value, c_.value, d.value)
        return out

    def dscale(self):
        return lib().or_dscale(self.h)

    def set_motion(self, v, imass_by_type):
        """Velocities by id (None: zero) and inverse masses by type for advance()."""
        vv = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
        im = np.ascontiguousarray(imass_by_type, np.float64)
        if lib().or_set_motion(self.h, _p(vv), _p(im)) < 0:
            raise RuntimeError("or_set_motion failed")

    def advance(self, dt, nthreads=1):
        """Kick-drift with the last nonbond()'s forces; returns how many
        particles left their cell (they are not migrated)."""
        out = lib().or_advance(self.h, float(dt), nthreads)
        if out < 0:
            raise RuntimeError("or_advance needs set_motion() and a nonbond() first")
        return out

    def task_counts(self):
        """Pairs evaluated by each task (tasks() order) in the last nonbond()."""
        nt = lib().or_ntasks(self.h, None, None)
        out = np.zeros(max(nt, 1), np.int64)
        lib().or_task_counts(self.h, _p(out, C.c_longlong))
        return out[:nt]

    def cell_x(self, cid, count):
        """Cell-relative FP64 positions of cell cid in storage order."""
        out = np.zeros((max(count, 1), 3))
        lib().or_cell_x(self.h, cid, _p(out))
        return out[:count]

    def sortlists(self):
        """All cells' 13 sort lists, cells in id order (13*count words each)."""
        n = lib().or_sortlists(self.h, None)
        out = np.zeros(max(n, 1), np.uint32)
        lib().or_sortlists(self.h, _p(out, C.c_uint))
        return out[:n]

    def close(self):
        if self.h:
            lib().or_free(self.h)
            self.h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sort_desc(words):
    a = np.ascontiguousarray(words, np.uint32).copy()
    lib().or_sort_desc(_p(a, C.c_uint), len(a))
    return a


def key(x, axis, dscale, bias):
    xx = np.ascontiguousarray(x, np.float64)
    ax = np.ascontiguousarray(axis, np.float64)
    return lib().or_key(_p(xx), _p(ax), dscale, bias)


def evaluate(table, r2):
    e = C.c_double()
    f = C.c_double()
    alpha = np.ascontiguousarray(table["alpha"], np.float64)
    c = np.ascontiguousarray(table["c"], np.float64)
    lib().or_eval(table["n"], _p(alpha), _p(c), float(r2), C.byref(e), C.byref(f))
    return e.value, f.value


def from_system(sys):
    """Oracle over a systems.System (tables taken as data from its host tables)."""
    import systems
    return Oracle(sys.origin, sys.dim, sys.L, sys.cutoff, sys.x, sys.types, systems.tables(sys), sys.excl)
