"""ctypes mirror of the mdcore C API for the MI355X nonbonded engine.

The C library (lib/libmdcore_gpu.so, built by the Makefile next to this file)
keeps the reference's host API -- engine_init / engine_addtype /
engine_addpot / engine_load / engine_start / engine_step / engine_unload and
the potential_create_* tables (reference src/engine.h:319-368,
src/potential.h:82-95) -- and the C-ABI of the GPU path
(include/engine_gpu.h, replacing engine_cuda_* of src/engine.h:381-398).
This module only declares those symbols for Python callers; every
computation happens in the library.  There is no Python or CPU fallback:
loading fails loudly if the library is missing.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# MDCORE_LIB selects another build of the same library (e.g. a tuning variant).
LIBPATH = Path(os.environ.get("MDCORE_LIB", HERE / "lib" / "libmdcore_gpu.so"))

# engine flags (include/mdcore.h, reference src/engine.h:56-72)
engine_flag_none = 0
engine_flag_gpu = 4
engine_flag_cuda = 4
engine_flag_nullpart = 32768
engine_flag_projwindow = 65536
engine_flag_verlet = 16
engine_flag_shake = 1024
engine_flag_verlet_pseudo = 256
engine_flag_mlp = 131072
engine_flag_fullshell = 262144
engine_flag_halfshell = 524288
engine_gpu_half_mean = 48
engine_flag_nosplit = 1048576
engine_flag_unpaired = 2097152
engine_flag_unsorted = 512

space_periodic_full = 7

engine_gpu_sync_parts_up = 1
engine_gpu_sync_forces_down = 2
engine_gpu_sync_sort_down = 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)
_llp = C.POINTER(C.c_longlong)
_fp = C.POINTER(C.c_float)

_lib = None


class MdcoreError(RuntimeError):
    pass


def library():
    """Load lib/libmdcore_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise MdcoreError(f"{LIBPATH} is missing: build it with `make -C {HERE}`")
    lib = C.CDLL(str(LIBPATH), mode=C.RTLD_GLOBAL)
    sig = {
        "errs_gettext": (C.c_int, [C.c_char_p, C.c_int]),
        "errs_clear": (C.c_int, []),
        "engine_alloc": (C.c_void_p, []),
        "engine_free": (None, [C.c_void_p]),
        "engine_sizeof": (C.c_ulong, []),
        "engine_setdt": (C.c_int, [C.c_void_p, C.c_double]),
        "engine_setflags": (C.c_int, [C.c_void_p, C.c_uint, C.c_uint]),
        "engine_init": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_uint, C.c_int, C.c_uint]),
        "engine_addtype": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_char_p, C.c_char_p]),
        "engine_addpot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
        "engine_load": (C.c_int, [C.c_void_p, _dp, _dp, _ip, _ip, _ip, _dp, _up, C.c_int]),
        "engine_unload": (C.c_int, [C.c_void_p, _dp, _dp, _ip, _ip, _ip, _dp, _up, _dp, C.c_int]),
        "engine_start": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "engine_step": (C.c_int, [C.c_void_p]),
        "engine_shuffle": (C.c_int, [C.c_void_p]),
        "engine_nonbond_eval": (C.c_int, [C.c_void_p]),
        "engine_exclusion_add": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "engine_exclusion_shrink": (C.c_int, [C.c_void_p]),
        "engine_finalize": (C.c_int, [C.c_void_p]),
        "engine_getforces": (C.c_int, [C.c_void_p, _dp, C.c_int]),
        "engine_getpositions": (C.c_int, [C.c_void_p, _dp, C.c_int]),
        "engine_getvelocities": (C.c_int, [C.c_void_p, _dp, C.c_int]),
        "engine_setvelocities": (C.c_int, [C.c_void_p, _dp, C.c_int]),
        "engine_getenergies": (C.c_int, [C.c_void_p, _dp, _llp]),
        "engine_getsortlist": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _up]),
        "engine_getsortlists": (C.c_longlong, [C.c_void_p, _up]),
        "engine_getcells": (C.c_int, [C.c_void_p, _ip, _ip]),
        "engine_getlocal": (C.c_int, [C.c_void_p, _dp, _ip, C.c_int]),
        "engine_getmaxdx": (C.c_int, [C.c_void_p, _dp, _dp]),
        "engine_verlet_update": (C.c_int, [C.c_void_p]),
        "engine_rigid_add": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
        "engine_rigid_eval": (C.c_int, [C.c_void_p]),
        "engine_getrigid": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _dp]),
        "engine_read_psf": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "engine_read_cpf": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double]),
        "engine_read_xplor": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]),
        "engine_getpot": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int]),
        "engine_gettypeinfo": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_char_p, C.c_char_p]),
        "engine_gettopology": (C.c_int, [C.c_void_p, C.c_int, _ip]),
        "engine_gettype": (C.c_int, [C.c_void_p, C.c_char_p]),
        "engine_gettype2": (C.c_int, [C.c_void_p, C.c_char_p]),
        "engine_advance": (C.c_int, [C.c_void_p]),
        "engine_gettasks": (C.c_int, [C.c_void_p, _ip, C.c_int]),
        "engine_prepare": (C.c_int, [C.c_void_p]),
        "engine_gpu_setdevice": (C.c_int, [C.c_void_p, C.c_int]),
        "engine_gpu_load": (C.c_int, [C.c_void_p]),
        "engine_nonbond_gpu": (C.c_int, [C.c_void_p]),
        "engine_gpu_sync": (C.c_int, [C.c_void_p, C.c_uint]),
        "engine_gpu_finalize": (C.c_int, [C.c_void_p]),
        "engine_gpu_time": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
        "engine_gpu_results": (C.c_int, [C.c_void_p, _dp, _llp]),
        "engine_gpu_steps": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
        "engine_mlp_pairs": (C.c_int, [C.c_void_p, _ip, _ip, C.c_int]),
        "engine_mlp_init": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_ulonglong]),
        "engine_mlp_params": (C.c_int, [C.c_void_p, _fp, C.c_int]),
        "engine_mlp_train": (C.c_int, [C.c_void_p, _fp, _ip, _fp, _fp, C.c_longlong, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_ulonglong, _fp]),
        "engine_mlp_grad": (C.c_int, [C.c_void_p, _fp, _ip, _fp, _fp, C.c_longlong, C.c_double, C.c_double, _fp,
                                      _dp]),
        "engine_mlp_grad_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_longlong, C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
        "engine_mlp_eval": (C.c_int, [C.c_void_p, _fp, _ip, C.c_longlong, _fp, _fp]),
        "potential_create_LJ126": (C.c_void_p, [C.c_double] * 5),
        "potential_create_LJ126_switch": (C.c_void_p, [C.c_double] * 6),
        "potential_create_LJ126_Coulomb": (C.c_void_p, [C.c_double] * 6),
        "potential_create_LJ126_Ewald": (C.c_void_p, [C.c_double] * 7),
        "potential_create_Ewald": (C.c_void_p, [C.c_double] * 5),
        "potential_create_Coulomb": (C.c_void_p, [C.c_double] * 4),
        "potential_create_harmonic": (C.c_void_p, [C.c_double] * 5),
        "potential_create_LJ126_RF": (C.c_void_p, [C.c_double] * 8),
        "potential_create_LJ126_switch_RF": (C.c_void_p, [C.c_double] * 9),
        "potential_getdata": (C.c_int, [C.c_void_p, _ip, _dp, _dp, _dp, C.c_int]),
        "potential_eval": (None, [C.c_void_p, C.c_double, _dp, _dp]),
        "potential_free": (None, [C.c_void_p]),
        "potential_eval_r": (C.c_int, [C.c_void_p, _dp, C.c_longlong, _dp, _dp]),
        "md_sample_pairs": (C.c_longlong, [_dp, _ip, C.c_int, _dp, C.c_double, C.c_longlong, C.c_uint64, _dp, _ip,
                                           _ip]),
        "md_gen_fcc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp]),
        "md_gen_maxwell": (C.c_int, [C.c_int, _dp, C.c_double, C.c_uint64, _dp]),
        "md_gen_water": (C.c_int, [C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_uint64, _dp]),
        "md_gen_rsa": (C.c_int, [C.c_int, _dp, C.c_double, C.c_uint64, C.c_longlong, _dp]),
        "md_gen_types": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, _ip]),
        "md_gen_uniform": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_uint64, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def errors():
    buf = C.create_string_buffer(8192)
    library().errs_gettext(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc, what):
    if rc is None or (isinstance(rc, int) and rc < 0):
        msg = errors()
        library().errs_clear()
        raise MdcoreError(f"{what} failed ({rc}):\n{msg}")
    return rc


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


# ---------------------------------------------------------------- potentials

class Potential:
    """A host table built by one of the potential_create_* constructors."""

    def __init__(self, handle, desc):
        if not handle:
            msg = errors()
            library().errs_clear()
            raise MdcoreError(f"potential {desc} failed:\n{msg}")
        self.handle = C.c_void_p(handle)
        self.desc = desc

    def data(self):
        lib = library()
        n = C.c_int(0)
        ab = np.zeros(2)
        alpha = np.zeros(4)
        check(lib.potential_getdata(self.handle, C.byref(n), ptr(ab), ptr(alpha), None, 0), "potential_getdata")
        c = np.zeros((n.value + 1) * 8)
        check(lib.potential_getdata(self.handle, None, None, None, ptr(c), len(c)), "potential_getdata")
        return {"n": n.value, "a": ab[0], "b": ab[1], "alpha": alpha, "c": c}

    def eval(self, r2):
        e = C.c_double(0.0)
        f = C.c_double(0.0)
        library().potential_eval(self.handle, float(r2), C.byref(e), C.byref(f))
        return e.value, f.value

    def eval_r(self, r):
        """E and dE/dr at the distances r (FP64 table evaluation in the library)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        e = np.zeros_like(r)
        d = np.zeros_like(r)
        check(library().potential_eval_r(self.handle, ptr(r), len(r), ptr(e), ptr(d)), "potential_eval_r")
        return e, d


def LJ126(a, b, A, B, tol):
    return Potential(library().potential_create_LJ126(a, b, A, B, tol), f"LJ126({a},{b},{A},{B},{tol})")


def LJ126_switch(a, b, A, B, s, tol):
    return Potential(library().potential_create_LJ126_switch(a, b, A, B, s, tol), "LJ126_switch")


def LJ126_Coulomb(a, b, A, B, q, tol):
    return Potential(library().potential_create_LJ126_Coulomb(a, b, A, B, q, tol), "LJ126_Coulomb")


def LJ126_Ewald(a, b, A, B, q, kappa, tol):
    return Potential(library().potential_create_LJ126_Ewald(a, b, A, B, q, kappa, tol), "LJ126_Ewald")


def Ewald(a, b, q, kappa, tol):
    return Potential(library().potential_create_Ewald(a, b, q, kappa, tol), "Ewald")


def Coulomb(a, b, q, tol):
    return Potential(library().potential_create_Coulomb(a, b, q, tol), "Coulomb")


def harmonic(a, b, K, r0, tol):
    return Potential(library().potential_create_harmonic(a, b, K, r0, tol), "harmonic")


def LJ126_RF(a, b, A, B, q, ke, eps_rf, tol):
    return Potential(library().potential_create_LJ126_RF(a, b, A, B, q, ke, eps_rf, tol),
                     f"LJ126_RF({a},{b},{A},{B},{q},{ke},{eps_rf},{tol})")


def LJ126_switch_RF(a, b, A, B, s, q, ke, eps_rf, tol):
    return Potential(library().potential_create_LJ126_switch_RF(a, b, A, B, s, q, ke, eps_rf, tol),
                     f"LJ126_switch_RF({a},{b},{A},{B},{s},{q},{ke},{eps_rf},{tol})")


# ---------------------------------------------------------------- engine

class Engine:
    """One struct engine (engine_alloc + engine_init)."""

    def __init__(self, origin, dim, L, cutoff, max_type, flags=engine_flag_gpu, period=space_periodic_full):
        lib = library()
        self.lib = lib
        self.e = C.c_void_p(lib.engine_alloc())
        if not self.e:
            raise MdcoreError("engine_alloc failed")
        o = np.asarray(origin, dtype=np.float64)
        d = np.asarray(dim, dtype=np.float64)
        l = np.asarray(L, dtype=np.float64)
        check(lib.engine_init(self.e, ptr(o), ptr(d), ptr(l), float(cutoff), period, int(max_type), flags),
              "engine_init")
        self.flags = flags
        self.N = 0
        self._pots = []

    def addtype(self, mass, charge=0.0, name="X"):
        return check(self.lib.engine_addtype(self.e, mass, charge, name.encode(), name.encode()), "engine_addtype")

    def addpot(self, pot, i, j):
        self._pots.append(pot)
        check(self.lib.engine_addpot(self.e, pot.handle, i, j), "engine_addpot")

    def load(self, x, types, v=None, ids=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(types, dtype=np.int32)
        N = len(t)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        ii = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        check(self.lib.engine_load(self.e, ptr(x), ptr(vv), ptr(t, C.c_int), ptr(ii, C.c_int), None, None, None, N),
              "engine_load")
        self.N += N

    def exclusion_add(self, i, j):
        check(self.lib.engine_exclusion_add(self.e, int(i), int(j)), "engine_exclusion_add")

    def exclusions_add(self, pairs):
        for i, j in np.asarray(pairs, dtype=np.int64).reshape(-1, 2):
            check(self.lib.engine_exclusion_add(self.e, int(i), int(j)), "engine_exclusion_add")

    def exclusion_shrink(self):
        check(self.lib.engine_exclusion_shrink(self.e), "engine_exclusion_shrink")

    def start(self, nr_runners=1):
        check(self.lib.engine_start(self.e, nr_runners, nr_runners), "engine_start")

    def step(self):
        check(self.lib.engine_step(self.e), "engine_step")

    def setflags(self, set=0, clear=0):
        check(self.lib.engine_setflags(self.e, set, clear), "engine_setflags")

    def setdt(self, dt):
        check(self.lib.engine_setdt(self.e, dt), "engine_setdt")

    def shuffle(self):
        check(self.lib.engine_shuffle(self.e), "engine_shuffle")

    def prepare(self):
        check(self.lib.engine_prepare(self.e), "engine_prepare")

    def nonbond(self):
        """One nonbonded pass (space_prepare + engine_nonbond_gpu), no integration."""
        self.prepare()
        check(self.lib.engine_nonbond_eval(self.e), "engine_nonbond_eval")

    def forces(self):
        f = np.zeros(3 * self.N)
        check(self.lib.engine_getforces(self.e, ptr(f), self.N), "engine_getforces")
        return f.reshape(-1, 3)

    def positions(self):
        x = np.zeros(3 * self.N)
        check(self.lib.engine_getpositions(self.e, ptr(x), self.N), "engine_getpositions")
        return x.reshape(-1, 3)

    def velocities(self):
        v = np.zeros(3 * self.N)
        check(self.lib.engine_getvelocities(self.e, ptr(v), self.N), "engine_getvelocities")
        return v.reshape(-1, 3)

    def set_velocities(self, v):
        vv = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        if vv.size != 3 * self.N:
            raise ValueError("set_velocities: need 3 values per particle")
        check(self.lib.engine_setvelocities(self.e, ptr(vv), self.N), "engine_setvelocities")

    def energies(self):
        en = np.zeros(4)
        cnt = np.zeros(2, dtype=np.int64)
        check(self.lib.engine_getenergies(self.e, ptr(en), ptr(cnt, C.c_longlong)), "engine_getenergies")
        return {"epot": en[0], "epot_nonbond": en[1], "epot_exclusion": en[2], "virial": en[3],
                "pairs": int(cnt[0]), "exclusions": int(cnt[1])}

    def gpu_results(self):
        out = np.zeros(3)
        cnt = np.zeros(3, dtype=np.int64)
        check(self.lib.engine_gpu_results(self.e, ptr(out), ptr(cnt, C.c_longlong)), "engine_gpu_results")
        return {"epot": out[0], "epot_exclusion": out[1], "virial": out[2], "pairs": int(cnt[0]),
                "exclusions": int(cnt[1]), "pairs_refwindow": int(cnt[2])}

    def sync(self, what):
        check(self.lib.engine_gpu_sync(self.e, what), "engine_gpu_sync")

    def gpu_time(self, steps, warmup):
        ms = np.zeros(4)
        check(self.lib.engine_gpu_time(self.e, steps, warmup, ptr(ms)), "engine_gpu_time")
        return {"sort_ms": ms[0], "nonbond_ms": ms[1], "reduce_ms": ms[2], "total_ms": ms[3]}

    def gpu_steps(self, nsteps):
        """nsteps engine_step calls with the particles resident on the device
        (engine_gpu_steps); returns the per-step pass energy and kinetic energy."""
        epot = np.zeros(max(nsteps, 1))
        ekin = np.zeros(max(nsteps, 1))
        check(self.lib.engine_gpu_steps(self.e, nsteps, ptr(epot), ptr(ekin)), "engine_gpu_steps")
        return epot[:nsteps], ekin[:nsteps]

    def local(self):
        """By particle id: cell-relative positions as stored, and cell ids."""
        x = np.zeros(3 * max(self.N, 1))
        cell = np.zeros(max(self.N, 1), dtype=np.int32)
        check(self.lib.engine_getlocal(self.e, ptr(x), ptr(cell, C.c_int), self.N), "engine_getlocal")
        return x[:3 * self.N].reshape(-1, 3), cell[:self.N]

    def maxdx(self, ncells):
        """(s->maxdx, per-cell maxdx, s->verlet_rebuild)."""
        m = np.zeros(1)
        cm = np.zeros(max(ncells, 1))
        rb = check(self.lib.engine_getmaxdx(self.e, ptr(m), ptr(cm)), "engine_getmaxdx")
        return float(m[0]), cm[:ncells], rb

    # ------------------------------------------------------------ readers

    def read_psf(self, psf, pdb):
        """engine_read_psf on the two files; particles join the engine."""
        fp, fd = os.open(str(psf), os.O_RDONLY), os.open(str(pdb), os.O_RDONLY)
        try:
            check(self.lib.engine_read_psf(self.e, fp, fd), "engine_read_psf")
        finally:
            os.close(fp)
            os.close(fd)
        self.N = check(self.lib.engine_getcells(self.e, None, None), "engine_getcells")

    def read_cpf(self, path, kappa, tol, rigidH=0, switchlen=0.0):
        f = os.open(str(path), os.O_RDONLY)
        try:
            check(self.lib.engine_read_cpf(self.e, f, kappa, tol, int(rigidH), switchlen), "engine_read_cpf")
        finally:
            os.close(f)

    def read_xplor(self, path, kappa, tol, rigidH=0):
        f = os.open(str(path), os.O_RDONLY)
        try:
            check(self.lib.engine_read_xplor(self.e, f, kappa, tol, int(rigidH)), "engine_read_xplor")
        finally:
            os.close(f)

    def pot(self, i, j):
        """The table of type pair (i, j), or None (borrowed: owned by the engine)."""
        h = self.lib.engine_getpot(self.e, i, j)
        return Potential(h, f"pair {i} {j}") if h else None

    def typeinfo(self, k):
        out = np.zeros(4)
        n1 = C.create_string_buffer(64)
        n2 = C.create_string_buffer(64)
        nt = check(self.lib.engine_gettypeinfo(self.e, k, ptr(out), n1, n2), "engine_gettypeinfo")
        return {"mass": out[0], "charge": out[1], "eps": out[2], "rmin": out[3], "name": n1.value.decode(),
                "name2": n2.value.decode(), "ntypes": nt}

    def topology(self, which):
        """0: bonds (i, j); 1: angles (i, j, k); 2: dihedrals (i, j, k, l, pid)."""
        width = (2, 3, 5)[which]
        n = check(self.lib.engine_gettopology(self.e, which, None), "engine_gettopology")
        out = np.zeros(max(n, 1) * width, dtype=np.int32)
        check(self.lib.engine_gettopology(self.e, which, ptr(out, C.c_int)), "engine_gettopology")
        return out[:n * width].reshape(-1, width)

    def rigid_add(self, i, j, d):
        check(self.lib.engine_rigid_add(self.e, int(i), int(j), float(d)), "engine_rigid_add")

    def rigids_add(self, triples):
        """(i, j, d) distance constraints, in order."""
        for i, j, d in triples:
            self.rigid_add(i, j, d)

    def rigid_eval(self):
        check(self.lib.engine_rigid_eval(self.e), "engine_rigid_eval")

    def nrigids(self):
        return check(self.lib.engine_getrigid(self.e, -1, None, None, None), "engine_getrigid")

    def rigid(self, rid):
        """(particle ids, constraint index pairs, d^2) of group rid."""
        parts = np.zeros(1 + 10, dtype=np.int32)
        ij = np.zeros(60, dtype=np.int32)
        d2 = np.zeros(30)
        nc = check(self.lib.engine_getrigid(self.e, rid, ptr(parts, C.c_int), ptr(ij, C.c_int), ptr(d2)),
                   "engine_getrigid")
        return parts[1:1 + parts[0]].copy(), ij[:2 * nc].reshape(-1, 2).copy(), d2[:nc].copy()

    def verlet_update(self):
        check(self.lib.engine_verlet_update(self.e), "engine_verlet_update")

    def advance(self):
        check(self.lib.engine_advance(self.e), "engine_advance")

    def cells(self, ncells):
        counts = np.zeros(ncells, dtype=np.int32)
        ids = np.zeros(max(self.N, 1), dtype=np.int32)
        n = check(self.lib.engine_getcells(self.e, ptr(counts, C.c_int), ptr(ids, C.c_int)), "engine_getcells")
        return counts, ids[:n]

    def tasks(self):
        n = check(self.lib.engine_gettasks(self.e, None, 0), "engine_gettasks")
        out = np.zeros((max(n, 1), 4), dtype=np.int32)
        check(self.lib.engine_gettasks(self.e, ptr(out, C.c_int), n), "engine_gettasks")
        return out[:n]

    def sortlist(self, cid, sid, count):
        out = np.zeros(max(count, 1), dtype=np.uint32)
        check(self.lib.engine_getsortlist(self.e, cid, sid, ptr(out, C.c_uint)), "engine_getsortlist")
        return out[:count]

    def sortlists(self):
        """Every cell's host sortlist block (13*count words), cells in id order."""
        n = self.lib.engine_getsortlists(self.e, None)
        if n < 0:
            check(int(n), "engine_getsortlists")
        out = np.zeros(max(n, 1), dtype=np.uint32)
        self.lib.engine_getsortlists(self.e, ptr(out, C.c_uint))
        return out[:n]

    # ------------------------------------------------------------ MLP surrogate

    def mlp_pairs(self):
        """Species pairs (ti, tj) in pair-id order."""
        n = check(self.lib.engine_mlp_pairs(self.e, None, None, 0), "engine_mlp_pairs")
        ti = np.zeros(max(n, 1), np.int32)
        tj = np.zeros(max(n, 1), np.int32)
        check(self.lib.engine_mlp_pairs(self.e, ptr(ti, C.c_int), ptr(tj, C.c_int), n), "engine_mlp_pairs")
        return ti[:n], tj[:n]

    def mlp_init(self, r_lo, r_hi, escale, rswitch, seed=1):
        """rswitch[p] >= r_hi[p]: no switch; otherwise the pair's energy is
        switched off between rswitch[p] and r_hi[p]."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (r_lo, r_hi, escale, rswitch)]
        return check(self.lib.engine_mlp_init(self.e, *[ptr(v) for v in a], seed), "engine_mlp_init")

    def mlp_params(self, theta=None):
        """Get (theta None) or set the fp32 parameter vector."""
        n = check(self.lib.engine_mlp_params(self.e, None, 1), "engine_mlp_params")
        if theta is None:
            out = np.zeros(n, np.float32)
            check(self.lib.engine_mlp_params(self.e, ptr(out, C.c_float), 1), "engine_mlp_params")
            return out
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.size == n
        check(self.lib.engine_mlp_params(self.e, ptr(t, C.c_float), 0), "engine_mlp_params")
        return t

    @staticmethod
    def _samples(r, pair, E=None, F=None):
        r = np.ascontiguousarray(r, dtype=np.float32)
        pair = np.ascontiguousarray(pair, dtype=np.int32)
        E = None if E is None else np.ascontiguousarray(E, dtype=np.float32)
        F = None if F is None else np.ascontiguousarray(F, dtype=np.float32)
        return r, pair, E, F

    def mlp_train(self, r, pair, E, F, steps, batch_tiles=512, lr0=3e-3, lr1=3e-5, wE=1.0, wF=1.0, seed=1):
        r, pair, E, F = self._samples(r, pair, E, F)
        hist = np.zeros(steps, np.float32)
        check(self.lib.engine_mlp_train(self.e, ptr(r, C.c_float), ptr(pair, C.c_int), ptr(E, C.c_float),
                                        ptr(F, C.c_float), len(r), steps, batch_tiles, lr0, lr1, wE, wF, seed,
                                        ptr(hist, C.c_float)), "engine_mlp_train")
        return hist

    def mlp_grad(self, r, pair, E, F, wE=1.0, wF=1.0):
        r, pair, E, F = self._samples(r, pair, E, F)
        n = check(self.lib.engine_mlp_params(self.e, None, 1), "engine_mlp_params")
        g = np.zeros(n, np.float32)
        loss = np.zeros(1)
        check(self.lib.engine_mlp_grad(self.e, ptr(r, C.c_float), ptr(pair, C.c_int), ptr(E, C.c_float),
                                       ptr(F, C.c_float), len(r), wE, wF, ptr(g, C.c_float), ptr(loss)),
              "engine_mlp_grad")
        return g, float(loss[0])

    def mlp_grad_noalloc(self, r, pair, E, F, n, grad, loss, wE=1.0, wF=1.0):
        """engine_mlp_grad_noalloc on device buffers given as addresses
        (r, E, F: float32; pair: int32; grad: float32 of the parameter count;
        loss: one float32).  Asynchronous on the engine's stream."""
        check(self.lib.engine_mlp_grad_noalloc(self.e, C.c_void_p(r), C.c_void_p(pair), C.c_void_p(E), C.c_void_p(F),
                                               int(n), wE, wF, C.c_void_p(grad), C.c_void_p(loss)),
              "engine_mlp_grad_noalloc")

    def mlp_eval(self, r, pair):
        r, pair, _, _ = self._samples(r, pair)
        E = np.zeros(len(r), np.float32)
        F = np.zeros(len(r), np.float32)
        check(self.lib.engine_mlp_eval(self.e, ptr(r, C.c_float), ptr(pair, C.c_int), len(r), ptr(E, C.c_float),
                                       ptr(F, C.c_float)), "engine_mlp_eval")
        return E, F

    def close(self):
        if self.e:
            self.lib.engine_free(self.e)
            self.e = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- workloads

def gen_fcc(nx, ny, nz, a, offset):
    x = np.zeros(4 * nx * ny * nz * 3)
    check(library().md_gen_fcc(nx, ny, nz, a, offset, ptr(x)), "md_gen_fcc")
    return x.reshape(-1, 3)


def gen_maxwell(masses, kT, seed):
    m = np.ascontiguousarray(masses, dtype=np.float64)
    v = np.zeros(3 * len(m))
    check(library().md_gen_maxwell(len(m), ptr(m), kT, seed, ptr(v)), "md_gen_maxwell")
    return v.reshape(-1, 3)


def gen_water(nmol, box, nl, roh, theta, jitter, seed):
    b = np.asarray(box, dtype=np.float64)
    x = np.zeros(9 * nmol)
    check(library().md_gen_water(nmol, ptr(b), nl[0], nl[1], nl[2], roh, theta, jitter, seed, ptr(x)),
          "md_gen_water")
    return x.reshape(-1, 3)


def gen_rsa(N, box, rmin, seed, max_trials=None):
    b = np.asarray(box, dtype=np.float64)
    x = np.zeros(3 * N)
    n = check(library().md_gen_rsa(N, ptr(b), rmin, seed, max_trials or 200 * N, ptr(x)), "md_gen_rsa")
    if n != N:
        raise MdcoreError(f"RSA placed only {n} of {N} particles")
    return x.reshape(-1, 3)


def gen_types(N, ntypes, first, seed):
    t = np.zeros(N, dtype=np.int32)
    check(library().md_gen_types(N, ntypes, first, seed, ptr(t, C.c_int)), "md_gen_types")
    return t


def sample_pairs(x, types, box, cutoff, nmax, seed):
    """In-cutoff pair distances of a periodic configuration: (r, type_i, type_j)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    t = np.ascontiguousarray(types, dtype=np.int32)
    b = np.ascontiguousarray(box, dtype=np.float64)
    r = np.zeros(nmax)
    ti = np.zeros(nmax, np.int32)
    tj = np.zeros(nmax, np.int32)
    n = library().md_sample_pairs(ptr(x), ptr(t, C.c_int), len(t), ptr(b), cutoff, nmax, seed, ptr(r),
                                  ptr(ti, C.c_int), ptr(tj, C.c_int))
    if n < 0:
        raise MdcoreError("md_sample_pairs: bad input (box must hold 3 cutoffs per axis)")
    return r[:n], ti[:n], tj[:n]


def gen_uniform(n, lo, hi, seed):
    out = np.zeros(n)
    check(library().md_gen_uniform(n, lo, hi, seed, ptr(out)), "md_gen_uniform")
    return out
