"""The benchmark systems of the nonbonded path, built with the C library.

  argon(n)      Lennard-Jones argon on an fcc lattice (n^3 unit cells, a=0.565 nm,
                eps=0.996 kJ/mol, sigma=0.3405 nm, m=39.948, rc=1.0 nm, table
                LJ126(0.275, 1.0, 4 eps sigma^12, 4 eps sigma^6, 1e-3));
                n=20 gives the 32,000-atom 11.3 nm box.
  water(nmol)   SPC/E-like 3-site water (O-H 0.1 nm, H-O-H 109.47 deg,
                q_O=-0.8476, q_H=+0.4238, O-O LJ A=2.6378e-6, B=2.6192e-3) on a
                lattice with random orientations, reaction-field Coulomb
                (conducting boundary) + LJ tables, rc=1.0 nm, intramolecular
                exclusions; nmol=33,333 in a 10 nm box is the 100k-atom system.
  mixture(N)    20 species, sigma~U[0.25,0.40] nm, eps~U[0.1,1.0] kJ/mol,
                q~U[-0.8,0.8] shifted to neutral, Lorentz-Berthelot mixing,
                LJ + reaction field for all 210 species pairs, density 100/nm^3,
                random sequential addition with 0.15 nm minimum separation,
                rc=1.2 nm, 0.6 nm cells; N=1,048,576 is the 1M-atom system.
Units: nm, ps, amu, kJ/mol, e; Coulomb constant 138.935458 kJ mol^-1 nm e^-2.
"""
import math
from dataclasses import dataclass, field

import numpy as np

import mdcore as md

KE = 138.935458
KB = 0.0083144626


@dataclass
class System:
    name: str
    origin: np.ndarray
    dim: np.ndarray
    L: np.ndarray
    cutoff: float
    x: np.ndarray
    types: np.ndarray          # engine type ids (1..nspecies; 0 is the null type)
    masses: np.ndarray         # per engine type (index 0 unused)
    pots: list                 # (ti, tj, Potential)
    excl: np.ndarray = None    # (M, 2) particle id pairs
    v: np.ndarray = None
    meta: dict = field(default_factory=dict)
    rigid: list = None         # (i, j, d) distance constraints

    @property
    def N(self):
        return len(self.types)

    @property
    def max_type(self):
        return len(self.masses)


def argon(n=20, a=0.565, seed=6178, T=100.0, tol=1e-3, jitter=0.0):
    """jitter > 0 displaces every atom by U[-jitter/2, jitter/2] per axis."""
    eps, sigma, mass = 0.996, 0.3405, 39.948
    A = 4 * eps * sigma ** 12
    B = 4 * eps * sigma ** 6
    x = md.gen_fcc(n, n, n, a, a / 4)
    box = n * a
    if jitter > 0:
        x = np.mod(x + md.gen_uniform(x.size, -0.5 * jitter, 0.5 * jitter, seed + 1).reshape(x.shape), box)
    N = len(x)
    pot = md.LJ126(0.275, 1.0, A, B, tol)
    v = md.gen_maxwell(np.full(N, mass), KB * T, seed)
    return System("argon", np.zeros(3), np.full(3, box), np.full(3, 1.0), 1.0, x, np.ones(N, np.int32),
                  np.array([0.0, mass]), [(1, 1, pot)], None, v,
                  {"eps": eps, "sigma": sigma, "A": A, "B": B, "box": box})


def water(nmol=33333, box=10.0, lattice=(33, 33, 31), seed=1234, jitter=0.02, a=0.05, eps_rf=0.0, tol=1e-3,
          cutoff=1.0, cell=1.0, rigid=False):
    """rigid: O-H1, O-H2 (0.1 nm) and H1-H2 distance constraints per
    molecule, in that order (SPC/E geometry)."""
    qO, qH = -0.8476, 0.4238
    A, B = 2.6378e-6, 2.6192e-3
    x = md.gen_water(nmol, [box] * 3, lattice, 0.1, math.radians(109.47), jitter, seed)
    N = 3 * nmol
    types = np.tile(np.array([1, 2, 2], np.int32), nmol)
    pOO = md.LJ126_RF(a, cutoff, A, B, qO * qO, KE, eps_rf, tol)
    pOH = md.LJ126_RF(a, cutoff, 0.0, 0.0, qO * qH, KE, eps_rf, tol)
    pHH = md.LJ126_RF(a, cutoff, 0.0, 0.0, qH * qH, KE, eps_rf, tol)
    m = np.arange(nmol, dtype=np.int64) * 3
    excl = np.concatenate([np.stack([m, m + 1], 1), np.stack([m, m + 2], 1), np.stack([m + 1, m + 2], 1)])
    cons = None
    if rigid:
        dhh = 2 * 0.1 * math.sin(math.radians(109.47) / 2)
        cons = [c for k in range(nmol) for c in ((3 * k, 3 * k + 1, 0.1), (3 * k, 3 * k + 2, 0.1),
                                                (3 * k + 1, 3 * k + 2, dhh))]
    return System("water", np.zeros(3), np.full(3, box), np.full(3, cell), cutoff, x, types,
                  np.array([0.0, 15.9994, 1.008]), [(1, 1, pOO), (1, 2, pOH), (2, 2, pHH)], excl, None,
                  {"qO": qO, "qH": qH, "box": box, "nmol": nmol}, cons)


def mixture_species(nspecies=20, seed=20):
    sig = md.gen_uniform(nspecies, 0.25, 0.40, seed)
    eps = md.gen_uniform(nspecies, 0.1, 1.0, seed + 1)
    q = md.gen_uniform(nspecies, -0.8, 0.8, seed + 2)
    mass = md.gen_uniform(nspecies, 10.0, 40.0, seed + 3)
    return sig, eps, q, mass


def mixture(N=1048576, density=100.0, seed=7, rmin=0.15, cutoff=1.2, cell=0.6, nspecies=20, a=0.12, tol=1e-3,
            eps_rf=0.0, switch=None):
    """switch: None for plain LJ + reaction field tables; a distance s < cutoff
    switches the LJ part off between s and the cutoff (force-continuous)."""
    box = (N / density) ** (1.0 / 3.0)
    sig, eps, q, mass = mixture_species(nspecies, seed)
    types = md.gen_types(N, nspecies, 1, seed + 10)
    counts = np.bincount(types - 1, minlength=nspecies)
    q = q - np.dot(counts, q) / N  # count-weighted shift: sum_i q_i = 0
    x = md.gen_rsa(N, [box] * 3, rmin, seed + 20)
    pots = mixture_pots(sig, eps, q, cutoff, a, eps_rf, tol, switch)
    return System("mixture", np.zeros(3), np.full(3, box), np.full(3, cell), cutoff, x, types,
                  np.concatenate([[0.0], mass]), pots, None, None,
                  {"box": box, "sigma": sig, "eps": eps, "q": q, "density": density, "a": a, "eps_rf": eps_rf,
                   "switch": switch})


def mixture_pots(sig, eps, q, cutoff, a, eps_rf, tol, switch=None):
    """LJ 12-6 (Lorentz-Berthelot) + reaction-field tables of every species
    pair; with switch = s the LJ part is switched off from s to the cutoff."""
    pots = []
    n = len(sig)
    for s in range(n):
        for t in range(s, n):
            sg = 0.5 * (sig[s] + sig[t])
            ep = math.sqrt(eps[s] * eps[t])
            A = 4 * ep * sg ** 12
            B = 4 * ep * sg ** 6
            if switch is None:
                pots.append((s + 1, t + 1, md.LJ126_RF(a, cutoff, A, B, q[s] * q[t], KE, eps_rf, tol)))
            else:
                pots.append((s + 1, t + 1, md.LJ126_switch_RF(a, cutoff, A, B, switch, q[s] * q[t], KE, eps_rf,
                                                              tol)))
    return pots


def with_tolerance(sys, tol):
    """The same mixture with its tables rebuilt at another tolerance."""
    m = sys.meta
    pots = mixture_pots(m["sigma"], m["eps"], m["q"], sys.cutoff, m["a"], m["eps_rf"], tol, m.get("switch"))
    return System(sys.name, sys.origin, sys.dim, sys.L, sys.cutoff, sys.x, sys.types, sys.masses, pots, sys.excl,
                  sys.v, dict(m, tol=tol))


def build_engine(sys, flags=md.engine_flag_gpu, start=True):
    """engine_init + addtype + addpot + load (+ exclusions) + engine_start."""
    e = md.Engine(sys.origin, sys.dim, sys.L, sys.cutoff, sys.max_type - 1, flags)
    for t in range(1, sys.max_type):
        e.addtype(float(sys.masses[t]), 0.0, f"T{t}")
    for ti, tj, p in sys.pots:
        e.addpot(p, ti, tj)
    e.load(sys.x, sys.types, sys.v)
    if sys.excl is not None and len(sys.excl):
        e.exclusions_add(sys.excl)
        e.exclusion_shrink()
    if sys.rigid:
        e.rigids_add(sys.rigid)
    if start:
        e.start()
    return e


def tables(sys):
    """Host FP64 tables of a system as flat arrays (pind by engine type pair)."""
    mt = sys.max_type
    pind = -np.ones(mt * mt, dtype=np.int32)
    ns, alphas, cs, offs = [], [], [], []
    off = 0
    index = {}  # one table per distinct potential object
    for ti, tj, p in sys.pots:
        k = index.get(id(p))
        if k is None:
            d = p.data()
            k = index[id(p)] = len(ns)
            ns.append(d["n"])
            alphas.append(d["alpha"])
            cs.append(d["c"])
            offs.append(off)
            off += len(d["c"])
        pind[ti * mt + tj] = k
        pind[tj * mt + ti] = k
    return {"max_type": mt, "pind": pind, "n": np.array(ns, np.int32), "alpha": np.concatenate(alphas),
            "c": np.concatenate(cs), "off": np.array(offs, np.int32)}


# Stages (dt in ps, steps, steps between rescales) that take the 1M mixture's
# random-addition start, whose closest pairs sit deep in each other's
# repulsion, to a state that integrates at 1 fs.
MIXTURE_EQUILIBRATION = ((2e-5, 100, 10), (1e-4, 200, 20), (5e-4, 300, 50), (1e-3, 400, 100))


def equilibrate(eng, sysm, T=300.0, stages=MIXTURE_EQUILIBRATION, seed=17, log=None):
    """Device-resident steps from Maxwell velocities at T, the velocities
    rescaled to T after every `every` steps (the reference's drivers rescale
    p->v toward the target temperature, examples/argon/argon.c:286-299),
    with the time step raised stage by stage.  Returns the temperature and
    potential energy at the end of each stage; leaves the engine's dt at the
    last stage's."""
    dof = 3 * sysm.N - 3
    eng.set_velocities(md.gen_maxwell(sysm.masses[sysm.types], KB * T, seed))
    out = []
    for dt, nsteps, every in stages:
        eng.setdt(dt)
        for _ in range(max(1, nsteps // every)):
            ep, ek = eng.gpu_steps(every)
            Tcur = 2.0 * ek[-1] / (dof * KB)
            if not (np.isfinite(Tcur) and np.isfinite(ep[-1])):
                raise RuntimeError(f"equilibration diverged at dt {dt}: T {Tcur}, epot {ep[-1]}")
            if Tcur > 0.0:
                eng.set_velocities(eng.velocities() * math.sqrt(T / Tcur))
            if log:
                log(dt, Tcur, float(ep[-1]))
        out.append((dt, Tcur, float(ep[-1])))
    return out
