"""A system set up entirely by the readers, on the GPU: 1,000 TIP3P waters
written as PSF / PDB / CHARMM parameter files, read with engine_read_psf and
engine_read_cpf (real-space Ewald tables, kappa 3 nm^-1, rigid hydrogens),
exclusions from the angles and the rigid constraints as the reference's
jac.c builds them.

- one pass against the FP64 oracle on the engine's own tables: pair counts
  bit-exact, forces to 1e-5 relative RMSE, energies to 1e-6;
- 20 device-resident steps keep every constraint within tol_rigid.
"""
import numpy as np
import pytest

import mdcore as md
import pyoracle as po
from gpuutil import rel_rmse
from test_io import LJ

pytestmark = pytest.mark.gpu

BOX, CUTOFF, NMOL = 3.1, 0.9, 1000


def write_inputs(tmp_path):
    x = 10.0 * md.gen_water(NMOL, [BOX] * 3, (10, 10, 10), 0.1, np.radians(109.47), 0.02, 5)  # Angstrom
    psf = ["PSF ", "", "       1 !NTITLE", "* TIP3P box", "", f"{3 * NMOL:8d} !NATOM"]
    for m in range(NMOL):
        for k, (an, ty, q, ms) in enumerate((("OH2", "OT", -0.834, 15.9994), ("H1", "HT", 0.417, 1.008),
                                             ("H2", "HT", 0.417, 1.008))):
            psf.append(f"{3 * m + k + 1:8d} W    {m + 1:<5d} TIP3 {an:<4s} {ty:<4s} {q:10.6f}    {ms:10.4f}           0")
    bonds = [(3 * m + 1, 3 * m + a) for m in range(NMOL) for a in (2, 3)] + [(3 * m + 2, 3 * m + 3) for m in range(NMOL)]
    angles = [(3 * m + 2, 3 * m + 1, 3 * m + 3) for m in range(NMOL)]
    for title, items, per in (("NBOND: bonds", bonds, 4), ("NTHETA: angles", angles, 3),
                              ("NPHI: dihedrals", [], 2), ("NIMPHI: impropers", [], 2)):
        psf += ["", f"{len(items):8d} !{title}"]
        for s in range(0, len(items), per):
            psf.append("".join(f"{i:8d}" for it in items[s:s + per] for i in it))
    (tmp_path / "w.psf").write_text("\n".join(psf) + "\n\n")
    pdb = ["REMARK TIP3P box"]
    for k in range(3 * NMOL):
        an = ("OH2", "H1", "H2")[k % 3]
        pdb.append(f"ATOM {k + 1:6d}  {an:<4s}TIP3 {k // 3 + 1:4d}    {x[k, 0]:8.3f}{x[k, 1]:8.3f}{x[k, 2]:8.3f}"
                   "  1.00  0.00      W")
    pdb.append("END")
    (tmp_path / "w.pdb").write_text("\n".join(pdb) + "\n")
    cpf = ["* TIP3P", "*", "BONDS", "OT   HT    450.0   0.9572", "HT   HT      0.0   1.5139", "",
           "ANGLES", "HT   OT   HT     55.0   104.52", "", "DIHEDRALS", "", "IMPROPER", "",
           "NONBONDED nbxmod  5 atom cdiel shift vatom vdistance vswitch cutnb 14.0"]
    for ty in ("OT", "HT"):
        eps, rh = LJ[ty]
        cpf.append(f"{ty:<6s} 0.000000  {eps:9.6f}  {rh:9.6f}")
    cpf += ["", "HBOND CUTHB 0.5", "END"]
    (tmp_path / "w.inp").write_text("\n".join(cpf) + "\n")
    return tmp_path / "w.psf", tmp_path / "w.pdb", tmp_path / "w.inp"


def build(tmp_path):
    psf, pdb, inp = write_inputs(tmp_path)
    e = md.Engine([0, 0, 0], [BOX] * 3, [CUTOFF] * 3, CUTOFF, 4, md.engine_flag_gpu)
    e.read_psf(psf, pdb)
    e.read_cpf(inp, 3.0, 1e-4, rigidH=1, switchlen=0.7)
    assert e.nrigids() == NMOL and len(e.topology(0)) == 0
    excl = [tuple(a[[0, 2]]) for a in e.topology(1)]
    for r in range(e.nrigids()):
        parts, ij, _ = e.rigid(r)
        excl += [(parts[i], parts[j]) for i, j in ij]
    e.exclusions_add(np.array(excl))
    e.exclusion_shrink()
    e.start()
    return e, np.array(sorted(set(tuple(sorted(p)) for p in excl)))


def oracle_tables(e, mt):
    """The engine's own pair tables as oracle input."""
    pind = -np.ones(mt * mt, np.int32)
    seen, ns, alphas, cs, offs, off = {}, [], [], [], [], 0
    for i in range(mt):
        for j in range(mt):
            p = e.pot(i, j)
            if p is None:
                continue
            key = p.handle.value
            if key not in seen:
                d = p.data()
                seen[key] = len(ns)
                ns.append(d["n"])
                alphas.append(d["alpha"])
                cs.append(d["c"])
                offs.append(off)
                off += len(d["c"])
            pind[i * mt + j] = seen[key]
    return {"max_type": mt, "pind": pind, "n": np.array(ns, np.int32), "alpha": np.concatenate(alphas),
            "c": np.concatenate(cs), "off": np.array(offs, np.int32)}


def test_charmm_water_pass_matches_oracle(tmp_path):
    e, excl = build(tmp_path)
    e.shuffle()
    e.nonbond()
    mt = 5  # engine types: the null type, OT, HT, room for two more (max_type 4 + 1)
    x = e.positions()
    names = {e.typeinfo(k)["name"]: k for k in range(e.typeinfo(0)["ntypes"])}
    types = np.where(np.arange(e.N) % 3 == 0, names["OT"], names["HT"]).astype(np.int32)
    o = po.Oracle(np.zeros(3), np.full(3, BOX), np.full(3, CUTOFF), CUTOFF, x, types, oracle_tables(e, mt), excl)
    r = o.nonbond(po.OR_WINDOW_REF)
    en = e.energies()
    assert en["pairs"] == r["pairs"] and en["exclusions"] == r["exclusions"]
    assert rel_rmse(e.forces(), r["f"]) <= 1e-5
    assert en["epot"] == pytest.approx(r["epot"] + r["epot_exclusion"], rel=1e-6)
    print(f"pairs {r['pairs']}, exclusions {r['exclusions']}, epot {en['epot']:.3f}", flush=True)
    e.close()


def test_charmm_water_resident_steps_keep_constraints(tmp_path):
    e, _ = build(tmp_path)
    e.setdt(0.002)
    ep, ek = e.gpu_steps(20)
    assert np.all(np.isfinite(ep)) and np.all(np.isfinite(ek))
    x = e.positions()
    worst = 0.0
    for rid in range(e.nrigids()):
        parts, ij, d2 = e.rigid(rid)
        for (i, j), dd in zip(ij, d2):
            dv = x[parts[i]] - x[parts[j]]
            dv -= BOX * np.round(dv / BOX)
            worst = max(worst, abs(dd - dv @ dv))
    print(f"largest |d^2 - r^2| after 20 steps: {worst:.2e}", flush=True)
    assert worst < 1e-6 * (1 + 1e-6)
    e.close()
