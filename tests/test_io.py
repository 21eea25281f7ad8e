"""Structure and parameter readers (engine_read_psf / engine_read_cpf /
engine_read_xplor, engine_io.c), checked against what the file contents fix:

- PSF: one type per atom type and charge, masses and charges as written,
  the topology lists as written, PDB positions = 0.1 x Angstrom, wrapped;
- CHARMM / XPLOR NONBONDED: every pair table has its Lennard-Jones minimum
  at r0 = Rmin/2_i + Rmin/2_j (CHARMM; XPLOR: the reference reads 0.05 x the
  given length) with depth -sqrt(eps_i eps_j) x 4.184 kJ/mol, plus the
  electrostatic term of the chosen kappa (shifted Coulomb, real-space Ewald
  or none), and starts where |V| = engine_maxKcutoff |V(r0)|;
- rigidH: bonds with a hydrogen type become rigid constraints of the BOND
  length and leave the bond list.

The reference's own JAC inputs (DHFR, 23,558 atoms, CHARMM22) are read when
the reference tree is present; the test is skipped otherwise.
"""
import math
import os
from pathlib import Path

import numpy as np
import pytest
from scipy.special import erfc

import mdcore as md

ESCALE = 0.079577471545947667882  # potential_escale, 1/(4 pi)

# TIP3P water and a methane, CHARMM names and charges
ATOMS = [  # (resid, resname, atomname, type, charge, mass, x, y, z in Angstrom)
    (1, "TIP3", "OH2", "OT", -0.834, 15.9994, 2.000, 2.000, 2.000),
    (1, "TIP3", "H1", "HT", 0.417, 1.008, 2.957, 2.000, 2.000),
    (1, "TIP3", "H2", "HT", 0.417, 1.008, 1.760, 2.927, 2.000),
    (2, "TIP3", "OH2", "OT", -0.834, 15.9994, 10.000, 10.500, 11.000),
    (2, "TIP3", "H1", "HT", 0.417, 1.008, 10.957, 10.500, 11.000),
    (2, "TIP3", "H2", "HT", 0.417, 1.008, 9.760, 11.427, 11.000),
    (3, "MET", "C", "CT3", -0.27, 12.011, 15.000, 5.000, 5.000),
    (3, "MET", "H1", "HA", 0.09, 1.008, 16.111, 5.000, 5.000),
    (3, "MET", "H2", "HA", 0.09, 1.008, 14.630, 6.047, 5.000),
    (3, "MET", "H3", "HA", 0.09, 1.008, 14.630, 4.476, 5.907),
    (3, "MET", "H4", "HA", 0.09, 1.008, 14.630, 4.476, 4.093),
    (4, "ION", "NA", "SOD", 1.0, 22.99, -3.000, 25.000, 7.000),  # outside the box: wraps
]
BONDS = [(1, 2), (1, 3), (2, 3), (4, 5), (4, 6), (5, 6), (7, 8), (7, 9), (7, 10), (7, 11)]
ANGLES = [(2, 1, 3), (5, 4, 6), (8, 7, 9), (8, 7, 10), (8, 7, 11), (9, 7, 10), (9, 7, 11), (10, 7, 11)]
DIHEDRALS = [(8, 7, 9, 10)]
LJ = {  # CHARMM: epsilon (kcal/mol, negative), Rmin/2 (Angstrom)
    "OT": (-0.1521, 1.7682), "HT": (-0.046, 0.2245), "CT3": (-0.08, 2.06), "HA": (-0.022, 1.32),
    "SOD": (-0.0469, 1.36375),
}
BOND_R0 = {("OT", "HT"): 0.9572, ("HT", "HT"): 1.5139, ("CT3", "HA"): 1.111}
BOX, CUTOFF = 2.0, 0.9


def write_psf(path):
    lines = ["PSF ", "", "       2 !NTITLE", "* synthetic test system", "* water, methane, an ion", "",
             f"{len(ATOMS):8d} !NATOM"]
    for k, (rid, rn, an, ty, q, m, *_x) in enumerate(ATOMS, 1):
        lines.append(f"{k:8d} SYN  {rid:<4d} {rn:<4s} {an:<4s} {ty:<4s} {q:10.6f}    {m:10.4f}           0")
    for title, items, per in (("NBOND: bonds", BONDS, 4), ("NTHETA: angles", ANGLES, 3),
                              ("NPHI: dihedrals", DIHEDRALS, 2), ("NIMPHI: impropers", [], 2)):
        lines += ["", f"{len(items):8d} !{title}"]
        for s in range(0, len(items), per):
            lines.append("".join(f"{i:8d}" for item in items[s:s + per] for i in item))
    Path(path).write_text("\n".join(lines) + "\n\n")


def write_pdb(path):
    lines = ["REMARK synthetic"]
    for k, (rid, rn, an, ty, q, m, x, y, z) in enumerate(ATOMS, 1):
        lines.append(f"ATOM {k:6d}  {an:<4s}{rn:<4s} {rid:4d}    {x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00      SYN")
    lines.append("END")
    Path(path).write_text("\n".join(lines) + "\n")


def write_cpf(path):
    lines = ["* synthetic CHARMM parameters", "*", "", "BONDS",
             "OT   HT    450.0   0.9572 ! water", "HT   HT      0.0   1.5139", "CT3  HA    322.0   1.1110",
             "", "ANGLES", "HT   OT   HT     55.0   104.52", "HA   CT3  HA     35.5   108.40    5.40   1.80200",
             "", "DIHEDRALS", "X    CT3  CT3  X     0.1600  3     0.00", "",
             "IMPROPER", "HA   X    X    CT3   0.0  0  0.00", "",
             "NONBONDED nbxmod  5 atom cdiel shift vatom vdistance vswitch cutnb 14.0 ctofnb 12.0 eps 1.0",
             "!atom  ignored    epsilon      Rmin/2   ignored   eps,1-4       Rmin/2,1-4"]
    for ty, (eps, rh) in LJ.items():
        lines.append(f"{ty:<6s} 0.000000  {eps:9.6f}  {rh:9.6f}")
    lines += ["", "HBOND CUTHB 0.5", "", "END"]
    Path(path).write_text("\n".join(lines) + "\n")


def write_xplor(path):
    lines = ["REMARKS synthetic XPLOR parameters", "set echo=false end", "{ a comment in braces",
             " BOND  OT   HT    450.0   0.9572 ! water", " BOND  HT   HT      0.0   1.5139",
             " BOND  CT3  HA    322.0   1.1110", " ANGLE HT   OT   HT     55.0   104.52",
             " DIHEDRAL  X  CT3  CT3  X   MULTIPLE= 2   0.16  3  0.00", "{ NONBONDED nbxmod 5 atom -"]
    for ty, (eps, rh) in LJ.items():
        lines.append(f" NONBONDED  {ty:<6s} {abs(eps):8.4f}  {2 * rh:8.4f}  {abs(eps):8.4f}  {2 * rh:8.4f}")
    lines.append("set echo=true end")
    Path(path).write_text("\n".join(lines) + "\n")


def engine(max_type=10):
    return md.Engine([0, 0, 0], [BOX] * 3, [CUTOFF] * 3, CUTOFF, max_type, md.engine_flag_gpu)


@pytest.fixture()
def files(tmp_path):
    paths = {k: tmp_path / f"syn.{k}" for k in ("psf", "pdb", "inp", "xplor")}
    write_psf(paths["psf"])
    write_pdb(paths["pdb"])
    write_cpf(paths["inp"])
    write_xplor(paths["xplor"])
    return paths


def type_ids(e):
    out = {}
    n = e.typeinfo(0)["ntypes"]
    for k in range(n):
        t = e.typeinfo(k)
        out[(t["name"], round(t["charge"], 6))] = k
    return out


def test_psf_pdb(files):
    e = engine()
    e.read_psf(files["psf"], files["pdb"])
    assert e.N == len(ATOMS)
    ids = type_ids(e)
    # null type + one type per (atom type, charge)
    assert len(ids) == 1 + len({(a[3], a[4]) for a in ATOMS})
    for rid, rn, an, ty, q, m, *_ in ATOMS:
        t = e.typeinfo(ids[(ty, round(q, 6))])
        assert t["mass"] == m and t["charge"] == q and t["name2"] == f"{ty}{q:.6f}"
    x = e.positions()
    want = np.mod(0.1 * np.array([a[6:9] for a in ATOMS]), BOX)
    np.testing.assert_allclose(x, want, rtol=0, atol=1e-12)
    assert np.array_equal(e.topology(0), np.array(BONDS) - 1)
    assert np.array_equal(e.topology(1), np.array(ANGLES) - 1)
    d = e.topology(2)
    assert np.array_equal(d[:, :4], np.array(DIHEDRALS) - 1) and np.all(d[:, 4] == -1)
    e.close()


def expected_v(r, ti, tj, kappa, shifted=True):
    """Pair potential at r < switchlen from the file's parameters; the
    inner-radius search uses the unshifted Coulomb term (shifted=False)."""
    (ei, hi), (ej, hj) = LJ[ti[0]], LJ[tj[0]]
    K = math.sqrt(4.184 * ei * 4.184 * ej)
    r0 = 0.1 * (hi + hj)
    lj = K * ((r0 / r) ** 12 - 2 * (r0 / r) ** 6)
    q = ti[1] * tj[1]
    if kappa < 0:
        return lj
    if kappa == 0:
        return lj + ESCALE * q * (1.0 / r - (1.0 / CUTOFF if shifted else 0.0))
    return lj + q * ESCALE * erfc(kappa * r) / r


@pytest.mark.parametrize("kappa", [-1.0, 0.0, 3.0])
def test_cpf_tables(files, kappa):
    e = engine()
    e.read_psf(files["psf"], files["pdb"])
    e.read_cpf(files["inp"], kappa, 1e-5, rigidH=0, switchlen=0.7)
    ids = type_ids(e)
    keys = sorted(k for k in ids if k[0] in LJ)
    checked = 0
    for a in range(len(keys)):
        for b in range(a, len(keys)):
            ta, tb = keys[a], keys[b]
            p = e.pot(ids[ta], ids[tb])
            assert p is not None, (ta, tb)
            r0 = 0.1 * (LJ[ta[0]][1] + LJ[tb[0]][1])
            v0 = expected_v(r0, ta, tb, kappa, shifted=False)
            d = p.data()
            # the inner radius: |V(a)| = 2 |V(r0)| within the 1e-5 nm bisection
            va = expected_v(d["a"], ta, tb, kappa, shifted=False)
            vb = expected_v(d["a"] + 2e-5, ta, tb, kappa, shifted=False)
            assert abs(vb) <= 2 * abs(v0) * (1 + 1e-9) <= abs(va) * (1 + 1e-9) or r0 <= d["a"] + 2e-5, (ta, tb)
            if r0 < d["a"]:
                continue
            for r in (r0, 0.5 * (r0 + 0.7), 0.65):
                if r >= d["b"] or r < d["a"]:
                    continue
                got = p.eval_r(np.array([r]))[0][0]
                want = expected_v(r, ta, tb, kappa)
                assert abs(got - want) <= 1e-4 * max(1.0, abs(want)), (ta, tb, r, got, want)
                checked += 1
    assert checked >= 10
    e.close()


def test_cpf_rigid_hydrogens(files):
    e = engine()
    e.read_psf(files["psf"], files["pdb"])
    e.read_cpf(files["inp"], 0.0, 1e-4, rigidH=1, switchlen=0.7)
    assert len(e.topology(0)) == 0  # every bond involves a hydrogen type
    assert e.nrigids() == 3
    groups = {}
    for r in range(3):
        parts, ij, d2 = e.rigid(r)
        groups[tuple(sorted(parts.tolist()))] = sorted(np.sqrt(d2).round(6).tolist())
    assert groups[(0, 1, 2)] == [0.09572, 0.09572, 0.15139]
    assert groups[(3, 4, 5)] == [0.09572, 0.09572, 0.15139]
    assert groups[(6, 7, 8, 9, 10)] == [0.1111] * 4
    e.close()


def test_xplor_tables(files):
    e = engine()
    e.read_psf(files["psf"], files["pdb"])
    e.read_xplor(files["xplor"], -1.0, 1e-5, rigidH=1)
    ids = type_ids(e)
    assert e.nrigids() == 3 and len(e.topology(0)) == 0
    t = e.typeinfo(ids[("OT", -0.834)])
    # eps as given (kcal/mol -> kJ/mol); length x 0.05 (the reference's reading)
    assert t["eps"] == pytest.approx(4.184 * 0.1521) and t["rmin"] == pytest.approx(0.05 * 2 * 1.7682)
    p = e.pot(ids[("OT", -0.834)], ids[("OT", -0.834)])
    r0 = 2 * t["rmin"]
    got = p.eval_r(np.array([r0]))[0][0]
    assert got == pytest.approx(-4.184 * 0.1521, rel=1e-4)
    e.close()


def test_reader_errors(files, tmp_path):
    bad = tmp_path / "bad.inp"
    bad.write_text("* no sections\nANGLES\nX X X 1 2\n")
    e = engine()
    e.read_psf(files["psf"], files["pdb"])
    with pytest.raises(md.MdcoreError):
        e.read_cpf(bad, 0.0, 1e-4)
    notpsf = tmp_path / "bad.psf"
    notpsf.write_text("PDB\n")
    e2 = engine()
    with pytest.raises(md.MdcoreError):
        e2.read_psf(notpsf, files["pdb"])
    e.close()
    e2.close()


JAC = Path("/root/reference/examples/jac")


@pytest.mark.skipif(not (JAC / "5dhfr_cube.psf").exists(), reason="reference example inputs not present")
def test_reference_jac_inputs():
    """The reference's DHFR benchmark inputs, set up as its jac.c does
    (62.23 A cube, rc 0.9 nm, kappa 3.0, tol 1e-4, rigid hydrogens, switch at
    0.7 nm), read in place."""
    e = md.Engine([0, 0, 0], [6.223] * 3, [0.9] * 3, 0.9, 100, md.engine_flag_gpu)
    e.read_psf(JAC / "5dhfr_cube.psf", JAC / "5dhfr_cube.pdb")
    assert e.N == 23558
    nbond = len(e.topology(0))
    header = [l for l in open(JAC / "5dhfr_cube.psf") if "!NBOND" in l][0]
    assert nbond == int(header.split()[0])
    ntypes = e.typeinfo(0)["ntypes"]
    assert 64 < ntypes <= 101  # more than the 6 type bits of the smallest record
    e.read_cpf(JAC / "par_all22_prot.inp", 3.0, 1e-4, rigidH=1, switchlen=0.7)
    nrig = e.nrigids()
    ncons = sum(len(e.rigid(r)[2]) for r in range(nrig))
    assert ncons + len(e.topology(0)) == nbond  # each removed bond became a constraint
    assert nrig > 7000  # 7,023 waters plus the protein's X-H groups
    # every type with a Lennard-Jones well has a table with every other
    with_lj = [k for k in range(1, ntypes) if e.typeinfo(k)["eps"] != 0.0]
    missing = [(i, j) for i in with_lj for j in with_lj if e.pot(i, j) is None]
    assert not missing
    # spot check: the water oxygen pair at its minimum
    ids = type_ids(e)
    o = ids[("OT", -0.834)]
    t = e.typeinfo(o)
    r0 = 2 * t["rmin"]
    got = e.pot(o, o).eval_r(np.array([r0]))[0][0]
    want = -abs(t["eps"]) + 0.834 ** 2 * ESCALE * erfc(3.0 * r0) / r0
    assert got == pytest.approx(want, rel=1e-3)
    e.close()
