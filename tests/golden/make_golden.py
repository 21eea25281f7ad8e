"""Write tests/golden/tables.npz from the reference's own table code.

The tables are built by oracle/_ref/libmdref.so, which is the reference's
unmodified potential.c/errs.c/cell.c compiled by oracle/Makefile.  Each case
stores the constructor arguments and the reference's n, alpha[4] and
coefficients, plus 64 (r^2, e, f) samples of the reference's
potential_eval.  Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
import pyoracle as po  # noqa: E402

KE = 138.935458

# (name, constructor, arguments); arguments follow the reference signatures.
CASES = [
    ("argon_LJ126", "LJ126", (0.275, 1.0, 4 * 0.996 * 0.3405 ** 12, 4 * 0.996 * 0.3405 ** 6, 1e-3)),
    ("argon_c_LJ126", "LJ126", (0.275, 1.0, 9.5075e-6, 6.1545e-3, 1e-3)),
    ("bulk_LJ126_Coulomb", "LJ126_Coulomb", (0.1, 1.0, 2.6378e-6, 2.6192e-3, 0.71842576, 1e-3)),
    ("Coulomb", "Coulomb", (0.1, 1.0, 0.71842576, 1e-3)),
    ("Ewald", "Ewald", (0.1, 1.0, -0.35921288, 3.0, 1e-4)),
    ("LJ126_Ewald", "LJ126_Ewald", (0.1, 1.0, 2.6378e-6, 2.6192e-3, 0.71842576, 3.0, 1e-4)),
    ("LJ126_switch", "LJ126_switch", (0.275, 1.0, 9.677e-6, 6.209e-3, 0.8, 1e-3)),
    ("harmonic", "harmonic", (0.05, 0.2, 1000.0, 0.1, 1e-4)),
    ("water_OO_RF", "LJ126_RF", (0.05, 1.0, 2.6378e-6, 2.6192e-3, 0.8476 * 0.8476, KE, 0.0, 1e-3)),
    ("water_OH_RF", "LJ126_RF", (0.05, 1.0, 0.0, 0.0, -0.8476 * 0.4238, KE, 0.0, 1e-3)),
    ("water_HH_RF", "LJ126_RF", (0.05, 1.0, 0.0, 0.0, 0.4238 * 0.4238, KE, 0.0, 1e-3)),
    ("mixture_RF_a", "LJ126_RF", (0.12, 1.2, 4 * 0.5 * 0.4 ** 12, 4 * 0.5 * 0.4 ** 6, 0.3, KE, 0.0, 1e-3)),
    ("mixture_RF_b", "LJ126_RF", (0.12, 1.2, 4 * 0.1 * 0.25 ** 12, 4 * 0.1 * 0.25 ** 6, -0.64, KE, 0.0, 1e-3)),
]


def ref_build(kind, args):
    r = po.ref()
    fn = {
        "LJ126": r.potential_create_LJ126,
        "LJ126_switch": r.potential_create_LJ126_switch,
        "LJ126_Coulomb": r.potential_create_LJ126_Coulomb,
        "LJ126_Ewald": r.potential_create_LJ126_Ewald,
        "Ewald": r.potential_create_Ewald,
        "Coulomb": r.potential_create_Coulomb,
        "harmonic": r.potential_create_harmonic,
        "LJ126_RF": r.ref_potential_create_LJ126_RF,
    }[kind]
    return po.ref_table(fn(*args))


def main():
    out = {}
    rng = np.random.default_rng(1234)
    for name, kind, args in CASES:
        t = ref_build(kind, args)
        a, b = args[0], args[1]
        r = np.sort(rng.uniform(a, b, 64))
        r2 = r * r
        ef = np.array([po.ref_eval(t, x) for x in r2])
        out[f"{name}/kind"] = np.array(kind)
        out[f"{name}/args"] = np.array(args, dtype=np.float64)
        out[f"{name}/n"] = np.array(t["n"])
        out[f"{name}/alpha"] = t["alpha"]
        out[f"{name}/c"] = t["c"]
        out[f"{name}/r2"] = r2
        out[f"{name}/ef"] = ef
        print(f"{name:20s} n={t['n']}")
    np.savez_compressed(Path(__file__).with_name("tables.npz"), **out)


if __name__ == "__main__":
    main()
