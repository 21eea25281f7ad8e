"""Rigid constraints on the host (engine_rigid_add, engine_rigid.c:78-230;
rigid_eval_shake / rigid_eval_pshake, rigid.c:81-523), pinned by what the
mathematics fixes:

- one constraint: SHAKE moves the pair along the bond before the drift to
  the root of a quadratic (closed form), keeping the centre of mass;
- a water molecule (three constraints): SHAKE and P-SHAKE both bring every
  residue |d^2 - r^2| below tol_rigid, keep the centre of mass, set
  v += (x_new - x) / dt, and agree with each other to the tolerance;
  P-SHAKE adds m (x_new - x) / dt^2 to the force;
- group bookkeeping: constraints sharing particles merge into one group.
"""
import numpy as np
import pytest

import mdcore as md


def engine(x, v, masses, types, dt, flags=md.engine_flag_none, box=3.0):
    e = md.Engine([0, 0, 0], [box] * 3, [1.0] * 3, 1.0, len(masses), flags)
    for m in masses:
        e.addtype(m)
    e.load(np.asarray(x, float), np.asarray(types, np.int32), np.asarray(v, float))
    e.setdt(dt)
    return e


def test_single_constraint_closed_form():
    rng = np.random.default_rng(1)
    m = np.array([15.9994, 1.008])
    x = np.array([[1.20, 1.30, 1.40], [1.29, 1.33, 1.43]])
    v = rng.normal(scale=2.0, size=(2, 3))
    dt, d = 0.002, 0.1
    e = engine(x, v, m, [0, 1], dt, md.engine_flag_shake)
    e.rigid_add(0, 1, d)
    e.rigid_eval()
    xn, vn = e.positions(), e.velocities()
    # closed form: x_i' = x_i + l m_j/(m_i+m_j) r_old, x_j' = x_j - l m_i/(m_i+m_j) r_old,
    # |r + l r_old|^2 = d^2 with r = x_i - x_j, r_old = (x_i - dt v_i) - (x_j - dt v_j)
    r = x[0] - x[1]
    ro = (x[0] - dt * v[0]) - (x[1] - dt * v[1])
    a, b, c = ro @ ro, 2 * r @ ro, r @ r - d * d
    lam = (-b + np.sqrt(b * b - 4 * a * c)) / (2 * a)
    want = np.array([x[0] + lam * m[1] / m.sum() * ro, x[1] - lam * m[0] / m.sum() * ro])
    # SHAKE stops at |d^2 - r^2| < 1e-6, i.e. within ~5e-6 nm of the root
    assert np.abs(xn - want).max() < 1e-5
    assert abs(np.sum((xn[0] - xn[1]) ** 2) - d * d) < 1e-6
    # every update moves along r_old: the displacements are parallel to it
    for k in range(2):
        dx = xn[k] - x[k]
        assert np.linalg.norm(np.cross(dx, ro)) <= 1e-9 * np.linalg.norm(dx) * np.linalg.norm(ro)
    np.testing.assert_allclose(m @ xn, m @ x, rtol=1e-14)
    np.testing.assert_allclose(vn, v + (xn - x) / dt, rtol=1e-9, atol=1e-9)
    e.close()


def water_molecule(seed):
    rng = np.random.default_rng(seed)
    roh, hoh = 0.1, np.radians(109.47)
    o = np.array([1.5, 1.5, 1.5])
    h1 = o + roh * np.array([1.0, 0.0, 0.0])
    h2 = o + roh * np.array([np.cos(hoh), np.sin(hoh), 0.0])
    x = np.array([o, h1, h2])
    v = rng.normal(scale=1.5, size=(3, 3))
    dt = 0.002
    # positions after a drift with small random forces: start from the
    # rigid geometry at t, drift with v plus a perturbation
    xt = x + dt * v + rng.normal(scale=2e-4, size=(3, 3))
    dhh = 2 * roh * np.sin(hoh / 2)
    return xt, v, dt, [(0, 1, roh), (0, 2, roh), (1, 2, dhh)]


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_water_shake_and_pshake(seed):
    m = np.array([15.9994, 1.008])
    types = [0, 1, 1]
    mm = m[types]
    x, v, dt, cons = water_molecule(seed)
    out = {}
    for name, flags in (("shake", md.engine_flag_shake), ("pshake", md.engine_flag_none)):
        e = engine(x, v, m, types, dt, flags)
        e.rigids_add(cons)
        assert e.nrigids() == 1
        e.rigid_eval()
        xn, vn, fn = e.positions(), e.velocities(), e.forces()
        for i, j, d in cons:
            assert abs(np.sum((xn[i] - xn[j]) ** 2) - d * d) < 1e-6, name
        np.testing.assert_allclose(mm @ xn, mm @ x, rtol=1e-13)
        np.testing.assert_allclose(vn, v + (xn - x) / dt, rtol=1e-9, atol=1e-9)
        if name == "pshake":
            np.testing.assert_allclose(fn, mm[:, None] * (xn - x) / dt ** 2, rtol=1e-9, atol=1e-6)
        else:
            assert np.abs(fn).max() == 0.0
        out[name] = xn
        e.close()
    assert np.abs(out["shake"] - out["pshake"]).max() < 1e-5


def test_groups_merge():
    x = np.array([[0.5 + 0.1 * k, 1.0, 1.0] for k in range(6)])
    e = engine(x, np.zeros((6, 3)), [1.0], [0] * 6, 0.002)
    e.rigids_add([(0, 1, 0.1), (2, 3, 0.1), (4, 5, 0.1)])
    assert e.nrigids() == 3
    e.rigid_add(1, 2, 0.1)  # joins the first two groups
    assert e.nrigids() == 2
    groups = sorted(sorted(e.rigid(r)[0].tolist()) for r in range(2))
    assert groups == [[0, 1, 2, 3], [4, 5]]
    e.rigid_add(3, 0, 0.3)  # both already in one group: a constraint only
    parts, ij, d2 = e.rigid([r for r in range(2) if len(e.rigid(r)[0]) == 4][0])
    assert len(ij) == 4 and np.allclose(d2, [0.01, 0.01, 0.01, 0.09])
    with pytest.raises(md.MdcoreError):
        e.rigid_add(2, 2, 0.1)
    e.close()
