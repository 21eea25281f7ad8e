"""Training data and training driver of the MLP surrogate (engine_mlp_*).

The surrogate replaces the table evaluation of potential_eval.h:67-104 for
pairs inside each species pair's domain [r_lo, rc]; the tables stay the
measure (targets, force error) and the fallback (pairs below r_lo).

  pair_domains   per species pair: r_lo where |E| first exceeds a cap
                 (the repulsive core stays on the table), r_hi = rc, the
                 energy scale sE of E = sE sinh(y) and the switch start
                 (r_hi: no switch)
  make_samples   half uniform in r per species pair, half the pair
                 distances of configurations (the plain engine's own
                 trajectory), targets E and dE/dr from the FP64 tables
  train          engine_mlp_init + engine_mlp_train on the GPU
  nve_drift      energy drift of device-resident NVE steps with the table
                 or with a network trained on the equilibrated state
"""
import numpy as np

import mdcore as md


def pair_potentials(eng, sysm):
    """Potential object of every species pair id (engine_mlp_pairs order)."""
    ti, tj = eng.mlp_pairs()
    table = {}
    for a, b, p in sysm.pots:
        table[(a, b)] = p
        table[(b, a)] = p
    return [table[(int(a), int(b))] for a, b in zip(ti, tj)], ti, tj


def pair_domains(eng, sysm, ecap=1.0e4, escale=1.0, switch=None, ngrid=20000):
    """r_lo[p]: smallest r such that |E| <= ecap on [r, rc] (the table's a at
    least); switch: None, or the distance from which the network's energy is
    switched off to zero at rc."""
    pots, _, _ = pair_potentials(eng, sysm)
    rc = float(sysm.cutoff)
    lo, hi = [], []
    for p in pots:
        a = p.data()["a"]
        r = np.linspace(a, rc, ngrid)
        e, _ = p.eval_r(r)
        bad = np.nonzero(np.abs(e) > ecap)[0]
        lo.append(r[bad[-1] + 1] if len(bad) else a)
        hi.append(rc)
    n = len(pots)
    rsw = np.array(hi) if switch is None else np.full(n, float(switch))
    return np.array(lo), np.array(hi), np.full(n, escale), rsw


def configuration_pairs(eng, sysm, x, nmax, seed, r_lo):
    """Pair distances of configuration x mapped to species pair ids, r >= r_lo[p]."""
    _, ti, tj = pair_potentials(eng, sysm)
    mt = sysm.max_type
    pid = -np.ones((mt, mt), np.int64)
    for k, (a, b) in enumerate(zip(ti, tj)):
        pid[a, b] = pid[b, a] = k
    r, a, b = md.sample_pairs(x, sysm.types, sysm.dim, float(sysm.cutoff), nmax, seed)
    p = pid[a, b]
    keep = (p >= 0) & (r >= r_lo[np.maximum(p, 0)])
    return r[keep], p[keep]


def make_samples(eng, sysm, n, r_lo, r_hi, configs=(), seed=1):
    """n samples: n/2 uniform in r over every species pair's domain, n/2 from
    the pair distances of the given configurations (if any).  Returns
    (r, pair, E, dE/dr) with FP64 table targets."""
    pots, _, _ = pair_potentials(eng, sysm)
    npair = len(pots)
    rng = np.random.default_rng(seed)
    nu = n // 2 if len(configs) else n
    p_u = rng.integers(0, npair, nu)
    r_u = r_lo[p_u] + (r_hi[p_u] - r_lo[p_u]) * rng.random(nu)
    rs, ps = [r_u], [p_u]
    if len(configs):
        need = n - nu
        per = need // len(configs) + 1
        for k, x in enumerate(configs):
            r_c, p_c = configuration_pairs(eng, sysm, x, 4 * per, seed + 17 * k + 1, r_lo)
            if len(r_c) > per:
                sel = rng.choice(len(r_c), per, replace=False)
                r_c, p_c = r_c[sel], p_c[sel]
            rs.append(r_c)
            ps.append(p_c)
    r = np.concatenate(rs)[:n]
    p = np.concatenate(ps)[:n]
    E = np.zeros_like(r)
    F = np.zeros_like(r)
    for k in range(npair):
        m = p == k
        if m.any():
            E[m], F[m] = pots[k].eval_r(r[m])
    return r, p.astype(np.int32), E, F


def train(eng, sysm, samples, domains, steps=4000, batch_tiles=512, lr0=3e-3, lr1=3e-5, seed=1):
    """engine_mlp_init + engine_mlp_train; returns the per-step loss."""
    r_lo, r_hi, es, rsw = domains
    eng.mlp_init(r_lo, r_hi, es, rsw, seed)
    r, p, E, F = samples
    return eng.mlp_train(r, p, E, F, steps, batch_tiles, lr0, lr1, 1.0, 1.0, seed)


def nve_energies(eng, masses, steps, dt):
    """Run engine_step `steps` times (kick-drift, leapfrog velocities) and
    return the total energy at each integer step:
    PE(x_n) + (KE(v_{n-1/2}) + KE(v_{n+1/2})) / 2."""
    eng.setdt(dt)
    m = np.asarray(masses)[:, None]
    v_prev = eng.velocities()
    out = []
    for _ in range(steps):
        eng.step()
        pe = eng.energies()["epot"]
        v = eng.velocities()
        ke = 0.25 * float(np.sum(m * (v_prev ** 2 + v ** 2)))
        out.append(pe + ke)
        v_prev = v
    return np.array(out)


def drift_per_ns(etot, dt):
    """Least-squares slope of the total energy, per ns, relative to |<E>|,
    and in kJ/mol per ns."""
    t = np.arange(len(etot)) * dt
    slope = np.polyfit(t, etot, 1)[0] * 1000.0
    return slope / abs(float(np.mean(etot))), slope


def nve_drift(sysm, mlp=False, switch=None, equil=1000, steps=10000, dt=0.002, train_samples=1 << 20,
              train_steps=10000, seed=7, sample_seed=5, train_seed=11):
    """Energy drift of `steps` device-resident NVE steps (engine_gpu_steps)
    after `equil` steps with the system's tables.  With mlp, the run after
    the equilibration evaluates the pairs with a network trained on the
    tables (half uniform-in-r samples, half the equilibrated configuration's
    pair distances), switched off to zero from `switch` to the cutoff if
    given (seed: initial weights; sample_seed: training samples;
    train_seed: batch order).  Total energy at step n: PE(x_n) + (KE(v_{n-1/2}) + KE(v_{n+1/2}))/2;
    drift = least-squares slope per ns, relative to |<E>| (whole run and
    each half) and in kJ/mol."""
    import dataclasses
    import time
    import systems
    t0 = time.perf_counter()
    eng = systems.build_engine(sysm, md.engine_flag_gpu)
    eng.setdt(dt)
    eng.gpu_steps(equil)
    if mlp:
        s_eq = dataclasses.replace(sysm, x=eng.positions(), v=eng.velocities())
        eng.close()
        eng = systems.build_engine(s_eq, md.engine_flag_gpu)
        eng.setdt(dt)
        dom = pair_domains(eng, s_eq, switch=switch)
        S = make_samples(eng, s_eq, train_samples, dom[0], dom[1], configs=[s_eq.x], seed=sample_seed)
        eng.mlp_init(*dom, seed=seed)
        eng.mlp_train(*S, steps=train_steps, batch_tiles=512, lr0=1e-2, lr1=1e-5, seed=train_seed)
        eng.setflags(md.engine_flag_mlp, 0)
    ep, ek = eng.gpu_steps(steps)
    eng.close()
    et = ep[1:] + 0.5 * (ek[:-1] + ek[1:])
    rel, absd = drift_per_ns(et, dt)
    h = len(et) // 2
    return {"drift_rel_per_ns": rel, "drift_kJ_mol_per_ns": absd,
            "drift_rel_per_ns_halves": [drift_per_ns(et[:h], dt)[0], drift_per_ns(et[h:], dt)[0]],
            "e_mean": float(et.mean()), "e_std": float(et.std()),
            "temperature_K": float(2 * ek.mean() / (3 * sysm.N * systems.KB)),
            "wall_s": round(time.perf_counter() - t0, 1)}
