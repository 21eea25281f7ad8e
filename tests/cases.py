"""Synthetic systems shared by the CPU oracle tests and the GPU parity tests."""
import numpy as np

import mdcore as md
import systems


def face_pairs(ncell=4, h=(0.96, 0.98, 1.0), cutoff=0.95, density=20.0, npairs=24, gap=2e-6, seed=29):
    """Uniform random atoms in ncell^3 cells of h (span 1), plus planted
    pairs straddling a cell face: atom P sits `gap` inside the far face of
    cell c along axis k, atom Q `gap` inside the near face of the neighbour
    c + e_k, 0.25 nm apart across it.  The cell sizes put the fractional
    part of dscale*h_k above one half on every axis (dscale = 32767/(2|h|),
    runner_dosort.c:118-119), so P's key exceeds Q's by floor(dscale*h_k) + 1
    = dnshift + 1: the reference window's unsigned break
    (runner_dopair.c:462) wraps at the first i and the whole pair task is
    skipped, although P and Q are in range.  (In cubic cells dscale*h =
    32767/(2*sqrt(3)) = 9459.02 and the keys can differ by at most dnshift.)"""
    rng = np.random.default_rng(seed)
    h = np.asarray(h, float)
    dim = ncell * h
    N0 = int(density * np.prod(dim))
    x = [md.gen_uniform(3 * N0, 0.0, 1.0, seed).reshape(-1, 3) * dim]
    for k in range(npairs):
        ax = k % 3
        c = rng.integers(0, ncell, 3).astype(float)
        p = c * h + rng.uniform(0.3, 0.7, 3) * h
        q = p + 0.25 / np.sqrt(2.0)
        p[ax] = (c[ax] + 1) * h[ax] - gap
        q[ax] = (c[ax] + 1) * h[ax] + gap
        x.append(np.mod(np.stack([p, q]), dim))
    x = np.concatenate(x)
    pot = md.LJ126(0.05, cutoff, 1e-6, 1e-3, 1e-3)
    return systems.System("face_pairs", np.zeros(3), dim, 0.999 * h, cutoff, x, np.ones(len(x), np.int32),
                          np.array([0.0, 1.0]), [(1, 1, pot)])
