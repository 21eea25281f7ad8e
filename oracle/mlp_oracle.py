"""FP64 reference of the MLP surrogate (test infrastructure only).

The product's network lives in mdcore-history_amd/csrc/mlp_tile.h (forward)
and mlp_gpu.hip (backward, Adam).  This module restates it in numpy, in
float64, so the GPU's forward values and parameter gradients can be checked.

The network has one output per pair, y(r); the force is its exact
derivative, carried forward through both layers (forward mode), so the
surrogate is conservative by construction:

  x     = (rn, irn)  = ((2r - r_lo - r_hi) / (r_hi - r_lo), r_lo / r)
  dx/dr = (drn, dirn) = (2 / (r_hi - r_lo), -r_lo / r^2)
  z1    = W1 x + b1 + E1 emb[p]      t1 = W1 dx/dr         (dz1/dr)
  h1    = silu(z1)                   g1 = silu'(z1) t1      (dh1/dr)
  z2    = W2 h1 + b2                 t2 = W2 g1
  h2    = silu(z2)                   g2 = silu'(z2) t2
  y     = W3[p] . h2 + b3[p]         yd = W3[p] . g2        (dy/dr)
  E     = S(r) (sE sinh(y) - shift[p])
  dE/dr = S'(r) (sE sinh(y) - shift[p]) + S(r) sE cosh(y) yd

S is the switch of potential_switch (the reference's potential.c:69-83,
derivative 86-100) from r_sw[p] to r_hi[p]; r_sw >= r_hi means S = 1.
shift[p] makes the unswitched energy at r_hi equal to the table's there.

Training loss per sample, in the output's own coordinate y (the switch and
shift are not part of it; they are exact functions applied afterwards):
  wE (y - a)^2 + wF (yd - ad)^2 / (ad^2 + 1/c^2),
  a = asinh(E*/sE),  ad = F*/sqrt(sE^2 + E*^2) = da/dr for the target's
  dE*/dr = F*,  c = (r_hi - r_lo) / 2.

`rounding=True` applies bf16 round-to-nearest-even exactly where the GPU
feeds its MFMA operands: W2 on every product; h1 and g1 on the forward
products as a pair hi + lo (hi = bf16(v), lo = bf16(v - hi)); dz2 and dt2
on the backward products, and the hi halves of h1 and g1 in dW2.  The
backward pass treats the rounding as the identity, as the GPU does.  With `rounding=False` this is the exact
gradient of the smooth network, which the tests pin against central finite
differences.  Nothing in the product imports this file.
"""
import numpy as np

H = 64
EMB = 8
BLOCKS = ["W1", "B1", "E1", "W2", "B2", "EMB", "W3", "B3", "N"]


def offsets(npair):
    o = {}
    o["W1"] = 0
    o["B1"] = o["W1"] + 2 * H
    o["E1"] = o["B1"] + H
    o["W2"] = o["E1"] + H * EMB
    o["B2"] = o["W2"] + H * H
    o["EMB"] = o["B2"] + H
    o["W3"] = o["EMB"] + EMB * npair
    o["B3"] = o["W3"] + H * npair
    o["N"] = o["B3"] + npair
    return o


def unpack(theta, npair):
    o = offsets(npair)
    t = np.asarray(theta, dtype=np.float64)
    assert t.size == o["N"], (t.size, o["N"])
    return {
        "W1": t[o["W1"]:o["B1"]].reshape(H, 2),
        "b1": t[o["B1"]:o["E1"]],
        "E1": t[o["E1"]:o["W2"]].reshape(H, EMB),
        "W2": t[o["W2"]:o["B2"]].reshape(H, H),
        "b2": t[o["B2"]:o["EMB"]],
        "emb": t[o["EMB"]:o["W3"]].reshape(npair, EMB),
        "W3": t[o["W3"]:o["B3"]].reshape(npair, H),
        "b3": t[o["B3"]:o["N"]].reshape(npair),
    }


def pack(P, npair):
    o = offsets(npair)
    t = np.zeros(o["N"])
    t[o["W1"]:o["B1"]] = P["W1"].ravel()
    t[o["B1"]:o["E1"]] = P["b1"]
    t[o["E1"]:o["W2"]] = P["E1"].ravel()
    t[o["W2"]:o["B2"]] = P["W2"].ravel()
    t[o["B2"]:o["EMB"]] = P["b2"]
    t[o["EMB"]:o["W3"]] = P["emb"].ravel()
    t[o["W3"]:o["B3"]] = P["W3"].ravel()
    t[o["B3"]:o["N"]] = P["b3"].ravel()
    return t


def bf16(x):
    """Round to bfloat16 (round to nearest, ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def d2silu(z):
    s = sigmoid(z)
    return s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s))


def switch(r, a, b):
    """S(r) and dS/dr of the switch from a to b (S = 1 below a, 0 above b):
    (b^2 - r^2)^2 (b^2 + 2 r^2 - 3 a^2) / (b^2 - a^2)^3; a >= b gives S = 1."""
    r = np.asarray(r, np.float64)
    a = np.broadcast_to(np.asarray(a, np.float64), r.shape)
    b = np.broadcast_to(np.asarray(b, np.float64), r.shape)
    S = np.ones_like(r)
    dS = np.zeros_like(r)
    on = (a < b) & (r > a)
    ra, aa, bb = r[on], a[on], b[on]
    r2, a2, b2 = ra * ra, aa * aa, bb * bb
    q = b2 - r2
    den = (b2 - a2) ** 3
    S[on] = np.where(ra < bb, q * q * (b2 + 2 * r2 - 3 * a2) / den, 0.0)
    dS[on] = np.where(ra < bb, (-4 * ra * q * (b2 + 2 * r2 - 3 * a2) + q * q * 4 * ra) / den, 0.0)
    return S, dS


def forward(theta, npair, sc, r, pair, rounding=True, shift=None):
    """sc: (npair, 4) = (r_lo, r_hi, sE, r_sw).  Returns a dict with y, yd,
    E, dEdr and the activations."""
    P = unpack(theta, npair)
    r = np.asarray(r, dtype=np.float64)
    p = np.asarray(pair, dtype=np.int64)
    sc = np.asarray(sc, dtype=np.float64)
    lo, hi = sc[p, 0], sc[p, 1]
    rn = (2 * r - lo - hi) / (hi - lo)
    irn = lo / r
    drn = 2.0 / (hi - lo)
    dirn = -lo / (r * r)
    b1e = P["b1"][None, :] + P["emb"] @ P["E1"].T          # (npair, H)
    z1 = rn[:, None] * P["W1"][None, :, 0] + irn[:, None] * P["W1"][None, :, 1] + b1e[p]
    t1 = drn[:, None] * P["W1"][None, :, 0] + dirn[:, None] * P["W1"][None, :, 1]
    h1 = silu(z1)
    g1 = dsilu(z1) * t1
    W2 = P["W2"]
    h1hi, g1hi = h1, g1
    if rounding:
        h1hi, g1hi = bf16(h1), bf16(g1)
        h1 = h1hi + bf16(np.float32(h1) - np.float32(h1hi))
        g1 = g1hi + bf16(np.float32(g1) - np.float32(g1hi))
        W2 = bf16(W2)
    z2 = h1 @ W2.T + P["b2"][None, :]
    t2 = g1 @ W2.T
    h2 = silu(z2)
    g2 = dsilu(z2) * t2
    y = np.einsum("nh,nh->n", h2, P["W3"][p]) + P["b3"][p]
    yd = np.einsum("nh,nh->n", g2, P["W3"][p])
    sh = np.zeros(len(p)) if shift is None else np.asarray(shift, np.float64)[p]
    S, dS = switch(r, sc[p, 3], hi)
    eu = sc[p, 2] * np.sinh(y) - sh
    return {"P": P, "rn": rn, "irn": irn, "drn": drn, "dirn": dirn, "z1": z1, "t1": t1, "h1": h1, "g1": g1,
            "h1hi": h1hi, "g1hi": g1hi,
            "W2": W2, "z2": z2, "t2": t2, "h2": h2, "g2": g2, "y": y, "yd": yd, "E": S * eu,
            "dEdr": dS * eu + S * sc[p, 2] * np.cosh(y) * yd, "p": p}


def shifts(theta, npair, sc, e_table_at_rhi, rounding=True):
    """shift[p] = sE sinh(y_p(r_hi)) - E_table_p(r_hi): the unswitched
    network energy at r_hi then equals the table's."""
    sc = np.asarray(sc, np.float64)
    p = np.arange(npair)
    f = forward(theta, npair, sc, sc[:, 1], p, rounding)
    return sc[:, 2] * np.sinh(f["y"]) - np.asarray(e_table_at_rhi, np.float64)


def targets(sc, p, E, F):
    """The loss's targets in y: asinh(E/sE) and its r-derivative."""
    sE = np.asarray(sc, np.float64)[p, 2]
    E = np.asarray(E, np.float64)
    F = np.asarray(F, np.float64)
    return np.arcsinh(E / sE), F / np.sqrt(sE * sE + E * E)


def loss_and_grad(theta, npair, sc, r, pair, E, F, wE=1.0, wF=1.0, rounding=True):
    """Mean loss and its gradient with respect to theta (flat, float64)."""
    f = forward(theta, npair, sc, r, pair, rounding)
    P, p = f["P"], f["p"]
    sc = np.asarray(sc, dtype=np.float64)
    n = len(p)
    a, ad = targets(sc, p, E, F)
    c = 0.5 * (sc[p, 1] - sc[p, 0])
    q = 1.0 / (ad * ad + 1.0 / (c * c))
    d0 = f["y"] - a
    d1 = f["yd"] - ad
    loss = np.mean(wE * d0 ** 2 + wF * q * d1 ** 2)
    dy = 2 * wE * d0 / n
    dyd = 2 * wF * q * d1 / n
    G = {k: np.zeros_like(v) for k, v in P.items()}
    np.add.at(G["W3"], p, dy[:, None] * f["h2"] + dyd[:, None] * f["g2"])
    np.add.at(G["b3"], p, dy)
    dh2 = dy[:, None] * P["W3"][p]
    dg2 = dyd[:, None] * P["W3"][p]
    dz2 = dh2 * dsilu(f["z2"]) + dg2 * f["t2"] * d2silu(f["z2"])
    dt2 = dg2 * dsilu(f["z2"])
    G["b2"] = dz2.sum(0)
    dz2m = bf16(dz2) if rounding else dz2
    dt2m = bf16(dt2) if rounding else dt2
    G["W2"] = dz2m.T @ f["h1hi"] + dt2m.T @ f["g1hi"]
    dh1 = dz2m @ f["W2"]
    dg1 = dt2m @ f["W2"]
    dz1 = dh1 * dsilu(f["z1"]) + dg1 * f["t1"] * d2silu(f["z1"])
    dt1 = dg1 * dsilu(f["z1"])
    G["W1"][:, 0] = dz1.T @ f["rn"] + dt1.T @ (f["drn"] * np.ones(n))
    G["W1"][:, 1] = dz1.T @ f["irn"] + dt1.T @ f["dirn"]
    gb1e = np.zeros((npair, H))
    np.add.at(gb1e, p, dz1)
    G["b1"] = gb1e.sum(0)
    G["E1"] = gb1e.T @ P["emb"]
    G["emb"] = gb1e @ P["E1"]
    return loss, pack(G, npair)
