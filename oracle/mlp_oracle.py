"""FP64 reference of the MLP surrogate (test infrastructure only).

The product's network lives in mdcore-history_amd/csrc/mlp_tile.h (forward)
and mlp_gpu.hip (backward, Adam).  This module restates it in numpy, in
float64, so the GPU's forward values and parameter gradients can be checked:

  x     = (rn, irn) = ((2r - r_lo - r_hi) / (r_hi - r_lo), r_lo / r)
  z1    = W1 x + b1 + E1 emb[p]
  h1    = silu(z1)                   (rounded to bf16 where the GPU rounds)
  z2    = W2 h1 + b2                 (W2 rounded to bf16)
  h2    = silu(z2)
  y_o   = W3[p, o] . h2 + b3[p, o]
  E     = sE sinh(y0),  dE/dr = sF sinh(y1)
  loss  = mean(wE (y0 - asinh(E*/sE))^2 + wF (y1 - asinh(F*/sF))^2)

`rounding=True` applies bf16 round-to-nearest-even exactly where the GPU
feeds its MFMA operands (h1 and W2 on the forward product; dz2 and W2 on
the two backward products); the backward pass treats the rounding as the
identity, as the GPU does.  With `rounding=False` this is the exact
gradient of the smooth network, which the tests pin against central finite
differences.  Nothing in the product imports this file.
"""
import numpy as np

H = 64
EMB = 8


def offsets(npair):
    o = {}
    o["W1"] = 0
    o["B1"] = o["W1"] + 2 * H
    o["E1"] = o["B1"] + H
    o["W2"] = o["E1"] + H * EMB
    o["B2"] = o["W2"] + H * H
    o["EMB"] = o["B2"] + H
    o["W3"] = o["EMB"] + EMB * npair
    o["B3"] = o["W3"] + 2 * H * npair
    o["N"] = o["B3"] + 2 * npair
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
        "W3": t[o["W3"]:o["B3"]].reshape(npair, 2, H),
        "b3": t[o["B3"]:o["N"]].reshape(npair, 2),
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


def forward(theta, npair, sc, r, pair, rounding=True):
    """sc: (npair, 4) = (r_lo, r_hi, sE, sF).  Returns a dict with y0, y1, E, dEdr and the activations."""
    P = unpack(theta, npair)
    r = np.asarray(r, dtype=np.float64)
    p = np.asarray(pair, dtype=np.int64)
    sc = np.asarray(sc, dtype=np.float64)
    lo, hi = sc[p, 0], sc[p, 1]
    rn = (2 * r - lo - hi) / (hi - lo)
    irn = lo / r
    b1e = P["b1"][None, :] + P["emb"] @ P["E1"].T          # (npair, H)
    z1 = rn[:, None] * P["W1"][None, :, 0] + irn[:, None] * P["W1"][None, :, 1] + b1e[p]
    h1 = silu(z1)
    W2 = P["W2"]
    if rounding:
        h1 = bf16(h1)
        W2 = bf16(W2)
    z2 = h1 @ W2.T + P["b2"][None, :]
    h2 = silu(z2)
    y0 = np.einsum("nh,nh->n", h2, P["W3"][p, 0]) + P["b3"][p, 0]
    y1 = np.einsum("nh,nh->n", h2, P["W3"][p, 1]) + P["b3"][p, 1]
    return {"P": P, "rn": rn, "irn": irn, "z1": z1, "h1": h1, "W2": W2, "z2": z2, "h2": h2, "y0": y0, "y1": y1,
            "E": sc[p, 2] * np.sinh(y0), "dEdr": sc[p, 3] * np.sinh(y1), "p": p}


def loss_and_grad(theta, npair, sc, r, pair, E, F, wE=1.0, wF=1.0, rounding=True):
    """Mean loss and its gradient with respect to theta (flat, float64)."""
    f = forward(theta, npair, sc, r, pair, rounding)
    P, p = f["P"], f["p"]
    sc = np.asarray(sc, dtype=np.float64)
    n = len(p)
    t0 = np.arcsinh(np.asarray(E, np.float64) / sc[p, 2])
    t1 = np.arcsinh(np.asarray(F, np.float64) / sc[p, 3])
    d0 = f["y0"] - t0
    d1 = f["y1"] - t1
    loss = np.mean(wE * d0 ** 2 + wF * d1 ** 2)
    dy0 = 2 * wE * d0 / n
    dy1 = 2 * wF * d1 / n
    G = {k: np.zeros_like(v) for k, v in P.items()}
    np.add.at(G["W3"][:, 0], p, dy0[:, None] * f["h2"])
    np.add.at(G["W3"][:, 1], p, dy1[:, None] * f["h2"])
    np.add.at(G["b3"][:, 0], p, dy0)
    np.add.at(G["b3"][:, 1], p, dy1)
    dh2 = dy0[:, None] * P["W3"][p, 0] + dy1[:, None] * P["W3"][p, 1]
    dz2 = dh2 * dsilu(f["z2"])
    G["b2"] = dz2.sum(0)
    dz2m = bf16(dz2) if rounding else dz2
    G["W2"] = dz2m.T @ f["h1"]
    dh1 = dz2m @ f["W2"]
    dz1 = dh1 * dsilu(f["z1"])
    G["W1"][:, 0] = dz1.T @ f["rn"]
    G["W1"][:, 1] = dz1.T @ f["irn"]
    gb1e = np.zeros((npair, H))
    np.add.at(gb1e, p, dz1)
    G["b1"] = gb1e.sum(0)
    G["E1"] = gb1e.T @ P["emb"]
    G["emb"] = gb1e @ P["E1"]
    return loss, pack(G, npair)
