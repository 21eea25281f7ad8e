/*
 * Per-species-pair MLP surrogate of the interpolated pair potential:
 * the device-side network shared by the training, sample-evaluation and
 * nonbonded kernels.  No reference counterpart (the reference only has the
 * piecewise-polynomial tables of potential_eval.h:67-104).
 *
 * Network, for a pair of species pair id p at distance r in [r_lo[p], r_hi[p]]:
 *   x     = (rn, irn) = ((2r - r_lo - r_hi) / (r_hi - r_lo), r_lo / r)
 *   z1    = W1 x + b1 + E1 emb[p]          (64 units; emb[p] has 8 dims, so
 *                                          b1 + E1 emb[p] = b1e[p] is a
 *                                          per-pair bias)
 *   h1    = silu(z1), split into bf16 hi + lo
 *   z2    = W2 h1 + b2                     (64x64, bf16 weights on MFMA,
 *                                          fp32 accumulation: W2 hi + W2 lo)
 *   h2    = silu(z2)
 *   y     = W3[p] . h2 + b3[p]             (one output per pair)
 * and, carried alongside (forward-mode differentiation), the exact
 * derivative yd = dy/dr:
 *   t1 = W1 dx/dr,  g1 = silu'(z1) t1 (bf16 hi + lo),  t2 = W2 g1,
 *   g2 = silu'(z2) t2,  yd = W3[p] . g2.
 * The pair's energy and force are
 *   E     = S(r) (sE sinh(y) - shift[p])
 *   dE/dr = S'(r) (sE sinh(y) - shift[p]) + S(r) sE cosh(y) yd,
 * so the force is the derivative of the energy the pass reports and the
 * dynamics conserve it (the table's force is likewise the derivative of its
 * energy polynomial, potential_eval.h:94-102).  sinh keeps the error
 * roughly relative over the many decades a pair energy spans between its
 * repulsive wall and its tail.  S is the reference's switch
 * (potential.c:69-100) from r_sw[p] to r_hi[p], or 1 when r_sw >= r_hi;
 * shift[p] makes sE sinh(y(r_hi)) - shift[p] the table's energy at r_hi.
 *
 * Wave layout: one wave evaluates a tile of 16 pairs.  Lane l works on pair
 * column n = l % 16 and on the 16 units u(j) = 16 (j / 4) + 4 (l / 16) + j % 4,
 * j = 0..15.  That is exactly the row layout of the 16x16 output blocks of
 * v_mfma_f32_16x16x32_bf16 (lane l holds rows 4 (l / 16) + i of column
 * l % 16), so layer-1 values computed per lane are already the B operands
 * of layer 2 (one for the values, one for the derivatives) and layer-2
 * results are already in place for the head: no LDS traffic in the forward
 * pass.  The K order of an MFMA only has to agree between A and B; the
 * weights are stored pre-permuted to the order
 * k(kb, g, e) = 16 (2 kb + e / 4) + 4 g + e % 4 that the registers carry.
 */
#ifndef MLP_TILE_H
#define MLP_TILE_H

#include <hip/hip_runtime.h>

#define MLP_H 64       /* hidden width */
#define MLP_EMB 8      /* species-pair embedding */
#define MLP_TILE 16    /* pairs per MFMA tile */

typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

/* Parameter vector theta (fp32 master copy), np = number of species pairs. */
#define MLP_OFF_W1 0                         /* [64][2]  */
#define MLP_OFF_B1 (MLP_OFF_W1 + 2 * MLP_H)  /* [64]     */
#define MLP_OFF_E1 (MLP_OFF_B1 + MLP_H)      /* [64][8]  */
#define MLP_OFF_W2 (MLP_OFF_E1 + MLP_H * MLP_EMB) /* [64 out][64 in] */
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)   /* [64]     */
#define MLP_OFF_EMB (MLP_OFF_B2 + MLP_H)          /* [np][8]  */
#define MLP_OFF_W3(np) (MLP_OFF_EMB + MLP_EMB * (np))   /* [np][64] */
#define MLP_OFF_B3(np) (MLP_OFF_W3(np) + MLP_H * (np))  /* [np]     */
#define MLP_NPARAMS(np) (MLP_OFF_B3(np) + (np))

/* Device view of a published network (derived from theta after each update). */
struct mlp_net {
    const mlp_bf16x8 *w2f;   /* forward A operands  [4 mb][2 kb][64 lanes]  */
    const mlp_bf16x8 *w2b;   /* backward A operands [4 mb][2 kb][64 lanes]  */
    const float *w1;         /* [64][2]                                     */
    const float *b1e;        /* [np][64] = b1 + E1 emb[p]                   */
    const float *b2;         /* [64]                                        */
    const float *w3;         /* [np][64]                                    */
    const float *b3;         /* [np]                                        */
    const float4 *sc;        /* [np] = (r_lo, r_hi, sE, r_sw)               */
    const float *shift;      /* [np]                                        */
    const float *rlo2;       /* [mt*mt] per type pair: sc[p].x^2 (p its pair) or +inf */
};

__device__ __forceinline__ int mlp_unit(int j, int lane) { return 16 * (j >> 2) + 4 * (lane >> 4) + (j & 3); }

__device__ __forceinline__ float mlp_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

/* silu(z) = z s and silu'(z) = s (1 + z (1 - s)) = s + silu(z) (1 - s),
   s = sigmoid(z), from one exponential and one reciprocal. */
__device__ __forceinline__ void mlp_silu2(float z, float &h, float &d) {
    const float s = mlp_sigmoid(z);
    h = z * s;
    d = s + h * (1.0f - s);
}

/* silu''(z) = s (1 - s) (2 + z (1 - 2 s)) */
__device__ __forceinline__ void mlp_silu3(float z, float &h, float &d, float &dd) {
    const float s = mlp_sigmoid(z);
    h = z * s;
    d = s + h * (1.0f - s);
    dd = s * (1.0f - s) * (2.0f + z * (1.0f - 2.0f * s));
}

/* Network inputs of one pair and their r-derivatives. */
__device__ __forceinline__ void mlp_inputs(const float4 sc, float r, float &rn, float &irn, float &drn, float &dirn) {
    const float w = sc.y - sc.x;
    rn = (2.0f * r - sc.x - sc.y) / w;
    irn = sc.x / r;
    drn = 2.0f / w;
    dirn = -irn / r;
}

/* Forward pass of a 16-pair tile with its r-derivative.  Every lane passes
   the inputs (mlp_inputs) and p of pair l % 16 (lanes l, l+16, l+32, l+48
   pass the same pair).  Results: y and yd = dy/dr of that pair in every
   lane.  w1 points at layer 1's weights [64][2] (N.w1 or a copy).  With
   KEEP, the lane's z1/t1/z2/t2 and the hi halves of the bf16
   h1/g1 operands are kept for the backward pass (its dW2 product takes the
   hi halves); without it (inference) each 16-unit block goes layer 2 ->
   SiLU -> head as soon as its MFMAs are done, so few registers stay live.  With
   REG, layer 2's A operands and biases come from registers the caller
   loaded once (mlp_load_layer2) instead of memory.  All variants do the
   same arithmetic in the same order, so they give the same bits. */
template <bool KEEP, bool REG>
__device__ __forceinline__ void mlp_forward16_in(const mlp_net &N, const float *w1, const mlp_bf16x8 *w2r,
                                                 const float4 *b2r, float rn, float irn, float drn, float dirn, int p,
                                                 int lane, float *z1,
                                                 float *t1, mlp_bf16x8 *h1b, mlp_bf16x8 *g1b, float *z2, float *t2,
                                                 float &y, float &yd) {
    const float *b1e = N.b1e + (size_t)p * MLP_H;
    mlp_bf16x8 hb[2], hl[2], gb[2], gl[2];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int u = mlp_unit(j, lane);
        const float2 w = *(const float2 *)(w1 + 2 * u);
        const float z = fmaf(w.y, irn, fmaf(w.x, rn, b1e[u]));
        const float t = fmaf(w.y, dirn, w.x * drn);
        if (KEEP) {
            z1[j] = z;
            t1[j] = t;
        }
        float h, d;
        mlp_silu2(z, h, d);
        const float g = d * t;
        /* each layer-2 input as a pair of bf16 values hi + lo (16 bits of
           mantissa): the network is then a smooth function of r to FP32
           rounding, and its force the derivative of its energy */
        const __bf16 hh = (__bf16)h, gh = (__bf16)g;
        hb[j >> 3][j & 7] = hh;
        hl[j >> 3][j & 7] = (__bf16)(h - (float)hh);
        gb[j >> 3][j & 7] = gh;
        gl[j >> 3][j & 7] = (__bf16)(g - (float)gh);
    }
    if (KEEP) {
        h1b[0] = hb[0];
        h1b[1] = hb[1];
        g1b[0] = gb[0];
        g1b[1] = gb[1];
    }
    const float *w3 = N.w3 + (size_t)p * MLP_H;
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        const mlp_bf16x8 a0 = REG ? w2r[mb * 2 + 0] : N.w2f[(mb * 2 + 0) * 64 + lane];
        const mlp_bf16x8 a1 = REG ? w2r[mb * 2 + 1] : N.w2f[(mb * 2 + 1) * 64 + lane];
        const int u0 = 16 * mb + 4 * (lane >> 4);
        const float4 bb = REG ? b2r[mb] : *(const float4 *)(N.b2 + u0);
        /* the biases start the accumulation */
        mlp_f4 acc = {bb.x, bb.y, bb.z, bb.w}, act = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, hb[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, hb[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, hl[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, hl[1], acc, 0, 0, 0);
        act = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, gb[0], act, 0, 0, 0);
        act = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, gb[1], act, 0, 0, 0);
        act = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, gl[0], act, 0, 0, 0);
        act = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, gl[1], act, 0, 0, 0);
        const float4 wh = *(const float4 *)(w3 + u0);
        const float zz[4] = {acc[0], acc[1], acc[2], acc[3]};
        const float ww[4] = {wh.x, wh.y, wh.z, wh.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float h, d;
            mlp_silu2(zz[i], h, d);
            if (KEEP) {
                z2[4 * mb + i] = zz[i];
                t2[4 * mb + i] = act[i];
            }
            s0 += h * ww[i];
            s1 += d * act[i] * ww[i];
        }
    }
    s0 += __shfl_xor(s0, 16);
    s1 += __shfl_xor(s1, 16);
    s0 += __shfl_xor(s0, 32);
    s1 += __shfl_xor(s1, 32);
    /* REG (the pair pass): the caller adds b3[p], once per pair */
    y = REG ? s0 : s0 + N.b3[p];
    yd = s1;
}

/* The same from the pair distance: the inputs of pair p at r first. */
template <bool KEEP, bool REG>
__device__ __forceinline__ void mlp_forward16_t(const mlp_net &N, const mlp_bf16x8 *w2r, const float4 *b2r, float r,
                                                int p, int lane, float *z1, float *t1, mlp_bf16x8 *h1b,
                                                mlp_bf16x8 *g1b, float *z2, float *t2, float &y, float &yd) {
    float rn, irn, drn, dirn;
    mlp_inputs(N.sc[p], r, rn, irn, drn, dirn);
    mlp_forward16_in<KEEP, REG>(N, N.w1, w2r, b2r, rn, irn, drn, dirn, p, lane, z1, t1, h1b, g1b, z2, t2, y, yd);
}

/* Training forward: keeps the activations of this lane's 16 units. */
__device__ __forceinline__ void mlp_forward16(const mlp_net &N, float r, int p, int lane, float z1[16], float t1[16],
                                              mlp_bf16x8 h1b[2], mlp_bf16x8 g1b[2], float z2[16], float t2[16],
                                              float &y, float &yd) {
    mlp_forward16_t<true, false>(N, nullptr, nullptr, r, p, lane, z1, t1, h1b, g1b, z2, t2, y, yd);
}

/* Inference forward. */
__device__ __forceinline__ void mlp_infer16(const mlp_net &N, float r, int p, int lane, float &y, float &yd) {
    mlp_forward16_t<false, false>(N, nullptr, nullptr, r, p, lane, nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, y, yd);
}

/* Layer 2's operands of this lane, for a run of tiles: w2r[mb * 2 + kb]
   = A operand (mb, kb), b2r[mb] = biases of the lane's 4 rows of block mb
   (32 + 16 VGPRs; no reload per tile). */
__device__ __forceinline__ void mlp_load_layer2(const mlp_net &N, int lane, mlp_bf16x8 w2r[8], float4 b2r[4]) {
#pragma unroll
    for (int i = 0; i < 8; i++) w2r[i] = N.w2f[i * 64 + lane];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) b2r[mb] = *(const float4 *)(N.b2 + 16 * mb + 4 * (lane >> 4));
}

/* Inference forward with layer 2's operands from mlp_load_layer2, the
   pair's inputs (mlp_inputs) computed by the caller and layer 1's weights
   w1 (N.w1 or a copy of it, e.g. in LDS).  y comes without the output bias:
   y + N.b3[p] is the network's output. */
__device__ __forceinline__ void mlp_infer16_reg(const mlp_net &N, const float *w1, const mlp_bf16x8 w2r[8],
                                                const float4 b2r[4], float rn, float irn, float drn, float dirn, int p,
                                                int lane, float &y, float &yd) {
    mlp_forward16_in<false, true>(N, w1, w2r, b2r, rn, irn, drn, dirn, p, lane, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, y, yd);
}

/* Energy and dE/dr of one pair from the network outputs y, yd (see the
   header: switch S from sc.w to sc.y, shift). */
__device__ __forceinline__ void mlp_energy_force(const float4 sc, float shift, float r, float y, float yd, float &e,
                                                 float &dedr) {
    const float a = __expf(y), ia = __builtin_amdgcn_rcpf(a);
    const float eu = sc.z * (0.5f * (a - ia)) - shift;
    const float ed = sc.z * (0.5f * (a + ia)) * yd;
    e = eu;
    dedr = ed;
    if (r > sc.w && sc.w < sc.y) {
        if (r < sc.y) {
            const float b2 = sc.y * sc.y, a2 = sc.w * sc.w, r2 = r * r;
            const float q = b2 - r2, m = b2 + 2.0f * r2 - 3.0f * a2;
            const float den = b2 - a2, iden = 1.0f / (den * den * den);
            const float S = q * q * m * iden, dS = 4.0f * r * q * (q - m) * iden;
            e = S * eu;
            dedr = dS * eu + S * ed;
        } else {
            e = 0.0f;
            dedr = 0.0f;
        }
    }
}

#endif /* MLP_TILE_H */
