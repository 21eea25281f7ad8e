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
 *   h1    = silu(z1), rounded to bf16
 *   z2    = W2 h1 + b2                     (64x64, bf16 weights on MFMA,
 *                                          fp32 accumulation)
 *   h2    = silu(z2)
 *   y_o   = W3[p][o] . h2 + b3[p][o]       (o = 0, 1: per-pair heads)
 *   E     = sE[p] sinh(y_0),  dE/dr = sF[p] sinh(y_1)
 * The sinh output keeps the error roughly relative over the many decades a
 * pair energy spans between its repulsive wall and its tail; training
 * targets are asinh(E / sE) and asinh(dE/dr / sF).
 *
 * Wave layout: one wave evaluates a tile of 16 pairs.  Lane l works on pair
 * column n = l % 16 and on the 16 units u(j) = 16 (j / 4) + 4 (l / 16) + j % 4,
 * j = 0..15.  That is exactly the row layout of the 16x16 output blocks of
 * v_mfma_f32_16x16x32_bf16 (lane l holds rows 4 (l / 16) + i of column
 * l % 16), so layer-1 values computed per lane are already the B operand of
 * layer 2 and layer-2 results are already in place for the heads: no LDS
 * traffic in the forward pass.  The K order of an MFMA only has to agree
 * between A and B; the weights are stored pre-permuted to the order
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
#define MLP_OFF_W3(np) (MLP_OFF_EMB + MLP_EMB * (np))       /* [np][2][64] */
#define MLP_OFF_B3(np) (MLP_OFF_W3(np) + 2 * MLP_H * (np))  /* [np][2]     */
#define MLP_NPARAMS(np) (MLP_OFF_B3(np) + 2 * (np))

/* Device view of a published network (derived from theta after each update). */
struct mlp_net {
    const mlp_bf16x8 *w2f;   /* forward A operands  [4 mb][2 kb][64 lanes]  */
    const mlp_bf16x8 *w2b;   /* backward A operands [4 mb][2 kb][64 lanes]  */
    const float *w1;         /* [64][2]                                     */
    const float *b1e;        /* [np][64] = b1 + E1 emb[p]                   */
    const float *b2;         /* [64]                                        */
    const float *w3;         /* [np][2][64]                                 */
    const float *b3;         /* [np][2]                                     */
    const float4 *sc;        /* [np] = (r_lo, r_hi, sE, sF)                 */
};

__device__ __forceinline__ int mlp_unit(int j, int lane) { return 16 * (j >> 2) + 4 * (lane >> 4) + (j & 3); }

__device__ __forceinline__ float mlp_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ float mlp_silu(float z) { return z * mlp_sigmoid(z); }

/* d silu / dz = s (1 + z (1 - s)), s = sigmoid(z) */
__device__ __forceinline__ float mlp_dsilu(float z) {
    const float s = mlp_sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

/* Network inputs of one pair. */
__device__ __forceinline__ void mlp_inputs(const float4 sc, float r, float &rn, float &irn) {
    rn = (2.0f * r - sc.x - sc.y) / (sc.y - sc.x);
    irn = sc.x / r;
}

/* Forward pass of a 16-pair tile.  Every lane passes the (r, p) of pair
   l % 16 (lanes l, l+16, l+32, l+48 pass the same pair).  Results: y0, y1
   of that pair in every lane.  With KEEP, the lane's z1/z2/h2 and the bf16
   h1 operands are kept for the backward pass; without it (inference) each
   16-unit block goes layer 2 -> SiLU -> heads as soon as its MFMA is done,
   so few registers stay live.  With REG, layer 2's A operands and biases
   come from registers the caller loaded once (mlp_load_layer2) instead of
   memory.  All variants do the same arithmetic in the same order, so they
   give the same bits. */
template <bool KEEP, bool REG>
__device__ __forceinline__ void mlp_forward16_t(const mlp_net &N, const mlp_bf16x8 *w2r, const float4 *b2r, float r,
                                                int p, int lane, float *z1, mlp_bf16x8 *h1b, float *z2, float *h2,
                                                float &y0, float &y1) {
    const float4 sc = N.sc[p];
    float rn, irn;
    mlp_inputs(sc, r, rn, irn);
    const float *b1e = N.b1e + (size_t)p * MLP_H;
    mlp_bf16x8 hb[2];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int u = mlp_unit(j, lane);
        const float2 w = *(const float2 *)(N.w1 + 2 * u);
        const float z = w.x * rn + w.y * irn + b1e[u];
        if (KEEP) z1[j] = z;
        hb[j >> 3][j & 7] = (__bf16)mlp_silu(z);
    }
    if (KEEP) {
        h1b[0] = hb[0];
        h1b[1] = hb[1];
    }
    const float *w3 = N.w3 + (size_t)p * 2 * MLP_H;
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(REG ? w2r[mb * 2 + 0] : N.w2f[(mb * 2 + 0) * 64 + lane], hb[0],
                                                      acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(REG ? w2r[mb * 2 + 1] : N.w2f[(mb * 2 + 1) * 64 + lane], hb[1],
                                                      acc, 0, 0, 0);
        const int u0 = 16 * mb + 4 * (lane >> 4);
        const float4 bb = REG ? b2r[mb] : *(const float4 *)(N.b2 + u0);
        float hh[4];
        hh[0] = mlp_silu(acc[0] + bb.x);
        hh[1] = mlp_silu(acc[1] + bb.y);
        hh[2] = mlp_silu(acc[2] + bb.z);
        hh[3] = mlp_silu(acc[3] + bb.w);
        if (KEEP) {
            z2[4 * mb] = acc[0] + bb.x;
            z2[4 * mb + 1] = acc[1] + bb.y;
            z2[4 * mb + 2] = acc[2] + bb.z;
            z2[4 * mb + 3] = acc[3] + bb.w;
#pragma unroll
            for (int i = 0; i < 4; i++) h2[4 * mb + i] = hh[i];
        }
        const float4 a = *(const float4 *)(w3 + u0);
        const float4 b = *(const float4 *)(w3 + MLP_H + u0);
        s0 += hh[0] * a.x + hh[1] * a.y + hh[2] * a.z + hh[3] * a.w;
        s1 += hh[0] * b.x + hh[1] * b.y + hh[2] * b.z + hh[3] * b.w;
    }
    s0 += __shfl_xor(s0, 16);
    s1 += __shfl_xor(s1, 16);
    s0 += __shfl_xor(s0, 32);
    s1 += __shfl_xor(s1, 32);
    y0 = s0 + N.b3[2 * p];
    y1 = s1 + N.b3[2 * p + 1];
}

/* Training forward: keeps the activations of this lane's 16 units. */
__device__ __forceinline__ void mlp_forward16(const mlp_net &N, float r, int p, int lane, float z1[16],
                                              mlp_bf16x8 h1b[2], float z2[16], float h2[16], float &y0, float &y1) {
    mlp_forward16_t<true, false>(N, nullptr, nullptr, r, p, lane, z1, h1b, z2, h2, y0, y1);
}

/* Inference forward. */
__device__ __forceinline__ void mlp_infer16(const mlp_net &N, float r, int p, int lane, float &y0, float &y1) {
    mlp_forward16_t<false, false>(N, nullptr, nullptr, r, p, lane, nullptr, nullptr, nullptr, nullptr, y0, y1);
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

/* Inference forward with layer 2's operands from mlp_load_layer2. */
__device__ __forceinline__ void mlp_infer16_reg(const mlp_net &N, const mlp_bf16x8 w2r[8], const float4 b2r[4], float r,
                                                int p, int lane, float &y0, float &y1) {
    mlp_forward16_t<false, true>(N, w2r, b2r, r, p, lane, nullptr, nullptr, nullptr, nullptr, y0, y1);
}

__device__ __forceinline__ float mlp_sinh(float y) {
    const float a = __expf(y);
    return 0.5f * (a - __builtin_amdgcn_rcpf(a));
}

/* Energy and (dE/dr)/r of one pair from the network outputs. */
__device__ __forceinline__ void mlp_energy_force(const float4 sc, float r, float y0, float y1, float &e, float &f) {
    e = sc.z * mlp_sinh(y0);
    f = sc.w * mlp_sinh(y1) / r;
}

#endif /* MLP_TILE_H */
