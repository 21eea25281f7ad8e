/*
 * MLP surrogate of the interpolated pair potential (no reference
 * counterpart; the reference evaluates only the tables of
 * potential_eval.h:67-104, which stay the measure of its error and the fallback here).
 *
 * Training runs on the GPU with this file's own kernels:
 *   k_mlp_grad     forward (values and r-derivatives) + backward of tiles
 *                  of 16 samples of one species pair: the loss holds the
 *                  derivative output yd, so the backward pass runs through
 *                  the tangent path as well (second derivatives of SiLU).
 *                  Layer 2 and its backward products (dH1 = W2^T dZ2,
 *                  dG1 = W2^T dT2, dW2 = dZ2 H1^T + dT2 G1^T) run on
 *                  v_mfma_f32_16x16x32_bf16 with bf16 operands and fp32
 *                  accumulation; the rest is fp32 VALU.
 *   k_mlp_wsum     fixed-order sum of the per-wave partial rows of the
 *                  shared weights (W2, W1, b2);
 *   k_mlp_tsum     fixed-order sum of the per-tile rows of the per-pair
 *                  parameters (heads W3, b3, the bias gradient of b1e) and
 *                  of the loss, tiles in batch order;
 *   k_mlp_chain    gradient of b1, E1 and the embeddings from the gradient
 *                  of the per-pair bias b1e = b1 + E1 emb[p];
 *   k_mlp_adam     Adam on the fp32 master parameters;
 *   k_mlp_publish  derived device operands: b1e and the bf16 W2 operands in
 *                  the lane order of the MFMA tiles.
 * Every gradient sum runs in a fixed order (no float atomics), so a
 * gradient, and a whole training run, gives the same bits every time.
 * k_mlp_eval evaluates the network on arbitrary (r, pair) samples.
 * The forward pass and the lane layout are in mlp_tile.h.
 *
 * engine_mlp_grad_noalloc is engine_mlp_grad on device buffers with its
 * scratch in the file-scope array `scratch`: it allocates nothing.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gpu_state.h"

#define MLP_WPART (MLP_H * MLP_H + 3 * MLP_H) /* per-wave partial: W2, W1, b2 */
/* Waves of a gradient launch (both entry points, so that they sum the same
   partial rows in the same order) */
#define MLP_GRID_MAX 256
/* Per-tile row: [0,64) dW3 of the tile's pair, [64] db3, [65,129) gradient
   of the pair's b1e, [129] the tile's loss (times 1/n), [130,132) padding. */
#define MLP_TROW 132
#define MLP_TR_B3 64
#define MLP_TR_B1E 65
#define MLP_TR_LOSS 129
#define MLP_TCHUNK 4096 /* tiles per gradient launch (rows of the per-tile buffer) */

/* Scratch of engine_mlp_grad_noalloc, one copy per device for the life of
   the process (see scratch_init): the bad-pair flag, the tile list, the
   b1e gradient, the per-wave and the per-tile partial rows.  At most
   MLP_NA_TILES tiles of 16 samples, MLP_NA_PAIRS species pairs and
   MLP_NA_GRID waves per gradient launch: 6.8 MB. */
#define MLP_NA_TILES 4096
#define MLP_NA_PAIRS 1024
#define MLP_NA_GRID MLP_GRID_MAX
#define MLP_NA_MAXDEV 64
#define NA_OFF_ERR 0
#define NA_OFF_TILES 64
#define NA_OFF_GB1E (NA_OFF_TILES + MLP_NA_TILES)
#define NA_OFF_WPART (NA_OFF_GB1E + MLP_NA_PAIRS * MLP_H)
#define NA_OFF_TPART (NA_OFF_WPART + MLP_NA_GRID * MLP_WPART)
#define NA_SIZE (NA_OFF_TPART + MLP_NA_TILES * MLP_TROW)
static __device__ __align__(256) float scratch[NA_SIZE];

/* Device address of each device's copy of scratch (the array's name is a
   host-side stand-in and must never be passed to a kernel). */
static float *scratch_ptr[MLP_NA_MAXDEV];

/* Look up this device's copy of scratch once, with the device current. */
static int scratch_init(int dev) {
    if (dev < 0 || dev >= MLP_NA_MAXDEV) return gerror_(engine_err_range, "device id beyond the scratch table");
    if (scratch_ptr[dev] != NULL) return engine_err_ok;
    void *ptr = NULL;
    HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    scratch_ptr[dev] = (float *)ptr;
    return engine_err_ok;
}

/* ------------------------------------------------------------ kernels */

__global__ __launch_bounds__(256) void k_mlp_fill(float *__restrict__ a, long long n, float v) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        a[k] = v;
}

__global__ __launch_bounds__(256) void k_mlp_iota(int *__restrict__ a, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) a[k] = k;
}

/* K index carried by element e of lane group g in K-block kb. */
__device__ __forceinline__ int mlp_kperm(int kb, int g, int e) { return 16 * (2 * kb + (e >> 2)) + 4 * g + (e & 3); }

__global__ __launch_bounds__(256) void k_mlp_publish(const float *__restrict__ theta, int np, float *__restrict__ b1e,
                                                     mlp_bf16x8 *__restrict__ w2f, mlp_bf16x8 *__restrict__ w2b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 512) {
        const int blk = t >> 6, l = t & 63, mb = blk >> 1, kb = blk & 1, g = l >> 4;
        const int row = 16 * mb + (l & 15);
        mlp_bf16x8 a, b;
        for (int e = 0; e < 8; e++) {
            const int k = mlp_kperm(kb, g, e);
            a[e] = (__bf16)theta[MLP_OFF_W2 + row * MLP_H + k]; /* W2[out = row][in = k]   */
            b[e] = (__bf16)theta[MLP_OFF_W2 + k * MLP_H + row]; /* W2^T: [in = row][out = k] */
        }
        w2f[t] = a;
        w2b[t] = b;
    }
    if (t < np * MLP_H) {
        const int p = t / MLP_H, u = t % MLP_H;
        float acc = theta[MLP_OFF_B1 + u];
        for (int e = 0; e < MLP_EMB; e++) acc += theta[MLP_OFF_E1 + u * MLP_EMB + e] * theta[MLP_OFF_EMB + p * MLP_EMB + e];
        b1e[t] = acc;
    }
}

/* Sum over the 16 pair columns of a lane group. */
__device__ __forceinline__ float mlp_sum16(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

/* One wave per block; each block walks tiles of 16 samples s = 16 t + n,
   n = 0..15, which all share species pair sp[16 t].  Loss per sample:
   wE (y - a)^2 + wF (yd - ad)^2 / (ad^2 + 1/c^2), a = asinh(E/sE),
   ad = F/sqrt(sE^2 + E^2) its r-derivative for the target's dE/dr = F,
   c = (r_hi - r_lo)/2: the derivative's error relative where |ad| is large
   (the repulsive wall), in units of 1/c where it is small; gradients are of
   the sum times inv_n.  Backward, per sample:
     dy, dyd            -> dW3 = dy h2 + dyd g2, db3 = dy,
     dz2 = w3 (dy silu'(z2) + dyd t2 silu''(z2)),  dt2 = w3 dyd silu'(z2),
     dh1 = W2^T dz2,  dg1 = W2^T dt2                       (MFMA),
     dz1 = dh1 silu'(z1) + dg1 t1 silu''(z1),  dt1 = dg1 silu'(z1),
     dW1 = dz1 x^T + dt1 (dx/dr)^T,  db1e = dz1,
     dW2 = dz2 h1^T + dt2 g1^T: one MFMA K of 32 per tile, the 16 value
     columns and the 16 derivative columns (through LDS).
   Sums over the shared weights go to this wave's partial row, those of the
   tile's pair to the tile's row of tpart (index ti of tiles[]).  A pair id
   outside [0, np) is read as pair 0 and sets *err. */
__global__ __launch_bounds__(64) void k_mlp_grad(struct mlp_net N, const float *__restrict__ sr,
                                                 const int *__restrict__ sp, const float *__restrict__ tE,
                                                 const float *__restrict__ tF, const int *__restrict__ tiles,
                                                 int ntiles, float wE, float wF, float inv_n, int np,
                                                 float *__restrict__ tpart, float *__restrict__ partial,
                                                 float *__restrict__ err) {
    __shared__ __attribute__((aligned(16))) __bf16 h1t[MLP_H][32]; /* [in unit][h1 of 16 samples, g1 of 16]   */
    __shared__ __attribute__((aligned(16))) __bf16 dzt[MLP_H][32]; /* [out unit][dz2 of 16 samples, dt2 of 16] */
    const int lane = threadIdx.x, n = lane & 15, g = lane >> 4;
    mlp_f4 dw2[16];
    float dw1a[16], dw1b[16], db2[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        dw2[q] = (mlp_f4){0.f, 0.f, 0.f, 0.f};
        dw1a[q] = dw1b[q] = db2[q] = 0.f;
    }
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const int tile = tiles[ti];
        const int s = 16 * tile + n;
        int p = sp[16 * tile];
        if ((unsigned)p >= (unsigned)np) {
            if (lane == 0) *err = 1.0f;
            p = 0;
        }
        float *trow = tpart + (size_t)ti * MLP_TROW;
        const float r = sr[s];
        float z1[16], t1[16], z2[16], t2[16], y, yd;
        mlp_bf16x8 h1b[2], g1b[2];
        mlp_forward16(N, r, p, lane, z1, t1, h1b, g1b, z2, t2, y, yd);
        const float4 sc = N.sc[p];
        const float E = tE[s], F = tF[s];
        const float c = 0.5f * (sc.y - sc.x);
        const float ad = F / sqrtf(sc.z * sc.z + E * E);
        const float d0 = y - asinhf(E / sc.z);
        const float d1 = yd - ad;
        const float q = 1.0f / (ad * ad + 1.0f / (c * c));
        {
            const float lt = mlp_sum16(g == 0 ? wE * d0 * d0 + wF * q * d1 * d1 : 0.f);
            if (lane == 0) trow[MLP_TR_LOSS] = lt * inv_n;
        }
        const float dy = 2.0f * wE * d0 * inv_n;
        const float dyd = 2.0f * wF * q * d1 * inv_n;

        /* Head: dW3[p], db3[p]; dZ2, dT2. */
        const float *w3 = N.w3 + (size_t)p * MLP_H;
        float dz2[16], dt2[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int u = mlp_unit(j, lane);
            const float w = w3[u];
            float h, d, dd;
            mlp_silu3(z2[j], h, d, dd);
            dz2[j] = w * (dy * d + dyd * t2[j] * dd);
            dt2[j] = w * dyd * d;
            db2[j] += dz2[j];
            const float gw = mlp_sum16(dy * h + dyd * d * t2[j]);
            if (n == 0) trow[u] = gw;
        }
        {
            const float gb = mlp_sum16(dy);
            if (lane == 0) trow[MLP_TR_B3] = gb;
        }

        /* dH1 = W2^T dZ2 and dG1 = W2^T dT2 on MFMA; dZ1, dT1. */
        mlp_bf16x8 dzb[2], dtb[2];
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                dzb[kb][e] = (__bf16)dz2[8 * kb + e];
                dtb[kb][e] = (__bf16)dt2[8 * kb + e];
            }
        float rn, irn, drn, dirn;
        mlp_inputs(sc, r, rn, irn, drn, dirn);
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const mlp_bf16x8 a0 = N.w2b[(mb * 2 + 0) * 64 + lane], a1 = N.w2b[(mb * 2 + 1) * 64 + lane];
            mlp_f4 acc = {0.f, 0.f, 0.f, 0.f}, act = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, dzb[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, dzb[1], acc, 0, 0, 0);
            act = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, dtb[0], act, 0, 0, 0);
            act = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, dtb[1], act, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int j = 4 * mb + i;
                float h, d, dd;
                mlp_silu3(z1[j], h, d, dd);
                const float dz1 = acc[i] * d + act[i] * t1[j] * dd;
                const float dt1 = act[i] * d;
                dw1a[j] += dz1 * rn + dt1 * drn;
                dw1b[j] += dz1 * irn + dt1 * dirn;
                const float gsum = mlp_sum16(dz1);
                if (n == 0) trow[MLP_TR_B1E + mlp_unit(j, lane)] = gsum;
            }
        }

        /* Stage H1|G1 and dZ2|dT2 (bf16) for dW2. */
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int u = mlp_unit(j, lane);
            h1t[u][n] = h1b[j >> 3][j & 7];
            h1t[u][16 + n] = g1b[j >> 3][j & 7];
            dzt[u][n] = dzb[j >> 3][j & 7];
            dzt[u][16 + n] = dtb[j >> 3][j & 7];
        }
        __syncthreads();
        /* dW2[m][k] += sum over the 32 columns of dZ2|dT2[m] H1|G1[k]. */
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const mlp_bf16x8 a = *(const mlp_bf16x8 *)&dzt[16 * mb + n][8 * g];
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                const mlp_bf16x8 b = *(const mlp_bf16x8 *)&h1t[16 * kb + n][8 * g];
                dw2[mb * 4 + kb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, dw2[mb * 4 + kb], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    /* This wave's sums of the shared weights go to its own partial row
       (summed over waves in a fixed order by k_mlp_wsum): W2, then W1, b2. */
    float *row = partial + (size_t)blockIdx.x * MLP_WPART;
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
            for (int i = 0; i < 4; i++) row[(16 * mb + 4 * g + i) * MLP_H + 16 * kb + n] = dw2[mb * 4 + kb][i];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int u = mlp_unit(j, lane);
        const float a = mlp_sum16(dw1a[j]), b = mlp_sum16(dw1b[j]), c = mlp_sum16(db2[j]);
        if (n == 0) {
            row[MLP_H * MLP_H + 2 * u] = a;
            row[MLP_H * MLP_H + 2 * u + 1] = b;
            row[MLP_H * MLP_H + 2 * MLP_H + u] = c;
        }
    }
}

/* Sum the per-wave partial rows of k_mlp_grad into the gradient.  Block b
   takes columns [64 b, 64 b + 64); its wave s sums rows s, s + MLP_WSUM_SPLIT,
   ... in order, and the MLP_WSUM_SPLIT sums are then added in wave order, so
   every run gives the same bits.  (One thread per column summing all rows
   left 17 blocks each waiting on ~500 dependent loads: 113 us per step.) */
#define MLP_WSUM_SPLIT 16
__global__ __launch_bounds__(64 * MLP_WSUM_SPLIT) void k_mlp_wsum(const float *__restrict__ partial, int nwaves,
                                                                  float *__restrict__ grad) {
    __shared__ float part[MLP_WSUM_SPLIT][64];
    const int col = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int t = blockIdx.x * 64 + col;
    float acc = 0.f;
    if (t < MLP_WPART)
        for (int w = s; w < nwaves; w += MLP_WSUM_SPLIT) acc += partial[(size_t)w * MLP_WPART + t];
    part[s][col] = acc;
    __syncthreads();
    if (s != 0 || t >= MLP_WPART) return;
    acc = 0.f;
    for (int k = 0; k < MLP_WSUM_SPLIT; k++) acc += part[k][col];
    if (t < MLP_H * MLP_H)
        grad[MLP_OFF_W2 + t] += acc;
    else if (t < MLP_H * MLP_H + 2 * MLP_H)
        grad[MLP_OFF_W1 + t - MLP_H * MLP_H] += acc;
    else
        grad[MLP_OFF_B2 + t - MLP_H * MLP_H - 2 * MLP_H] += acc;
}

/* Add the per-tile rows of k_mlp_grad: block p < np sums the rows of the
   tiles of pair p, in tile order, into the gradients of its heads and
   b1e; block np sums the tiles' losses (thread k takes tiles k, k + 256,
   ... in order, then a fixed tree).  A pair id outside [0, np) marks no
   tile; with *err set the loss becomes NaN. */
__global__ __launch_bounds__(256) void k_mlp_tsum(const float *__restrict__ tpart, const int *__restrict__ tiles,
                                                  int ntiles, const int *__restrict__ sp, int np,
                                                  float *__restrict__ grad, float *__restrict__ gb1e,
                                                  float *__restrict__ loss, const float *__restrict__ err) {
    __shared__ unsigned char mine[MLP_TCHUNK];
    __shared__ float red[256];
    const int p = blockIdx.x, c = threadIdx.x;
    if (p < np) {
        for (int t = c; t < ntiles; t += 256) mine[t] = sp[16 * tiles[t]] == p;
        __syncthreads();
        if (c < MLP_TR_LOSS) {
            float acc = 0.f;
            for (int t = 0; t < ntiles; t++)
                if (mine[t]) acc += tpart[(size_t)t * MLP_TROW + c];
            if (c < MLP_TR_B3)
                grad[MLP_OFF_W3(np) + p * MLP_H + c] += acc;
            else if (c == MLP_TR_B3)
                grad[MLP_OFF_B3(np) + p] += acc;
            else
                gb1e[p * MLP_H + c - MLP_TR_B1E] += acc;
        }
    } else {
        float acc = 0.f;
        for (int t = c; t < ntiles; t += 256) acc += tpart[(size_t)t * MLP_TROW + MLP_TR_LOSS];
        red[c] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (c < w) red[c] += red[c + w];
            __syncthreads();
        }
        if (c == 0) *loss = *err != 0.f ? __builtin_nanf("") : *loss + red[0];
    }
}

/* b1e[p][u] = b1[u] + sum_e E1[u][e] emb[p][e]: gradients of b1, E1, emb,
   each a fixed-order sum. */
__global__ __launch_bounds__(64 * MLP_WSUM_SPLIT) void k_mlp_chain(const float *__restrict__ theta, int np,
                                                                   const float *__restrict__ gb1e,
                                                                   float *__restrict__ grad) {
    /* Outputs t: [0, 64) b1, then 64 x 8 E1, then np x 8 emb.  Block b takes
       outputs [64 b, 64 b + 64); wave s sums the terms k = s, s + 16, ... of
       each output's sum, and the 16 sums are added in wave order. */
    __shared__ float part[MLP_WSUM_SPLIT][64];
    const int col = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int t = blockIdx.x * 64 + col;
    const int nb1 = MLP_H, ne1 = MLP_H * MLP_EMB, nout = nb1 + ne1 + np * MLP_EMB;
    float acc = 0.f;
    if (t < nb1) {
        for (int p = s; p < np; p += MLP_WSUM_SPLIT) acc += gb1e[p * MLP_H + t];
    } else if (t < nb1 + ne1) {
        const int q = t - nb1, u = q / MLP_EMB, e = q % MLP_EMB;
        for (int p = s; p < np; p += MLP_WSUM_SPLIT) acc += gb1e[p * MLP_H + u] * theta[MLP_OFF_EMB + p * MLP_EMB + e];
    } else if (t < nout) {
        const int q = t - nb1 - ne1, p = q / MLP_EMB, e = q % MLP_EMB;
        for (int u = s; u < MLP_H; u += MLP_WSUM_SPLIT) acc += gb1e[p * MLP_H + u] * theta[MLP_OFF_E1 + u * MLP_EMB + e];
    }
    part[s][col] = acc;
    __syncthreads();
    if (s != 0 || t >= nout) return;
    acc = 0.f;
    for (int k = 0; k < MLP_WSUM_SPLIT; k++) acc += part[k][col];
    if (t < nb1)
        grad[MLP_OFF_B1 + t] += acc;
    else if (t < nb1 + ne1)
        grad[MLP_OFF_E1 + t - nb1] += acc;
    else
        grad[MLP_OFF_EMB + t - nb1 - ne1] += acc;
}

__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ theta, const float *__restrict__ grad,
                                                  float *__restrict__ m, float *__restrict__ v, int n, float lr,
                                                  float b1, float b2, float c1, float c2, float eps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = grad[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    theta[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

__global__ __launch_bounds__(64) void k_mlp_eval(struct mlp_net N, const float *__restrict__ r,
                                                 const int *__restrict__ pair, long long n, float *__restrict__ E,
                                                 float *__restrict__ F) {
    const int lane = threadIdx.x;
    for (long long tile = blockIdx.x; 16 * tile < n; tile += gridDim.x) {
        const long long s = 16 * tile + (lane & 15);
        const long long sv = s < n ? s : n - 1;
        const int p = pair[sv];
        const float rr = r[sv];
        float y, yd;
        mlp_infer16(N, rr, p, lane, y, yd);
        if (lane < 16 && s < n) {
            float e, dedr;
            mlp_energy_force(N.sc[p], N.shift[p], rr, y, yd, e, dedr);
            E[s] = e;
            F[s] = dedr;
        }
    }
}

/* y of every pair at its r_hi (for the energy shifts). */
__global__ __launch_bounds__(64) void k_mlp_rhi(struct mlp_net N, int np, float *__restrict__ yout) {
    const int lane = threadIdx.x;
    const int q = 16 * blockIdx.x + (lane & 15);
    const int p = q < np ? q : np - 1;
    float y, yd;
    mlp_infer16(N, N.sc[p].y, p, lane, y, yd);
    if (lane < 16 && q < np) yout[q] = y;
}

/* ------------------------------------------------------------ host side */

static unsigned long long mlp_rng(unsigned long long *s) {
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double mlp_normal(unsigned long long *s) {
    double u1 = ((mlp_rng(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = ((mlp_rng(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void gpu_mlp_free(struct gpu_state *g) {
    (void)hipFree(g->d_theta);
    (void)hipFree(g->d_grad);
    (void)hipFree(g->d_adam_m);
    (void)hipFree(g->d_adam_v);
    (void)hipFree(g->d_gb1e);
    (void)hipFree(g->d_b1e);
    (void)hipFree(g->d_loss);
    (void)hipFree(g->d_w2f);
    (void)hipFree(g->d_w2b);
    (void)hipFree(g->d_msc);
    (void)hipFree(g->d_mrlo2);
    g->d_mrlo2 = NULL;
    (void)hipFree(g->d_msh);
    (void)hipFree(g->d_mshy);
    free(g->h_msc);
    g->h_msc = NULL;
    g->d_msh = g->d_mshy = NULL;
    (void)hipFree(g->d_wpart);
    (void)hipFree(g->d_tpart);
    (void)hipFree(g->d_err);
    g->d_wpart = g->d_tpart = g->d_err = NULL;
    (void)hipFree(g->d_sr);
    (void)hipFree(g->d_sp);
    (void)hipFree(g->d_sE);
    (void)hipFree(g->d_sF);
    (void)hipFree(g->d_tiles);
    g->d_theta = g->d_grad = g->d_adam_m = g->d_adam_v = g->d_gb1e = g->d_b1e = g->d_loss = NULL;
    g->d_w2f = g->d_w2b = NULL;
    g->d_msc = NULL;
    g->d_sr = g->d_sE = g->d_sF = NULL;
    g->d_sp = g->d_tiles = NULL;
    g->d_samples_cap = 0;
    g->mlp_np = g->mlp_nparams = 0;
    g->mlp_step = 0;
}

static struct gpu_state *mlp_state(struct engine *e, int need_net) {
    struct gpu_state *g;
    if (e == NULL) {
        gerror_(engine_err_null, "NULL engine");
        return NULL;
    }
    g = (struct gpu_state *)e->gpu;
    if (g == NULL) {
        gerror_(engine_err_nogpu, "engine_gpu_load was not called");
        return NULL;
    }
    if (need_net && g->mlp_np == 0) {
        gerror_(engine_err_range, "engine_mlp_init was not called");
        return NULL;
    }
    if (hipSetDevice(g->device) != hipSuccess) {
        gerror_(engine_err_gpu, "hipSetDevice failed");
        return NULL;
    }
    return g;
}

static int mlp_publish(struct gpu_state *g) {
    const int nthr = g->mlp_np * MLP_H > 512 ? g->mlp_np * MLP_H : 512;
    hipLaunchKernelGGL(k_mlp_publish, dim3((nthr + 255) / 256), dim3(256), 0, g->stream, g->d_theta, g->mlp_np,
                       g->d_b1e, g->d_w2f, g->d_w2b);
    HIPCHECK(hipGetLastError());
    return engine_err_ok;
}

/* Energy shifts: shift[p] = sE sinh(y_p(r_hi)) - E_p(r_hi), with E_p the
   FP64 table of species pair p (potential_eval_r), so that the network's
   unswitched energy at r_hi is the table's.  Run after every change of the
   parameters outside a training step. */
static int mlp_update_shift(struct engine *e, struct gpu_state *g) {
    const int np = g->mlp_np;
    float *y = (float *)malloc(sizeof(float) * np), *sh = (float *)malloc(sizeof(float) * np);
    int p, rc = engine_err_ok;
    if (y == NULL || sh == NULL) {
        free(y);
        free(sh);
        return gerror_(engine_err_malloc, "malloc failed");
    }
    hipLaunchKernelGGL(k_mlp_rhi, dim3((np + 15) / 16), dim3(64), 0, g->stream, gpu_mlp_net(g), np, g->d_mshy);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess ||
        hipMemcpy(y, g->d_mshy, sizeof(float) * np, hipMemcpyDeviceToHost) != hipSuccess)
        rc = gerror_(engine_err_gpu, "energy shift evaluation failed");
    for (p = 0; p < np && rc == engine_err_ok; p++) {
        const struct potential *pot = e->p[g->h_pair_ti[p] * g->max_type + g->h_pair_tj[p]];
        double ee = 0.0, ff = 0.0;
        if (pot == NULL) {
            rc = gerror_(engine_err_range, "no table for a species pair");
            break;
        }
        /* the network is evaluated at r_hi rounded to FP32, which may lie
           just past the table's last node b; the table is read at b */
        const double r = (double)g->h_msc[p].y < pot->b ? (double)g->h_msc[p].y : pot->b;
        if (potential_eval_r(pot, &r, 1, &ee, &ff) < 0) rc = gerror_(engine_err_range, "table evaluation failed");
        sh[p] = (float)(g->h_msc[p].z * sinh((double)y[p]) - ee);
    }
    if (rc == engine_err_ok && hipMemcpy(g->d_msh, sh, sizeof(float) * np, hipMemcpyHostToDevice) != hipSuccess)
        rc = gerror_(engine_err_gpu, "energy shift upload failed");
    free(y);
    free(sh);
    return rc;
}

/* Grow the device sample buffers to hold n samples (and n/16 tile ids). */
static int mlp_samples(struct gpu_state *g, long long n) {
    if (n <= g->d_samples_cap) return engine_err_ok;
    (void)hipFree(g->d_sr);
    (void)hipFree(g->d_sp);
    (void)hipFree(g->d_sE);
    (void)hipFree(g->d_sF);
    (void)hipFree(g->d_tiles);
    g->d_samples_cap = 0;
    HIPCHECK(hipMalloc(&g->d_sr, sizeof(float) * n));
    HIPCHECK(hipMalloc(&g->d_sp, sizeof(int) * n));
    HIPCHECK(hipMalloc(&g->d_sE, sizeof(float) * n));
    HIPCHECK(hipMalloc(&g->d_sF, sizeof(float) * n));
    HIPCHECK(hipMalloc(&g->d_tiles, sizeof(int) * (n / 16 + 1)));
    g->d_samples_cap = n;
    return engine_err_ok;
}

extern "C" int engine_mlp_pairs(struct engine *e, int *ti, int *tj, int max) {
    struct gpu_state *g = mlp_state(e, 0);
    int k;
    if (g == NULL) return engine_err;
    for (k = 0; k < g->ntab && k < max; k++) {
        if (ti) ti[k] = g->h_pair_ti[k];
        if (tj) tj[k] = g->h_pair_tj[k];
    }
    return g->ntab;
}

extern "C" int engine_mlp_init(struct engine *e, const double *r_lo, const double *r_hi, const double *escale,
                               const double *rswitch, unsigned long long seed) {
    struct gpu_state *g = mlp_state(e, 0);
    int np, nparams, k;
    float *theta;
    float4 *sc;
    if (g == NULL) return engine_err;
    if (r_lo == NULL || r_hi == NULL || escale == NULL || rswitch == NULL)
        return gerror_(engine_err_null, "NULL pair scales");
    np = g->ntab;
    if (np < 1) return gerror_(engine_err_range, "no potentials");
    for (k = 0; k < np; k++)
        if (!(r_lo[k] > 0.0 && r_hi[k] > r_lo[k] && escale[k] > 0.0 && rswitch[k] > r_lo[k] && isfinite(rswitch[k])))
            return gerror_(engine_err_range, "pair scales must satisfy 0 < r_lo < r_hi, r_lo < r_switch, escale > 0");
    gpu_mlp_free(g);
    nparams = MLP_NPARAMS(np);
    theta = (float *)malloc(sizeof(float) * nparams);
    sc = (float4 *)malloc(sizeof(float4) * np);
    if (theta == NULL || sc == NULL) {
        free(theta);
        free(sc);
        return gerror_(engine_err_malloc, "malloc failed");
    }
    /* Initial weights: normal, scaled by fan-in (SiLU gain for the hidden layers). */
    {
        unsigned long long s = seed;
        for (k = 0; k < nparams; k++) theta[k] = 0.0f;
        for (k = 0; k < 2 * MLP_H; k++) theta[MLP_OFF_W1 + k] = (float)(mlp_normal(&s) * 1.0);
        for (k = 0; k < MLP_H * MLP_EMB; k++) theta[MLP_OFF_E1 + k] = (float)(mlp_normal(&s) * sqrt(1.0 / MLP_EMB));
        for (k = 0; k < MLP_H * MLP_H; k++) theta[MLP_OFF_W2 + k] = (float)(mlp_normal(&s) * sqrt(2.0 / MLP_H));
        for (k = 0; k < MLP_EMB * np; k++) theta[MLP_OFF_EMB + k] = (float)mlp_normal(&s);
        for (k = 0; k < MLP_H * np; k++) theta[MLP_OFF_W3(np) + k] = (float)(mlp_normal(&s) * sqrt(1.0 / MLP_H));
    }
    for (k = 0; k < np; k++) sc[k] = make_float4((float)r_lo[k], (float)r_hi[k], (float)escale[k], (float)rswitch[k]);
    g->mlp_np = np;
    g->mlp_nparams = nparams;
    g->mlp_step = 0;
    if (hipMalloc(&g->d_theta, sizeof(float) * nparams) != hipSuccess ||
        hipMalloc(&g->d_grad, sizeof(float) * nparams) != hipSuccess ||
        hipMalloc(&g->d_adam_m, sizeof(float) * nparams) != hipSuccess ||
        hipMalloc(&g->d_adam_v, sizeof(float) * nparams) != hipSuccess ||
        hipMalloc(&g->d_gb1e, sizeof(float) * np * MLP_H) != hipSuccess ||
        hipMalloc(&g->d_b1e, sizeof(float) * np * MLP_H) != hipSuccess ||
        hipMalloc(&g->d_w2f, sizeof(mlp_bf16x8) * 512) != hipSuccess ||
        hipMalloc(&g->d_w2b, sizeof(mlp_bf16x8) * 512) != hipSuccess ||
        hipMalloc(&g->d_msc, sizeof(float4) * np) != hipSuccess ||
        hipMalloc(&g->d_msh, sizeof(float) * np) != hipSuccess ||
        hipMalloc(&g->d_mshy, sizeof(float) * np) != hipSuccess ||
        hipMalloc(&g->d_wpart, sizeof(float) * MLP_WPART * MLP_GRID_MAX) != hipSuccess ||
        hipMalloc(&g->d_tpart, sizeof(float) * MLP_TROW * MLP_TCHUNK) != hipSuccess ||
        hipMalloc(&g->d_err, sizeof(float)) != hipSuccess ||
        hipMemcpy(g->d_theta, theta, sizeof(float) * nparams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_msc, sc, sizeof(float4) * np, hipMemcpyHostToDevice) != hipSuccess) {
        free(theta);
        free(sc);
        gpu_mlp_free(g);
        return gerror_(engine_err_gpu, "MLP allocation failed");
    }
    {
        /* r_lo^2 by type pair, so that the pair kernels' split between the
           tables and the network is one load; the same FP32 product as
           r_lo * r_lo on the device */
        const int mt = g->max_type;
        float *rl2 = (float *)malloc(sizeof(float) * mt * mt);
        if (rl2 == NULL) {
            free(theta);
            free(sc);
            gpu_mlp_free(g);
            return gerror_(engine_err_malloc, "malloc failed");
        }
        for (k = 0; k < mt * mt; k++) {
            const int p = g->h_tabof[k];
            rl2[k] = p >= 0 ? sc[p].x * sc[p].x : INFINITY;
        }
        const int bad = hipMalloc(&g->d_mrlo2, sizeof(float) * mt * mt) != hipSuccess ||
                        hipMemcpy(g->d_mrlo2, rl2, sizeof(float) * mt * mt, hipMemcpyHostToDevice) != hipSuccess;
        free(rl2);
        if (bad) {
            free(theta);
            free(sc);
            gpu_mlp_free(g);
            return gerror_(engine_err_gpu, "MLP allocation failed");
        }
    }
    free(theta);
    g->h_msc = sc;
    hipLaunchKernelGGL(k_mlp_fill, dim3(64), dim3(256), 0, g->stream, g->d_adam_m, (long long)nparams, 0.0f);
    hipLaunchKernelGGL(k_mlp_fill, dim3(64), dim3(256), 0, g->stream, g->d_adam_v, (long long)nparams, 0.0f);
    hipLaunchKernelGGL(k_mlp_fill, dim3(1), dim3(64), 0, g->stream, g->d_err, 1LL, 0.0f);
    if (mlp_publish(g) < 0) return engine_err_gpu;
    if (scratch_init(g->device) < 0) return engine_err_gpu;
    if (mlp_update_shift(e, g) < 0) return engine_err;
    return nparams;
}

extern "C" int engine_mlp_params(struct engine *e, float *theta, int get) {
    struct gpu_state *g = mlp_state(e, 1);
    if (g == NULL) return engine_err;
    if (theta == NULL) return g->mlp_nparams;
    HIPCHECK(hipStreamSynchronize(g->stream));
    if (get) {
        HIPCHECK(hipMemcpy(theta, g->d_theta, sizeof(float) * g->mlp_nparams, hipMemcpyDeviceToHost));
    } else {
        HIPCHECK(hipMemcpy(g->d_theta, theta, sizeof(float) * g->mlp_nparams, hipMemcpyHostToDevice));
        if (mlp_publish(g) < 0) return engine_err_gpu;
        if (mlp_update_shift(e, g) < 0) return engine_err;
    }
    return g->mlp_nparams;
}

/* Check samples: pair ids in range, r > 0, finite targets. */
static int mlp_check(const struct gpu_state *g, const float *r, const int *pair, long long n) {
    for (long long k = 0; k < n; k++)
        if (pair[k] < 0 || pair[k] >= g->mlp_np || !(r[k] > 0.0f))
            return gerror_(engine_err_range, "sample with pair id out of range or r <= 0");
    return engine_err_ok;
}

/* Buffers of one gradient evaluation: the gradient and loss outputs and the
   scratch (b1e gradient, per-wave and per-tile partial rows, bad-pair flag),
   and the largest grid of the k_mlp_grad launches. */
struct mlp_bufs {
    float *grad, *loss, *gb1e, *wpart, *tpart, *err;
    int grid_max;
};

/* One gradient evaluation over the tiles tiles[0..ntiles) (device array) of
   the samples (device arrays sr, sp, sE, sF): zero the gradients, then per
   chunk of up to MLP_TCHUNK tiles k_mlp_grad, k_mlp_wsum and k_mlp_tsum,
   then k_mlp_chain.  The mean loss adds into *B->loss. */
static int mlp_gradient(struct gpu_state *g, const struct mlp_bufs *B, const float *sr, const int *sp,
                        const float *sE, const float *sF, const int *tiles, int ntiles, float wE, float wF) {
    const int np = g->mlp_np;
    const float inv_n = 1.0f / (16.0f * ntiles);
    if (ntiles < 1) return gerror_(engine_err_range, "empty batch");
    hipLaunchKernelGGL(k_mlp_fill, dim3(64), dim3(256), 0, g->stream, B->grad, (long long)g->mlp_nparams, 0.0f);
    hipLaunchKernelGGL(k_mlp_fill, dim3(16), dim3(256), 0, g->stream, B->gb1e, (long long)np * MLP_H, 0.0f);
    for (int c0 = 0; c0 < ntiles; c0 += MLP_TCHUNK) {
        const int nt = ntiles - c0 < MLP_TCHUNK ? ntiles - c0 : MLP_TCHUNK;
        const int grid = nt < B->grid_max ? nt : B->grid_max;
        hipLaunchKernelGGL(k_mlp_grad, dim3(grid), dim3(64), 0, g->stream, gpu_mlp_net(g), sr, sp, sE, sF, tiles + c0,
                           nt, wE, wF, inv_n, np, B->tpart, B->wpart, B->err);
        hipLaunchKernelGGL(k_mlp_wsum, dim3((MLP_WPART + 63) / 64), dim3(64 * MLP_WSUM_SPLIT), 0, g->stream,
                           B->wpart, grid, B->grad);
        hipLaunchKernelGGL(k_mlp_tsum, dim3(np + 1), dim3(256), 0, g->stream, B->tpart, tiles + c0, nt, sp, np, B->grad,
                           B->gb1e, B->loss, B->err);
    }
    hipLaunchKernelGGL(k_mlp_chain, dim3((MLP_H + MLP_H * MLP_EMB + np * MLP_EMB + 63) / 64), dim3(64 * MLP_WSUM_SPLIT),
                       0, g->stream, g->d_theta, np, B->gb1e, B->grad);
    HIPCHECK(hipGetLastError());
    return engine_err_ok;
}

static struct mlp_bufs mlp_engine_bufs(struct gpu_state *g, float *loss) {
    struct mlp_bufs B = {g->d_grad, loss, g->d_gb1e, g->d_wpart, g->d_tpart, g->d_err, MLP_GRID_MAX};
    return B;
}

extern "C" int engine_mlp_grad(struct engine *e, const float *r, const int *pair, const float *E, const float *F,
                               long long n, double wE, double wF, float *grad, double *loss) {
    struct gpu_state *g = mlp_state(e, 1);
    long long k;
    int *tiles;
    float lossf = 0.0f;
    if (g == NULL) return engine_err;
    if (r == NULL || pair == NULL || E == NULL || F == NULL || grad == NULL) return gerror_(engine_err_null, "NULL");
    if (n < 16 || n % 16 != 0 || n / 16 > (1 << 24)) return gerror_(engine_err_range, "n must be a positive multiple of 16");
    if (mlp_check(g, r, pair, n) < 0) return engine_err_range;
    for (k = 0; k < n; k++)
        if (pair[k] != pair[k - k % 16]) return gerror_(engine_err_range, "each tile of 16 samples must share a pair id");
    if (mlp_samples(g, n) < 0) return engine_err_gpu;
    tiles = (int *)malloc(sizeof(int) * (n / 16));
    if (tiles == NULL) return gerror_(engine_err_malloc, "malloc failed");
    for (k = 0; k < n / 16; k++) tiles[k] = (int)k;
    (void)hipFree(g->d_loss);
    g->d_loss = NULL;
    if (hipMalloc(&g->d_loss, sizeof(float)) != hipSuccess ||
        hipMemcpy(g->d_sr, r, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_sp, pair, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_sE, E, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_sF, F, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_tiles, tiles, sizeof(int) * (n / 16), hipMemcpyHostToDevice) != hipSuccess) {
        free(tiles);
        return gerror_(engine_err_gpu, "sample upload failed");
    }
    free(tiles);
    hipLaunchKernelGGL(k_mlp_fill, dim3(1), dim3(64), 0, g->stream, g->d_loss, 1LL, 0.0f);
    {
        const struct mlp_bufs B = mlp_engine_bufs(g, g->d_loss);
        if (mlp_gradient(g, &B, g->d_sr, g->d_sp, g->d_sE, g->d_sF, g->d_tiles, (int)(n / 16), (float)wE, (float)wF) < 0)
            return engine_err_gpu;
    }
    HIPCHECK(hipStreamSynchronize(g->stream));
    HIPCHECK(hipMemcpy(grad, g->d_grad, sizeof(float) * g->mlp_nparams, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(&lossf, g->d_loss, sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = lossf;
    return engine_err_ok;
}

extern "C" int engine_mlp_grad_noalloc(struct engine *e, const float *r, const int *pair, const float *E,
                                       const float *F, long long n, double wE, double wF, float *grad, float *loss) {
    struct gpu_state *g = mlp_state(e, 1);
    float *sc;
    if (g == NULL) return engine_err;
    if (r == NULL || pair == NULL || E == NULL || F == NULL || grad == NULL || loss == NULL)
        return gerror_(engine_err_null, "NULL device buffer");
    if (n < 16 || n % 16 != 0) return gerror_(engine_err_range, "n must be a positive multiple of 16");
    if (n / 16 > MLP_NA_TILES) return gerror_(engine_err_range, "n exceeds the scratch limit of 65536 samples");
    if (g->mlp_np > MLP_NA_PAIRS) return gerror_(engine_err_range, "more species pairs than the scratch holds (1024)");
    if (g->device < 0 || g->device >= MLP_NA_MAXDEV || (sc = scratch_ptr[g->device]) == NULL)
        return gerror_(engine_err_gpu, "scratch not set up on this device (engine_mlp_init)");
    const int ntiles = (int)(n / 16);
    const struct mlp_bufs B = {grad, loss, sc + NA_OFF_GB1E, sc + NA_OFF_WPART, sc + NA_OFF_TPART, sc + NA_OFF_ERR,
                               MLP_NA_GRID};
    int *tiles = (int *)(sc + NA_OFF_TILES);
    hipLaunchKernelGGL(k_mlp_iota, dim3(16), dim3(256), 0, g->stream, tiles, ntiles);
    hipLaunchKernelGGL(k_mlp_fill, dim3(1), dim3(64), 0, g->stream, sc + NA_OFF_ERR, 1LL, 0.0f);
    hipLaunchKernelGGL(k_mlp_fill, dim3(1), dim3(64), 0, g->stream, loss, 1LL, 0.0f);
    if (mlp_gradient(g, &B, r, pair, E, F, tiles, ntiles, (float)wE, (float)wF) < 0) return engine_err_gpu;
    return engine_err_ok;
}

extern "C" int engine_mlp_train(struct engine *e, const float *r, const int *pair, const float *E, const float *F,
                                long long n, int steps, int batch_tiles, double lr0, double lr1, double wE, double wF,
                                unsigned long long seed, float *loss_hist) {
    struct gpu_state *g = mlp_state(e, 1);
    long long k, ntiles = 0, m;
    long long *cnt, *start;
    float *hr, *hE, *hF;
    int *hp, *tiles;
    int np, step, pos;
    unsigned long long rs = seed;
    const float b1 = 0.9f, b2 = 0.999f;
    if (g == NULL) return engine_err;
    if (r == NULL || pair == NULL || E == NULL || F == NULL) return gerror_(engine_err_null, "NULL samples");
    if (n < 1 || steps < 1 || batch_tiles < 1) return gerror_(engine_err_range, "n, steps and batch must be positive");
    if (mlp_check(g, r, pair, n) < 0) return engine_err_range;
    np = g->mlp_np;
    /* Group the samples by pair into tiles of 16; a pair's last tile is
       padded with its own first samples. */
    cnt = (long long *)calloc(np + 1, sizeof(long long));
    start = (long long *)calloc(np + 1, sizeof(long long));
    if (cnt == NULL || start == NULL) {
        free(cnt);
        free(start);
        return gerror_(engine_err_malloc, "malloc failed");
    }
    for (k = 0; k < n; k++) cnt[pair[k]]++;
    for (k = 0; k < np; k++) ntiles += (cnt[k] + 15) / 16;
    if (ntiles > (1 << 26)) {
        free(cnt);
        free(start);
        return gerror_(engine_err_range, "too many samples");
    }
    m = 16 * ntiles;
    hr = (float *)malloc(sizeof(float) * m);
    hE = (float *)malloc(sizeof(float) * m);
    hF = (float *)malloc(sizeof(float) * m);
    hp = (int *)malloc(sizeof(int) * m);
    tiles = (int *)malloc(sizeof(int) * ntiles);
    if (hr == NULL || hE == NULL || hF == NULL || hp == NULL || tiles == NULL) {
        free(cnt); free(start); free(hr); free(hE); free(hF); free(hp); free(tiles);
        return gerror_(engine_err_malloc, "malloc failed");
    }
    {
        long long off = 0;
        long long *fill = (long long *)calloc(np + 1, sizeof(long long));
        for (k = 0; k < np; k++) {
            start[k] = off;
            off += 16 * ((cnt[k] + 15) / 16);
        }
        for (k = 0; k < n; k++) {
            const int p = pair[k];
            const long long d = start[p] + fill[p]++;
            hr[d] = r[k];
            hE[d] = E[k];
            hF[d] = F[k];
            hp[d] = p;
        }
        for (k = 0; k < np; k++)
            for (long long q = cnt[k]; q < 16 * ((cnt[k] + 15) / 16); q++) {
                const long long d = start[k] + q, src = start[k] + (q - cnt[k]) % cnt[k];
                hr[d] = hr[src];
                hE[d] = hE[src];
                hF[d] = hF[src];
                hp[d] = (int)k;
            }
        free(fill);
    }
    for (k = 0; k < ntiles; k++) tiles[k] = (int)k;
    free(cnt);
    free(start);
    if (mlp_samples(g, m) < 0) {
        free(hr); free(hE); free(hF); free(hp); free(tiles);
        return engine_err_gpu;
    }
    (void)hipFree(g->d_loss);
    g->d_loss = NULL;
    if (hipMalloc(&g->d_loss, sizeof(float) * steps) != hipSuccess ||
        hipMemcpy(g->d_sr, hr, sizeof(float) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_sp, hp, sizeof(int) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_sE, hE, sizeof(float) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_sF, hF, sizeof(float) * m, hipMemcpyHostToDevice) != hipSuccess) {
        free(hr); free(hE); free(hF); free(hp); free(tiles);
        return gerror_(engine_err_gpu, "sample upload failed");
    }
    free(hr); free(hE); free(hF); free(hp);
    hipLaunchKernelGGL(k_mlp_fill, dim3(16), dim3(256), 0, g->stream, g->d_loss, (long long)steps, 0.0f);
    if (batch_tiles > ntiles) batch_tiles = (int)ntiles;
    pos = (int)ntiles; /* forces a shuffle before the first step */
    for (step = 0; step < steps; step++) {
        int nb;
        double lr, c1, c2;
        if (pos >= ntiles) {
            for (k = ntiles - 1; k > 0; k--) {
                const long long j = (long long)(mlp_rng(&rs) % (unsigned long long)(k + 1));
                const int t = tiles[k];
                tiles[k] = tiles[j];
                tiles[j] = t;
            }
            /* the previous step's kernels may still read d_tiles */
            if (hipStreamSynchronize(g->stream) != hipSuccess ||
                hipMemcpy(g->d_tiles, tiles, sizeof(int) * ntiles, hipMemcpyHostToDevice) != hipSuccess) {
                free(tiles);
                return gerror_(engine_err_gpu, "tile upload failed");
            }
            pos = 0;
        }
        nb = (int)(ntiles - pos < batch_tiles ? ntiles - pos : batch_tiles);
        g->mlp_step += 1;
        lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + cos(3.141592653589793 * step / steps));
        c1 = 1.0 / (1.0 - pow(b1, (double)g->mlp_step));
        c2 = 1.0 / (1.0 - pow(b2, (double)g->mlp_step));
        {
            const struct mlp_bufs B = mlp_engine_bufs(g, g->d_loss + step);
            if (mlp_gradient(g, &B, g->d_sr, g->d_sp, g->d_sE, g->d_sF, g->d_tiles + pos, nb, (float)wE, (float)wF) < 0) {
                free(tiles);
                return engine_err_gpu;
            }
        }
        hipLaunchKernelGGL(k_mlp_adam, dim3((g->mlp_nparams + 255) / 256), dim3(256), 0, g->stream, g->d_theta,
                           g->d_grad, g->d_adam_m, g->d_adam_v, g->mlp_nparams, (float)lr, b1, b2, (float)c1,
                           (float)c2, 1e-8f);
        if (mlp_publish(g) < 0) {
            free(tiles);
            return engine_err_gpu;
        }
        pos += nb;
    }
    free(tiles);
    HIPCHECK(hipStreamSynchronize(g->stream));
    HIPCHECK(hipGetLastError());
    if (loss_hist) HIPCHECK(hipMemcpy(loss_hist, g->d_loss, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return mlp_update_shift(e, g);
}

extern "C" int engine_mlp_eval(struct engine *e, const float *r, const int *pair, long long n, float *E, float *F) {
    struct gpu_state *g = mlp_state(e, 1);
    float *dE = NULL, *dF = NULL;
    int grid;
    if (g == NULL) return engine_err;
    if (r == NULL || pair == NULL || E == NULL || F == NULL) return gerror_(engine_err_null, "NULL");
    if (n < 1) return gerror_(engine_err_range, "n must be positive");
    if (mlp_check(g, r, pair, n) < 0) return engine_err_range;
    if (mlp_samples(g, n) < 0) return engine_err_gpu;
    HIPCHECK(hipMemcpy(g->d_sr, r, sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(g->d_sp, pair, sizeof(int) * n, hipMemcpyHostToDevice));
    dE = g->d_sE;
    dF = g->d_sF;
    grid = (int)((n + 15) / 16 < 65536 ? (n + 15) / 16 : 65536);
    hipLaunchKernelGGL(k_mlp_eval, dim3(grid), dim3(64), 0, g->stream, gpu_mlp_net(g), g->d_sr, g->d_sp, n, dE, dF);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(g->stream));
    HIPCHECK(hipMemcpy(E, dE, sizeof(float) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(F, dF, sizeof(float) * n, hipMemcpyDeviceToHost));
    return engine_err_ok;
}
