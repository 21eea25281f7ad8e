/*
 * gfx950 nonbonded engine behind the mdcore C API (C-ABI in engine_gpu.h).
 *
 * One nonbonded pass is four launches on one stream:
 *   k_pack      cell-ordered FP64 positions -> float4 (x, y, z, type bits),
 *               id -> slot map, slot -> cell map;
 *   k_sort      runner_dosort (runner_dosort.c:97-161) for every cell and
 *               every flagged direction: 16-bit projection keys computed in
 *               FP64 without contraction, then the reference's descending
 *               quicksort (runner.c:110-174) replayed so that the sorted
 *               lists, ties included, are the reference's;
 *   k_nonbond   persistent kernel: each wave pulls an owner cell from one of
 *               8 device queues (one per XCD slab of cells), walks the cell's
 *               neighbour list (the reference's self/pair tasks that touch
 *               it) and accumulates the forces of its own atoms only, so no
 *               atomics and no cross-wave write conflicts; pairs are tested
 *               with the reference's sorted-window predicate
 *               (runner_dopair.c:455-478) on the sorted lists staged in LDS,
 *               the cutoff test is re-done in FP64 inside a thin band
 *               around rc^2, the table is evaluated in FP32
 *               (potential_eval.h:67-104), and per-atom sums are two-sum
 *               compensated; excluded pairs are subtracted afterwards as in
 *               exclusion.c:81-282;
 *   k_reduce    per-cell FP64 energies/virials/counts -> totals, fixed order.
 * Every pair is evaluated by both of its cells (once per owner); energy,
 * virial and counts are taken on the task's i side only.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "engine_gpu.h"

#define MD_NQUEUE 8
#define MD_MAX_CELL 1024
#define MD_MAX_TYPES 64

#define gerror_(id, msg) errs_register((id), (msg), __LINE__, __func__, __FILE__)

#define HIPCHECK(call)                                                             \
    do {                                                                           \
        hipError_t err_ = (call);                                                  \
        if (err_ != hipSuccess) {                                                  \
            errs_register(engine_err_gpu, hipGetErrorString(err_), __LINE__, __func__, __FILE__); \
            return engine_err_gpu;                                                 \
        }                                                                          \
    } while (0)

/* ------------------------------------------------------------ parameters */

struct nb_params {
    int ncells, nparts, nids, max_type, ntab, proj;
    int dmaxdist;
    double dscale, bias, cutoff2;
    double dim[3], h[3];
    float rc2f, band;
    /* dynamic LDS layout (byte offsets) */
    int lds_tab, lds_toff, lds_pos, lds_key;
};

/* Neighbour-list entry of an owner cell: x = other cell, y = sid | orient<<4
   | self<<5 (orient 0: owner is the task's i cell), z = dnshift with the
   reference window |shift|, w = dnshift with the projected shift.n. */

struct gpu_state {
    int device;
    hipStream_t stream;
    hipEvent_t ev[8];
    int ncells, ntab, max_type, nparts_cap, nids_cap, maxcount;
    int nparts, nids;
    struct nb_params P;
    int grid_nb, ncu;
    size_t lds_nb;
    /* static */
    int *d_pind, *d_toff;
    float4 *d_tab;
    float *d_coeffs;
    int4 *d_nbr;
    int *d_nbr_off;
    double *d_origin;
    int *d_loc;
    unsigned int *d_sortflags;
    int *d_excl_off, *d_excl_list;
    /* per pass */
    int *d_count, *d_offset;
    double *d_x;
    int *d_type, *d_id, *d_index_of_id, *d_cell_of;
    float4 *d_xf, *d_fhi, *d_flo;
    unsigned int *d_sortlist;
    unsigned short *d_keyof;
    double *d_cell_acc, *d_tot;
    unsigned long long *d_cell_cnt, *d_tot_cnt;
    int *d_queue;
    /* pinned host staging */
    double *h_x;
    int *h_type, *h_id, *h_count, *h_offset;
    float4 *h_fhi, *h_flo;
    double *h_tot;
    unsigned long long *h_tot_cnt;
};

/* ------------------------------------------------------------ device helpers */

__device__ __forceinline__ void acc2(float &hi, float &lo, float v) {
    /* hi + lo += v with the rounding error of hi + v kept in lo (two-sum). */
    float s = hi + v;
    float bp = s - hi;
    float err = (hi - (s - bp)) + (v - bp);
    hi = s;
    lo += err;
}

__device__ __forceinline__ double min_image(double s, double dim) {
    if (s * 2 > dim)
        s -= dim;
    else if (s * 2 < -dim)
        s += dim;
    return s;
}

/* Reference r^2 of a pair task (runner_dopair.c:465-486 with fptype.h:275-278):
   pix = x_i - shift, dx = pix - x_j, r2 = dx0^2 + dx1^2 + dx2^2, no FMA. */
__device__ __attribute__((noinline)) double r2_task64(const double *xi, const double *xj, double s0, double s1,
                                                      double s2) {
#pragma clang fp contract(off)
    double p0 = xi[0] - s0, p1 = xi[1] - s1, p2 = xi[2] - s2;
    double d0 = p0 - xj[0], d1 = p1 - xj[1], d2 = p2 - xj[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

/* Reference r^2 of an exclusion (exclusion.c:139-147): i is the smaller id. */
__device__ __attribute__((noinline)) double r2_excl64(const double *xi, const double *xj, const int *li,
                                                      const int *lj, const double *h) {
#pragma clang fp contract(off)
    double r2 = 0.0;
    for (int k = 0; k < 3; k++) {
        int sh = li[k] - lj[k];
        if (sh > 1)
            sh = -1;
        else if (sh < -1)
            sh = 1;
        double d = xi[k] - xj[k] + h[k] * sh;
        r2 += d * d;
    }
    return r2;
}

/* Table evaluation in FP32: e and f = (dE/dr)/r (potential_eval.h:67-104). */
__device__ __forceinline__ void pot_eval32(const float4 tab, const float *__restrict__ c0, float r2, float &e,
                                           float &f) {
    float ir = __builtin_amdgcn_rsqf(r2);
    float r = r2 * ir;
    float fi = fmaxf(0.0f, tab.x + r * (tab.y + r * tab.z));
    int ind = (int)fi;
    int n = __float_as_int(tab.w);
    ind = ind > n ? n : ind;
    const float4 *c = (const float4 *)(c0 + ind * 8);
    float4 a = c[0], b = c[1];
    float x = (r - a.x) * a.y;
    float ee = a.z * x + a.w;
    float eff = a.z;
    eff = eff * x + ee;
    ee = ee * x + b.x;
    eff = eff * x + ee;
    ee = ee * x + b.y;
    eff = eff * x + ee;
    ee = ee * x + b.z;
    eff = eff * x + ee;
    ee = ee * x + b.w;
    e = ee;
    f = eff * a.y * ir;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* ------------------------------------------------------------ k_pack */

__global__ __launch_bounds__(256) void k_pack(int nparts, const double *__restrict__ x, const int *__restrict__ type,
                                              const int *__restrict__ id, int ncells,
                                              const int *__restrict__ offset, float4 *__restrict__ xf,
                                              int *__restrict__ index_of_id, int *__restrict__ cell_of,
                                              int *__restrict__ queue) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a == 0)
        for (int q = 0; q < MD_NQUEUE; q++) queue[q] = 0;
    if (a < nparts) {
        xf[a] = make_float4((float)x[3 * a], (float)x[3 * a + 1], (float)x[3 * a + 2], __int_as_float(type[a]));
        index_of_id[id[a]] = a;
    }
    if (a < ncells)
        for (int k = offset[a]; k < offset[a + 1]; k++) cell_of[k] = a;
}

__global__ __launch_bounds__(256) void k_clear_ids(int nids, int *__restrict__ index_of_id) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nids) index_of_id[k] = -1;
}

/* ------------------------------------------------------------ k_sort */

/* Descending sort of (index<<16 | key) words by key: the reference's
   quicksort with selection sort below 16 elements and an explicit stack
   (runner.c:110-174).  The partition and push order decide the order of
   equal keys, so they are kept as they are. */
__device__ void sort_desc(unsigned int *a, int N, int *overflow) {
    int st_lo[16], st_hi[16];
    int qpos = 0;
    st_lo[0] = 0;
    st_hi[0] = N - 1;
    while (qpos >= 0) {
        int lo = st_lo[qpos], hi = st_hi[qpos];
        qpos -= 1;
        if (hi - lo < 15) {
            for (int i = lo; i < hi; i++) {
                int imax = i;
                unsigned int kmax = a[i] & 0xffff;
                for (int j = i + 1; j <= hi; j++) {
                    unsigned int kj = a[j] & 0xffff;
                    if (kj > kmax) {
                        imax = j;
                        kmax = kj;
                    }
                }
                if (imax != i) {
                    unsigned int t = a[imax];
                    a[imax] = a[i];
                    a[i] = t;
                }
            }
        } else {
            unsigned int pivot = a[(lo + hi) / 2] & 0xffff;
            int i = lo, j = hi;
            while (i <= j) {
                while ((a[i] & 0xffff) > pivot) i++;
                while ((a[j] & 0xffff) < pivot) j--;
                if (i <= j) {
                    if (i < j) {
                        unsigned int t = a[i];
                        a[i] = a[j];
                        a[j] = t;
                    }
                    i += 1;
                    j -= 1;
                }
            }
            if (qpos + 2 >= 16) {
                *overflow = 1;
                return;
            }
            if (j > (lo + hi) / 2) {
                if (lo < j) { qpos++; st_lo[qpos] = lo; st_hi[qpos] = j; }
                if (i < hi) { qpos++; st_lo[qpos] = i; st_hi[qpos] = hi; }
            } else {
                if (i < hi) { qpos++; st_lo[qpos] = i; st_hi[qpos] = hi; }
                if (lo < j) { qpos++; st_lo[qpos] = lo; st_hi[qpos] = j; }
            }
        }
    }
}

__constant__ double c_axis[39];

/* Projection key (runner_dosort.c:146-149): dscale*(bias + x.n) in FP64,
   left to right, no fused multiply-add, truncated to unsigned. */
__device__ __attribute__((noinline)) unsigned int sort_key(const double *x, int sid, double dscale, double bias) {
#pragma clang fp contract(off)
    double v = dscale * (bias + x[0] * c_axis[3 * sid] + x[1] * c_axis[3 * sid + 1] + x[2] * c_axis[3 * sid + 2]);
    return (unsigned int)v;
}

__global__ __launch_bounds__(256) void k_sort(struct nb_params P, const int *__restrict__ count,
                                              const int *__restrict__ offset,
                                              const unsigned int *__restrict__ sortflags,
                                              const double *__restrict__ x, unsigned int *__restrict__ sortlist,
                                              unsigned short *__restrict__ keyof, int *__restrict__ flag) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int cid = t / 13, sid = t - 13 * (t / 13);
    if (cid >= P.ncells) return;
    if (!((sortflags[cid] >> sid) & 1u)) return;
    int n = count[cid];
    if (n == 0) return;
    int off = offset[cid];
    unsigned int *list = sortlist + 13 * (size_t)off + (size_t)sid * n;
    for (int i = 0; i < n; i++) {
        unsigned int key = sort_key(&x[3 * (size_t)(off + i)], sid, P.dscale, P.bias);
        list[i] = ((unsigned int)i << 16) | key;
        keyof[(size_t)sid * P.nparts + off + i] = (unsigned short)key;
    }
    sort_desc(list, n, flag);
}

/* ------------------------------------------------------------ k_nonbond */

struct lane_acc {
    float hx, hy, hz, lx, ly, lz;
    double epot, vir, eexcl;
    unsigned int np, nref, nexcl;
};

__global__ __launch_bounds__(64) void k_nonbond(
    struct nb_params P, const int *__restrict__ count, const int *__restrict__ offset,
    const int4 *__restrict__ nbr, const int *__restrict__ nbr_off, const double *__restrict__ origin,
    const int *__restrict__ loc, const float4 *__restrict__ xf, const double *__restrict__ x64,
    const int *__restrict__ pid, const unsigned int *__restrict__ sortlist,
    const unsigned short *__restrict__ keyof, const int *__restrict__ pind_g, const float4 *__restrict__ tab_g,
    const int *__restrict__ toff_g, const float *__restrict__ coeffs, const int *__restrict__ excl_off,
    const int *__restrict__ excl_list, const int *__restrict__ index_of_id, const int *__restrict__ cell_of,
    float4 *__restrict__ fhi, float4 *__restrict__ flo, double *__restrict__ cell_acc,
    unsigned long long *__restrict__ cell_cnt, int *__restrict__ queue) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    int *s_pind = (int *)smem;
    float4 *s_tab = (float4 *)(smem + P.lds_tab);
    int *s_toff = (int *)(smem + P.lds_toff);
    float4 *s_pos = (float4 *)(smem + P.lds_pos);
    int2 *s_key = (int2 *)(smem + P.lds_key);
    const int lane = threadIdx.x;
    const int mt = P.max_type;

    for (int k = lane; k < mt * mt; k += 64) s_pind[k] = pind_g[k];
    for (int k = lane; k < P.ntab; k += 64) {
        s_tab[k] = tab_g[k];
        s_toff[k] = toff_g[k];
    }
    __syncthreads();

    int q = blockIdx.x % MD_NQUEUE, tried = 0;
    while (tried < MD_NQUEUE) {
        int lo_q = (int)(((long long)q * P.ncells) / MD_NQUEUE);
        int hi_q = (int)(((long long)(q + 1) * P.ncells) / MD_NQUEUE);
        int it = 0;
        if (lane == 0) it = atomicAdd(&queue[q], 1);
        it = __shfl(it, 0);
        if (lo_q + it >= hi_q) {
            q = (q + 1) % MD_NQUEUE;
            tried++;
            continue;
        }
        const int c = lo_q + it;
        const int nc = count[c];
        if (nc == 0) continue;
        const int offc = offset[c];
        const double oc0 = origin[3 * c], oc1 = origin[3 * c + 1], oc2 = origin[3 * c + 2];
        const int nchunk = (nc + 63) / 64;
        double cell_e = 0.0, cell_v = 0.0, cell_x = 0.0;
        unsigned long long cell_np = 0, cell_nref = 0, cell_nx = 0;

        for (int ch = 0; ch < nchunk; ch++) {
            const int i0 = (int)(((long long)ch * nc) / nchunk);
            const int i1 = (int)(((long long)(ch + 1) * nc) / nchunk);
            const int nsub = i1 - i0;
            const int NP = 64 / nsub;
            const int il = lane % nsub, part = lane / nsub;
            const bool active = part < NP;
            const int myidx = i0 + il; /* storage index in the owner cell */
            const int a = offc + myidx;
            const float4 xi = xf[a];
            const int ti = __float_as_int(xi.w);
            lane_acc A;
            A.hx = A.hy = A.hz = A.lx = A.ly = A.lz = 0.0f;
            A.epot = A.vir = A.eexcl = 0.0;
            A.np = A.nref = A.nexcl = 0;

            for (int k = nbr_off[c]; k < nbr_off[c + 1]; k++) {
                const int4 en = nbr[k];
                const int cj = en.x, sid = en.y & 15, orient = (en.y >> 4) & 1, self = (en.y >> 5) & 1;
                const int nj = count[cj];
                if (nj == 0) continue;
                const int offj = offset[cj];
                __syncthreads();
                if (self) {
                    for (int t = lane; t < nj; t += 64) {
                        s_pos[t] = xf[offj + t];
                        s_key[t] = make_int2(0, t);
                    }
                } else {
                    const unsigned int *sl = sortlist + 13 * (size_t)offj + (size_t)sid * nj;
                    for (int t = lane; t < nj; t += 64) {
                        unsigned int w = sl[orient ? t : nj - 1 - t];
                        int idx = (int)(w >> 16);
                        idx = idx < nj ? idx : nj - 1; /* sort wrote idx < nj; keep loads in bounds regardless */
                        s_pos[t] = xf[offj + idx];
                        s_key[t] = make_int2((int)(w & 0xffff), idx);
                    }
                }
                __syncthreads();

                /* Shift from the owner cell to the neighbour (minimum image). */
                const double s0 = min_image(origin[3 * cj] - oc0, P.dim[0]);
                const double s1 = min_image(origin[3 * cj + 1] - oc1, P.dim[1]);
                const double s2 = min_image(origin[3 * cj + 2] - oc2, P.dim[2]);
                const float sf0 = (float)s0, sf1 = (float)s1, sf2 = (float)s2;
                float le = 0.0f, lv = 0.0f, lfx = 0.0f, lfy = 0.0f, lfz = 0.0f;
                unsigned int lnp = 0, lnref = 0;

                if (self) {
                    if (active) {
                        for (int t = part; t < nj; t += NP) {
                            if (t == myidx) continue;
                            const float4 pj = s_pos[t];
                            const int tb = s_pind[ti * mt + __float_as_int(pj.w)];
                            if (tb < 0) continue;
                            const float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            bool in = r2 <= P.rc2f;
                            if (fabsf(r2 - P.rc2f) <= P.band)
                                in = r2_task64(&x64[3 * (size_t)a], &x64[3 * (size_t)(offj + t)], 0.0, 0.0, 0.0) <=
                                     P.cutoff2;
                            if (!in) continue;
                            float e, f;
                            pot_eval32(s_tab[tb], coeffs + s_toff[tb], r2, e, f);
                            lfx -= f * dx;
                            lfy -= f * dy;
                            lfz -= f * dz;
                            if (myidx > t) {
                                le += e;
                                lv -= f * r2;
                                lnp++;
                            }
                        }
                    }
                    lnref = lnp;
                } else {
                    const int dns = P.proj ? en.w : en.z;
                    const unsigned int *slc = sortlist + 13 * (size_t)offc + (size_t)sid * nc;
                    const int kfirst = (int)(slc[0] & 0xffff), klast = (int)(slc[nc - 1] & 0xffff);
                    const int s_imax = orient ? s_key[0].x : kfirst;
                    const int s_jmin = orient ? klast : s_key[0].x;
                    const bool live_ref = (s_jmin + en.z - s_imax) >= 0;
                    if ((P.proj || live_ref) && active) {
                        const int myk = keyof[(size_t)sid * P.nparts + a];
                        for (int t = part; t < nj; t += NP) {
                            const int2 kj = s_key[t];
                            const int gap = orient ? (myk + dns - kj.x) : (kj.x + dns - myk);
                            if (gap >= P.dmaxdist) break;
                            const float4 pj = s_pos[t];
                            const int tb = s_pind[ti * mt + __float_as_int(pj.w)];
                            if (tb < 0) continue;
                            const float dx = (xi.x - pj.x) - sf0, dy = (xi.y - pj.y) - sf1,
                                        dz = (xi.z - pj.z) - sf2;
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            bool in = r2 <= P.rc2f;
                            if (fabsf(r2 - P.rc2f) <= P.band) {
                                const double *xa = &x64[3 * (size_t)a];
                                const double *xb = &x64[3 * (size_t)(offj + kj.y)];
                                double r2d;
                                if (orient == 0)
                                    r2d = r2_task64(xa, xb, s0, s1, s2);
                                else
                                    r2d = r2_task64(xb, xa, min_image(oc0 - origin[3 * cj], P.dim[0]),
                                                    min_image(oc1 - origin[3 * cj + 1], P.dim[1]),
                                                    min_image(oc2 - origin[3 * cj + 2], P.dim[2]));
                                in = r2d <= P.cutoff2;
                            }
                            if (!in) continue;
                            float e, f;
                            pot_eval32(s_tab[tb], coeffs + s_toff[tb], r2, e, f);
                            lfx -= f * dx;
                            lfy -= f * dy;
                            lfz -= f * dz;
                            if (orient == 0) {
                                le += e;
                                lv -= f * r2;
                                lnp++;
                                lnref += (live_ref && (kj.x + en.z - myk) < P.dmaxdist) ? 1u : 0u;
                            }
                        }
                    }
                }
                acc2(A.hx, A.lx, lfx);
                acc2(A.hy, A.ly, lfy);
                acc2(A.hz, A.lz, lfz);
                A.epot += (double)le;
                A.vir += (double)lv;
                A.np += lnp;
                A.nref += lnref;
            }

            /* Excluded pairs: subtract what the pair pass added
               (exclusion.c:139-208), same FP32 arithmetic as above. */
            if (active && part == 0) {
                const int myid = pid[a];
                for (int k = excl_off[myid]; k < excl_off[myid + 1]; k++) {
                    const int oid = excl_list[k];
                    if (oid == myid) continue;
                    const int b = index_of_id[oid];
                    if (b < 0) continue;
                    const int cb = cell_of[b];
                    const float4 pj = xf[b];
                    const int tb = s_pind[ti * mt + __float_as_int(pj.w)];
                    if (tb < 0) continue;
                    const double s0 = min_image(origin[3 * cb] - oc0, P.dim[0]);
                    const double s1 = min_image(origin[3 * cb + 1] - oc1, P.dim[1]);
                    const double s2 = min_image(origin[3 * cb + 2] - oc2, P.dim[2]);
                    const float dx = (xi.x - pj.x) - (float)s0, dy = (xi.y - pj.y) - (float)s1,
                                dz = (xi.z - pj.z) - (float)s2;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    bool in = r2 <= P.rc2f;
                    if (fabsf(r2 - P.rc2f) <= P.band) {
                        const bool mine = myid < oid;
                        const int ia = mine ? a : b, ib = mine ? b : a;
                        const int ca = mine ? c : cb, cbb = mine ? cb : c;
                        in = r2_excl64(&x64[3 * (size_t)ia], &x64[3 * (size_t)ib], &loc[3 * ca], &loc[3 * cbb],
                                       P.h) <= P.cutoff2;
                    }
                    if (!in) continue;
                    float e, f;
                    pot_eval32(s_tab[tb], coeffs + s_toff[tb], r2, e, f);
                    acc2(A.hx, A.lx, f * dx);
                    acc2(A.hy, A.ly, f * dy);
                    acc2(A.hz, A.lz, f * dz);
                    if (myid < oid) {
                        A.eexcl -= (double)e;
                        A.nexcl++;
                    }
                }
            }

            /* Fold the partitions of each atom into partition 0, in order. */
            for (int p = 1; p < NP; p++) {
                const int src = (lane + p * nsub) & 63;
                const float ohx = __shfl(A.hx, src), ohy = __shfl(A.hy, src), ohz = __shfl(A.hz, src);
                const float olx = __shfl(A.lx, src), oly = __shfl(A.ly, src), olz = __shfl(A.lz, src);
                if (part == 0) {
                    acc2(A.hx, A.lx, ohx);
                    acc2(A.hy, A.ly, ohy);
                    acc2(A.hz, A.lz, ohz);
                    A.lx += olx;
                    A.ly += oly;
                    A.lz += olz;
                }
            }
            if (part == 0) {
                fhi[a] = make_float4(A.hx, A.hy, A.hz, 0.0f);
                flo[a] = make_float4(A.lx, A.ly, A.lz, 0.0f);
            }
            if (!active) {
                A.epot = A.vir = A.eexcl = 0.0;
                A.np = A.nref = A.nexcl = 0;
            }
            cell_e += wave_sum(A.epot);
            cell_v += wave_sum(A.vir);
            cell_x += wave_sum(A.eexcl);
            cell_np += (unsigned long long)wave_sum((int)A.np);
            cell_nref += (unsigned long long)wave_sum((int)A.nref);
            cell_nx += (unsigned long long)wave_sum((int)A.nexcl);
        }
        if (lane == 0) {
            cell_acc[3 * c] = cell_e;
            cell_acc[3 * c + 1] = cell_x;
            cell_acc[3 * c + 2] = cell_v;
            cell_cnt[3 * c] = cell_np;
            cell_cnt[3 * c + 1] = cell_nx;
            cell_cnt[3 * c + 2] = cell_nref;
        }
    }
}

/* ------------------------------------------------------------ k_reduce */

__global__ __launch_bounds__(1024) void k_reduce(int ncells, const double *__restrict__ cell_acc,
                                                 const unsigned long long *__restrict__ cell_cnt,
                                                 double *__restrict__ tot, unsigned long long *__restrict__ tot_cnt) {
    __shared__ double sd[3][1024];
    __shared__ unsigned long long sc[3][1024];
    const int t = threadIdx.x;
    double d0 = 0, d1 = 0, d2 = 0;
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    for (int k = t; k < ncells; k += 1024) {
        d0 += cell_acc[3 * k];
        d1 += cell_acc[3 * k + 1];
        d2 += cell_acc[3 * k + 2];
        c0 += cell_cnt[3 * k];
        c1 += cell_cnt[3 * k + 1];
        c2 += cell_cnt[3 * k + 2];
    }
    sd[0][t] = d0;
    sd[1][t] = d1;
    sd[2][t] = d2;
    sc[0][t] = c0;
    sc[1][t] = c1;
    sc[2][t] = c2;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s)
            for (int m = 0; m < 3; m++) {
                sd[m][t] += sd[m][t + s];
                sc[m][t] += sc[m][t + s];
            }
        __syncthreads();
    }
    if (t == 0)
        for (int m = 0; m < 3; m++) {
            tot[m] = sd[m][0];
            tot_cnt[m] = sc[m][0];
        }
}

/* ------------------------------------------------------------ host side */

static void gpu_free_pass(struct gpu_state *g) {
    (void)hipFree(g->d_count);
    (void)hipFree(g->d_offset);
    (void)hipFree(g->d_x);
    (void)hipFree(g->d_type);
    (void)hipFree(g->d_id);
    (void)hipFree(g->d_cell_of);
    (void)hipFree(g->d_xf);
    (void)hipFree(g->d_fhi);
    (void)hipFree(g->d_flo);
    (void)hipFree(g->d_sortlist);
    (void)hipFree(g->d_keyof);
    (void)hipHostFree(g->h_x);
    (void)hipHostFree(g->h_type);
    (void)hipHostFree(g->h_id);
    (void)hipHostFree(g->h_fhi);
    (void)hipHostFree(g->h_flo);
    g->d_count = g->d_offset = NULL;
    g->d_x = NULL;
    g->d_type = g->d_id = g->d_cell_of = NULL;
    g->d_xf = g->d_fhi = g->d_flo = NULL;
    g->d_sortlist = NULL;
    g->d_keyof = NULL;
    g->h_x = NULL;
    g->h_type = g->h_id = NULL;
    g->h_fhi = g->h_flo = NULL;
    g->nparts_cap = 0;
}

static int gpu_alloc_parts(struct gpu_state *g, int n) {
    size_t N;
    if (n <= g->nparts_cap) return 0;
    gpu_free_pass(g);
    N = (size_t)n + n / 8 + 64;
    HIPCHECK(hipMalloc(&g->d_x, sizeof(double) * 3 * N));
    HIPCHECK(hipMalloc(&g->d_type, sizeof(int) * N));
    HIPCHECK(hipMalloc(&g->d_id, sizeof(int) * N));
    HIPCHECK(hipMalloc(&g->d_cell_of, sizeof(int) * N));
    HIPCHECK(hipMalloc(&g->d_xf, sizeof(float4) * N));
    HIPCHECK(hipMalloc(&g->d_fhi, sizeof(float4) * N));
    HIPCHECK(hipMalloc(&g->d_flo, sizeof(float4) * N));
    HIPCHECK(hipMalloc(&g->d_sortlist, sizeof(unsigned int) * 13 * N));
    HIPCHECK(hipMalloc(&g->d_keyof, sizeof(unsigned short) * 13 * N));
    HIPCHECK(hipMalloc(&g->d_count, sizeof(int) * g->ncells));
    HIPCHECK(hipMalloc(&g->d_offset, sizeof(int) * (g->ncells + 1)));
    HIPCHECK(hipHostMalloc((void **)&g->h_x, sizeof(double) * 3 * N, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_type, sizeof(int) * N, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_id, sizeof(int) * N, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_fhi, sizeof(float4) * N, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_flo, sizeof(float4) * N, hipHostMallocDefault));
    g->nparts_cap = (int)N;
    return 0;
}

static int gpu_alloc_ids(struct gpu_state *g, int nids) {
    if (nids <= g->nids_cap) return 0;
    (void)hipFree(g->d_index_of_id);
    HIPCHECK(hipMalloc(&g->d_index_of_id, sizeof(int) * (size_t)nids));
    g->nids_cap = nids;
    return 0;
}

extern "C" int engine_gpu_setdevice(struct engine *e, int id) {
    int n = 0;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    HIPCHECK(hipGetDeviceCount(&n));
    if (id < 0 || id >= n) return gerror_(engine_err_range, "no such HIP device");
    HIPCHECK(hipSetDevice(id));
    if (e->gpu != NULL) ((struct gpu_state *)e->gpu)->device = id;
    return engine_err_ok;
}

extern "C" int engine_gpu_finalize(struct engine *e) {
    struct gpu_state *g;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    g = (struct gpu_state *)e->gpu;
    if (g == NULL) return engine_err_ok;
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->stream);
    gpu_free_pass(g);
    (void)hipFree(g->d_pind);
    (void)hipFree(g->d_toff);
    (void)hipFree(g->d_tab);
    (void)hipFree(g->d_coeffs);
    (void)hipFree(g->d_nbr);
    (void)hipFree(g->d_nbr_off);
    (void)hipFree(g->d_origin);
    (void)hipFree(g->d_loc);
    (void)hipFree(g->d_sortflags);
    (void)hipFree(g->d_excl_off);
    (void)hipFree(g->d_excl_list);
    (void)hipFree(g->d_index_of_id);
    (void)hipFree(g->d_cell_acc);
    (void)hipFree(g->d_cell_cnt);
    (void)hipFree(g->d_tot);
    (void)hipFree(g->d_tot_cnt);
    (void)hipFree(g->d_queue);
    (void)hipHostFree(g->h_count);
    (void)hipHostFree(g->h_offset);
    (void)hipHostFree(g->h_tot);
    (void)hipHostFree(g->h_tot_cnt);
    for (int k = 0; k < 8; k++) (void)hipEventDestroy(g->ev[k]);
    (void)hipStreamDestroy(g->stream);
    free(g);
    e->gpu = NULL;
    return engine_err_ok;
}

/* dnshift of a pair task (runner_dopair.c:426-452): with the reference's
   norm |shift| and with the projection shift.n on the pair's axis. */
static void task_dnshift(const struct space *s, int ti, int tj, int sid, double dscale, int *dns_ref,
                         int *dns_proj) {
    double sh[3], nshift, dot;
    int k;
    for (k = 0; k < 3; k++) {
        sh[k] = s->cells[tj].origin[k] - s->cells[ti].origin[k];
        if (sh[k] * 2 > s->dim[k])
            sh[k] -= s->dim[k];
        else if (sh[k] * 2 < -s->dim[k])
            sh[k] += s->dim[k];
    }
    nshift = sqrt(sh[0] * sh[0] + sh[1] * sh[1] + sh[2] * sh[2]);
    *dns_ref = (int)(dscale * nshift);
    dot = sh[0] * cell_shift[3 * sid] + sh[1] * cell_shift[3 * sid + 1] + sh[2] * cell_shift[3 * sid + 2];
    *dns_proj = (int)(dscale * dot);
}

extern "C" int engine_gpu_load(struct engine *e) {
    struct gpu_state *g;
    struct space *s;
    int mt, k, i, j, ntab = 0, ncoef = 0, cid;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    if (e->gpu != NULL) engine_gpu_finalize(e);
    s = &e->s;
    mt = e->max_type;
    if (mt > MD_MAX_TYPES) return gerror_(engine_err_nyi, "more than 64 particle types");
    g = (struct gpu_state *)calloc(1, sizeof(struct gpu_state));
    if (g == NULL) return gerror_(engine_err_malloc, "calloc failed");
    e->gpu = g;
    HIPCHECK(hipGetDevice(&g->device));
    HIPCHECK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    for (k = 0; k < 8; k++) HIPCHECK(hipEventCreate(&g->ev[k]));
    g->ncells = s->nr_cells;
    g->max_type = mt;

    /* Tables: one device table per distinct potential, FP32-rounded. */
    {
        struct potential **uniq = (struct potential **)calloc((size_t)mt * mt, sizeof(*uniq));
        int *pind = (int *)malloc(sizeof(int) * mt * mt);
        for (k = 0; k < mt * mt; k++) {
            struct potential *p = e->p[k];
            pind[k] = -1;
            if (p == NULL) continue;
            for (j = 0; j < ntab; j++)
                if (uniq[j] == p) break;
            if (j == ntab) {
                uniq[ntab++] = p;
                ncoef += (p->n + 1) * potential_chunk;
            }
            pind[k] = j;
        }
        float4 *tab = (float4 *)malloc(sizeof(float4) * (ntab > 0 ? ntab : 1));
        int *toff = (int *)malloc(sizeof(int) * (ntab > 0 ? ntab : 1));
        float *coef = (float *)malloc(sizeof(float) * (ncoef > 0 ? ncoef : 8));
        int pos = 0;
        for (j = 0; j < ntab; j++) {
            struct potential *p = uniq[j];
            int n = p->n;
            tab[j] = make_float4((float)p->alpha[0], (float)p->alpha[1], (float)p->alpha[2], 0.0f);
            memcpy(&tab[j].w, &n, sizeof(int));
            toff[j] = pos;
            for (k = 0; k < (n + 1) * potential_chunk; k++) coef[pos + k] = (float)p->c[k];
            pos += (n + 1) * potential_chunk;
        }
        g->ntab = ntab;
        HIPCHECK(hipMalloc(&g->d_pind, sizeof(int) * mt * mt));
        HIPCHECK(hipMalloc(&g->d_tab, sizeof(float4) * (ntab > 0 ? ntab : 1)));
        HIPCHECK(hipMalloc(&g->d_toff, sizeof(int) * (ntab > 0 ? ntab : 1)));
        HIPCHECK(hipMalloc(&g->d_coeffs, sizeof(float) * (ncoef > 0 ? ncoef : 8)));
        HIPCHECK(hipMemcpy(g->d_pind, pind, sizeof(int) * mt * mt, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_tab, tab, sizeof(float4) * (ntab > 0 ? ntab : 1), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_toff, toff, sizeof(int) * (ntab > 0 ? ntab : 1), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_coeffs, coef, sizeof(float) * (ncoef > 0 ? ncoef : 8), hipMemcpyHostToDevice));
        free(uniq);
        free(pind);
        free(tab);
        free(toff);
        free(coef);
    }

    /* Geometry constants of the sorted window (runner_dopair.c:424-428). */
    {
        struct nb_params *P = &g->P;
        memset(P, 0, sizeof(*P));
        P->ncells = s->nr_cells;
        P->max_type = mt;
        P->ntab = ntab;
        P->proj = (e->flags & engine_flag_projwindow) ? 1 : 0;
        P->bias = sqrt(s->h[0] * s->h[0] + s->h[1] * s->h[1] + s->h[2] * s->h[2]);
        P->dscale = (double)32767 / (2 * P->bias);
        P->dmaxdist = (int)(2 + P->dscale * (s->cutoff + 0.0 + 0.0));
        P->cutoff2 = s->cutoff * s->cutoff;
        for (k = 0; k < 3; k++) {
            P->dim[k] = s->dim[k];
            P->h[k] = s->h[k];
        }
        P->rc2f = (float)P->cutoff2;
        /* FP32 r^2 of cell-relative coordinates is within ~2e-6 rc^2 of the
           FP64 value; pairs inside this band are decided in FP64. */
        P->band = 2.0e-5f * P->rc2f;
        HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_axis), cell_shift, sizeof(double) * 39));
    }

    /* Owner neighbour lists from the task list. */
    {
        int *cnt = (int *)calloc(s->nr_cells + 1, sizeof(int));
        int *fill;
        int4 *nbr;
        int total = 0;
        for (k = 0; k < s->nr_tasks; k++) {
            struct task *t = &s->tasks[k];
            if (t->type == task_type_self)
                cnt[t->i]++;
            else if (t->type == task_type_pair) {
                cnt[t->i]++;
                cnt[t->j]++;
            }
        }
        for (cid = 0; cid < s->nr_cells; cid++) total += cnt[cid];
        nbr = (int4 *)malloc(sizeof(int4) * (total > 0 ? total : 1));
        fill = (int *)malloc(sizeof(int) * (s->nr_cells + 1));
        fill[0] = 0;
        for (cid = 0; cid < s->nr_cells; cid++) fill[cid + 1] = fill[cid] + cnt[cid];
        int *off = (int *)malloc(sizeof(int) * (s->nr_cells + 1));
        memcpy(off, fill, sizeof(int) * (s->nr_cells + 1));
        for (k = 0; k < s->nr_tasks; k++) {
            struct task *t = &s->tasks[k];
            if (t->type == task_type_self) {
                nbr[fill[t->i]++] = make_int4(t->i, 1 << 5, 0, 0);
            } else if (t->type == task_type_pair) {
                int dr, dp;
                task_dnshift(s, t->i, t->j, t->flags, g->P.dscale, &dr, &dp);
                nbr[fill[t->i]++] = make_int4(t->j, t->flags, dr, dp);
                nbr[fill[t->j]++] = make_int4(t->i, t->flags | (1 << 4), dr, dp);
            }
        }
        HIPCHECK(hipMalloc(&g->d_nbr, sizeof(int4) * (total > 0 ? total : 1)));
        HIPCHECK(hipMalloc(&g->d_nbr_off, sizeof(int) * (s->nr_cells + 1)));
        HIPCHECK(hipMemcpy(g->d_nbr, nbr, sizeof(int4) * (total > 0 ? total : 1), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_nbr_off, off, sizeof(int) * (s->nr_cells + 1), hipMemcpyHostToDevice));
        free(cnt);
        free(fill);
        free(off);
        free(nbr);
    }

    /* Cell origins, locations and sort masks. */
    {
        double *org = (double *)malloc(sizeof(double) * 3 * s->nr_cells);
        int *loc = (int *)malloc(sizeof(int) * 3 * s->nr_cells);
        unsigned int *sf = (unsigned int *)malloc(sizeof(unsigned int) * s->nr_cells);
        for (cid = 0; cid < s->nr_cells; cid++) {
            for (k = 0; k < 3; k++) {
                org[3 * cid + k] = s->cells[cid].origin[k];
                loc[3 * cid + k] = s->cells[cid].loc[k];
            }
            sf[cid] = s->cells[cid].sortflags;
        }
        HIPCHECK(hipMalloc(&g->d_origin, sizeof(double) * 3 * s->nr_cells));
        HIPCHECK(hipMalloc(&g->d_loc, sizeof(int) * 3 * s->nr_cells));
        HIPCHECK(hipMalloc(&g->d_sortflags, sizeof(unsigned int) * s->nr_cells));
        HIPCHECK(hipMemcpy(g->d_origin, org, sizeof(double) * 3 * s->nr_cells, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_loc, loc, sizeof(int) * 3 * s->nr_cells, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_sortflags, sf, sizeof(unsigned int) * s->nr_cells, hipMemcpyHostToDevice));
        free(org);
        free(loc);
        free(sf);
    }

    /* Exclusions as per-particle partner lists, by particle id. */
    {
        int maxid = -1;
        for (k = 0; k < s->size_parts; k++)
            if (s->partlist[k] != NULL) maxid = k;
        for (k = 0; k < e->nr_exclusions; k++) {
            if (e->exclusions[k].i > maxid) maxid = e->exclusions[k].i;
            if (e->exclusions[k].j > maxid) maxid = e->exclusions[k].j;
        }
        int nids = maxid + 1;
        int *xo = (int *)calloc(nids + 2, sizeof(int));
        int *xl = (int *)malloc(sizeof(int) * (2 * e->nr_exclusions + 1));
        int *fill = (int *)malloc(sizeof(int) * (nids + 1));
        for (k = 0; k < e->nr_exclusions; k++) {
            if (e->exclusions[k].i < 0 || e->exclusions[k].j < 0) continue;
            xo[e->exclusions[k].i + 1]++;
            xo[e->exclusions[k].j + 1]++;
        }
        for (k = 0; k < nids; k++) xo[k + 1] += xo[k];
        memcpy(fill, xo, sizeof(int) * (nids + 1));
        for (k = 0; k < e->nr_exclusions; k++) {
            i = e->exclusions[k].i;
            j = e->exclusions[k].j;
            if (i < 0 || j < 0) continue;
            xl[fill[i]++] = j;
            xl[fill[j]++] = i;
        }
        HIPCHECK(hipMalloc(&g->d_excl_off, sizeof(int) * (nids + 1)));
        HIPCHECK(hipMalloc(&g->d_excl_list, sizeof(int) * (2 * e->nr_exclusions + 1)));
        HIPCHECK(hipMemcpy(g->d_excl_off, xo, sizeof(int) * (nids + 1), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_excl_list, xl, sizeof(int) * (2 * e->nr_exclusions + 1), hipMemcpyHostToDevice));
        free(xo);
        free(xl);
        free(fill);
        g->nids = nids;
        if (gpu_alloc_ids(g, nids > 0 ? nids : 1) < 0) return engine_err_gpu;
    }

    HIPCHECK(hipMalloc(&g->d_cell_acc, sizeof(double) * 3 * s->nr_cells));
    HIPCHECK(hipMalloc(&g->d_cell_cnt, sizeof(unsigned long long) * 3 * s->nr_cells));
    HIPCHECK(hipMalloc(&g->d_tot, sizeof(double) * 4));
    HIPCHECK(hipMalloc(&g->d_tot_cnt, sizeof(unsigned long long) * 4));
    HIPCHECK(hipMalloc(&g->d_queue, sizeof(int) * (MD_NQUEUE + 1)));
    HIPCHECK(hipHostMalloc((void **)&g->h_count, sizeof(int) * s->nr_cells, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_offset, sizeof(int) * (s->nr_cells + 1), hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_tot, sizeof(double) * 4, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_tot_cnt, sizeof(unsigned long long) * 4, hipHostMallocDefault));

    {
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, g->device));
        g->ncu = prop.multiProcessorCount;
    }
    g->lds_nb = 0;
    return engine_err_ok;
}

/* Pack host particles in cell order and upload them (no kernels). */
static int gpu_upload(struct engine *e) {
    struct gpu_state *g = (struct gpu_state *)e->gpu;
    struct space *s = &e->s;
    int cid, k, n = 0, maxc = 0;
    for (cid = 0; cid < s->nr_cells; cid++) {
        int c = s->cells[cid].count;
        if (c > maxc) maxc = c;
        n += c;
    }
    if (maxc > MD_MAX_CELL) return gerror_(engine_err_maxparts, "more than 1024 particles in a cell");
    if (maxc >= 65536) return gerror_(engine_err_maxparts, "sortlist index exceeds 16 bits");
    if (gpu_alloc_parts(g, n > 0 ? n : 1) < 0) return engine_err_gpu;
    int maxid = -1;
    for (cid = 0; cid < s->nr_cells; cid++) {
        struct cell *c = &s->cells[cid];
        g->h_count[cid] = c->count;
        g->h_offset[cid] = (cid == 0) ? 0 : g->h_offset[cid - 1] + s->cells[cid - 1].count;
        for (k = 0; k < c->count; k++) {
            int a = g->h_offset[cid] + k;
            g->h_x[3 * a] = c->parts[k].x[0];
            g->h_x[3 * a + 1] = c->parts[k].x[1];
            g->h_x[3 * a + 2] = c->parts[k].x[2];
            g->h_type[a] = c->parts[k].type;
            g->h_id[a] = c->parts[k].id;
            if (c->parts[k].id > maxid) maxid = c->parts[k].id;
        }
    }
    g->h_offset[s->nr_cells] = n;
    if (maxid + 1 > g->nids) return gerror_(engine_err_range, "particle id beyond the ids known at engine_gpu_load");
    g->nparts = n;
    g->maxcount = maxc;
    g->P.nparts = n;
    g->P.nids = g->nids;
    HIPCHECK(hipMemcpyAsync(g->d_x, g->h_x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_type, g->h_type, sizeof(int) * n, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_id, g->h_id, sizeof(int) * n, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_count, g->h_count, sizeof(int) * s->nr_cells, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_offset, g->h_offset, sizeof(int) * (s->nr_cells + 1), hipMemcpyHostToDevice,
                            g->stream));
    return engine_err_ok;
}

/* Launch the four kernels; optional events around each. */
static int gpu_launch(struct gpu_state *g, int timed) {
    struct nb_params P = g->P;
    const int mt = g->max_type;
    int nb;
    size_t lds = 0;
    lds += sizeof(int) * mt * mt;
    lds = (lds + 15) & ~(size_t)15;
    P.lds_tab = (int)lds;
    lds += sizeof(float4) * (g->ntab > 0 ? g->ntab : 1);
    P.lds_toff = (int)lds;
    lds += sizeof(int) * (g->ntab > 0 ? g->ntab : 1);
    lds = (lds + 15) & ~(size_t)15;
    P.lds_pos = (int)lds;
    lds += sizeof(float4) * (g->maxcount > 0 ? g->maxcount : 1);
    P.lds_key = (int)lds;
    lds += sizeof(int2) * (g->maxcount > 0 ? g->maxcount : 1);
    if (lds > 64 * 1024) return gerror_(engine_err_maxparts, "LDS staging exceeds 64 KiB");
    if (lds != g->lds_nb) {
        /* Persistent grid: as many single-wave blocks as can be resident. */
        int per_cu = 0;
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond, 64, lds));
        if (per_cu < 1) per_cu = 1;
        g->grid_nb = g->ncu * per_cu;
        g->lds_nb = lds;
    }

    if (timed) HIPCHECK(hipEventRecord(g->ev[0], g->stream));
    nb = (g->nids + 255) / 256;
    if (nb > 0) hipLaunchKernelGGL(k_clear_ids, dim3(nb), dim3(256), 0, g->stream, g->nids, g->d_index_of_id);
    nb = ((g->nparts > g->ncells ? g->nparts : g->ncells) + 255) / 256;
    hipLaunchKernelGGL(k_pack, dim3(nb), dim3(256), 0, g->stream, g->nparts, g->d_x, g->d_type, g->d_id, g->ncells,
                       g->d_offset, g->d_xf, g->d_index_of_id, g->d_cell_of, g->d_queue);
    nb = (g->ncells * 13 + 255) / 256;
    hipLaunchKernelGGL(k_sort, dim3(nb), dim3(256), 0, g->stream, P, g->d_count, g->d_offset, g->d_sortflags,
                       g->d_x, g->d_sortlist, g->d_keyof, g->d_queue + MD_NQUEUE);
    if (timed) HIPCHECK(hipEventRecord(g->ev[1], g->stream));
    hipLaunchKernelGGL(k_nonbond, dim3(g->grid_nb), dim3(64), lds, g->stream, P, g->d_count, g->d_offset, g->d_nbr,
                       g->d_nbr_off, g->d_origin, g->d_loc, g->d_xf, g->d_x, g->d_id, g->d_sortlist, g->d_keyof,
                       g->d_pind, g->d_tab, g->d_toff, g->d_coeffs, g->d_excl_off, g->d_excl_list,
                       g->d_index_of_id, g->d_cell_of, g->d_fhi, g->d_flo, g->d_cell_acc, g->d_cell_cnt,
                       g->d_queue);
    if (timed) HIPCHECK(hipEventRecord(g->ev[2], g->stream));
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, g->stream, g->ncells, g->d_cell_acc, g->d_cell_cnt, g->d_tot,
                       g->d_tot_cnt);
    if (timed) HIPCHECK(hipEventRecord(g->ev[3], g->stream));
    HIPCHECK(hipGetLastError());
    return engine_err_ok;
}

extern "C" int engine_gpu_sync(struct engine *e, unsigned int what) {
    struct gpu_state *g;
    struct space *s;
    int cid;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    g = (struct gpu_state *)e->gpu;
    if (g == NULL) return gerror_(engine_err_nogpu, "engine_gpu_load was not called");
    s = &e->s;
    HIPCHECK(hipSetDevice(g->device));
    if (what & engine_gpu_sync_parts_up) {
        if (gpu_upload(e) < 0) return engine_err_gpu;
        HIPCHECK(hipStreamSynchronize(g->stream));
    }
    if (what & engine_gpu_sync_sort_down) {
        unsigned int *tmp = (unsigned int *)malloc(sizeof(unsigned int) * 13 * (size_t)(g->nparts > 0 ? g->nparts : 1));
        if (tmp == NULL) return gerror_(engine_err_malloc, "malloc failed");
        if (hipMemcpy(tmp, g->d_sortlist, sizeof(unsigned int) * 13 * (size_t)g->nparts, hipMemcpyDeviceToHost) !=
            hipSuccess) {
            free(tmp);
            return gerror_(engine_err_gpu, "sortlist download failed");
        }
        for (cid = 0; cid < s->nr_cells; cid++) {
            struct cell *c = &s->cells[cid];
            if (c->count != g->h_count[cid]) {
                free(tmp);
                return gerror_(engine_err_range, "cells changed since the last pass");
            }
            memcpy(c->sortlist, &tmp[13 * (size_t)g->h_offset[cid]], sizeof(unsigned int) * 13 * c->count);
        }
        free(tmp);
    }
    if (what & engine_gpu_sync_forces_down) {
        HIPCHECK(hipMemcpyAsync(g->h_fhi, g->d_fhi, sizeof(float4) * g->nparts, hipMemcpyDeviceToHost, g->stream));
        HIPCHECK(hipMemcpyAsync(g->h_flo, g->d_flo, sizeof(float4) * g->nparts, hipMemcpyDeviceToHost, g->stream));
        HIPCHECK(hipMemcpyAsync(g->h_tot, g->d_tot, sizeof(double) * 3, hipMemcpyDeviceToHost, g->stream));
        HIPCHECK(hipMemcpyAsync(g->h_tot_cnt, g->d_tot_cnt, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost,
                                g->stream));
        HIPCHECK(hipStreamSynchronize(g->stream));
    }
    return engine_err_ok;
}

extern "C" int engine_nonbond_gpu(struct engine *e) {
    struct gpu_state *g;
    struct space *s;
    int cid, k, flag = 0;
    double t0, t1;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    g = (struct gpu_state *)e->gpu;
    if (g == NULL) return gerror_(engine_err_nogpu, "engine_gpu_load was not called");
    s = &e->s;
    HIPCHECK(hipSetDevice(g->device));
    if (gpu_upload(e) < 0) return engine_err_gpu;
    HIPCHECK(hipMemsetAsync(g->d_queue + MD_NQUEUE, 0, sizeof(int), g->stream));
    if (gpu_launch(g, 1) < 0) return engine_err_gpu;
    HIPCHECK(hipMemcpyAsync(g->h_fhi, g->d_fhi, sizeof(float4) * g->nparts, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipMemcpyAsync(g->h_flo, g->d_flo, sizeof(float4) * g->nparts, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipMemcpyAsync(g->h_tot, g->d_tot, sizeof(double) * 3, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipMemcpyAsync(g->h_tot_cnt, g->d_tot_cnt, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost,
                            g->stream));
    HIPCHECK(hipMemcpyAsync(&flag, g->d_queue + MD_NQUEUE, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    if (flag) return gerror_(engine_err_range, "sort stack overflow");
    {
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, g->ev[0], g->ev[3]));
        e->timers[engine_timer_gpu_kernels] = ms * 1e-3;
    }
    t0 = 0.0;
    (void)t0;
    (void)t1;
    for (cid = 0; cid < s->nr_cells; cid++) {
        struct cell *c = &s->cells[cid];
        for (k = 0; k < c->count; k++) {
            int a = g->h_offset[cid] + k;
            c->parts[k].f[0] += (double)g->h_fhi[a].x + (double)g->h_flo[a].x;
            c->parts[k].f[1] += (double)g->h_fhi[a].y + (double)g->h_flo[a].y;
            c->parts[k].f[2] += (double)g->h_fhi[a].z + (double)g->h_flo[a].z;
        }
    }
    s->epot_nonbond += g->h_tot[0];
    s->epot_exclusion += g->h_tot[1];
    s->epot += g->h_tot[0] + g->h_tot[1];
    s->virial += g->h_tot[2];
    s->nr_pairs += (long long)g->h_tot_cnt[0];
    s->nr_excl_evaluated += (long long)g->h_tot_cnt[1];
    return engine_err_ok;
}

extern "C" int engine_gpu_time(struct engine *e, int nsteps, int nwarm, double *ms) {
    struct gpu_state *g;
    double acc[4] = {0, 0, 0, 0};
    float t;
    int k;
    if (e == NULL || ms == NULL) return gerror_(engine_err_null, "NULL argument");
    g = (struct gpu_state *)e->gpu;
    if (g == NULL || g->nparts_cap == 0) return gerror_(engine_err_nogpu, "no resident particles");
    HIPCHECK(hipSetDevice(g->device));
    for (k = 0; k < nwarm; k++)
        if (gpu_launch(g, 0) < 0) return engine_err_gpu;
    HIPCHECK(hipStreamSynchronize(g->stream));
    for (k = 0; k < nsteps; k++) {
        if (gpu_launch(g, 1) < 0) return engine_err_gpu;
        HIPCHECK(hipEventSynchronize(g->ev[3]));
        HIPCHECK(hipEventElapsedTime(&t, g->ev[0], g->ev[1]));
        acc[0] += t;
        HIPCHECK(hipEventElapsedTime(&t, g->ev[1], g->ev[2]));
        acc[1] += t;
        HIPCHECK(hipEventElapsedTime(&t, g->ev[2], g->ev[3]));
        acc[2] += t;
        HIPCHECK(hipEventElapsedTime(&t, g->ev[0], g->ev[3]));
        acc[3] += t;
    }
    for (k = 0; k < 4; k++) ms[k] = (nsteps > 0) ? acc[k] / nsteps : 0.0;
    HIPCHECK(hipMemcpyAsync(g->h_tot, g->d_tot, sizeof(double) * 3, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipMemcpyAsync(g->h_tot_cnt, g->d_tot_cnt, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost,
                            g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    return engine_err_ok;
}

extern "C" int engine_gpu_results(struct engine *e, double *out, long long *counts) {
    struct gpu_state *g;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    g = (struct gpu_state *)e->gpu;
    if (g == NULL) return gerror_(engine_err_nogpu, "engine_gpu_load was not called");
    if (out) {
        out[0] = g->h_tot[0];
        out[1] = g->h_tot[1];
        out[2] = g->h_tot[2];
    }
    if (counts) {
        counts[0] = (long long)g->h_tot_cnt[0];
        counts[1] = (long long)g->h_tot_cnt[1];
        counts[2] = (long long)g->h_tot_cnt[2];
    }
    return engine_err_ok;
}
