/*
 * gfx950 nonbonded engine behind the mdcore C API (C-ABI in engine_gpu.h).
 *
 * One nonbonded pass is a chain of launches on one stream:
 *   k_pack      cell-ordered FP64 positions -> float4 (x, y, z, type bits),
 *               id -> slot map, slot -> cell map;
 *   k_sort_rank runner_dosort (runner_dosort.c:97-161) for every cell and
 *               every flagged direction: 16-bit projection keys computed in
 *               FP64 without contraction; particles with a key of their own
 *               are placed by rank, and a list with equal keys is sorted by
 *               the reference's descending quicksort (runner.c:110-174)
 *               replayed, so that the sorted lists, ties included, are the
 *               reference's (k_sort: the same per list in global memory, for
 *               cells of more than 480 particles).  Cells of more than 64
 *               particles get a wave per direction and a wave-parallel
 *               restatement of that quicksort (sort_desc_wave); smaller
 *               cells one wave that sorts all their lists (sort_lists_wave);
 *   k_nonbond   persistent kernel: each wave pulls an owner cell from one of
 *               8 device queues (one per XCD slab of cells), walks the cell's
 *               neighbour list (the reference's self/pair tasks that touch
 *               it) and accumulates the forces of its own atoms only, so no
 *               atomics on forces and no cross-wave write conflicts.  Lanes
 *               are (owner atom, partition); per neighbour cell the positions
 *               and sorted words are prefetched one neighbour ahead and the
 *               cell is staged in LDS in window order as packed records
 *               (x, y, z, key | index << 16 | type << tshift).  Each lane walks
 *               its share of the sorted window (runner_dopair.c:455-478),
 *               keeps every record with r^2 up to the top of a thin band
 *               around rc^2 in a lane-private LDS list, then evaluates them
 *               two at a time, the band's pairs decided in FP64 first: FP32 table evaluation
 *               (potential_eval.h:67-104), forces summed in registers and
 *               folded into two-sum (hi, lo) accumulators per neighbour.
 *               Excluded pairs are subtracted afterwards as in
 *               exclusion.c:81-282.  The queues hand out work items: owner
 *               cells in 6x6x6 blocks, each split over up to 8 neighbour
 *               groups when there are few, densely filled cells;
 *   k_nonbond_half (and the surrogate k_nonbond_mlp_half): each pair task
 *               is visited once, at its i cell; the opposite forces go into
 *               LDS copies of the visited cell and out to a per-visit slot;
 *   k_fold      per atom, the groups' partial forces and the half-shell
 *               slots, added in a fixed order;
 *   k_reduce_a/b per-item FP64 energies/virials/counts -> totals, fixed order.
 * Staged neighbour records hold their positions moved into the owner's
 * frame (x + shift, once per record), so a candidate's distance is one
 * subtraction per axis.  In the full-shell kernels every pair is evaluated
 * by both of its cells (once per owner); the two owners' FP32 distances are
 * opposite to rounding (each side rounds its own shifted position), so the
 * forces sum to zero to FP32 rounding; energy, virial and counts are taken
 * on the task's i side only.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gpu_state.h"
#include "hostpack.h"

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
__device__ __forceinline__ double r2_task64(const double *xi, const double *xj, double s0, double s1, double s2) {
#pragma clang fp contract(off)
    double p0 = xi[0] - s0, p1 = xi[1] - s1, p2 = xi[2] - s2;
    double d0 = p0 - xj[0], d1 = p1 - xj[1], d2 = p2 - xj[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

/* Reference r^2 of an exclusion (exclusion.c:139-147): i is the smaller id. */
__device__ __forceinline__ double r2_excl64(const double *xi, const double *xj, const int *li, const int *lj,
                                            const double *h) {
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

/* Table evaluation in FP32: e and f = (dE/dr)/r (potential_eval.h:67-104).
   hdr = (alpha0, alpha1, alpha2, bits(n | chunk0 << 9)). */
__device__ __forceinline__ void pot_eval32(const float4 hdr, const float4 *__restrict__ c4, float r2, float &e,
                                           float &f) {
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * ir;
    const float fi = fmaxf(0.0f, hdr.x + r * (hdr.y + r * hdr.z));
    const unsigned int w = __float_as_uint(hdr.w);
    const int n = (int)(w & 511u);
    int ind = (int)fi;
    ind = ind > n ? n : ind;
    const float4 *c = c4 + 2 * ((int)(w >> 9) + ind);
    const float4 a = c[0], b = c[1];
    const float x = (r - a.x) * a.y;
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

/* Buffer-resource loads: 32-bit byte offsets instead of 64-bit addresses
   (word 3 of the descriptor as for every gfx9 target); an offset past the
   declared size reads zeros. */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, bytes, 0x00020000);
}

__device__ __forceinline__ float4 ld_f4(__amdgpu_buffer_rsrc_t r, int byte_off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0);
    float4 out;
    __builtin_memcpy(&out, &v, sizeof(out));
    return out;
}

/* pot_eval32 with the coefficients behind a buffer resource. */
__device__ __forceinline__ void pot_eval32r(const float4 hdr, __amdgpu_buffer_rsrc_t rc, float r2, float &e,
                                            float &f) {
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * ir;
    const float fi = fmaxf(0.0f, hdr.x + r * (hdr.y + r * hdr.z));
    const unsigned int w = __float_as_uint(hdr.w);
    const int n = (int)(w & 511u);
    int ind = (int)fi;
    ind = ind > n ? n : ind;
    const int off = 32 * ((int)(w >> 9) + ind);
    const float4 a = ld_f4(rc, off), b = ld_f4(rc, off + 16);
    const float x = (r - a.x) * a.y;
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

/* pot_eval32r for two pairs at once, without branches: both coefficient
   rows are requested before either polynomial is evaluated, so the two
   table lookups overlap.  A pair with use_* false (no hit, or no table for
   its types) reads interval 0 of the first table instead and returns
   e = f = 0; the arithmetic of a used pair is that of pot_eval32r. */
__device__ __forceinline__ void pot_eval32r2(const float4 ha, const float4 hb, bool use_a, bool use_b,
                                             __amdgpu_buffer_rsrc_t rc, float r2a, float r2b, float &ea,
                                             float &fa, float &eb, float &fb) {
    const float ira = __builtin_amdgcn_rsqf(r2a), irb = __builtin_amdgcn_rsqf(r2b);
    const float ra = r2a * ira, rb = r2b * irb;
    const unsigned int wa = use_a ? __float_as_uint(ha.w) : 0u;
    const unsigned int wb = use_b ? __float_as_uint(hb.w) : 0u;
    int ia = (int)fmaxf(0.0f, ha.x + ra * (ha.y + ra * ha.z));
    int ib = (int)fmaxf(0.0f, hb.x + rb * (hb.y + rb * hb.z));
    ia = ia > (int)(wa & 511u) ? (int)(wa & 511u) : ia;
    ib = ib > (int)(wb & 511u) ? (int)(wb & 511u) : ib;
    const int offa = 32 * ((int)(wa >> 9) + ia), offb = 32 * ((int)(wb >> 9) + ib);
    const float4 a0 = ld_f4(rc, offa), a1 = ld_f4(rc, offa + 16);
    const float4 b0 = ld_f4(rc, offb), b1 = ld_f4(rc, offb + 16);
    auto horner = [](float r, float ir, const float4 c0, const float4 c1, float &e, float &f) {
        const float x = (r - c0.x) * c0.y;
        float ee = c0.z * x + c0.w;
        float eff = c0.z;
        eff = eff * x + ee;
        ee = ee * x + c1.x;
        eff = eff * x + ee;
        ee = ee * x + c1.y;
        eff = eff * x + ee;
        ee = ee * x + c1.z;
        eff = eff * x + ee;
        ee = ee * x + c1.w;
        e = ee;
        f = eff * c0.y * ir;
    };
    float e1, f1, e2, f2;
    horner(ra, ira, a0, a1, e1, f1);
    horner(rb, irb, b0, b1, e2, f2);
    ea = use_a ? e1 : 0.f;
    fa = use_a ? f1 : 0.f;
    eb = use_b ? e2 : 0.f;
    fb = use_b ? f2 : 0.f;
}

/* The pair kernel's workgroup is a single wavefront, and the LDS executes a
   wavefront's instructions in order: data written by one lane is visible to
   another lane's later read once this wave's LDS operations have completed.
   This waits for LDS only (lgkmcnt(0)) and keeps the compiler from moving
   memory accesses across it; unlike a workgroup barrier it does not drain
   outstanding global loads, so prefetches stay in flight. */
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* Packed staged record: key (16 bits) | index in cell (10 bits) | type (6 bits). */
__device__ __forceinline__ unsigned int rec_key(unsigned int w) { return w & 0xffffu; }
__device__ __forceinline__ int rec_idx(unsigned int w, int imask) { return (int)((w >> 16) & (unsigned int)imask); }
__device__ __forceinline__ int rec_type(unsigned int w, int tshift) { return (int)(w >> tshift); }

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

#define MD_SORT_STACK 16

/* Descending sort of (index<<16 | key) words by key: the reference's
   quicksort with selection sort below 16 elements and an explicit stack
   (runner.c:110-174).  The partition and push order decide the order of
   equal keys, so they are kept as they are. */
template <typename A, typename S>
__device__ __forceinline__ void sort_desc_t(A a, int N, S st_lo, S st_hi, int *overflow) {
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
            if (qpos + 2 >= MD_SORT_STACK) {
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

__device__ void sort_desc(unsigned int *a, int N, int *overflow) {
    int st_lo[MD_SORT_STACK], st_hi[MD_SORT_STACK];
    sort_desc_t(a, N, st_lo, st_hi, overflow);
}

__constant__ double c_axis[39];

/* Projection key (runner_dosort.c:146-149): dscale*(bias + x.n) in FP64,
   left to right, no fused multiply-add, truncated to unsigned. */
__device__ __forceinline__ unsigned int sort_key(const double *x, int sid, double dscale, double bias) {
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
    const size_t seg = 13 * (size_t)off + (size_t)sid * n;
    unsigned int *list = sortlist + seg;
    for (int i = 0; i < n; i++) {
        unsigned int key = sort_key(&x[3 * (size_t)(off + i)], sid, P.dscale, P.bias);
        list[i] = ((unsigned int)i << 16) | key;
        keyof[(size_t)sid * P.nparts + off + i] = (unsigned short)key;
    }
    sort_desc(list, n, flag);
}

/* The quicksort of sort_desc_t, run by a whole wave on one LDS list with the
   same result, element for element.
   - Segments below 16 elements (selection sort): each step takes the first
     maximal key of the rest, the max of key << 6 | (63 - index) over the
     leaf's lanes (a DPP reduction within the first row of 16 lanes), and
     swaps it to the front with lane reads.
   - Partition (pivot p = middle key): the serial scan's k-th swap pairs the
     k-th element from the left with key <= p (ltab[k]) and the k-th from the
     right with key >= p (rtab[k]), for k = 1..K while ltab[k] <= rtab[k];
     both tables come from ballots and prefix counts.  The scan then stops
     with i = ltab[K] + 1, j = rtab[K] - 1 if these have crossed, else at
     i = min(ltab[K+1], rtab[K]), j = max(rtab[K+1], ltab[K]) (the swapped
     elements stop the pointers), without a further swap.
   The stack (same pushes and overflow test as sort_desc_t) lives in LDS and
   is written by lane 0.  ltab / rtab: maxn + 1 entries each. */

/* Largest value over lanes 0..15, wave-uniform: four DPP row shifts (lane l
   takes the max of lanes l-1, l-2..l-3, l-4..l-7, l-8..l-15 of its row;
   lanes shifted in from outside the row read 0), then lane 15. */
__device__ __forceinline__ unsigned int row0_max_u32(unsigned int v) {
    unsigned int u;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true); /* row_shr:1 */
    v = u > v ? u : v;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true); /* row_shr:2 */
    v = u > v ? u : v;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true); /* row_shr:4 */
    v = u > v ? u : v;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true); /* row_shr:8 */
    v = u > v ? u : v;
    return (unsigned int)__builtin_amdgcn_readlane((int)v, 15);
}

/* One partition step of sort_desc_t on a[lo..hi] (hi - lo >= 15, at most
   64 R elements) by a whole wave, pivot p = the middle key; returns the
   scan's final (i, j).  ltab / rtab: hi - lo + 2 entries each. */
template <int R>
__device__ __forceinline__ int2 part_wave(unsigned int *a, int lo, int hi, unsigned short *ltab,
                                          unsigned short *rtab, int lane) {
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned int p = a[(lo + hi) / 2] & 0xffffu;
    unsigned int wv[R];
    bool isl[R], isr[R];
    int nR = 0;
    for (int r = 0; r < R; r++) {
        const int x = lo + lane + 64 * r;
        const bool valid = x <= hi;
        wv[r] = valid ? a[x] : 0u;
        isl[r] = valid && (wv[r] & 0xffffu) <= p;
        isr[r] = valid && (wv[r] & 0xffffu) >= p;
        nR += __builtin_popcountll(__builtin_amdgcn_ballot_w64(isr[r]));
    }
    int cl = 0, cr = 0;
    for (int r = 0; r < R; r++) {
        const int x = lo + lane + 64 * r;
        const unsigned long long ml = __builtin_amdgcn_ballot_w64(isl[r]);
        const unsigned long long mr = __builtin_amdgcn_ballot_w64(isr[r]);
        if (isl[r]) ltab[cl + __builtin_popcountll(ml & below) + 1] = (unsigned short)x;
        if (isr[r]) rtab[nR - cr - __builtin_popcountll(mr & below)] = (unsigned short)x;
        cl += __builtin_popcountll(ml);
        cr += __builtin_popcountll(mr);
    }
    const int nL = cl, m = nL < nR ? nL : nR;
    wave_lds_sync();
    int K = 0;
    int li[R], ri[R];
    for (int r = 0; r < R; r++) {
        const int k = 1 + lane + 64 * r;
        li[r] = k <= m ? ltab[k] : 0;
        ri[r] = k <= m ? rtab[k] : 0;
        K += __builtin_popcountll(__builtin_amdgcn_ballot_w64(k <= m && li[r] <= ri[r]));
    }
    /* the K swaps touch distinct elements, each lane its own */
    for (int r = 0; r < R; r++) {
        const int k = 1 + lane + 64 * r;
        if (k <= K && li[r] < ri[r]) {
            const unsigned int va = a[li[r]], vb = a[ri[r]];
            a[li[r]] = vb;
            a[ri[r]] = va;
        }
    }
    wave_lds_sync();
    const int lK = ltab[K], rK = rtab[K];
    int i = lK + 1, j = rK - 1;
    if (i <= j) {
        const int l1 = K + 1 <= nL ? (int)ltab[K + 1] : 1 << 30;
        const int r1 = K + 1 <= nR ? (int)rtab[K + 1] : -(1 << 30);
        i = l1 < rK ? l1 : rK;
        j = r1 > lK ? r1 : lK;
    }
    return make_int2(i, j);
}

/* Two partition steps of part_wave at once, one per half wave: lanes 32h..
   32h+31 partition a[lo..hi] (16 to 32 elements; lo > hi leaves the half
   idle) with their half's tables ltab / rtab (33 entries each); lo, hi,
   ltab and rtab are the lane's half's.  Returns the half's final (i, j). */
__device__ __forceinline__ int2 part_wave2(unsigned int *a, int lo, int hi, unsigned short *ltab,
                                           unsigned short *rtab, int lane) {
    const unsigned long long half = (lane & 32) ? 0xffffffff00000000ull : 0x00000000ffffffffull;
    const unsigned long long below = ((1ull << lane) - 1ull) & half;
    const int x = lo + (lane & 31);
    const bool valid = x <= hi;
    const unsigned int p = a[(lo + hi) >> 1] & 0xffffu;
    const unsigned int w = valid ? a[x] : 0u;
    const bool isl = valid && (w & 0xffffu) <= p, isr = valid && (w & 0xffffu) >= p;
    const unsigned long long ml = __builtin_amdgcn_ballot_w64(isl) & half;
    const unsigned long long mr = __builtin_amdgcn_ballot_w64(isr) & half;
    const int nL = __builtin_popcountll(ml), nR = __builtin_popcountll(mr);
    if (isl) ltab[__builtin_popcountll(ml & below) + 1] = (unsigned short)x;
    if (isr) rtab[nR - __builtin_popcountll(mr & below)] = (unsigned short)x;
    const int m = nL < nR ? nL : nR;
    wave_lds_sync();
    const int k = 1 + (lane & 31);
    const int li = k <= m ? ltab[k] : 0, ri = k <= m ? rtab[k] : 0;
    const int K = __builtin_popcountll(__builtin_amdgcn_ballot_w64(k <= m && li <= ri) & half);
    if (k <= K && li < ri) {
        const unsigned int va = a[li], vb = a[ri];
        a[li] = vb;
        a[ri] = va;
    }
    wave_lds_sync();
    const int lK = ltab[K], rK = rtab[K];
    int i = lK + 1, j = rK - 1;
    if (i <= j) {
        const int l1 = K + 1 <= nL ? (int)ltab[K + 1] : 1 << 30;
        const int r1 = K + 1 <= nR ? (int)rtab[K + 1] : -(1 << 30);
        i = l1 < rK ? l1 : rK;
        j = r1 > lK ? r1 : lK;
    }
    return make_int2(i, j);
}

/* R: up to 64 R elements per partition. */
template <int R>
__device__ void sort_desc_wave(unsigned int *a, int N, int *st_lo, int *st_hi, unsigned short *ltab,
                               unsigned short *rtab, int *overflow, int lane) {
    int q = 0;
    if (lane == 0) {
        st_lo[0] = 0;
        st_hi[0] = N - 1;
    }
    wave_lds_sync();
    int pops = 0;
    while (q >= 0) {
        const int lo = st_lo[q], hi = st_hi[q];
        q -= 1;
        /* the serial sort pops at most 2N segments, all inside [0, N): anything
           else is reported as a sort failure instead of looping on */
        if (++pops > 2 * N + 2 || lo < 0 || hi >= N || lo > hi) {
            if (lane == 0) *overflow = 1;
            return;
        }
        if (hi - lo < 15) {
            const int len = hi - lo + 1;
            unsigned int w = lane < len ? a[lo + lane] : 0u;
            /* the leaf lies in lanes 0..14: a row-0 max and lane reads */
            for (int i = 0; i < len - 1; i++) {
                const unsigned int pk =
                    (lane >= i && lane < len) ? (((w & 0xffffu) << 6) | (unsigned int)(63 - lane)) : 0u;
                const int im = 63 - (int)(row0_max_u32(pk) & 63u);
                const unsigned int wi = (unsigned int)__builtin_amdgcn_readlane((int)w, i);
                const unsigned int wm = (unsigned int)__builtin_amdgcn_readlane((int)w, im);
                if (lane == i)
                    w = wm;
                else if (lane == im)
                    w = wi;
            }
            if (lane < len) a[lo + lane] = w;
            wave_lds_sync();
            continue;
        }
        const int2 ij = part_wave<R>(a, lo, hi, ltab, rtab, lane);
        const int i = ij.x, j = ij.y;
        if (q + 2 >= MD_SORT_STACK) {
            if (lane == 0) *overflow = 1;
            return;
        }
        const int mid = (lo + hi) / 2;
        const bool left_first = j > mid;
        /* pushes in sort_desc_t's order: [lo, j] then [i, hi] when j > mid */
        const bool pa = left_first ? lo < j : i < hi, pb = left_first ? i < hi : lo < j;
        const int alo = left_first ? lo : i, ahi = left_first ? j : hi;
        const int blo = left_first ? i : lo, bhi = left_first ? hi : j;
        if (pa) {
            q++;
            if (lane == 0) {
                st_lo[q] = alo;
                st_hi[q] = ahi;
            }
        }
        if (pb) {
            q++;
            if (lane == 0) {
                st_lo[q] = blo;
                st_hi[q] = bhi;
            }
        }
        wave_lds_sync();
    }
}

/* Largest value over the lane's row of 16 lanes, in every lane of the row:
   max with the row rotated by 1, 2, 4 and 8 lanes (DPP row_ror). */
__device__ __forceinline__ unsigned int row_max_all(unsigned int v) {
    unsigned int u;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xf, 0xf, false); /* row_ror:1 */
    v = u > v ? u : v;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xf, 0xf, false); /* row_ror:2 */
    v = u > v ? u : v;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xf, 0xf, false); /* row_ror:4 */
    v = u > v ? u : v;
    u = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xf, 0xf, false); /* row_ror:8 */
    v = u > v ? u : v;
    return v;
}

/* sort_desc_t on every list l of a cell with bit l of tm set (the lists of
   n <= 64 words at a + l * n), by one wave, with the same result.  A
   partition or a selection sort only touches its own segment, so the
   segments are taken in any order:
   - segments of 33 or more elements are partitioned by the whole wave
     (part_wave), those of 16 to 32 two at a time, one per half wave
     (part_wave2); their parts are queued again;
   - segments of 2..15 elements (the serial sort's selection-sort leaves)
     are sorted four at a time, one per row of 16 lanes: each step takes the
     first maximal key of the rest, the row max of
     key << 12 | (15 - row lane) << 8 | particle index
     (particle indices stay below 256), and swaps it with the step's lane,
     whose word every lane of the row gets from a second row max.
   Segment queues: segq (416 words: segments above 32 elements from the
   front, 16 to 32 from the middle) and leafq (leafcap words), lo | hi << 16
   in list-array positions.  Partition tables: tb, 136 entries.  Reports a
   queue overflow or a runaway loop in *overflow (the host turns it into a
   sort failure). */
__device__ void sort_lists_wave(unsigned int *a, int n, unsigned int tm, unsigned int *segq, unsigned int *leafq,
                                int leafcap, unsigned short *tb, unsigned int *__restrict__ out, int *overflow,
                                int lane) {
    constexpr int QHALF = 13 * MD_SORT_STACK;
    int nbig = 0, nsmall = 0, nl = 0;
    bool bad = false;
    auto push = [&](int lo, int hi) {
        const unsigned int w = (unsigned int)lo | ((unsigned int)hi << 16);
        if (hi - lo >= 32) {
            if (nbig < QHALF) {
                if (lane == 0) segq[nbig] = w;
                nbig++;
            } else
                bad = true;
        } else if (hi - lo >= 15) {
            if (nsmall < QHALF) {
                if (lane == 0) segq[QHALF + nsmall] = w;
                nsmall++;
            } else
                bad = true;
        } else if (hi > lo) {
            if (nl < leafcap) {
                if (lane == 0) leafq[nl] = w;
                nl++;
            } else
                bad = true;
        }
    };
    /* Only the marked elements (keys that others of their list share) are
       written from here on; rank placed the others.  A segment without a
       marked element is dropped.  With one, that element's place is the
       segment's start plus the number of greater keys in it (the others'
       keys differ from its own).  With more, the segment is queued, and its
       leaves write all of their elements in the end. */
    auto place = [&](int lo, int hi) {
        const int x = lo + lane;
        const bool valid = x <= hi;
        const unsigned int w = valid ? a[x] : 0u;
        const unsigned long long mk = __builtin_amdgcn_ballot_w64(valid && (w >> 31) != 0u);
        if (mk == 0ull) return;
        if ((mk & (mk - 1ull)) == 0ull) {
            const unsigned int wt = (unsigned int)__builtin_amdgcn_readlane((int)w, (int)__builtin_ctzll(mk));
            const int g = __builtin_popcountll(__builtin_amdgcn_ballot_w64(valid && (w & 0xffffu) > (wt & 0xffffu)));
            if (lane == 0) out[lo + g] = wt & 0x7fffffffu;
            return;
        }
        push(lo, hi);
    };
    /* the elements a partition leaves between its two parts are in place */
    auto settle = [&](int lo, int hi) {
        const int x = lo + lane;
        if (x <= hi) {
            const unsigned int w = a[x];
            if ((w >> 31) != 0u) out[x] = w & 0x7fffffffu;
        }
    };
    for (unsigned int t = tm; t; t &= t - 1) {
        const int l = __builtin_ctz(t);
        push(l * n, l * n + n - 1);
    }
    wave_lds_sync();
    /* each partition splits a segment into two strictly smaller ones */
    int parts = 0;
    while ((nbig > 0 || nsmall > 0) && !bad) {
        if (nbig > 0) {
            nbig--;
            const unsigned int sg = segq[nbig];
            const int lo = (int)(sg & 0xffffu), hi = (int)(sg >> 16);
            if (++parts > 13 * n) {
                bad = true;
                break;
            }
            const int2 ij = part_wave<1>(a, lo, hi, tb, tb + 68, lane);
            wave_lds_sync();
            place(lo, ij.y);
            place(ij.x, hi);
            settle(ij.y + 1, ij.x - 1);
        } else {
            const bool two = nsmall >= 2;
            const unsigned int s0 = segq[QHALF + nsmall - 1];
            const unsigned int s1 = two ? segq[QHALF + nsmall - 2] : 1u;
            nsmall -= two ? 2 : 1;
            parts += two ? 2 : 1;
            if (parts > 13 * n) {
                bad = true;
                break;
            }
            const unsigned int sm = (lane & 32) ? s1 : s0;
            const int lo = (int)(sm & 0xffffu), hi = (int)(sm >> 16);
            unsigned short *lt = tb + ((lane & 32) ? 68 : 0);
            const int2 ij = part_wave2(a, lo, hi, lt, lt + 34, lane);
            const int i0 = __builtin_amdgcn_readlane(ij.x, 0), j0 = __builtin_amdgcn_readlane(ij.y, 0);
            const int i1 = __builtin_amdgcn_readlane(ij.x, 32), j1 = __builtin_amdgcn_readlane(ij.y, 32);
            wave_lds_sync();
            place((int)(s0 & 0xffffu), j0);
            place(i0, (int)(s0 >> 16));
            settle(j0 + 1, i0 - 1);
            if (two) {
                place((int)(s1 & 0xffffu), j1);
                place(i1, (int)(s1 >> 16));
                settle(j1 + 1, i1 - 1);
            }
        }
        wave_lds_sync();
    }
    if (bad) {
        if (lane == 0) *overflow = 1;
        return;
    }
    const int row = lane >> 4, lr = lane & 15;
    for (int b = 0; b < nl; b += 4) {
        const bool has = b + row < nl;
        const unsigned int sg = has ? leafq[b + row] : 0u;
        const int lo = (int)(sg & 0xffffu), len = has ? (int)(sg >> 16) - lo + 1 : 0;
        unsigned int w = lr < len ? a[lo + lr] : 0u;
        int steps = __builtin_amdgcn_readlane(len, 0);
        steps = max(steps, __builtin_amdgcn_readlane(len, 16));
        steps = max(steps, __builtin_amdgcn_readlane(len, 32));
        steps = max(steps, __builtin_amdgcn_readlane(len, 48));
        for (int i = 0; i < steps - 1; i++) {
            const bool in = lr >= i && lr < len;
            const unsigned int pk =
                in ? ((w & 0xffffu) << 12) | ((unsigned int)(15 - lr) << 8) | ((w >> 16) & 0xffu) : 0u;
            const unsigned int mx = row_max_all(pk);
            const unsigned int wi = row_max_all(lr == i ? w : 0u);
            const int im = 15 - (int)((mx >> 8) & 15u);
            const unsigned int wm = ((mx & 0xffu) << 16) | (mx >> 12);
            if (i < len - 1) {
                if (lr == i)
                    w = wm;
                else if (lr == im)
                    w = wi;
            }
        }
        if (lr < len) out[lo + lr] = w & 0x7fffffffu;
    }
}

/* Sort by rank: one workgroup per cell, all 13 directions at once.  A key that no
   other particle of its list shares has exactly (number of greater keys)
   particles before it in any descending order, the reference's quicksort
   included, so those particles are placed by counting.  Only the order of
   equal keys depends on the quicksort's partitions: a list with equal keys
   is sorted again with the reference's algorithm (runner.c:110-174), in LDS,
   and written whole.  Cells of at most MD_SORT_ONEWAVE_MAX particles get one
   wave, which sorts all such lists of its cell together (sort_lists_wave);
   larger cells get 13 waves, wave l sorting list l (sort_desc_wave).  For
   cells of at most MD_SORT_LDS_MAX particles.  LDS: 3 * maxn positions,
   13 * maxn list words, 13 quicksort stacks, 13 pairs of partition tables
   and 136 more entries (the one-wave sort keeps its segment queues there). */
#define MD_SORT_LDS_MAX 480 /* 63.5 KiB of LDS at 480 */
#define MD_SORT_ONEWAVE_MAX 64
__global__ __launch_bounds__(1024) void k_sort_rank(struct nb_params P, const int *__restrict__ count,
                                                  const int *__restrict__ offset,
                                                  const unsigned int *__restrict__ sortflags,
                                                  const double *__restrict__ x, unsigned int *__restrict__ sortlist,
                                                  unsigned short *__restrict__ keyof, int *__restrict__ flag,
                                                  int maxn) {
    extern __shared__ double sh_rank[];
    double *xs = sh_rank;
    unsigned int *lists = (unsigned int *)(sh_rank + 3 * maxn);
    int *stk = (int *)(lists + 13 * maxn);
    unsigned short *tabs = (unsigned short *)(stk + 13 * 2 * MD_SORT_STACK);
    __shared__ int tie[13];
    const int c = blockIdx.x, lane = threadIdx.x;
    const int n = count[c];
    if (n == 0) return;
    const unsigned int fl = sortflags[c];
    const int off = offset[c];
    if (lane < 13) tie[lane] = 0;
    for (int k = lane; k < 3 * n; k += (int)blockDim.x) xs[k] = x[3 * (size_t)off + k];
    __syncthreads();
    for (int t = lane; t < 13 * n; t += (int)blockDim.x) {
        const int sid = t / n, i = t - sid * n;
        if (!((fl >> sid) & 1u)) continue;
        const unsigned int key = sort_key(&xs[3 * i], sid, P.dscale, P.bias);
        lists[sid * n + i] = ((unsigned int)i << 16) | key;
        keyof[(size_t)sid * P.nparts + off + i] = (unsigned short)key;
    }
    __syncthreads();
    unsigned int *out = sortlist + 13 * (size_t)off;
    for (int t = lane; t < 13 * n; t += (int)blockDim.x) {
        const int sid = t / n, i = t - sid * n;
        if (!((fl >> sid) & 1u)) continue;
        const unsigned int *L = lists + sid * n;
        const unsigned int w = L[i], key = w & 0xffffu;
        int gt = 0, eq = 0;
        for (int j = 0; j < n; j++) {
            const unsigned int kj = L[j] & 0xffffu;
            gt += kj > key;
            eq += kj == key;
        }
        if (eq == 1) {
            out[sid * n + gt] = w;
        } else {
            tie[sid] = 1;
            lists[sid * n + i] = w | 0x80000000u; /* marked: shares its key */
        }
    }
    __syncthreads();
    /* Lists with a repeated key: the reference's quicksort, by 13 waves one
       list each, or by the cell's one wave all lists together. */
    const int wave = lane >> 6, nwave = (int)blockDim.x >> 6, wl = lane & 63;
    if (nwave > 1) {
        for (int l = wave; l < 13; l += nwave) {
            if (!((fl >> l) & 1u) || !tie[l]) continue;
            unsigned int *L = lists + l * n;
            int *st = stk + l * 2 * MD_SORT_STACK;
            unsigned short *tb = tabs + (size_t)l * 2 * (maxn + 1);
            if (n <= 192)
                sort_desc_wave<3>(L, n, st, st + MD_SORT_STACK, tb, tb + maxn + 1, flag, wl);
            else
                sort_desc_wave<8>(L, n, st, st + MD_SORT_STACK, tb, tb + maxn + 1, flag, wl);
            for (int i = wl; i < n; i += 64) out[l * n + i] = L[i] & 0x7fffffffu;
        }
    } else {
        unsigned int tm = 0;
        for (int l = 0; l < 13; l++) tm |= (unsigned int)(tie[l] != 0) << l;
        tm &= fl;
        if (tm) {
            /* segment queue: the stacks' space; partition tables and leaf
               queue: the partition tables' space (26 (maxn + 1) + 136
               entries) */
            sort_lists_wave(lists, n, tm, (unsigned int *)stk, (unsigned int *)(tabs + 136), 13 * (maxn + 1), tabs,
                            out, flag, wl);
        }
    }
}

/* ------------------------------------------------------------ k_nonbond */

/* Window bound of a pair task (runner_dopair.c:428): 2 + dscale (cutoff +
   maxdx_i + maxdx_j) in FP64, left to right, no contraction, truncated; i
   and j in task order. */
__device__ __forceinline__ int pair_dmaxdist(double dscale, double cutoff, double mi, double mj) {
#pragma clang fp contract(off)
    return (int)(2 + dscale * (cutoff + mi + mj));
}

struct nb_out {
    float4 *fhi, *flo;
    double *cell_acc;
    unsigned long long *cell_cnt;
    /* half-shell mode: per visit (neighbour-list entry), the forces on the
       visited cell's atoms, jcap slots per entry */
    float4 *jpart;
    int jcap;
    /* split owners (nsplit > 1): group g's forces on atom a at [g * nparts + a] */
    float4 *phi, *plo;
};

/* True if p holds on any lane (scalar test of the wave's ballot mask). */
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

/* Compile-time tag of a visit kind (self / owner is i / owner is j). */
template <int V>
struct visit_kind {
    static constexpr int value = V;
};

/* Exclusive prefix sum of v over the wave. */
__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x - v;
}

/* Evaluate the dense pair list [0, T) with the network, 16 pairs per tile;
   f = (dE/dr)/r and e of pair q go to dn_f[q], dn_e[q].  dE/dr is the
   network's own r-derivative, so the forces conserve the energies reported.
   The list goes in chunks of 64 pairs: lane l first computes the inputs of
   pair c + l (square root and four divisions, done once per pair instead of
   in all four lanes that carry its column of a tile) and parks them in the
   list's own slots (rn in dn_r2, irn in dn_f, d irn / dr in dn_e; d rn / dr
   stays in lane l's register), then the chunk's tiles run, each leaving y
   (without the output bias) and dy/dr of its pairs in dn_f and dn_e, and
   lane l turns those of pair c + l into its energy and force.  dn_r2 does not keep r^2.  w1s: the wave's
   LDS copy of layer 1's weights (read there, every tile's layer-1 loads
   leave the vector memory path: 40.9 -> 39.9 ms on the 1M mixture).  All
   64 lanes must call this together (T is wave-uniform). */
__device__ __forceinline__ void mlp_dense_eval(const struct mlp_net &N, const float *w1s, float *dn_r2, const int *dn_p,
                                               float *dn_f, float *dn_e, int T, int lane) {
    /* layer 2's operands stay in registers for the whole list: reloading
       them per tile cost 13 % of the 1M mixture's surrogate pass */
    mlp_bf16x8 w2r[8];
    float4 b2r[4];
    mlp_load_layer2(N, lane, w2r, b2r);
    for (int c = 0; c < T; c += 64) {
        const int q = c + lane;
        float r = 1.0f, drn = 0.0f;
        if (q < T) {
            r = sqrtf(dn_r2[q]);
            float rn, irn, dirn;
            mlp_inputs(N.sc[dn_p[q]], r, rn, irn, drn, dirn);
            dn_r2[q] = rn;
            dn_f[q] = irn;
            dn_e[q] = dirn;
        }
        wave_lds_sync();
        const int nt = min(4, (T - c + 15) >> 4);
        for (int t = 0; t < nt; t++) {
            const int qt = c + 16 * t + (lane & 15);
            const int qv = qt < T ? qt : T - 1;
            const float drt = __shfl(drn, qv - c);
            /* keep the network operands' loads inside the loop: hoisting them
               out keeps ~100 registers live through the whole pair kernel */
            asm volatile("" ::: "memory");
            float y, yd;
            mlp_infer16_reg(N, w1s, w2r, b2r, dn_r2[qv], dn_f[qv], drt, dn_e[qv], dn_p[qv], lane, y, yd);
            if (lane < 16 && qt < T) {
                dn_f[qt] = y;
                dn_e[qt] = yd;
            }
        }
        wave_lds_sync();
        if (q < T) {
            const int p = dn_p[q];
            float e, dedr;
            mlp_energy_force(N.sc[p], N.shift[p], r, dn_f[q] + N.b3[p], dn_e[q], e, dedr);
            dn_f[q] = dedr / r;
            dn_e[q] = e;
        }
    }
}

#define MD_HCAP 32     /* table mode: hits buffered per lane before they are evaluated */
#define MD_HCAP_WIDE 28 /* the same in the 5-wave builds: their LDS then allows 20 waves per CU */
#define MD_SPLIT_ITEMS_PER_CU 32 /* pair-kernel work items per CU below which owners are split */
#define MD_SPLIT_MAX 8           /* at most this many neighbour groups per owner cell */
#define MD_HCAP_MLP 8  /* MLP mode: hits per lane per compaction round */
#define MD_XCAP_MLP 8  /* MLP mode: excluded partners per lane per round */

/* The pair kernel body, for both evaluation modes:
   MLP = false  every hit is evaluated from the FP32 tables by its lane;
   MLP = true   hits below the species pair's r_lo (or without a network)
                are evaluated from the tables, the others are compacted
                wave-wide into a dense LDS pair list and evaluated by the
                surrogate network 16 pairs per MFMA tile; each lane then
                adds its own contiguous run of results in order, so the
                result stays deterministic.  Excluded pairs take the same
                route, so a pair and its exclusion cancel exactly. */
template <bool MLP, bool HALF, bool PAIRED = false, bool REFCNT = true, int HC = MD_HCAP>
__device__ __forceinline__ void nonbond_body(
    unsigned char *smem, struct nb_params P, const int2 *__restrict__ cnt_off, const int4 *__restrict__ nbr,
    const int *__restrict__ nbr_off, const float4 *__restrict__ nbs, const double *__restrict__ origin,
    const int *__restrict__ loc,
    const float4 *__restrict__ xf, const double *__restrict__ x64, const int *__restrict__ pid,
    const unsigned int *__restrict__ sortlist, const unsigned short *__restrict__ keyof,
    const float4 *__restrict__ ptab, const float4 *__restrict__ coeffs4, const int *__restrict__ excl_off,
    const int *__restrict__ excl_list, const int *__restrict__ index_of_id, const int *__restrict__ cell_of,
    struct nb_out O, int *__restrict__ queue, const struct mlp_net &N, const int *__restrict__ ptype) {
    constexpr int HCAP = MLP ? MD_HCAP_MLP : HC;
    float4 *raw = (float4 *)smem;                                        /* [64]        */
    unsigned short *own_key = (unsigned short *)(smem + P.lds_own_key);  /* [13*64]     */
    int *kfl = (int *)(smem + P.lds_kfl);                                /* [26]        */
    float4 *stage = (float4 *)(smem + P.lds_stage);                      /* [maxn]      */
    unsigned short *hitl = (unsigned short *)(smem + P.lds_hit);         /* [HCAP*64]   */
    /* MLP mode: dense pair list (r^2, species pair) and its results (f, e),
       and the excluded partners of each lane. */
    float *dn_r2 = (float *)(smem + P.lds_dense);                        /* [HCAP*64]   */
    int *dn_p = (int *)(smem + P.lds_dense + 4 * 64 * MD_HCAP_MLP);
    float *dn_f = (float *)(smem + P.lds_dense + 8 * 64 * MD_HCAP_MLP);
    float *dn_e = (float *)(smem + P.lds_dense + 12 * 64 * MD_HCAP_MLP);
    int *exb = (int *)(smem + P.lds_dense + 16 * 64 * MD_HCAP_MLP);      /* [XCAP*64]   */
    /* MLP mode: layer 1's weights [64][2], copied once per wave */
    float *w1s = (float *)(smem + P.lds_dense + 16 * 64 * MD_HCAP_MLP + 4 * 64 * MD_XCAP_MLP);
    /* half-shell mode: forces on the visited cell's records, 3 per record */
    float *jf = (float *)(smem + P.lds_jf);                              /* [4][3*maxn] */
    const int lane = threadIdx.x;
    const int mt = P.max_type;
    if constexpr (MLP) {
        w1s[lane] = N.w1[lane];
        w1s[64 + lane] = N.w1[64 + lane];
        wave_lds_sync();
    }
    /* A candidate beyond rc2hi is out.  rc2hi is rc2f + band rounded down:
       then every kept r2 has r2 - rc2f <= band exactly (the difference of
       two floats this close is exact), so the evaluation's band test
       catches every kept pair that is not certainly inside, and FP64
       decides it.  (Rounded up, a pair at r2 = rc2hi could slip past both.) */
    float rc2hi = P.rc2f + P.band;
    if (rc2hi - P.rc2f > P.band) rc2hi = __uint_as_float(__float_as_uint(rc2hi) - 1u);
    const __amdgpu_buffer_rsrc_t rs_ptab = make_rsrc(ptab, P.ptab_bytes);
    const __amdgpu_buffer_rsrc_t rs_coef = make_rsrc(coeffs4, P.coef_bytes);

    if (HALF) /* the visited cell's force copies start at zero; each visit's write-out clears them again */
        for (int z = lane; z < 4 * 3 * P.maxn; z += 64) jf[z] = 0.f;
    /* Work items: (owner cell, neighbour group), nsplit groups per cell. */
    const int nitems = P.ncells * P.nsplit;
    int q = blockIdx.x % MD_NQUEUE, tried = 0;
    while (tried < MD_NQUEUE) {
        const int lo_q = (int)(((long long)q * nitems) / MD_NQUEUE);
        const int hi_q = (int)(((long long)(q + 1) * nitems) / MD_NQUEUE);
        int it = 0;
        if (lane == 0) it = atomicAdd(&queue[q], 1);
        it = __shfl(it, 0);
        if (lo_q + it >= hi_q) {
            q = (q + 1) % MD_NQUEUE;
            tried++;
            continue;
        }
        /* Wave-uniform from here: the neighbour list is read with scalar loads. */
        const int item = __builtin_amdgcn_readfirstlane(lo_q + it);
        const int ci = item / P.nsplit, grp = item - ci * P.nsplit;
        const int c = P.order[ci];
        const int2 co_c = cnt_off[c];
        const int nc = co_c.x, offc = co_c.y;
        if (nc == 0) {
            /* An empty cell still owns its slots of the per-item sums. */
            if (lane < 3) {
                O.cell_acc[3 * item + lane] = 0.0;
                O.cell_cnt[3 * item + lane] = 0ull;
            }
            continue;
        }
        const double oc0 = origin[3 * c], oc1 = origin[3 * c + 1], oc2 = origin[3 * c + 2];
        /* this item's share of the owner's neighbour entries */
        const int kc = nbr_off[c], nall = nbr_off[c + 1] - kc;
        const int k0 = kc + (grp * nall) / P.nsplit, nn = kc + ((grp + 1) * nall) / P.nsplit - k0;

        /* Owner metadata: first/last key per direction. */
        wave_lds_sync();
        if (lane < 26) {
            const int sid = lane % 13;
            kfl[lane] = (int)(sortlist[13 * (size_t)offc + (size_t)sid * nc + (lane < 13 ? 0 : nc - 1)] & 0xffffu);
        }

        const int nchunk = (nc + 63) / 64;
        for (int ch = 0; ch < nchunk; ch++) {
            const int i0 = (int)(((long long)ch * nc) / nchunk);
            const int nsub = (int)(((long long)(ch + 1) * nc) / nchunk) - i0;
            /* Lane = (atom il, partition part): every lane works.  Atom il has
               NPi = (63 - il) / nsub + 1 partitions (lanes il, il + nsub, ...),
               each walking every NPi-th record of the atom's window. */
            const int il = lane % nsub, part = lane / nsub;
            const int NPi = (63 - il) / nsub + 1;
            const int NPmax = 63 / nsub + 1;
            const bool active = part < NPi;
            const int myidx = i0 + il; /* storage index in the owner cell */
            const int a = offc + myidx;
            const float4 xi = xf[a];
            const int ti = __float_as_int(xi.w);

            wave_lds_sync();
            if (lane < nsub)
                for (int sid = 0; sid < 13; sid++)
                    own_key[sid * 64 + lane] = keyof[(size_t)sid * P.nparts + offc + i0 + lane];
            wave_lds_sync();

            /* Visits: one neighbour entry or, in the paired kernels, the two
               entries of opposite offsets +d and -d, which share the sort
               direction and so the owner's keys.  The first of the two is the
               one whose pair task has the owner as its i cell (bit 6 marks
               it, bit 7 its partner).  The owner atoms' windows into the two
               cells are long and short in turn, so that each lane's walk
               through both is about as long as any other's.  Segment A of
               the staged records is the first entry's cell, segment B the
               second's. */
            int kbeg = 0;
            if constexpr (PAIRED)
                if (nn > 0 && (nbr[k0].y & 128)) kbeg = 1; /* opened by its partner in the previous group */

            /* Prefetch a visit's first 64 staged positions: lane t < nA holds
               atom t of A (unsorted position) and A's window word t, lane
               nA + u the same for atom u of B. */
            float4 pfx = make_float4(0.f, 0.f, 0.f, 0.f);
            unsigned int pfw = 0;
            auto prefetch = [&](int kv) {
                const int4 ea = nbr[k0 + kv];
                const int2 ca = cnt_off[ea.x];
                bool sB = false;
                int2 cs = ca;
                int4 es = ea;
                if constexpr (PAIRED) {
                    if (ea.y & 64) {
                        const int4 eb = nbr[k0 + kv + 1];
                        const int2 cb = cnt_off[eb.x];
                        sB = lane >= ca.x;
                        cs = sB ? cb : ca;
                        es = sB ? eb : ea;
                    }
                }
                const int u = sB ? lane - ca.x : lane;
                if (u < cs.x) {
                    pfx = xf[cs.y + u];
                    if (!(es.y & 32)) {
                        const int sid = es.y & 15, orient = (es.y >> 4) & 1;
                        pfw = sortlist[13 * (size_t)cs.y + (size_t)sid * cs.x + (orient ? u : cs.x - 1 - u)];
                    }
                }
            };
            if (kbeg < nn) prefetch(kbeg);

            float hx = 0.f, hy = 0.f, hz = 0.f, lx = 0.f, ly = 0.f, lz = 0.f;
            double epot = 0.0, vir = 0.0, eexcl = 0.0;
            unsigned int np = 0, nref = 0, nexcl = 0;

            for (int k = kbeg, step = 1; k < nn; k += step) {
                const int4 en = nbr[k0 + k];
                const bool paired = PAIRED && (en.y & 64);
                step = paired ? 2 : 1;
                const int4 enb = paired ? nbr[k0 + k + 1] : en;
                const int2 cjc = cnt_off[en.x];
                const int cj = en.x, nA = cjc.x, offj = cjc.y;
                int nB = 0, offB = 0;
                if (paired) {
                    const int2 cb = cnt_off[enb.x];
                    nB = cb.x;
                    offB = cb.y;
                }
                const int nj = nA + nB;
                const int sid = en.y & 15, orient = (en.y >> 4) & 1, self = (en.y >> 5) & 1;

                /* Shifts from the owner cell to the neighbours (minimum image),
                   rounded to FP32 on the host; the FP64 value is recomputed
                   where needed. */
                const float4 sfv = nbs[k0 + k];
                const float4 sgv = paired ? nbs[k0 + k + 1] : sfv;

                /* Stage the neighbours, each in the order its window is
                   walked: ascending keys when the owner is the task's i cell,
                   descending otherwise, storage order for the self task.
                   Record = (x + s, y + s, z + s, key' | index << 16 | type << tshift):
                   the position moved into the owner's frame by its cell's
                   shift s, so that a candidate's distance is one subtraction
                   per axis, and key' = key, or 0xffff - key when the owner is
                   the j cell. */
                wave_lds_sync();
                if (lane < nj) raw[lane] = pfx;
                wave_lds_sync();
                for (int t = lane; t < nj; t += 64) {
                    const bool sB = t >= nA;
                    const int u = sB ? t - nA : t, ns = sB ? nB : nA, os = sB ? offB : offj, rb = sB ? nA : 0;
                    const int ori = sB ? 1 : orient;
                    const float4 sh = sB ? sgv : sfv;
                    if (self) {
                        const float4 p = t < 64 ? pfx : xf[offj + t];
                        stage[t] = make_float4(p.x + sh.x, p.y + sh.y, p.z + sh.z,
                                               __uint_as_float(((unsigned int)t << 16) |
                                                               ((unsigned int)__float_as_int(p.w) << P.tshift)));
                    } else {
                        const unsigned int w =
                            t < 64 ? pfw : sortlist[13 * (size_t)os + (size_t)sid * ns + (ori ? u : ns - 1 - u)];
                        const int idx = (int)(w >> 16) < ns ? (int)(w >> 16) : ns - 1;
                        const float4 p = rb + idx < 64 ? raw[rb + idx] : xf[os + idx];
                        const unsigned int kk = ori ? 0xffffu - (w & 0xffffu) : (w & 0xffffu);
                        stage[t] = make_float4(p.x + sh.x, p.y + sh.y, p.z + sh.z,
                                               __uint_as_float(kk | ((unsigned int)idx << 16) |
                                                               ((unsigned int)__float_as_int(p.w) << P.tshift)));
                    }
                }
                wave_lds_sync();

                /* Prefetch the next visit while this one is processed. */
                if (k + step < nn) prefetch(k + step);
                if (nj == 0) continue;

                int myk = 0, dns = 0;
                bool live_ref = true, run = true;
                if (!self) {
                    dns = P.proj ? en.w : en.z;
                    if (nA > 0) {
                        const int kr = (int)(__float_as_uint(stage[0].w) & 0xffffu);
                        const int kstage0 = orient ? 0xffff - kr : kr; /* staged keys are complemented when orient */
                        const int s_imax = orient ? kstage0 : kfl[sid];
                        const int s_jmin = orient ? kfl[13 + sid] : kstage0;
                        live_ref = (s_jmin + en.z - s_imax) >= 0;
                        run = P.proj || live_ref;
                    } else {
                        live_ref = run = false;
                    }
                    myk = own_key[sid * 64 + il];
                }
                /* Segment B: the owner is the j cell of its task. */
                int kmaxB = 0;
                bool runB = false;
                if (paired && nB > 0) {
                    const int kr = (int)(__float_as_uint(stage[nA].w) & 0xffffu);
                    runB = P.proj || (kfl[13 + sid] + enb.z - (0xffff - kr)) >= 0;
                    const int dmdB = P.maxdx != nullptr ? pair_dmaxdist(P.dscale, P.cutoff, P.maxdx[enb.x], P.maxdx[c])
                                                        : P.dmaxdist;
                    kmaxB = dmdB - (myk + (P.proj ? enb.w : enb.z) - 0xffff);
                }
                /* Window (runner_dopair.c:455-478): record t is admitted while
                   key_j + dnshift - key_i < dmaxdist (owner is i) or
                   key_i + dnshift - key_j < dmaxdist (owner is j).  With the
                   staged key k' = key (owner is i) or 0xffff - key (owner is
                   j), both read k' < kmax, and k' ascends along the staged
                   order, so a lane's window ends at its first failing record. */
                /* With pseudo-Verlet sort reuse the window widens by both
                   cells' displacement since the last sort. */
                const int dmd = (P.maxdx != nullptr && !self)
                                    ? pair_dmaxdist(P.dscale, P.cutoff, P.maxdx[orient ? en.x : c],
                                                    P.maxdx[orient ? c : en.x])
                                    : P.dmaxdist;
                const int kmax = orient ? dmd - (myk + dns - 0xffff) : dmd - (dns - myk);

                float lfx = 0.f, lfy = 0.f, lfz = 0.f, le = 0.f, lv = 0.f;
                unsigned int lnp = 0, lnref = 0;

                /* The reference's FP64 cutoff test (runner_dopair.c:465-489) of
                   this lane's atom and the record with word w of cell cn
                   (storage offset on), in the task's own orientation: the
                   owner is the task's i cell (own_i) or its j cell. */
                auto in_cut64 = [&](int cn, int on, unsigned int w, bool own_i) -> bool {
                    const double *xa = &x64[3 * (size_t)a];
                    const double *xb = &x64[3 * (size_t)(on + rec_idx(w, P.imask))];
                    const double *o_own = &origin[3 * c], *o_nb = &origin[3 * cn];
                    const double r2d =
                        own_i ? r2_task64(xa, xb, min_image(o_nb[0] - o_own[0], P.dim[0]),
                                          min_image(o_nb[1] - o_own[1], P.dim[1]),
                                          min_image(o_nb[2] - o_own[2], P.dim[2]))
                              : r2_task64(xb, xa, min_image(o_own[0] - o_nb[0], P.dim[0]),
                                          min_image(o_own[1] - o_nb[1], P.dim[1]),
                                          min_image(o_own[2] - o_nb[2], P.dim[2]));
                    return r2d <= P.cutoff2;
                };

                /* One visit, specialised by kind: 0 self task, 1 owner is the
                   task's i cell (energies, virial and counts are taken here),
                   2 owner is the j cell (forces only). */
                auto visit = [&](auto kind) {
                    constexpr int VK = decltype(kind)::value;
                    int t = part;
                    bool alive = active && run;
                    while (wave_any(alive)) {
                        /* Filter: up to HCAP records per lane, predicated; the
                           wave leaves as soon as no lane is inside its window. */
                        int h = 0;
                        for (int it = 0; it < HCAP; it++) {
                            const float4 pj = stage[t < nj ? t : nj - 1];
                            const unsigned int wj = __float_as_uint(pj.w);
                            alive = alive && t < nj;
                            if (VK != 0) alive = alive && (int)rec_key(wj) < kmax;
                            /* half shell: a self task takes each pair once, j < i */
                            if (HALF && VK == 0) alive = alive && t < myidx;
                            const bool cand = VK == 0 ? (alive && t != myidx) : alive;
                            const float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            bool hit;
                            if constexpr (MLP) {
                                hit = cand && r2 <= P.rc2f;
                                const bool near = cand && fabsf(r2 - P.rc2f) <= P.band;
                                if (wave_any(near))
                                    if (near) hit = in_cut64(cj, offj, wj, VK != 2);
                            } else {
                                /* keep everything up to the band's upper edge; the
                                   evaluation decides the band's pairs in FP64 */
                                hit = cand && r2 <= rc2hi;
                            }
                            hitl[h * 64 + lane] = (unsigned short)t;
                            h += hit ? 1 : 0;
                            t += NPi;
                            if (!wave_any(alive)) break;
                        }

                        if constexpr (!MLP) {
                            /* Evaluate the kept hits, two per lane per round.  In
                               half-shell mode each lane starts its list at il mod h,
                               so that lanes of one partition (which kept hits from
                               the same records) add to different records at once. */
                            const int rot = (HALF && h > 0) ? il % h : 0;
                            for (int s = 0; wave_any(s < h); s += 2) {
                                const bool va = s < h, vb = s + 1 < h;
                                int sa = s + rot, sb = s + 1 + rot;
                                if (sa >= h) sa -= h;
                                if (sb >= h) sb -= h;
                                const int ta = va ? hitl[sa * 64 + lane] : 0;
                                const int tb = vb ? hitl[sb * 64 + lane] : ta;
                                const float4 pa = stage[ta], pb = stage[tb];
                                const unsigned int wa = __float_as_uint(pa.w), wb = __float_as_uint(pb.w);
                                const float4 ha = ld_f4(rs_ptab, 16 * (ti * mt + rec_type(wa, P.tshift)));
                                const float4 hb = ld_f4(rs_ptab, 16 * (ti * mt + rec_type(wb, P.tshift)));
                                const float dxa = xi.x - pa.x, dya = xi.y - pa.y, dza = xi.z - pa.z;
                                const float dxb = xi.x - pb.x, dyb = xi.y - pb.y, dzb = xi.z - pb.z;
                                const float r2a = dxa * dxa + dya * dya + dza * dza;
                                const float r2b = dxb * dxb + dyb * dyb + dzb * dzb;
                                bool ua = va && __float_as_int(ha.w) != -1;
                                bool ub = vb && __float_as_int(hb.w) != -1;
                                /* pairs within the band of the cutoff: FP64 decides */
                                const bool na = ua && fabsf(r2a - P.rc2f) <= P.band;
                                const bool nb = ub && fabsf(r2b - P.rc2f) <= P.band;
                                if (wave_any(na || nb)) {
                                    if (na) ua = in_cut64(cj, offj, wa, VK != 2);
                                    if (nb) ub = in_cut64(cj, offj, wb, VK != 2);
                                }
                                float ea, fa, eb, fb;
                                pot_eval32r2(ha, hb, ua, ub, rs_coef, r2a, r2b, ea, fa, eb, fb);
                                lfx = fmaf(-fb, dxb, fmaf(-fa, dxa, lfx));
                                lfy = fmaf(-fb, dyb, fmaf(-fa, dya, lfy));
                                lfz = fmaf(-fb, dzb, fmaf(-fa, dza, lfz));
                                if constexpr (HALF) {
                                    /* the opposite force on the record's atom, into this
                                       lane's copy (lane mod 4) of the visit's array */
                                    float *jc = jf + (lane & 3) * 3 * P.maxn;
                                    if (ua) {
                                        atomicAdd(&jc[3 * ta], fa * dxa);
                                        atomicAdd(&jc[3 * ta + 1], fa * dya);
                                        atomicAdd(&jc[3 * ta + 2], fa * dza);
                                    }
                                    if (ub) {
                                        atomicAdd(&jc[3 * tb], fb * dxb);
                                        atomicAdd(&jc[3 * tb + 1], fb * dyb);
                                        atomicAdd(&jc[3 * tb + 2], fb * dzb);
                                    }
                                }
                                if (VK != 2) {
                                    const bool ca = ua && (VK == 1 || myidx > rec_idx(wa, P.imask));
                                    const bool cb = ub && (VK == 1 || myidx > rec_idx(wb, P.imask));
                                    if (ca) {
                                        le += ea;
                                        lv -= fa * r2a;
                                        lnp++;
                                        if (REFCNT && P.proj) lnref += (VK == 0 ||
                                                  (live_ref && ((int)rec_key(wa) + en.z - myk) < dmd)) ? 1u : 0u;
                                    }
                                    if (cb) {
                                        le += eb;
                                        lv -= fb * r2b;
                                        lnp++;
                                        if (REFCNT && P.proj) lnref += (VK == 0 ||
                                                  (live_ref && ((int)rec_key(wb) + en.z - myk) < dmd)) ? 1u : 0u;
                                    }
                                }
                            }
                        } else {
                            /* Below the pair's r_lo: table, by this lane.  Otherwise
                               keep the hit for the dense list. */
                            int nd = 0;
                            for (int s = 0; s < h; s++) {
                                const int ta = hitl[s * 64 + lane];
                                const float4 pa = stage[ta];
                                const unsigned int wa = __float_as_uint(pa.w);
                                const float dxa = xi.x - pa.x, dya = xi.y - pa.y, dza = xi.z - pa.z;
                                const float r2a = dxa * dxa + dya * dya + dza * dza;
                                /* r_lo^2 is +inf for a type pair without potential, so
                                   only the few hits below r_lo load the table header */
                                if (r2a >= N.rlo2[ti * mt + rec_type(wa, P.tshift)]) {
                                    hitl[nd * 64 + lane] = (unsigned short)ta;
                                    nd++;
                                    continue;
                                }
                                const float4 ha = ptab[ti * mt + rec_type(wa, P.tshift)];
                                if (__float_as_int(ha.w) == -1) continue;
                                float ea, fa;
                                pot_eval32(ha, coeffs4, r2a, ea, fa);
                                lfx -= fa * dxa;
                                lfy -= fa * dya;
                                lfz -= fa * dza;
                                if constexpr (HALF) {
                                    float *jc = jf + (lane & 3) * 3 * P.maxn;
                                    atomicAdd(&jc[3 * ta], fa * dxa);
                                    atomicAdd(&jc[3 * ta + 1], fa * dya);
                                    atomicAdd(&jc[3 * ta + 2], fa * dza);
                                }
                                if (VK != 2 && (VK == 1 || myidx > rec_idx(wa, P.imask))) {
                                    le += ea;
                                    lv -= fa * r2a;
                                    lnp++;
                                    if (REFCNT && P.proj) lnref += (VK == 0 ||
                                              (live_ref && ((int)rec_key(wa) + en.z - myk) < dmd)) ? 1u : 0u;
                                }
                            }
                            /* Dense list: this lane's hits at [off, off + nd). */
                            const int off = wave_excl_scan(nd, lane);
                            const int T = __shfl(off + nd, 63);
                            for (int s = 0; s < nd; s++) {
                                const float4 pa = stage[hitl[s * 64 + lane]];
                                const float dxa = xi.x - pa.x, dya = xi.y - pa.y, dza = xi.z - pa.z;
                                dn_r2[off + s] = dxa * dxa + dya * dya + dza * dza;
                                dn_p[off + s] = ptype[ti * mt + rec_type(__float_as_uint(pa.w), P.tshift)];
                            }
                            wave_lds_sync();
                            mlp_dense_eval(N, w1s, dn_r2, dn_p, dn_f, dn_e, T, lane);
                            wave_lds_sync();
                            for (int s = 0; s < nd; s++) {
                                const int ta = hitl[s * 64 + lane];
                                const float4 pa = stage[ta];
                                const unsigned int wa = __float_as_uint(pa.w);
                                const float dxa = xi.x - pa.x, dya = xi.y - pa.y, dza = xi.z - pa.z;
                                const float r2a = dxa * dxa + dya * dya + dza * dza;
                                const float fa = dn_f[off + s], ea = dn_e[off + s];
                                lfx -= fa * dxa;
                                lfy -= fa * dya;
                                lfz -= fa * dza;
                                if constexpr (HALF) {
                                    float *jc = jf + (lane & 3) * 3 * P.maxn;
                                    atomicAdd(&jc[3 * ta], fa * dxa);
                                    atomicAdd(&jc[3 * ta + 1], fa * dya);
                                    atomicAdd(&jc[3 * ta + 2], fa * dza);
                                }
                                if (VK != 2 && (VK == 1 || myidx > rec_idx(wa, P.imask))) {
                                    le += ea;
                                    lv -= fa * r2a;
                                    lnp++;
                                    if (REFCNT && P.proj) lnref += (VK == 0 ||
                                              (live_ref && ((int)rec_key(wa) + en.z - myk) < dmd)) ? 1u : 0u;
                                }
                            }
                            wave_lds_sync();
                        }
                    }
                };
                /* A paired visit: each lane walks its share of its atom's window
                   in A (the owner is the i cell: energies, virial and counts
                   are taken there), then, from its first record past that
                   window, its share of the window in B (forces only). */
                auto visit_pair = [&]() {
                    int t = part, seg = run ? 0 : 1;
                    bool alive = active && (run || runB);
                    while (wave_any(alive)) {
                        int h = 0;
                        for (int it = 0; it < HCAP; it++) {
                            const int base = seg ? nA : 0, ns = seg ? nB : nA, km = seg ? kmaxB : kmax;
                            const int tt = base + (t < ns ? t : (ns > 0 ? ns - 1 : 0));
                            const float4 pj = stage[tt];
                            const unsigned int wj = __float_as_uint(pj.w);
                            const bool inwin = alive && t < ns && (int)rec_key(wj) < km;
                            const bool sw = alive && !inwin && seg == 0;
                            alive = inwin || (sw && runB);
                            const float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            const bool hit = inwin && r2 <= rc2hi;
                            hitl[h * 64 + lane] = (unsigned short)tt;
                            h += hit ? 1 : 0;
                            if (sw) {
                                t = part;
                                seg = 1;
                            } else if (inwin)
                                t += NPi;
                            if (!wave_any(alive)) break;
                        }
                        /* Evaluate the kept hits, two per lane per round. */
                        for (int s = 0; wave_any(s < h); s += 2) {
                            const bool va = s < h, vb = s + 1 < h;
                            const int ta = va ? hitl[s * 64 + lane] : 0;
                            const int tb = vb ? hitl[(s + 1) * 64 + lane] : ta;
                            const float4 pa = stage[ta], pb = stage[tb];
                            const unsigned int wa = __float_as_uint(pa.w), wb = __float_as_uint(pb.w);
                            const float4 ha = ld_f4(rs_ptab, 16 * (ti * mt + rec_type(wa, P.tshift)));
                            const float4 hb = ld_f4(rs_ptab, 16 * (ti * mt + rec_type(wb, P.tshift)));
                            const bool ia = ta < nA, ib = tb < nA;
                            const float dxa = xi.x - pa.x, dya = xi.y - pa.y, dza = xi.z - pa.z;
                            const float dxb = xi.x - pb.x, dyb = xi.y - pb.y, dzb = xi.z - pb.z;
                            const float r2a = dxa * dxa + dya * dya + dza * dza;
                            const float r2b = dxb * dxb + dyb * dyb + dzb * dzb;
                            bool ua = va && __float_as_int(ha.w) != -1;
                            bool ub = vb && __float_as_int(hb.w) != -1;
                            const bool na = ua && fabsf(r2a - P.rc2f) <= P.band;
                            const bool nb = ub && fabsf(r2b - P.rc2f) <= P.band;
                            if (wave_any(na || nb)) {
                                if (na) ua = ia ? in_cut64(cj, offj, wa, true) : in_cut64(enb.x, offB, wa, false);
                                if (nb) ub = ib ? in_cut64(cj, offj, wb, true) : in_cut64(enb.x, offB, wb, false);
                            }
                            float ea, fa, eb, fb;
                            pot_eval32r2(ha, hb, ua, ub, rs_coef, r2a, r2b, ea, fa, eb, fb);
                            lfx = fmaf(-fb, dxb, fmaf(-fa, dxa, lfx));
                            lfy = fmaf(-fb, dyb, fmaf(-fa, dya, lfy));
                            lfz = fmaf(-fb, dzb, fmaf(-fa, dza, lfz));
                            if (ua && ia) {
                                le += ea;
                                lv -= fa * r2a;
                                lnp++;
                                if (REFCNT && P.proj) lnref += (live_ref && ((int)rec_key(wa) + en.z - myk) < dmd) ? 1u : 0u;
                            }
                            if (ub && ib) {
                                le += eb;
                                lv -= fb * r2b;
                                lnp++;
                                if (REFCNT && P.proj) lnref += (live_ref && ((int)rec_key(wb) + en.z - myk) < dmd) ? 1u : 0u;
                            }
                        }
                    }
                };
                if constexpr (PAIRED) {
                    if (paired)
                        visit_pair();
                    else if (self)
                        visit(visit_kind<0>{});
                    else if (orient == 0)
                        visit(visit_kind<1>{});
                    else
                        visit(visit_kind<2>{});
                } else {
                    if (self)
                        visit(visit_kind<0>{});
                    else if (orient == 0)
                        visit(visit_kind<1>{});
                    else
                        visit(visit_kind<2>{});
                }
                if constexpr (HALF) {
                    /* this visit's forces on the visited atoms, by slot; later
                       chunks of the owner add to the first one's */
                    wave_lds_sync();
                    for (int q = lane; q < nj; q += 64) {
                        const int idx = rec_idx(__float_as_uint(stage[q].w), P.imask);
                        /* sum the 4 copies, and leave them zero for the next visit */
                        float vv[3];
                        for (int m = 0; m < 3; m++) {
                            float *j0 = jf + 3 * q + m;
                            const int s = 3 * P.maxn;
                            vv[m] = (j0[0] + j0[s]) + (j0[2 * s] + j0[3 * s]);
                            j0[0] = j0[s] = j0[2 * s] = j0[3 * s] = 0.f;
                        }
                        float4 v = make_float4(vv[0], vv[1], vv[2], 0.f);
                        float4 *dst = &O.jpart[(size_t)(k0 + k) * O.jcap + idx];
                        if (ch > 0) {
                            const float4 o = *dst;
                            v = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, 0.f);
                        }
                        *dst = v;
                    }
                }
                acc2(hx, lx, lfx);
                acc2(hy, ly, lfy);
                acc2(hz, lz, lfz);
                epot += (double)le;
                vir += (double)lv;
                np += lnp;
                nref += (REFCNT && P.proj) ? lnref : lnp;
            }

            /* Excluded pairs: subtract what the pair pass added
               (exclusion.c:139-208), same FP32 arithmetic as the pair pass. */
            if constexpr (!MLP) {
                if (active && part == 0 && grp == 0) {
                    const int myid = pid[a];
                    for (int k = excl_off[myid]; k < excl_off[myid + 1]; k++) {
                        const int oid = excl_list[k];
                        if (oid == myid) continue;
                        const int b = index_of_id[oid];
                        if (b < 0) continue;
                        const int cb = cell_of[b];
                        const float4 pj = xf[b];
                        const float4 hdr = ptab[ti * mt + __float_as_int(pj.w)];
                        if (__float_as_int(hdr.w) == -1) continue;
                        const float ex = (float)min_image(origin[3 * cb] - oc0, P.dim[0]);
                        const float ey = (float)min_image(origin[3 * cb + 1] - oc1, P.dim[1]);
                        const float ez = (float)min_image(origin[3 * cb + 2] - oc2, P.dim[2]);
                        const float dx = (xi.x - pj.x) - ex, dy = (xi.y - pj.y) - ey, dz = (xi.z - pj.z) - ez;
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
                        float e, ff;
                        pot_eval32(hdr, coeffs4, r2, e, ff);
                        acc2(hx, lx, ff * dx);
                        acc2(hy, ly, ff * dy);
                        acc2(hz, lz, ff * dz);
                        if (myid < oid) {
                            eexcl -= (double)e;
                            nexcl++;
                        }
                    }
                }
            } else {
                const bool xl = active && part == 0 && grp == 0;
                const int myid = xl ? pid[a] : 0;
                int kx = xl ? excl_off[myid] : 0;
                const int kend = xl ? excl_off[myid + 1] : 0;
                /* dx and r^2 of this lane's atom to slot b, as in the pair pass. */
                auto excl_geom = [&](int b, float &dx, float &dy, float &dz) {
                    const int cb = cell_of[b];
                    const float4 pj = xf[b];
                    dx = (xi.x - pj.x) - (float)min_image(origin[3 * cb] - oc0, P.dim[0]);
                    dy = (xi.y - pj.y) - (float)min_image(origin[3 * cb + 1] - oc1, P.dim[1]);
                    dz = (xi.z - pj.z) - (float)min_image(origin[3 * cb + 2] - oc2, P.dim[2]);
                    return __float_as_int(pj.w);
                };
                while (__any(kx < kend)) {
                    int nd = 0;
                    while (kx < kend && nd < MD_XCAP_MLP) {
                        const int oid = excl_list[kx++];
                        if (oid == myid) continue;
                        const int b = index_of_id[oid];
                        if (b < 0) continue;
                        float dx, dy, dz;
                        const int tb = excl_geom(b, dx, dy, dz);
                        const float4 hdr = ptab[ti * mt + tb];
                        if (__float_as_int(hdr.w) == -1) continue;
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        bool in = r2 <= P.rc2f;
                        if (fabsf(r2 - P.rc2f) <= P.band) {
                            const int cb = cell_of[b];
                            const bool mine = myid < oid;
                            const int ia = mine ? a : b, ib = mine ? b : a;
                            const int ca = mine ? c : cb, cbb = mine ? cb : c;
                            in = r2_excl64(&x64[3 * (size_t)ia], &x64[3 * (size_t)ib], &loc[3 * ca], &loc[3 * cbb],
                                           P.h) <= P.cutoff2;
                        }
                        if (!in) continue;
                        if (r2 >= N.rlo2[ti * mt + tb]) {
                            exb[nd * 64 + lane] = b;
                            nd++;
                            continue;
                        }
                        float e, ff;
                        pot_eval32(hdr, coeffs4, r2, e, ff);
                        acc2(hx, lx, ff * dx);
                        acc2(hy, ly, ff * dy);
                        acc2(hz, lz, ff * dz);
                        if (myid < oid) {
                            eexcl -= (double)e;
                            nexcl++;
                        }
                    }
                    const int off = wave_excl_scan(nd, lane);
                    const int T = __shfl(off + nd, 63);
                    for (int s = 0; s < nd; s++) {
                        float dx, dy, dz;
                        const int tb = excl_geom(exb[s * 64 + lane], dx, dy, dz);
                        dn_r2[off + s] = dx * dx + dy * dy + dz * dz;
                        dn_p[off + s] = ptype[ti * mt + tb];
                    }
                    wave_lds_sync();
                    mlp_dense_eval(N, w1s, dn_r2, dn_p, dn_f, dn_e, T, lane);
                    wave_lds_sync();
                    for (int s = 0; s < nd; s++) {
                        const int b = exb[s * 64 + lane];
                        float dx, dy, dz;
                        excl_geom(b, dx, dy, dz);
                        const float ff = dn_f[off + s], e = dn_e[off + s];
                        acc2(hx, lx, ff * dx);
                        acc2(hy, ly, ff * dy);
                        acc2(hz, lz, ff * dz);
                        if (myid < pid[b]) {
                            eexcl -= (double)e;
                            nexcl++;
                        }
                    }
                    wave_lds_sync();
                }
            }

            /* Fold the partitions of each atom into partition 0, in order. */
            for (int p = 1; p < NPmax; p++) {
                const int src = (lane + p * nsub) & 63;
                const float ohx = __shfl(hx, src), ohy = __shfl(hy, src), ohz = __shfl(hz, src);
                const float olx = __shfl(lx, src), oly = __shfl(ly, src), olz = __shfl(lz, src);
                if (part == 0 && p < NPi) {
                    acc2(hx, lx, ohx);
                    acc2(hy, ly, ohy);
                    acc2(hz, lz, ohz);
                    lx += olx;
                    ly += oly;
                    lz += olz;
                }
            }
            if (part == 0) {
                if (P.nsplit == 1) {
                    O.fhi[a] = make_float4(hx, hy, hz, 0.0f);
                    O.flo[a] = make_float4(lx, ly, lz, 0.0f);
                } else {
                    O.phi[(size_t)grp * P.nparts + a] = make_float4(hx, hy, hz, 0.0f);
                    O.plo[(size_t)grp * P.nparts + a] = make_float4(lx, ly, lz, 0.0f);
                }
            }
            if (!active) {
                epot = vir = eexcl = 0.0;
                np = nref = nexcl = 0;
            }

            /* Per-item sums; this wave owns the item's slots, chunks add in order. */
            epot = wave_sum(epot);
            vir = wave_sum(vir);
            eexcl = wave_sum(eexcl);
            np = (unsigned int)wave_sum((int)np);
            nref = (unsigned int)wave_sum((int)nref);
            nexcl = (unsigned int)wave_sum((int)nexcl);
            if (lane == 0) {
                if (ch == 0) {
                    O.cell_acc[3 * item] = epot;
                    O.cell_acc[3 * item + 1] = eexcl;
                    O.cell_acc[3 * item + 2] = vir;
                    O.cell_cnt[3 * item] = np;
                    O.cell_cnt[3 * item + 1] = nexcl;
                    O.cell_cnt[3 * item + 2] = nref;
                } else {
                    O.cell_acc[3 * item] += epot;
                    O.cell_acc[3 * item + 1] += eexcl;
                    O.cell_acc[3 * item + 2] += vir;
                    O.cell_cnt[3 * item] += np;
                    O.cell_cnt[3 * item + 1] += nexcl;
                    O.cell_cnt[3 * item + 2] += nref;
                }
            }
        }
    }
}

#define NB_PARAMS                                                                                             \
    struct nb_params P, const int2 *__restrict__ cnt_off, const int4 *__restrict__ nbr,                       \
        const int *__restrict__ nbr_off, const float4 *__restrict__ nbs, const double *__restrict__ origin,   \
        const int *__restrict__ loc,                                                                         \
        const float4 *__restrict__ xf, const double *__restrict__ x64, const int *__restrict__ pid,          \
        const unsigned int *__restrict__ sortlist, const unsigned short *__restrict__ keyof,                \
        const float4 *__restrict__ ptab, const float4 *__restrict__ coeffs4, const int *__restrict__ excl_off, \
        const int *__restrict__ excl_list, const int *__restrict__ index_of_id, const int *__restrict__ cell_of, \
        struct nb_out O, int *__restrict__ queue
#define NB_ARGS                                                                                              \
    P, cnt_off, nbr, nbr_off, nbs, origin, loc, xf, x64, pid, sortlist, keyof, ptab, coeffs4, excl_off, excl_list, \
        index_of_id, cell_of, O, queue

/* Table mode.  The occupancy hint keeps it at 4 waves per SIMD (128
   VGPRs); the kernel is bound by instruction issue and latency, and 4 waves
   measured faster than 3 without the hint. */
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nonbond(NB_PARAMS) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const struct mlp_net N = {};
    nonbond_body<false, false>(smem, NB_ARGS, N, nullptr);
}

/* Table mode with paired visits (neighbour cells at opposite offsets
   together), under the 4- and the 5-wave register budgets.  Reference
   window only: they do not count the pairs the reference window would
   admit (with the reference window that count is the pair count). */
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nonbond_paired(NB_PARAMS) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const struct mlp_net N = {};
    nonbond_body<false, false, true, false>(smem, NB_ARGS, N, nullptr);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_nonbond_paired_wide(NB_PARAMS) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const struct mlp_net N = {};
    nonbond_body<false, false, true, false, MD_HCAP_WIDE>(smem, NB_ARGS, N, nullptr);
}

/* Table mode for systems with many more cells than resident waves: the same
   body under a 96-register budget (5 waves per SIMD), which pays off when
   every wave has cells to take; with few cells per wave the 4-wave build's
   lower latency per cell wins. */
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_nonbond_wide(NB_PARAMS) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const struct mlp_net N = {};
    nonbond_body<false, false, false, true, MD_HCAP_WIDE>(smem, NB_ARGS, N, nullptr);
}

/* Surrogate mode (engine_flag_mlp). */
__global__ __launch_bounds__(64) void k_nonbond_mlp(NB_PARAMS, struct mlp_net N, const int *__restrict__ ptype) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    nonbond_body<true, false>(smem, NB_ARGS, N, ptype);
}

/* Surrogate mode on the half shell: each pair's network evaluation once. */
__global__ __launch_bounds__(64) void k_nonbond_mlp_half(NB_PARAMS, struct mlp_net N, const int *__restrict__ ptype) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    nonbond_body<true, true>(smem, NB_ARGS, N, ptype);
}

/* Half shell: each pair task once, from its i cell, with the opposite
   forces on the j cell's atoms left per visit for k_fold. */
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nonbond_half(NB_PARAMS) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const struct mlp_net N = {};
    nonbond_body<false, true>(smem, NB_ARGS, N, nullptr);
}
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_nonbond_half_wide(NB_PARAMS) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const struct mlp_net N = {};
    nonbond_body<false, true>(smem, NB_ARGS, N, nullptr);
}

/* Second step of a split and/or half-shell pass, one wave per cell: each
   atom's force = its owner's (hi, lo), or with nsplit > 1 the groups'
   (hi, lo) in group order, plus (half shell) the forces its cell received as
   the j cell of other cells' tasks and in its own self task, in a fixed
   order (the self entry, then the incoming entries in ascending order),
   all added with two-sum. */
__global__ __launch_bounds__(64) void k_fold(int ncells, const int2 *__restrict__ cnt_off, int nsplit, int nparts,
                                             const float4 *__restrict__ phi, const float4 *__restrict__ plo, int half,
                                             const int *__restrict__ self_entry, const int *__restrict__ inv_off,
                                             const int *__restrict__ inv_list, const float4 *__restrict__ jpart,
                                             int jcap, float4 *__restrict__ fhi, float4 *__restrict__ flo) {
    const int c = blockIdx.x, lane = threadIdx.x;
    const int2 co = cnt_off[c];
    const int q0 = half ? inv_off[c] : 0, q1 = half ? inv_off[c + 1] : 0, es = half ? self_entry[c] : -1;
    for (int s = lane; s < co.x; s += 64) {
        const int a = co.y + s;
        const float4 h = nsplit > 1 ? phi[a] : fhi[a], l = nsplit > 1 ? plo[a] : flo[a];
        float hx = h.x, hy = h.y, hz = h.z, lx = l.x, ly = l.y, lz = l.z;
        for (int gp = 1; gp < nsplit; gp++) {
            const float4 oh = phi[(size_t)gp * nparts + a], ol = plo[(size_t)gp * nparts + a];
            acc2(hx, lx, oh.x);
            acc2(hy, ly, oh.y);
            acc2(hz, lz, oh.z);
            lx += ol.x;
            ly += ol.y;
            lz += ol.z;
        }
        if (es >= 0) {
            const float4 v = jpart[(size_t)es * jcap + s];
            acc2(hx, lx, v.x);
            acc2(hy, ly, v.y);
            acc2(hz, lz, v.z);
        }
        for (int q = q0; q < q1; q++) {
            const float4 v = jpart[(size_t)inv_list[q] * jcap + s];
            acc2(hx, lx, v.x);
            acc2(hy, ly, v.y);
            acc2(hz, lz, v.z);
        }
        fhi[a] = make_float4(hx, hy, hz, 0.0f);
        flo[a] = make_float4(lx, ly, lz, 0.0f);
    }
}

/* ------------------------------------------------------------ k_reduce */

/* Two steps, both in a fixed order (the totals are bitwise repeatable):
   k_reduce_a: block b sums the slots [b n / B, (b + 1) n / B) of the
   per-item FP64 sums and counts (strided per thread, then a tree);
   k_reduce_b: one block sums the B partials the same way. */
#define MD_RED_BLOCKS 128
__device__ __forceinline__ void block_tree3(double (*sd)[256], unsigned long long (*sc)[256], int t) {
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int m = 0; m < 3; m++) {
                sd[m][t] += sd[m][t + s];
                sc[m][t] += sc[m][t + s];
            }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_reduce_a(int n, const double *__restrict__ acc,
                                                  const unsigned long long *__restrict__ cnt,
                                                  double *__restrict__ pacc, unsigned long long *__restrict__ pcnt) {
    __shared__ double sd[3][256];
    __shared__ unsigned long long sc[3][256];
    const int t = threadIdx.x, b = blockIdx.x;
    const int lo = (int)(((long long)b * n) / gridDim.x), hi = (int)(((long long)(b + 1) * n) / gridDim.x);
    double d0 = 0, d1 = 0, d2 = 0;
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    for (int k = lo + t; k < hi; k += 256) {
        d0 += acc[3 * k];
        d1 += acc[3 * k + 1];
        d2 += acc[3 * k + 2];
        c0 += cnt[3 * k];
        c1 += cnt[3 * k + 1];
        c2 += cnt[3 * k + 2];
    }
    sd[0][t] = d0;
    sd[1][t] = d1;
    sd[2][t] = d2;
    sc[0][t] = c0;
    sc[1][t] = c1;
    sc[2][t] = c2;
    block_tree3(sd, sc, t);
    if (t < 3) {
        pacc[3 * b + t] = sd[t][0];
        pcnt[3 * b + t] = sc[t][0];
    }
}

__global__ __launch_bounds__(256) void k_reduce_b(int nb, const double *__restrict__ pacc,
                                                  const unsigned long long *__restrict__ pcnt,
                                                  double *__restrict__ tot, unsigned long long *__restrict__ tot_cnt) {
    __shared__ double sd[3][256];
    __shared__ unsigned long long sc[3][256];
    const int t = threadIdx.x;
    for (int m = 0; m < 3; m++) {
        double d = 0;
        unsigned long long c = 0;
        for (int k = t; k < nb; k += 256) {
            d += pacc[3 * k + m];
            c += pcnt[3 * k + m];
        }
        sd[m][t] = d;
        sc[m][t] = c;
    }
    block_tree3(sd, sc, t);
    if (t < 3) {
        tot[t] = sd[t][0];
        tot_cnt[t] = sc[t][0];
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
    (void)hipHostFree(g->h_v);
    g->d_count = g->d_offset = NULL;
    g->d_x = NULL;
    g->d_type = g->d_id = g->d_cell_of = NULL;
    g->d_xf = g->d_fhi = g->d_flo = NULL;
    g->d_sortlist = NULL;
    g->d_keyof = NULL;
    g->h_x = NULL;
    g->h_type = g->h_id = NULL;
    g->h_fhi = g->h_flo = NULL;
    g->h_v = NULL;
    g->nparts_cap = 0;
}

static int gpu_alloc_parts(struct gpu_state *g, int n) {
    size_t N;
    if (n <= g->nparts_cap) return 0;
    gpu_free_pass(g);
    g->sort_valid = 0;
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
    HIPCHECK(hipHostMalloc((void **)&g->h_v, sizeof(double) * 3 * N, hipHostMallocDefault));
    g->nparts_cap = (int)N;
    return 0;
}

static int gpu_alloc_ids(struct gpu_state *g, int nids) {
    if (nids <= g->nids_cap) return 0;
    (void)hipFree(g->d_index_of_id);
    (void)hipHostFree(g->h_ioi);
    g->h_ioi = NULL;
    HIPCHECK(hipMalloc(&g->d_index_of_id, sizeof(int) * (size_t)nids));
    HIPCHECK(hipHostMalloc((void **)&g->h_ioi, sizeof(int) * (size_t)nids, hipHostMallocDefault));
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
    gpu_mlp_free(g);
    gpu_steps_free(g);
    (void)hipFree(g->d_ptab);
    (void)hipFree(g->d_coeffs);
    (void)hipFree(g->d_ptype);
    free(g->h_tabof);
    free(g->h_pair_ti);
    free(g->h_pair_tj);
    (void)hipFree(g->d_nbr);
    (void)hipFree(g->d_nbs);
    (void)hipFree(g->d_nbr_off);
    (void)hipFree(g->d_nbr_p);
    (void)hipFree(g->d_nbs_p);
    (void)hipFree(g->d_nbr_h);
    (void)hipFree(g->d_nbs_h);
    (void)hipFree(g->d_nbr_off_h);
    (void)hipFree(g->d_self_entry);
    (void)hipFree(g->d_inv_off);
    (void)hipFree(g->d_inv_list);
    (void)hipFree(g->d_jpart);
    (void)hipFree(g->d_phi);
    (void)hipFree(g->d_plo);
    (void)hipFree(g->d_origin);
    (void)hipFree(g->d_loc);
    (void)hipFree(g->d_sortflags);
    (void)hipFree(g->d_excl_off);
    (void)hipFree(g->d_excl_list);
    (void)hipFree(g->d_index_of_id);
    (void)hipHostFree(g->h_ioi);
    (void)hipFree(g->d_cnt_off);
    (void)hipFree(g->d_cell_acc);
    (void)hipFree(g->d_cell_cnt);
    (void)hipFree(g->d_tot);
    (void)hipFree(g->d_red_acc);
    (void)hipFree(g->d_red_cnt);
    (void)hipFree(g->d_tot_cnt);
    (void)hipFree(g->d_queue);
    (void)hipFree(g->d_order);
    (void)hipHostFree(g->h_count);
    (void)hipFree(g->d_cell_maxdx);
    (void)hipHostFree(g->h_cell_maxdx);
    free(g->h_sort_count);
    (void)hipHostFree(g->h_offset);
    (void)hipHostFree(g->h_cnt_off);
    (void)hipHostFree(g->h_tot);
    (void)hipHostFree(g->h_tot_cnt);
    for (int k = 0; k < 8; k++) (void)hipEventDestroy(g->ev[k]);
    (void)hipStreamDestroy(g->stream);
    free(g);
    e->gpu = NULL;
    return engine_err_ok;
}

/* FP32 minimum-image shift from cell a's origin to cell b's: the same FP64
   arithmetic as min_image() on the device, rounded once. */
static float4 cell_shift32(const struct space *s, int a, int b) {
    float v[3];
    for (int k = 0; k < 3; k++) {
        double d = s->cells[b].origin[k] - s->cells[a].origin[k];
        if (d * 2 > s->dim[k])
            d -= s->dim[k];
        else if (d * 2 < -s->dim[k])
            d += s->dim[k];
        v[k] = (float)d;
    }
    return make_float4(v[0], v[1], v[2], 0.f);
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
    int mt, k, i, j, ntab = 0, nchunks = 0, cid;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    if (e->gpu != NULL) engine_gpu_finalize(e);
    s = &e->s;
    mt = e->max_type;
    if (mt > MD_MAX_TYPES) return gerror_(engine_err_nyi, "more than 1024 particle types");
    g = (struct gpu_state *)calloc(1, sizeof(struct gpu_state));
    if (g == NULL) return gerror_(engine_err_malloc, "calloc failed");
    e->gpu = g;
    HIPCHECK(hipGetDevice(&g->device));
    HIPCHECK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    for (k = 0; k < 8; k++) HIPCHECK(hipEventCreate(&g->ev[k]));
    g->ncells = s->nr_cells;
    g->max_type = mt;

    /* Tables: one device table per distinct potential, FP32-rounded, and a
       header per type pair.  A distinct potential is a species pair: its
       index (first appearance in type-pair order) is the pair id that the
       MLP surrogate uses. */
    {
        struct potential **uniq = (struct potential **)calloc((size_t)mt * mt, sizeof(*uniq));
        int *tabof = (int *)malloc(sizeof(int) * mt * mt);
        int *chunk0 = (int *)malloc(sizeof(int) * mt * mt);
        float4 *ptab = (float4 *)malloc(sizeof(float4) * mt * mt);
        float *coef = NULL;
        int bad = 0;
        if (uniq == NULL || tabof == NULL || chunk0 == NULL || ptab == NULL) bad = engine_err_malloc;
        for (k = 0; !bad && k < mt * mt; k++) {
            struct potential *p = e->p[k];
            tabof[k] = -1;
            if (p == NULL) continue;
            if (p->n < 1 || p->n > 511) {
                bad = engine_err_potential;
                break;
            }
            for (j = 0; j < ntab; j++)
                if (uniq[j] == p) break;
            if (j == ntab) {
                chunk0[ntab] = nchunks;
                uniq[ntab++] = p;
                nchunks += p->n + 1;
            }
            tabof[k] = j;
        }
        if (!bad && nchunks >= (1 << 22)) bad = engine_err_potential;
        if (!bad && (coef = (float *)malloc(sizeof(float) * 8 * (nchunks > 0 ? nchunks : 1))) == NULL)
            bad = engine_err_malloc;
        if (!bad) {
            for (j = 0; j < ntab; j++) {
                struct potential *p = uniq[j];
                for (k = 0; k < (p->n + 1) * potential_chunk; k++) coef[8 * chunk0[j] + k] = (float)p->c[k];
            }
            for (k = 0; k < mt * mt; k++) {
                if (tabof[k] < 0) {
                    int m1 = -1;
                    ptab[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                    memcpy(&ptab[k].w, &m1, sizeof(int));
                } else {
                    struct potential *p = uniq[tabof[k]];
                    unsigned int w = (unsigned int)p->n | ((unsigned int)chunk0[tabof[k]] << 9);
                    ptab[k] = make_float4((float)p->alpha[0], (float)p->alpha[1], (float)p->alpha[2], 0.f);
                    memcpy(&ptab[k].w, &w, sizeof(unsigned int));
                }
            }
            g->ntab = ntab;
            g->h_tabof = tabof;
            g->h_pair_ti = (int *)malloc(sizeof(int) * (ntab > 0 ? ntab : 1));
            g->h_pair_tj = (int *)malloc(sizeof(int) * (ntab > 0 ? ntab : 1));
            if (g->h_pair_ti == NULL || g->h_pair_tj == NULL) bad = engine_err_malloc;
            for (k = mt * mt - 1; !bad && k >= 0; k--)
                if (tabof[k] >= 0 && k / mt <= k % mt) {
                    g->h_pair_ti[tabof[k]] = k / mt;
                    g->h_pair_tj[tabof[k]] = k % mt;
                }
        }
        if (!bad) {
            if (hipMalloc(&g->d_ptab, sizeof(float4) * mt * mt) != hipSuccess ||
                hipMalloc(&g->d_coeffs, sizeof(float) * 8 * (nchunks > 0 ? nchunks : 1)) != hipSuccess ||
                hipMalloc(&g->d_ptype, sizeof(int) * mt * mt) != hipSuccess ||
                hipMemcpy(g->d_ptab, ptab, sizeof(float4) * mt * mt, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(g->d_coeffs, coef, sizeof(float) * 8 * (nchunks > 0 ? nchunks : 1),
                          hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(g->d_ptype, tabof, sizeof(int) * mt * mt, hipMemcpyHostToDevice) != hipSuccess)
                bad = engine_err_gpu;
        }
        free(uniq);
        free(chunk0);
        free(ptab);
        free(coef);
        if (bad) {
            if (g->h_tabof != tabof) free(tabof);
            return gerror_(bad, bad == engine_err_potential ? "table with more than 511 intervals or too large"
                                                            : "table upload failed");
        }
    }

    /* Geometry constants of the sorted window (runner_dopair.c:424-428). */
    {
        struct nb_params *P = &g->P;
        memset(P, 0, sizeof(*P));
        P->ptab_bytes = (int)(sizeof(float4) * mt * mt);
        P->coef_bytes = (int)(sizeof(float) * 8 * (nchunks > 0 ? nchunks : 1));
        P->ncells = s->nr_cells;
        P->max_type = mt;
        /* engine_flag_unsorted: the reference's unsorted runner (runner.c:801-804,
           runner_dopair.c:1049-1290) evaluates every pair within the cutoff,
           which is the pair set of the projected window */
        P->proj = (e->flags & (engine_flag_projwindow | engine_flag_unsorted)) ? 1 : 0;
        P->bias = sqrt(s->h[0] * s->h[0] + s->h[1] * s->h[1] + s->h[2] * s->h[2]);
        P->dscale = (double)32767 / (2 * P->bias);
        P->dmaxdist = (int)(2 + P->dscale * (s->cutoff + 0.0 + 0.0));
        P->cutoff = s->cutoff;
        /* Packed record: key (16 bits) | slot (ibits) | type (16 - ibits). */
        {
            int tbits = 6;
            while ((1 << tbits) < mt) tbits += 2;
            P->tshift = 32 - tbits;
            P->imask = (1 << (16 - tbits)) - 1;
            g->maxcell = 1 << (16 - tbits);
        }
        P->maxdx = NULL;
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
        int *fill, *off;
        int4 *nbr;
        float4 *nbs;
        int total = 0, maxnbr = 1;
        for (k = 0; k < s->nr_tasks; k++) {
            struct task *t = &s->tasks[k];
            if (t->type == task_type_self)
                cnt[t->i]++;
            else if (t->type == task_type_pair) {
                cnt[t->i]++;
                cnt[t->j]++;
            }
        }
        for (cid = 0; cid < s->nr_cells; cid++) {
            total += cnt[cid];
            if (cnt[cid] > maxnbr) maxnbr = cnt[cid];
        }
        nbr = (int4 *)malloc(sizeof(int4) * (total > 0 ? total : 1));
        fill = (int *)malloc(sizeof(int) * (s->nr_cells + 1));
        off = (int *)malloc(sizeof(int) * (s->nr_cells + 1));
        fill[0] = 0;
        for (cid = 0; cid < s->nr_cells; cid++) fill[cid + 1] = fill[cid] + cnt[cid];
        memcpy(off, fill, sizeof(int) * (s->nr_cells + 1));
        nbs = (float4 *)malloc(sizeof(float4) * (total > 0 ? total : 1));
        for (k = 0; k < s->nr_tasks; k++) {
            struct task *t = &s->tasks[k];
            if (t->type == task_type_self) {
                nbs[fill[t->i]] = make_float4(0.f, 0.f, 0.f, 0.f);
                nbr[fill[t->i]++] = make_int4(t->i, 1 << 5, 0, 0);
            } else if (t->type == task_type_pair) {
                int dr, dp;
                task_dnshift(s, t->i, t->j, t->flags, g->P.dscale, &dr, &dp);
                nbs[fill[t->i]] = cell_shift32(s, t->i, t->j);
                nbr[fill[t->i]++] = make_int4(t->j, t->flags, dr, dp);
                nbs[fill[t->j]] = cell_shift32(s, t->j, t->i);
                nbr[fill[t->j]++] = make_int4(t->i, t->flags | (1 << 4), dr, dp);
            }
        }
        g->maxnbr = maxnbr;
        g->P.maxnbr = maxnbr;
        HIPCHECK(hipMalloc(&g->d_nbr, sizeof(int4) * (total > 0 ? total : 1)));
        HIPCHECK(hipMalloc(&g->d_nbr_off, sizeof(int) * (s->nr_cells + 1)));
        HIPCHECK(hipMemcpy(g->d_nbr, nbr, sizeof(int4) * (total > 0 ? total : 1), hipMemcpyHostToDevice));
        HIPCHECK(hipMalloc(&g->d_nbs, sizeof(float4) * (total > 0 ? total : 1)));
        HIPCHECK(hipMemcpy(g->d_nbs, nbs, sizeof(float4) * (total > 0 ? total : 1), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(g->d_nbr_off, off, sizeof(int) * (s->nr_cells + 1), hipMemcpyHostToDevice));
        /* Paired order: per owner the self entry, then each entry whose task
           has the owner as its i cell followed by the entry at the opposite
           cell offset with the same sort direction (the owner is its task's
           j cell), then the entries left over, each list in task order. */
        {
            int4 *pn = (int4 *)malloc(sizeof(int4) * (total > 0 ? total : 1));
            float4 *ps = (float4 *)malloc(sizeof(float4) * (total > 0 ? total : 1));
            int *used = (int *)calloc(maxnbr + 1, sizeof(int));
            int *slot = (int *)malloc(sizeof(int) * 17 * 17 * 17);
            if (pn == NULL || ps == NULL || used == NULL || slot == NULL) {
                free(pn);
                free(ps);
                free(used);
                free(slot);
                return gerror_(engine_err_malloc, "malloc failed");
            }
            for (k = 0; k < 17 * 17 * 17; k++) slot[k] = -1;
            auto ofs_key = [&](const float4 v) {
                int d[3];
                const float sh[3] = {v.x, v.y, v.z};
                for (int m = 0; m < 3; m++) {
                    d[m] = (int)lrint(sh[m] / s->h[m]);
                    d[m] = d[m] < -8 ? -8 : (d[m] > 8 ? 8 : d[m]);
                }
                return ((d[0] + 8) * 17 + (d[1] + 8)) * 17 + (d[2] + 8);
            };
            for (cid = 0; cid < s->nr_cells; cid++) {
                const int b0 = off[cid], n = off[cid + 1] - off[cid];
                int q = b0;
                for (k = 0; k < n; k++) used[k] = 0;
                /* the entries the owner is the j cell of, by cell offset */
                for (k = 0; k < n; k++)
                    if (!(nbr[b0 + k].y & 32) && (nbr[b0 + k].y & 16)) {
                        const int key = ofs_key(nbs[b0 + k]);
                        if (slot[key] < 0) slot[key] = k;
                    }
                for (k = 0; k < n; k++)
                    if (nbr[b0 + k].y & 32) {
                        pn[q] = nbr[b0 + k];
                        ps[q++] = nbs[b0 + k];
                        used[k] = 1;
                    }
                for (k = 0; k < n; k++) {
                    const int4 ea = nbr[b0 + k];
                    if (used[k] || (ea.y & 48)) continue;
                    const float4 sa = nbs[b0 + k];
                    const int key = ofs_key(make_float4(-sa.x, -sa.y, -sa.z, 0.f));
                    const int m = slot[key];
                    used[k] = 1;
                    pn[q] = ea;
                    ps[q++] = sa;
                    if (m >= 0 && !used[m] && (nbr[b0 + m].y & 15) == (ea.y & 15)) {
                        used[m] = 1;
                        pn[q - 1].y |= 64;
                        pn[q] = nbr[b0 + m];
                        pn[q].y |= 128;
                        ps[q++] = nbs[b0 + m];
                    }
                }
                for (k = 0; k < n; k++)
                    if (!used[k]) {
                        pn[q] = nbr[b0 + k];
                        ps[q++] = nbs[b0 + k];
                    }
                for (k = 0; k < n; k++)
                    if (!(nbr[b0 + k].y & 32) && (nbr[b0 + k].y & 16)) slot[ofs_key(nbs[b0 + k])] = -1;
            }
            free(used);
            free(slot);
            if (hipMalloc(&g->d_nbr_p, sizeof(int4) * (total > 0 ? total : 1)) != hipSuccess ||
                hipMalloc(&g->d_nbs_p, sizeof(float4) * (total > 0 ? total : 1)) != hipSuccess ||
                hipMemcpy(g->d_nbr_p, pn, sizeof(int4) * (total > 0 ? total : 1), hipMemcpyHostToDevice) !=
                    hipSuccess ||
                hipMemcpy(g->d_nbs_p, ps, sizeof(float4) * (total > 0 ? total : 1), hipMemcpyHostToDevice) !=
                    hipSuccess) {
                free(pn);
                free(ps);
                return gerror_(engine_err_gpu, "neighbour list upload failed");
            }
            free(pn);
            free(ps);
        }
        free(cnt);
        free(fill);
        free(off);
        free(nbr);
        free(nbs);
    }

    /* Half-shell lists: the self task and each pair task once, at the task's
       i cell, in task order; each cell's self entry; and for each cell the
       pair entries that visit it as the task's j cell, in task order. */
    {
        const int nc = s->nr_cells;
        int *hcnt = (int *)calloc(nc + 1, sizeof(int)), *icnt = (int *)calloc(nc + 1, sizeof(int));
        int *hoff = (int *)malloc(sizeof(int) * (nc + 1)), *ioff = (int *)malloc(sizeof(int) * (nc + 1));
        int *hfill = (int *)malloc(sizeof(int) * (nc + 1)), *ifill = (int *)malloc(sizeof(int) * (nc + 1));
        int *self_e = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
        int htot = 0, itot = 0, bad = 0;
        int4 *nbr = NULL;
        float4 *nbs = NULL;
        int *ilist = NULL;
        if (hcnt == NULL || icnt == NULL || hoff == NULL || ioff == NULL || hfill == NULL || ifill == NULL ||
            self_e == NULL)
            bad = 1;
        for (k = 0; !bad && k < s->nr_tasks; k++) {
            const struct task *t = &s->tasks[k];
            if (t->type == task_type_self)
                hcnt[t->i]++;
            else if (t->type == task_type_pair) {
                hcnt[t->i]++;
                icnt[t->j]++;
            }
        }
        if (!bad) {
            hoff[0] = ioff[0] = 0;
            for (cid = 0; cid < nc; cid++) {
                hoff[cid + 1] = hoff[cid] + hcnt[cid];
                ioff[cid + 1] = ioff[cid] + icnt[cid];
                self_e[cid] = -1;
            }
            htot = hoff[nc];
            itot = ioff[nc];
            memcpy(hfill, hoff, sizeof(int) * (nc + 1));
            memcpy(ifill, ioff, sizeof(int) * (nc + 1));
            nbr = (int4 *)malloc(sizeof(int4) * (htot > 0 ? htot : 1));
            nbs = (float4 *)malloc(sizeof(float4) * (htot > 0 ? htot : 1));
            ilist = (int *)malloc(sizeof(int) * (itot > 0 ? itot : 1));
            if (nbr == NULL || nbs == NULL || ilist == NULL) bad = 1;
        }
        for (k = 0; !bad && k < s->nr_tasks; k++) {
            const struct task *t = &s->tasks[k];
            if (t->type == task_type_self) {
                const int en = hfill[t->i]++;
                self_e[t->i] = en;
                nbs[en] = make_float4(0.f, 0.f, 0.f, 0.f);
                nbr[en] = make_int4(t->i, 1 << 5, 0, 0);
            } else if (t->type == task_type_pair) {
                int dr, dp;
                const int en = hfill[t->i]++;
                task_dnshift(s, t->i, t->j, t->flags, g->P.dscale, &dr, &dp);
                nbs[en] = cell_shift32(s, t->i, t->j);
                nbr[en] = make_int4(t->j, t->flags, dr, dp);
                ilist[ifill[t->j]++] = en;
            }
        }
        if (!bad) {
            g->n_entries_h = htot;
            if (hipMalloc(&g->d_nbr_h, sizeof(int4) * (htot > 0 ? htot : 1)) != hipSuccess ||
                hipMalloc(&g->d_nbs_h, sizeof(float4) * (htot > 0 ? htot : 1)) != hipSuccess ||
                hipMalloc(&g->d_nbr_off_h, sizeof(int) * (nc + 1)) != hipSuccess ||
                hipMalloc(&g->d_self_entry, sizeof(int) * (nc > 0 ? nc : 1)) != hipSuccess ||
                hipMalloc(&g->d_inv_off, sizeof(int) * (nc + 1)) != hipSuccess ||
                hipMalloc(&g->d_inv_list, sizeof(int) * (itot > 0 ? itot : 1)) != hipSuccess ||
                hipMemcpy(g->d_nbr_h, nbr, sizeof(int4) * (htot > 0 ? htot : 1), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(g->d_nbs_h, nbs, sizeof(float4) * (htot > 0 ? htot : 1), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(g->d_nbr_off_h, hoff, sizeof(int) * (nc + 1), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(g->d_self_entry, self_e, sizeof(int) * (nc > 0 ? nc : 1), hipMemcpyHostToDevice) !=
                    hipSuccess ||
                hipMemcpy(g->d_inv_off, ioff, sizeof(int) * (nc + 1), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(g->d_inv_list, ilist, sizeof(int) * (itot > 0 ? itot : 1), hipMemcpyHostToDevice) !=
                    hipSuccess)
                bad = 2;
        }
        free(hcnt);
        free(icnt);
        free(hoff);
        free(ioff);
        free(hfill);
        free(ifill);
        free(self_e);
        free(nbr);
        free(nbs);
        free(ilist);
        if (bad) return gerror_(bad == 1 ? engine_err_malloc : engine_err_gpu, "half-shell lists");
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

    {
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, g->device));
        g->ncu = prop.multiProcessorCount;
    }
    /* Few, densely filled cells: split each owner's neighbour list into
       nsplit groups, so that the pair kernel has ~MD_SPLIT_ITEMS_PER_CU work
       items per CU.  (With few atoms per cell a visit is short and the extra
       items cost more than they spread.) */
    g->nsplit = 1;
    if (!(e->flags & engine_flag_nosplit) && s->nr_cells < MD_SPLIT_ITEMS_PER_CU * g->ncu &&
        (long long)s->nr_parts >= (long long)engine_gpu_half_mean * s->nr_cells) {
        g->nsplit = (MD_SPLIT_ITEMS_PER_CU * g->ncu + s->nr_cells - 1) / s->nr_cells;
        if (g->nsplit > MD_SPLIT_MAX) g->nsplit = MD_SPLIT_MAX;
    }
    g->P.nsplit = g->nsplit;
    {
        /* Owner cells in blocks of 6x6x6, blocks and the cells inside them in
           raster order: the owners a compute die works on at once, and their
           neighbours, stay close together in its L2. */
        int *ord = (int *)malloc(sizeof(int) * (size_t)s->nr_cells);
        if (ord == NULL) return gerror_(engine_err_malloc, "malloc failed");
        const int B = 6, nbx = (s->cdim[0] + B - 1) / B, nby = (s->cdim[1] + B - 1) / B,
                  nbz = (s->cdim[2] + B - 1) / B;
        int q = 0;
        for (int bx = 0; bx < nbx; bx++)
            for (int by = 0; by < nby; by++)
                for (int bz = 0; bz < nbz; bz++)
                    for (int x = bx * B; x < (bx + 1) * B && x < s->cdim[0]; x++)
                        for (int y = by * B; y < (by + 1) * B && y < s->cdim[1]; y++)
                            for (int z = bz * B; z < (bz + 1) * B && z < s->cdim[2]; z++)
                                ord[q++] = space_cellid(s, x, y, z);
        if (q != s->nr_cells) {
            free(ord);
            return gerror_(engine_err_range, "cell order");
        }
        HIPCHECK(hipMalloc(&g->d_order, sizeof(int) * (size_t)s->nr_cells));
        HIPCHECK(hipMemcpy(g->d_order, ord, sizeof(int) * (size_t)s->nr_cells, hipMemcpyHostToDevice));
        free(ord);
        g->P.order = g->d_order;
    }
    HIPCHECK(hipMalloc(&g->d_cnt_off, sizeof(int2) * s->nr_cells));
    HIPCHECK(hipMalloc(&g->d_cell_acc, sizeof(double) * 3 * (size_t)s->nr_cells * g->nsplit));
    HIPCHECK(hipMalloc(&g->d_cell_cnt, sizeof(unsigned long long) * 3 * (size_t)s->nr_cells * g->nsplit));
    HIPCHECK(hipMalloc(&g->d_tot, sizeof(double) * 4));
    HIPCHECK(hipMalloc(&g->d_red_acc, sizeof(double) * 3 * MD_RED_BLOCKS));
    HIPCHECK(hipMalloc(&g->d_red_cnt, sizeof(unsigned long long) * 3 * MD_RED_BLOCKS));
    HIPCHECK(hipMalloc(&g->d_tot_cnt, sizeof(unsigned long long) * 4));
    HIPCHECK(hipMalloc(&g->d_queue, sizeof(int) * (MD_NQUEUE + 1)));
    HIPCHECK(hipHostMalloc((void **)&g->h_count, sizeof(int) * s->nr_cells, hipHostMallocDefault));
    HIPCHECK(hipMalloc(&g->d_cell_maxdx, sizeof(double) * s->nr_cells));
    HIPCHECK(hipHostMalloc((void **)&g->h_cell_maxdx, sizeof(double) * s->nr_cells, hipHostMallocDefault));
    if ((g->h_sort_count = (int *)calloc(s->nr_cells, sizeof(int))) == NULL)
        return gerror_(engine_err_malloc, "malloc failed");
    g->sort_valid = 0;
    HIPCHECK(hipHostMalloc((void **)&g->h_offset, sizeof(int) * (s->nr_cells + 1), hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_cnt_off, sizeof(int2) * s->nr_cells, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_tot, sizeof(double) * 4, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&g->h_tot_cnt, sizeof(unsigned long long) * 4, hipHostMallocDefault));

    for (int k = 0; k < 8; k++) g->lds_nb[k] = 0;
    return engine_err_ok;
}

/* Pack host particles in cell order and upload them (no kernels). */
int gpu_upload(struct engine *e) {
    struct gpu_state *g = (struct gpu_state *)e->gpu;
    struct space *s = &e->s;
    int cid, n = 0, maxc = 0, maxid = -1;
    for (cid = 0; cid < s->nr_cells; cid++) {
        int c = s->cells[cid].count;
        if (c > maxc) maxc = c;
        n += c;
    }
    if (maxc > g->maxcell)
        return gerror_(engine_err_maxparts, "more particles in a cell than the packed record holds (1024 for up to "
                                            "64 types, 256 for up to 256, 64 for up to 1024)");
    if (gpu_alloc_parts(g, n > 0 ? n : 1) < 0) return engine_err_gpu;
    if (space_pack(s, g->max_type, g->h_count, g->h_offset, (int *)g->h_cnt_off, g->h_x, g->h_type, g->h_id, &maxc,
                   &maxid) < 0)
        return gerror_(engine_err_range, "particle type out of range");
    if (maxid + 1 > g->nids) return gerror_(engine_err_range, "particle id beyond the ids known at engine_gpu_load");
    g->nparts = n;
    g->maxcount = maxc;
    g->paired = !(e->flags & engine_flag_unpaired);
    /* default shell: the full shell, except for densely filled cells that
       are too few to fill the device (fewer than its CUs), or that the
       unpaired full-shell kernel would visit */
    g->half = (e->flags & engine_flag_halfshell)   ? 1
              : (e->flags & engine_flag_fullshell) ? 0
                                                   : (long long)n >= (long long)engine_gpu_half_mean * s->nr_cells &&
                                                         (!g->paired || s->nr_cells < g->ncu);
    /* surrogate mode: the network evaluation dominates a visit, so the half
       shell's single evaluation per pair wins at any occupancy */
    g->mlp_half = !(e->flags & engine_flag_fullshell);
    g->P.nparts = n;
    g->P.nids = g->nids;
    g->P.maxn = maxc > 0 ? maxc : 1;
    HIPCHECK(hipMemcpyAsync(g->d_x, g->h_x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_type, g->h_type, sizeof(int) * n, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_id, g->h_id, sizeof(int) * n, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_count, g->h_count, sizeof(int) * s->nr_cells, hipMemcpyHostToDevice, g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_offset, g->h_offset, sizeof(int) * (s->nr_cells + 1), hipMemcpyHostToDevice,
                            g->stream));
    HIPCHECK(hipMemcpyAsync(g->d_cnt_off, g->h_cnt_off, sizeof(int2) * s->nr_cells, hipMemcpyHostToDevice,
                            g->stream));
    return engine_err_ok;
}

/* Launch the kernels of one pass; optional events around each. */
/* One device pass on the resident particles; mlp selects the surrogate
   mode of the pair kernel. */
int gpu_launch(struct gpu_state *g, int timed, int mlp, int sort) {
    /* kernel: 0 table, 1 surrogate, 2 table with the 5-wave budget, 3 / 4
       the half-shell table kernels (4 / 5 waves), 5 the half-shell
       surrogate kernel, 6 / 7 the paired table kernels (4 / 5 waves; not
       for the projected window, whose pass also counts the reference
       window's pairs) */
    const int wide = g->ncells >= 4 * 16 * g->ncu;
    const int kern = mlp      ? (g->mlp_half ? 5 : 1)
                     : g->half ? (wide ? 4 : 3)
                     : (g->paired && !g->P.proj) ? (wide ? 7 : 6)
                                 : (wide ? 2 : 0);
    struct nb_params P = g->P;
    int nb;
    size_t lds = 0;
    const size_t maxn = (size_t)P.maxn;
    lds = sizeof(float4) * 64; /* raw */
    P.lds_own_key = (int)lds;
    lds += sizeof(unsigned short) * 13 * 64;
    lds = (lds + 15) & ~(size_t)15;
    P.lds_kfl = (int)lds;
    lds += sizeof(int) * 32;
    lds = (lds + 15) & ~(size_t)15;
    P.lds_stage = (int)lds;
    lds += sizeof(float4) * maxn * (kern >= 6 ? 2 : 1); /* paired visits stage two cells */
    P.lds_hit = (int)lds;
    lds += sizeof(unsigned short) * 64 * (mlp ? MD_HCAP_MLP : (kern == 2 || kern == 7) ? MD_HCAP_WIDE : MD_HCAP);
    lds = (lds + 15) & ~(size_t)15;
    P.lds_dense = (int)lds;
    if (mlp) lds += 16 * 64 * MD_HCAP_MLP + 4 * 64 * MD_XCAP_MLP + sizeof(float) * 2 * MLP_H;
    lds = (lds + 15) & ~(size_t)15;
    P.lds_jf = (int)lds;
    if (kern >= 3 && kern <= 5) lds += sizeof(float) * 4 * 3 * maxn;
    if (lds > 64 * 1024) return gerror_(engine_err_maxparts, "LDS staging exceeds 64 KiB");
    if (kern >= 3 && kern <= 5 && (g->d_jpart == NULL || g->jcap < (int)maxn)) {
        /* per half-shell entry, room for the visited cell's atoms; no
           pending pass may still use the old buffer */
        HIPCHECK(hipStreamSynchronize(g->stream));
        (void)hipFree(g->d_jpart);
        g->d_jpart = NULL;
        g->jcap = (int)maxn + 8;
        HIPCHECK(hipMalloc(&g->d_jpart, sizeof(float4) * (size_t)g->jcap * (g->n_entries_h > 0 ? g->n_entries_h : 1)));
    }
    if (g->nsplit > 1 && g->pcap < (size_t)g->nsplit * g->nparts) {
        /* per neighbour group, the owner forces of every atom; no pending
           pass may still use the old buffers */
        HIPCHECK(hipStreamSynchronize(g->stream));
        (void)hipFree(g->d_phi);
        (void)hipFree(g->d_plo);
        g->d_phi = g->d_plo = NULL;
        g->pcap = 0;
        HIPCHECK(hipMalloc(&g->d_phi, sizeof(float4) * (size_t)g->nsplit * g->nparts));
        HIPCHECK(hipMalloc(&g->d_plo, sizeof(float4) * (size_t)g->nsplit * g->nparts));
        g->pcap = (size_t)g->nsplit * g->nparts;
    }
    if (lds != g->lds_nb[kern]) {
        /* Persistent grid: as many single-wave blocks as can be resident. */
        int per_cu = 0;
        if (kern == 1)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond_mlp, 64, lds));
        else if (kern == 5)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond_mlp_half, 64, lds));
        else if (kern == 2)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond_wide, 64, lds));
        else if (kern == 3)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond_half, 64, lds));
        else if (kern == 4)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond_half_wide, 64, lds));
        else if (kern == 6)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond_paired, 64, lds));
        else if (kern == 7)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond_paired_wide, 64, lds));
        else
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbond, 64, lds));
        if (per_cu < 1) per_cu = 1;
        g->grid_nb[kern] = g->ncu * per_cu;
        g->lds_nb[kern] = lds;
    }
    if (mlp && g->mlp_np == 0) return gerror_(engine_err_range, "engine_flag_mlp without engine_mlp_init");
    if (!sort && !g->sort_valid) return gerror_(engine_err_range, "no sort lists to reuse");

    if (timed) HIPCHECK(hipEventRecord(g->ev[0], g->stream));
    nb = (g->nids + 255) / 256;
    if (nb > 0) hipLaunchKernelGGL(k_clear_ids, dim3(nb), dim3(256), 0, g->stream, g->nids, g->d_index_of_id);
    nb = ((g->nparts > g->ncells ? g->nparts : g->ncells) + 255) / 256;
    hipLaunchKernelGGL(k_pack, dim3(nb), dim3(256), 0, g->stream, g->nparts, g->d_x, g->d_type, g->d_id, g->ncells,
                       g->d_offset, g->d_xf, g->d_index_of_id, g->d_cell_of, g->d_queue);
    if (!sort) {
        /* pseudo-Verlet step between rebuilds: the sort lists of the last
           sort stay (runner.c:773) */
    } else if (g->maxcount <= MD_SORT_LDS_MAX) {
        const int maxn = g->maxcount > 0 ? g->maxcount : 1;
        const size_t sh = sizeof(double) * 3 * maxn + sizeof(unsigned int) * 13 * maxn +
                          sizeof(int) * 13 * 2 * MD_SORT_STACK + sizeof(unsigned short) * (13 * 2 * (maxn + 1) + 136);
        /* one wave per cell; for large cells (frequent repeated keys) one
           wave per direction */
        const int bs = maxn > MD_SORT_ONEWAVE_MAX ? 13 * 64 : 64;
        hipLaunchKernelGGL(k_sort_rank, dim3(g->ncells), dim3(bs), sh, g->stream, P, g->d_count, g->d_offset,
                           g->d_sortflags, g->d_x, g->d_sortlist, g->d_keyof, g->d_queue + MD_NQUEUE, maxn);
    } else {
        nb = (g->ncells * 13 + 255) / 256;
        hipLaunchKernelGGL(k_sort, dim3(nb), dim3(256), 0, g->stream, P, g->d_count, g->d_offset, g->d_sortflags,
                           g->d_x, g->d_sortlist, g->d_keyof, g->d_queue + MD_NQUEUE);
    }
    if (sort) g->sort_valid = 1;
    if (timed) HIPCHECK(hipEventRecord(g->ev[1], g->stream));
    struct nb_out O;
    O.fhi = g->d_fhi;
    O.flo = g->d_flo;
    O.cell_acc = g->d_cell_acc;
    O.cell_cnt = g->d_cell_cnt;
    O.jpart = g->d_jpart;
    O.jcap = g->jcap;
    O.phi = g->d_phi;
    O.plo = g->d_plo;
    /* no more waves than work items: the dispatcher then spreads them over
       all compute units */
    const int nitems = g->ncells * g->nsplit;
    const int grid = g->grid_nb[kern] < nitems ? g->grid_nb[kern] : nitems;
    if (kern == 3 || kern == 4) {
        hipLaunchKernelGGL(kern == 4 ? k_nonbond_half_wide : k_nonbond_half, dim3(grid), dim3(64), lds,
                           g->stream, P, g->d_cnt_off, g->d_nbr_h, g->d_nbr_off_h, g->d_nbs_h, g->d_origin, g->d_loc,
                           g->d_xf, g->d_x, g->d_id, g->d_sortlist, g->d_keyof, g->d_ptab,
                           (const float4 *)g->d_coeffs, g->d_excl_off, g->d_excl_list, g->d_index_of_id,
                           g->d_cell_of, O, g->d_queue);
    } else if (kern == 5)
        hipLaunchKernelGGL(k_nonbond_mlp_half, dim3(grid), dim3(64), lds, g->stream, P, g->d_cnt_off, g->d_nbr_h,
                           g->d_nbr_off_h, g->d_nbs_h, g->d_origin, g->d_loc, g->d_xf, g->d_x, g->d_id, g->d_sortlist,
                           g->d_keyof, g->d_ptab, (const float4 *)g->d_coeffs, g->d_excl_off, g->d_excl_list,
                           g->d_index_of_id, g->d_cell_of, O, g->d_queue, gpu_mlp_net(g), g->d_ptype);
    else if (kern == 1)
        hipLaunchKernelGGL(k_nonbond_mlp, dim3(grid), dim3(64), lds, g->stream, P, g->d_cnt_off, g->d_nbr,
                           g->d_nbr_off, g->d_nbs, g->d_origin, g->d_loc, g->d_xf, g->d_x, g->d_id, g->d_sortlist,
                           g->d_keyof, g->d_ptab, (const float4 *)g->d_coeffs, g->d_excl_off, g->d_excl_list,
                           g->d_index_of_id, g->d_cell_of, O, g->d_queue, gpu_mlp_net(g), g->d_ptype);
    else if (kern >= 6)
        hipLaunchKernelGGL(kern == 7 ? k_nonbond_paired_wide : k_nonbond_paired, dim3(grid), dim3(64), lds, g->stream,
                           P, g->d_cnt_off, g->d_nbr_p,
                           g->d_nbr_off, g->d_nbs_p, g->d_origin, g->d_loc, g->d_xf, g->d_x, g->d_id, g->d_sortlist,
                           g->d_keyof, g->d_ptab, (const float4 *)g->d_coeffs, g->d_excl_off, g->d_excl_list,
                           g->d_index_of_id, g->d_cell_of, O, g->d_queue);
    else
        hipLaunchKernelGGL(kern == 2 ? k_nonbond_wide : k_nonbond, dim3(grid), dim3(64), lds, g->stream,
                           P, g->d_cnt_off, g->d_nbr,
                           g->d_nbr_off, g->d_nbs, g->d_origin, g->d_loc, g->d_xf, g->d_x, g->d_id, g->d_sortlist,
                           g->d_keyof, g->d_ptab, (const float4 *)g->d_coeffs, g->d_excl_off, g->d_excl_list,
                           g->d_index_of_id, g->d_cell_of, O, g->d_queue);
    if ((kern >= 3 && kern <= 5) || g->nsplit > 1)
        hipLaunchKernelGGL(k_fold, dim3(g->ncells), dim3(64), 0, g->stream, g->ncells, g->d_cnt_off, g->nsplit,
                           g->nparts, (const float4 *)g->d_phi, (const float4 *)g->d_plo,
                           (kern >= 3 && kern <= 5) ? 1 : 0,
                           g->d_self_entry, g->d_inv_off, g->d_inv_list, (const float4 *)g->d_jpart, g->jcap,
                           g->d_fhi, g->d_flo);
    if (timed) HIPCHECK(hipEventRecord(g->ev[2], g->stream));
    hipLaunchKernelGGL(k_reduce_a, dim3(MD_RED_BLOCKS), dim3(256), 0, g->stream, g->ncells * g->nsplit,
                       g->d_cell_acc, g->d_cell_cnt, g->d_red_acc, g->d_red_cnt);
    hipLaunchKernelGGL(k_reduce_b, dim3(1), dim3(256), 0, g->stream, MD_RED_BLOCKS, g->d_red_acc, g->d_red_cnt,
                       g->d_tot, g->d_tot_cnt);
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

/* Pseudo-Verlet (engine_flag_verlet*): the pass sorts only on a rebuild
   step (runner.c:773) and widens each pair window by the cells' maxdx
   (runner_dopair.c:428).  Sets *sort and the device per-cell maxdx; a pass
   that reuses the sort lists needs the layout of the last sort. */
static int gpu_verlet(struct engine *e, struct gpu_state *g, int *sort) {
    struct space *s = &e->s;
    int cid;
    if (!(e->flags & engine_flag_verlet)) {
        g->P.maxdx = NULL;
        *sort = 1;
        return engine_err_ok;
    }
    for (cid = 0; cid < s->nr_cells; cid++) g->h_cell_maxdx[cid] = s->cells[cid].maxdx;
    HIPCHECK(hipMemcpyAsync(g->d_cell_maxdx, g->h_cell_maxdx, sizeof(double) * s->nr_cells, hipMemcpyHostToDevice,
                            g->stream));
    g->P.maxdx = g->d_cell_maxdx;
    *sort = s->verlet_rebuild != 0;
    if (!*sort) {
        if (!g->sort_valid) return gerror_(engine_err_range, "no sort lists to reuse");
        for (cid = 0; cid < s->nr_cells; cid++)
            if (s->cells[cid].count != g->h_sort_count[cid])
                return gerror_(engine_err_range, "cells changed since the last sort");
    }
    return engine_err_ok;
}

static double gpu_wall(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Counts of the layout the sort lists belong to. */
static void gpu_sorted(struct engine *e, struct gpu_state *g) {
    for (int cid = 0; cid < e->s.nr_cells; cid++) g->h_sort_count[cid] = e->s.cells[cid].count;
}

extern "C" int engine_nonbond_gpu(struct engine *e) {
    struct gpu_state *g;
    struct space *s;
    int flag = 0;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    g = (struct gpu_state *)e->gpu;
    if (g == NULL) return gerror_(engine_err_nogpu, "engine_gpu_load was not called");
    s = &e->s;
    HIPCHECK(hipSetDevice(g->device));
    double tw = gpu_wall();
    if (gpu_upload(e) < 0) return engine_err_gpu;
    e->timers[engine_timer_gpu_load] = gpu_wall() - tw;
    int sort = 1;
    if (gpu_verlet(e, g, &sort) < 0) return engine_err_range;
    HIPCHECK(hipMemsetAsync(g->d_queue + MD_NQUEUE, 0, sizeof(int), g->stream));
    if (gpu_launch(g, 1, (e->flags & engine_flag_mlp) != 0, sort) < 0) return engine_err_gpu;
    if (sort) gpu_sorted(e, g);
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
    tw = gpu_wall();
    space_add_forces_slots(s, g->h_offset, (const float *)g->h_fhi, (const float *)g->h_flo);
    e->timers[engine_timer_gpu_unload] = gpu_wall() - tw;
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
    /* the sort of a pass that engine_step would make, on the resident data */
    const int sort = !(e->flags & engine_flag_verlet) || e->s.verlet_rebuild || !g->sort_valid;
    for (k = 0; k < nwarm; k++)
        if (gpu_launch(g, 0, (e->flags & engine_flag_mlp) != 0, sort) < 0) return engine_err_gpu;
    HIPCHECK(hipStreamSynchronize(g->stream));
    for (k = 0; k < nsteps; k++) {
        if (gpu_launch(g, 1, (e->flags & engine_flag_mlp) != 0, sort) < 0) return engine_err_gpu;
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
