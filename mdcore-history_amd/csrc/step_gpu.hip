/*
 * Whole MD steps on the device: engine_gpu_steps.
 *
 * One step is what engine_step does on the host (engine.c:2070-2240 with
 * engine_advance 1945-2052 and space_shuffle_local), with the particles kept
 * on the GPU between steps:
 *   1. shuffle: particles outside their cell move to the neighbour cell;
 *   2. the nonbonded pass (k_pack, k_sort*, k_nonbond*, k_reduce);
 *   3. kick-drift v += f dt / m, x += dt v in FP64, then migration.
 * Migration keeps the host's order exactly, so the sorted lists, the pair
 * sets and the trajectory are bitwise those of the host-driven loop:
 *   - per source cell, in storage order, a leaving particle is appended to
 *     its departures and the cell's last particle takes its slot (k_migrate,
 *     one wave per cell, with the kick fused in);
 *   - new counts and offsets come from a fixed-order three-pass scan
 *     (k_scan_a/b/c), arrival counts from atomic increments (a sum, so the
 *     order of the increments does not matter);
 *   - stayers keep their relative order (k_stayers); each destination
 *     appends its arrivals by source cell and departure order (k_arrive).
 * The host's shuffle at the start of a step moves nothing when the previous
 * advance left every particle inside its cell; it is then skipped.
 * No reference counterpart as a device path: the reference's CUDA engine
 * returns to the host every step (runner_cuda.cu:2405-2430, 2514-2521).
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gpu_state.h"
#include "hostpack.h"

struct step_geom {
    double h[3];
    int cdim[3];
    int ncells;
};

#define MD_SCAN_CELLS 1024 /* cells per block of the count scan: 256 threads x 4 */

/* Leave code of a particle: 0 stays, else 1 + 9(dx+1) + 3(dy+1) + (dz+1). */
__device__ __forceinline__ int leave_code(const double *xa, const double *h) {
    int d[3], any = 0;
    for (int m = 0; m < 3; m++) {
        d[m] = (xa[m] >= h[m]) - (xa[m] < 0.0);
        any |= d[m];
    }
    return any ? 1 + 9 * (d[0] + 1) + 3 * (d[1] + 1) + (d[2] + 1) : 0;
}

/* One wave per cell.  With kick: v += f dt / m, x += dt v in the host's FP64
   order (engine_advance, engine.c:1945-2052), written back in place, and the
   cell's kinetic energy.  Then the departures in the host's order: slot pid
   is examined, a leaving particle is appended to the departures and the
   cell's last particle takes its slot (space_shuffle_local).  Slots
   (pid, cnt) always still hold their original particle, so the walk only
   tracks which original particle sits at pid.  Output per old slot:
   slot >= 0, the particle's slot among the stayers, or -1 - its departure
   index; per departure its source slot and destination cell. */
__global__ __launch_bounds__(64) void k_migrate(struct step_geom G, int kick, double dt,
                                                const double *__restrict__ mass, const double *__restrict__ imass,
                                                const float4 *__restrict__ fhi, const float4 *__restrict__ flo,
                                                const int *__restrict__ count, const int *__restrict__ offset,
                                                double *__restrict__ x, double *__restrict__ v,
                                                const int *__restrict__ type, int *__restrict__ cell_of,
                                                int *__restrict__ slot, int *__restrict__ qsrc,
                                                int *__restrict__ qdest, int *__restrict__ stay,
                                                int *__restrict__ ndep, int *__restrict__ narr,
                                                double *__restrict__ cell_ke) {
#pragma clang fp contract(off)
    __shared__ unsigned char code[MD_MAX_CELL];
    __shared__ short res[MD_MAX_CELL];
    __shared__ int sh_stay, sh_nd;
    const int c = blockIdx.x, lane = threadIdx.x;
    const int off = offset[c], n = count[c];
    double ke = 0.0;
    for (int k = lane; k < n; k += 64) {
        const int a = off + k;
        double xa[3] = {x[3 * a], x[3 * a + 1], x[3 * a + 2]};
        if (kick) {
            const int t = type[a];
            const double w = dt * imass[t];
            const float4 hi = fhi[a], lo = flo[a];
            /* the host adds the pass's force into a zeroed FP64 force */
            const double f[3] = {0.0 + ((double)hi.x + (double)lo.x), 0.0 + ((double)hi.y + (double)lo.y),
                                 0.0 + ((double)hi.z + (double)lo.z)};
            double vv[3];
            for (int m = 0; m < 3; m++) {
                vv[m] = v[3 * a + m] + f[m] * w;
                xa[m] = xa[m] + dt * vv[m];
                v[3 * a + m] = vv[m];
                x[3 * a + m] = xa[m];
            }
            ke += 0.5 * mass[t] * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
        }
        code[k] = (unsigned char)leave_code(xa, G.h);
        cell_of[a] = c;
    }
    __syncthreads();
    if (lane == 0) {
        int pid = 0, cnt = n, nd = 0, cur = 0;
        while (pid < cnt) {
            if (code[cur]) {
                res[cur] = (short)(-1 - nd);
                nd++;
                cnt--;
                cur = cnt; /* the last particle, still in its original slot, fills the hole */
            } else {
                res[cur] = (short)pid;
                pid++;
                cur = pid;
            }
        }
        sh_stay = cnt;
        sh_nd = nd;
    }
    __syncthreads();
    const int loc[3] = {c / (G.cdim[1] * G.cdim[2]), (c / G.cdim[2]) % G.cdim[1], c % G.cdim[2]};
    for (int k = lane; k < n; k += 64) {
        const int r = res[k];
        slot[off + k] = r;
        if (r < 0) {
            const int q = off + (-1 - r), cd = code[k] - 1;
            const int d[3] = {cd / 9 - 1, (cd / 3) % 3 - 1, cd % 3 - 1};
            const int dest = (((loc[0] + d[0] + G.cdim[0]) % G.cdim[0]) * G.cdim[1] +
                              (loc[1] + d[1] + G.cdim[1]) % G.cdim[1]) * G.cdim[2] +
                             (loc[2] + d[2] + G.cdim[2]) % G.cdim[2];
            qsrc[q] = k;
            qdest[q] = dest;
            atomicAdd(&narr[dest], 1);
        }
    }
    if (kick) {
        for (int s = 32; s > 0; s >>= 1) ke += __shfl_xor(ke, s, 64);
        if (lane == 0) cell_ke[c] = ke;
    }
    if (lane == 0) {
        stay[c] = sh_stay;
        ndep[c] = sh_nd;
    }
}

/* Count scan, pass 1: per block of MD_SCAN_CELLS cells the sum and maximum
   of the new counts (stayers + arrivals) and the sum of the cell energies. */
__global__ __launch_bounds__(256) void k_scan_a(int ncells, const int *__restrict__ stay,
                                                const int *__restrict__ narr, const double *__restrict__ cell_ke,
                                                int *__restrict__ bsum, int *__restrict__ bmax,
                                                double *__restrict__ bke) {
    __shared__ int ss[256], sm[256];
    __shared__ double sk[256];
    const int t = threadIdx.x, c0 = blockIdx.x * MD_SCAN_CELLS + 4 * t;
    int sum = 0, mx = 0;
    double ke = 0.0;
    for (int c = c0; c < c0 + 4 && c < ncells; c++) {
        const int n = stay[c] + narr[c];
        sum += n;
        mx = n > mx ? n : mx;
        if (cell_ke) ke += cell_ke[c];
    }
    ss[t] = sum;
    sm[t] = mx;
    sk[t] = ke;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] += ss[t + w];
            sm[t] = sm[t] > sm[t + w] ? sm[t] : sm[t + w];
            sk[t] += sk[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        bsum[blockIdx.x] = ss[0];
        bmax[blockIdx.x] = sm[0];
        bke[blockIdx.x] = sk[0];
    }
}

/* Pass 2, one thread: block bases, the total, the largest count, the
   kinetic energy and the pass totals of this step; clears the
   outside-the-cell flag. */
__global__ __launch_bounds__(64) void k_scan_b(int nblocks, int ncells, int *__restrict__ bsum,
                                               const int *__restrict__ bmax, const double *__restrict__ bke,
                                               int *__restrict__ offset, int *__restrict__ st,
                                               const double *__restrict__ tot, double *__restrict__ esteps) {
    if (threadIdx.x != 0) return;
    int run = 0, mx = 0;
    double ke = 0.0;
    for (int b = 0; b < nblocks; b++) {
        const int s = bsum[b];
        bsum[b] = run;
        run += s;
        mx = bmax[b] > mx ? bmax[b] : mx;
        ke += bke[b];
    }
    offset[ncells] = run;
    st[0] = mx;
    st[1] = 0;
    if (esteps) {
        esteps[0] = tot[0];
        esteps[1] = tot[1];
        esteps[2] = tot[2];
        esteps[3] = ke;
    }
}

/* Pass 3: counts, offsets and (count, offset) records of the new layout;
   clears the arrival counts for the next migration. */
__global__ __launch_bounds__(256) void k_scan_c(int ncells, const int *__restrict__ stay, int *__restrict__ narr,
                                                const int *__restrict__ bbase, int *__restrict__ count,
                                                int *__restrict__ offset, int2 *__restrict__ cnt_off) {
    __shared__ int ss[256];
    const int t = threadIdx.x, c0 = blockIdx.x * MD_SCAN_CELLS + 4 * t;
    int n[4] = {0, 0, 0, 0}, sum = 0;
    for (int k = 0; k < 4; k++)
        if (c0 + k < ncells) {
            n[k] = stay[c0 + k] + narr[c0 + k];
            sum += n[k];
        }
    ss[t] = sum;
    __syncthreads();
    for (int w = 1; w < 256; w <<= 1) {
        const int add = t >= w ? ss[t - w] : 0;
        __syncthreads();
        ss[t] += add;
        __syncthreads();
    }
    int run = bbase[blockIdx.x] + ss[t] - sum;
    for (int k = 0; k < 4; k++)
        if (c0 + k < ncells) {
            count[c0 + k] = n[k];
            offset[c0 + k] = run;
            cnt_off[c0 + k] = make_int2(n[k], run);
            narr[c0 + k] = 0;
            run += n[k];
        }
}

/* Stayers: old slot a to its slot in the new layout. */
__global__ __launch_bounds__(256) void k_stayers(int n, const int *__restrict__ cell_of, const int *__restrict__ slot,
                                                 const int *__restrict__ new_off, const double *__restrict__ x,
                                                 const double *__restrict__ v, const int *__restrict__ id,
                                                 const int *__restrict__ type, double *__restrict__ x2,
                                                 double *__restrict__ v2, int *__restrict__ id2,
                                                 int *__restrict__ type2) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int r = slot[a];
    if (r < 0) return;
    const int b = new_off[cell_of[a]] + r;
    for (int m = 0; m < 3; m++) {
        x2[3 * b + m] = x[3 * a + m];
        v2[3 * b + m] = v[3 * a + m];
    }
    id2[b] = id[a];
    type2[b] = type[a];
}

/* The wrapped coordinates l-1, l, l+1 of one axis in ascending order; a
   value equal to its predecessor is a repeat (fewer than 3 cells). */
__device__ __forceinline__ void sorted3(int l, int cdim, int *s) {
    int a = (l - 1 + cdim) % cdim, b = l, c = (l + 1) % cdim, t;
    if (a > b) t = a, a = b, b = t;
    if (b > c) t = b, b = c, c = t;
    if (a > b) t = a, a = b, b = t;
    s[0] = a;
    s[1] = b;
    s[2] = c;
}

/* Arrivals, one thread per destination cell: appended after the stayers in
   the host's order, by source cell id and then departure order (the source
   cells' ids ascend with their wrapped coordinates taken in lexicographic
   order).  Sets st[1] when an arrival is still outside its new cell. */
__global__ __launch_bounds__(128) void k_arrive(struct step_geom G, const int *__restrict__ old_off,
                                                const int *__restrict__ ndep, const int *__restrict__ qsrc,
                                                const int *__restrict__ qdest, const int *__restrict__ new_off,
                                                const int *__restrict__ stay, const double *__restrict__ x,
                                                const double *__restrict__ v, const int *__restrict__ id,
                                                const int *__restrict__ type, double *__restrict__ x2,
                                                double *__restrict__ v2, int *__restrict__ id2,
                                                int *__restrict__ type2, int *__restrict__ st) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G.ncells) return;
    const int loc[3] = {c / (G.cdim[1] * G.cdim[2]), (c / G.cdim[2]) % G.cdim[1], c % G.cdim[2]};
    int I[3], J[3], K[3];
    sorted3(loc[0], G.cdim[0], I);
    sorted3(loc[1], G.cdim[1], J);
    sorted3(loc[2], G.cdim[2], K);
    int b = new_off[c] + stay[c], outside = 0;
#pragma unroll
    for (int qi = 0; qi < 3; qi++) {
        if (qi > 0 && I[qi] == I[qi - 1]) continue;
#pragma unroll
        for (int qj = 0; qj < 3; qj++) {
            if (qj > 0 && J[qj] == J[qj - 1]) continue;
#pragma unroll
            for (int qk = 0; qk < 3; qk++) {
                if (qk > 0 && K[qk] == K[qk - 1]) continue;
                const int s = (I[qi] * G.cdim[1] + J[qj]) * G.cdim[2] + K[qk];
                const int off = old_off[s], nd = ndep[s];
                for (int q = off; q < off + nd; q++) {
                    if (qdest[q] != c) continue;
                    const int a = off + qsrc[q];
                    for (int m = 0; m < 3; m++) {
                        const double xm = x[3 * a + m];
                        const int d = (xm >= G.h[m]) - (xm < 0.0);
                        const double xn = xm - d * G.h[m];
                        x2[3 * b + m] = xn;
                        v2[3 * b + m] = v[3 * a + m];
                        outside |= (xn >= G.h[m]) | (xn < 0.0);
                    }
                    id2[b] = id[a];
                    type2[b] = type[a];
                    b++;
                }
            }
        }
    }
    if (outside) st[1] = 1;
}

/* Pseudo-Verlet advance, one wave per cell: with kick, v += f dt / m and
   x += dt v in place (engine_advance's Verlet branch, engine.c:1961-1985:
   no migration).  The cell's kinetic energy of the (new) velocities, and
   with oldx the largest squared displacement from the rebuild positions,
   ((0 + w0^2) + w1^2) + w2^2 with w = x - oldx as in engine_verlet_update
   (engine.c:222-233). */
__global__ __launch_bounds__(64) void k_advance_verlet(int kick, double dt, const double *__restrict__ mass,
                                                       const double *__restrict__ imass,
                                                       const float4 *__restrict__ fhi, const float4 *__restrict__ flo,
                                                       const int *__restrict__ count, const int *__restrict__ offset,
                                                       double *__restrict__ x, double *__restrict__ v,
                                                       const int *__restrict__ type, const double *__restrict__ oldx,
                                                       double *__restrict__ lmax2, double *__restrict__ cell_ke) {
#pragma clang fp contract(off)
    const int c = blockIdx.x, lane = threadIdx.x;
    const int off = offset[c], n = count[c];
    double ke = 0.0, lm = 0.0;
    for (int k = lane; k < n; k += 64) {
        const int a = off + k;
        double xa[3] = {x[3 * a], x[3 * a + 1], x[3 * a + 2]};
        if (kick) {
            const int t = type[a];
            const double w = dt * imass[t];
            const float4 hi = fhi[a], lo = flo[a];
            const double f[3] = {0.0 + ((double)hi.x + (double)lo.x), 0.0 + ((double)hi.y + (double)lo.y),
                                 0.0 + ((double)hi.z + (double)lo.z)};
            double vv[3];
            for (int m = 0; m < 3; m++) {
                vv[m] = v[3 * a + m] + f[m] * w;
                xa[m] = xa[m] + dt * vv[m];
                v[3 * a + m] = vv[m];
                x[3 * a + m] = xa[m];
            }
            ke += 0.5 * mass[t] * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
        } else {
            const double m = mass[type[a]], v0 = v[3 * a], v1 = v[3 * a + 1], v2 = v[3 * a + 2];
            ke += 0.5 * m * (v0 * v0 + v1 * v1 + v2 * v2);
        }
        if (oldx) {
            double dx = 0.0;
            for (int m = 0; m < 3; m++) {
                const double w = xa[m] - oldx[3 * a + m];
                dx += w * w;
            }
            lm = fmax(dx, lm);
        }
    }
    for (int s = 32; s > 0; s >>= 1) {
        ke += __shfl_xor(ke, s, 64);
        lm = fmax(lm, __shfl_xor(lm, s, 64));
    }
    if (lane == 0) {
        cell_ke[c] = ke;
        lmax2[c] = lm;
    }
}

/* Pass 1 of the pseudo-Verlet sums: per block of MD_SCAN_CELLS cells, the
   kinetic energy and the largest squared displacement. */
__global__ __launch_bounds__(256) void k_verlet_red_a(int ncells, const double *__restrict__ cell_ke,
                                                      const double *__restrict__ lmax2, double *__restrict__ bke,
                                                      double *__restrict__ bmx) {
    __shared__ double sk[256], sm[256];
    const int t = threadIdx.x, c0 = blockIdx.x * MD_SCAN_CELLS + 4 * t;
    double ke = 0.0, mx = 0.0;
    for (int c = c0; c < c0 + 4 && c < ncells; c++) {
        ke += cell_ke[c];
        mx = fmax(lmax2[c], mx);
    }
    sk[t] = ke;
    sm[t] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sk[t] += sk[t + w];
            sm[t] = fmax(sm[t], sm[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        bke[blockIdx.x] = sk[0];
        bmx[blockIdx.x] = sm[0];
    }
}

/* Pass 2, one thread: the step's energies (pass totals, kinetic) when es is
   given, and the largest squared displacement into *vmax. */
__global__ __launch_bounds__(64) void k_verlet_red_b(int nblocks, const double *__restrict__ bke,
                                                     const double *__restrict__ bmx, const double *__restrict__ tot,
                                                     double *__restrict__ es, double *__restrict__ vmax) {
    if (threadIdx.x != 0) return;
    double ke = 0.0, mx = 0.0;
    for (int b = 0; b < nblocks; b++) {
        ke += bke[b];
        mx = fmax(bmx[b], mx);
    }
    *vmax = mx;
    if (es) {
        es[0] = tot[0];
        es[1] = tot[1];
        es[2] = tot[2];
        es[3] = ke;
    }
}

/* Each cell's maxdx = sqrt(largest squared displacement), engine.c:231. */
__global__ __launch_bounds__(256) void k_cell_maxdx(int ncells, const double *__restrict__ lmax2,
                                                    double *__restrict__ cell_maxdx) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncells) cell_maxdx[c] = sqrt(lmax2[c]);
}

/* ------------------------------------------------------------ rigid */

#define MD_RIG_NP 4 /* particles per group on the device */
#define MD_RIG_NC 6 /* constraints per group on the device */

struct rigid_dev {
    int n;                  /* groups */
    const int *np, *nc;     /* per group */
    const int *parts;       /* [n][MD_RIG_NP] particle ids */
    const int *ij;          /* [n][MD_RIG_NC][2] indices into parts */
    const double *d2;       /* [n][MD_RIG_NC] */
    float *a;               /* [n][MD_RIG_NC^2] P-SHAKE coupling, kept */
};

/* Slot of each particle id and cell of each slot, current layout. */
__global__ __launch_bounds__(256) void k_layout(int n, int ncells, const int *__restrict__ id,
                                                const int *__restrict__ offset, int *__restrict__ slot_of_id,
                                                int *__restrict__ cell_of) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) slot_of_id[id[a]] = a;
    if (a < ncells)
        for (int k = offset[a]; k < offset[a + 1]; k++) cell_of[k] = a;
}

/* One thread per rigid group: rigid_eval_shake (shake != 0, rigid.c:81-277)
   or rigid_eval_pshake (rigid.c:279-523) with the host's FP64 (and, for the
   coupling matrix, FP32) arithmetic in the same order.  fcorr[3 id + j]
   receives P-SHAKE's force m (x_new - x) / dt^2; st[1] is set when a
   corrected position left its cell. */
__global__ __launch_bounds__(128) void k_rigid(struct rigid_dev R, int shake, int a_update, double dt, double tol,
                                               struct step_geom G, const double *__restrict__ mass,
                                               const int *__restrict__ slot_of_id, const int *__restrict__ cell_of,
                                               const int *__restrict__ type, double *__restrict__ x,
                                               double *__restrict__ v, double *__restrict__ fcorr,
                                               int *__restrict__ st) {
#pragma clang fp contract(off)
    const int rid = blockIdx.x * blockDim.x + threadIdx.x;
    if (rid >= R.n) return;
    const int np = R.np[rid], nc = R.nc[rid];
    int slot[MD_RIG_NP], cell[MD_RIG_NP], shift[3 * MD_RIG_NP], ci[MD_RIG_NC], cj[MD_RIG_NC];
    double m[MD_RIG_NP], xp[3 * MD_RIG_NP], xo[3 * MD_RIG_NP], res[MD_RIG_NC], d2[MD_RIG_NC];
    const double idt = 1.0 / dt;
    for (int k = 0; k < nc; k++) {
        ci[k] = R.ij[(rid * MD_RIG_NC + k) * 2];
        cj[k] = R.ij[(rid * MD_RIG_NC + k) * 2 + 1];
        d2[k] = R.d2[rid * MD_RIG_NC + k];
    }
    for (int k = 0; k < np; k++)
        if (slot_of_id[R.parts[rid * MD_RIG_NP + k]] < 0) return; /* not loaded: skipped, as on the host */
    for (int k = 0; k < np; k++) {
        const int a = slot_of_id[R.parts[rid * MD_RIG_NP + k]];
        slot[k] = a;
        cell[k] = cell_of[a];
        m[k] = mass[type[a]];
    }
    const int c0 = cell[0];
    const int l0[3] = {c0 / (G.cdim[1] * G.cdim[2]), (c0 / G.cdim[2]) % G.cdim[1], c0 % G.cdim[2]};
    for (int k = 0; k < np; k++) {
        const int ck = cell[k];
        const int lk[3] = {ck / (G.cdim[1] * G.cdim[2]), (ck / G.cdim[2]) % G.cdim[1], ck % G.cdim[2]};
        for (int j = 0; j < 3; j++) {
            int sh = 0;
            if (ck != c0) {
                sh = lk[j] - l0[j];
                if (sh > 1)
                    sh = -1;
                else if (sh < -1)
                    sh = 1;
            }
            shift[3 * k + j] = sh;
            const double xk = x[3 * slot[k] + j];
            xp[3 * k + j] = (ck != c0) ? xk + G.h[j] * sh : xk;
            xo[3 * k + j] = xp[3 * k + j] - dt * v[3 * slot[k] + j];
        }
    }
    if (shake) {
        double vc[3 * MD_RIG_NC], wvc[3 * MD_RIG_NC];
        for (int k = 0; k < nc; k++)
            for (int j = 0; j < 3; j++) {
                vc[3 * k + j] = xo[3 * ci[k] + j] - xo[3 * cj[k] + j];
                wvc[3 * k + j] = 1.0 / (m[ci[k]] + m[cj[k]]) * vc[3 * k + j];
            }
        for (int iter = 0; iter < rigid_maxiter; iter++) {
            double max_res = 0.0;
            for (int k = 0; k < nc; k++) {
                const int pid = ci[k], pjd = cj[k];
                res[k] = d2[k];
                for (int j = 0; j < 3; j++) res[k] -= (xp[3 * pid + j] - xp[3 * pjd + j]) * (xp[3 * pid + j] - xp[3 * pjd + j]);
                if (fabs(res[k]) > max_res) max_res = fabs(res[k]);
            }
            if (max_res < tol) break;
            for (int k = 0; k < nc; k++) {
                const int pid = ci[k], pjd = cj[k];
                const double lambda =
                    0.5 * res[k] /
                    ((xp[3 * pid] - xp[3 * pjd]) * vc[3 * k] + (xp[3 * pid + 1] - xp[3 * pjd + 1]) * vc[3 * k + 1] +
                     (xp[3 * pid + 2] - xp[3 * pjd + 2]) * vc[3 * k + 2]);
                for (int j = 0; j < 3; j++) {
                    const double w = lambda * wvc[3 * k + j];
                    xp[3 * pid + j] += w * m[pjd];
                    xp[3 * pjd + j] -= w * m[pid];
                }
            }
        }
    } else {
        float *A = R.a + rid * MD_RIG_NC * MD_RIG_NC;
        double vc[3 * MD_RIG_NC * MD_RIG_NP];
        for (int q = 0; q < 3 * nc * np; q++) vc[q] = 0.0;
        for (int k = 0; k < nc; k++) {
            const int pid = ci[k], pjd = cj[k];
            double dv[3];
            for (int j = 0; j < 3; j++) dv[j] = (xo[3 * pid + j] - xo[3 * pjd + j]) / (m[pid] + m[pjd]);
            for (int i = 0; i < nc; i++)
                if (A[k * nc + i] != 0.0f)
                    for (int j = 0; j < 3; j++) {
                        vc[i * np * 3 + pid * 3 + j] += A[k * nc + i] * dv[j] * m[pjd];
                        vc[i * np * 3 + pjd * 3 + j] += -A[k * nc + i] * dv[j] * m[pid];
                    }
        }
        int iter;
        for (iter = 0; iter < rigid_maxiter; iter++) {
            double max_res = 0.0;
            for (int k = 0; k < nc; k++) {
                const double *g = &vc[k * 3 * np];
                const int pid = ci[k], pjd = cj[k];
                res[k] = d2[k];
                for (int j = 0; j < 3; j++) res[k] -= (xp[3 * pid + j] - xp[3 * pjd + j]) * (xp[3 * pid + j] - xp[3 * pjd + j]);
                if (fabs(res[k]) > max_res) max_res = fabs(res[k]);
                const double lambda = 0.5 * res[k] /
                                      ((xp[3 * pid] - xp[3 * pjd]) * (g[3 * pid] - g[3 * pjd]) +
                                       (xp[3 * pid + 1] - xp[3 * pjd + 1]) * (g[3 * pid + 1] - g[3 * pjd + 1]) +
                                       (xp[3 * pid + 2] - xp[3 * pjd + 2]) * (g[3 * pid + 2] - g[3 * pjd + 2]));
                for (int j = 0; j < 3 * np; j++) xp[j] += lambda * g[j];
            }
            if (max_res < tol) break;
        }
        if (nc > 1 && (a_update || iter > rigid_pshake_refine)) {
            float a_new[MD_RIG_NC * MD_RIG_NC], tmp[MD_RIG_NC * MD_RIG_NC];
            float w, dx[3], max_alpha = 0.0f;
            for (int i = 0; i < nc; i++) {
                const double *gi = &vc[i * 3 * np];
                const int pid = ci[i], pjd = cj[i];
                for (int j = 0; j < 3; j++) dx[j] = xp[3 * pid + j] - xp[3 * pjd + j];
                w = -1.0 / (dx[0] * (gi[pid * 3] - gi[pjd * 3]) + dx[1] * (gi[pid * 3 + 1] - gi[pjd * 3 + 1]) +
                            dx[2] * (gi[pid * 3 + 2] - gi[pjd * 3 + 2]));
                for (int j = i + 1; j < nc; j++) {
                    const double *gj = &vc[j * 3 * np];
                    a_new[i + j * nc] = w * (dx[0] * (gj[pid * 3] - gj[pjd * 3]) + dx[1] * (gj[pid * 3 + 1] - gj[pjd * 3 + 1]) +
                                             dx[2] * (gj[pid * 3 + 2] - gj[pjd * 3 + 2]));
                    a_new[j + i * nc] = a_new[i + j * nc];
                    max_alpha = fmaxf(max_alpha, fabsf(a_new[i + j * nc]));
                }
                a_new[i + i * nc] = 1.0f;
            }
            if (max_alpha > rigid_pshake_maxalpha) {
                w = 0.1 / max_alpha;
                for (int i = 0; i < nc * nc; i++) a_new[i] *= w;
                for (int i = 0; i < nc; i++) a_new[i * nc + i] = 1.0f;
            }
            for (int i = 0; i < nc * nc; i++) tmp[i] = 0.0f;
            for (int j = 0; j < nc; j++)
                for (int k = 0; k < nc; k++)
                    for (int i = 0; i < nc; i++) tmp[j * nc + i] += A[k * nc + i] * a_new[j * nc + k];
            for (int i = 0; i < nc * nc; i++) A[i] = tmp[i];
        }
    }
    int outside = 0;
    for (int k = 0; k < np; k++) {
        const int a = slot[k], id = R.parts[rid * MD_RIG_NP + k];
        for (int j = 0; j < 3; j++) {
            const double xn = (cell[k] != c0) ? xp[3 * k + j] - G.h[j] * shift[3 * k + j] : xp[3 * k + j];
            if (!shake) fcorr[3 * id + j] = idt * idt * (xn - x[3 * a + j]) * m[k];
            v[3 * a + j] += idt * (xn - x[3 * a + j]);
            x[3 * a + j] = xn;
            outside |= (xn >= G.h[j]) | (xn < 0.0);
        }
    }
    if (outside) st[1] = 1;
}

/* ------------------------------------------------------------ host side */

struct step_bufs {
    double *v, *x2, *v2, *mass, *imass, *cell_ke, *bke, *esteps;
    double *oldx, *lmax2, *bmx, *vmax; /* pseudo-Verlet */
    /* rigid groups: device copies, slot of each id, P-SHAKE forces by id */
    int *rg_np, *rg_nc, *rg_parts, *rg_ij, *sid;
    double *rg_d2, *fcorr;
    float *rg_a;
    int nrig;
    int *id2, *type2, *slot, *qsrc, *qdest, *stay, *ndep, *narr, *bsum, *bmax, *st, *count2, *offset2;
    int nblocks;
};

static void step_free(struct step_bufs *b) {
    double *d[] = {b->v, b->x2, b->v2, b->mass, b->imass, b->cell_ke, b->bke, b->esteps,
                   b->oldx, b->lmax2, b->bmx, b->vmax, b->rg_d2, b->fcorr};
    int *i[] = {b->id2, b->type2, b->slot, b->qsrc, b->qdest, b->stay, b->ndep, b->narr, b->bsum,
                b->bmax, b->st, b->count2, b->offset2, b->rg_np, b->rg_nc, b->rg_parts, b->rg_ij, b->sid};
    (void)hipFree(b->rg_a);
    for (size_t k = 0; k < sizeof(d) / sizeof(d[0]); k++) (void)hipFree(d[k]);
    for (size_t k = 0; k < sizeof(i) / sizeof(i[0]); k++) (void)hipFree(i[k]);
    memset(b, 0, sizeof(*b));
}

static int step_alloc(struct step_bufs *b, size_t N, int ncells, int ntypes, int nsteps) {
    memset(b, 0, sizeof(*b));
    const size_t C = (size_t)ncells;
    b->nblocks = (ncells + MD_SCAN_CELLS - 1) / MD_SCAN_CELLS;
    const size_t B = (size_t)b->nblocks;
    if (hipMalloc(&b->v, sizeof(double) * 3 * N) != hipSuccess ||
        hipMalloc(&b->x2, sizeof(double) * 3 * N) != hipSuccess ||
        hipMalloc(&b->v2, sizeof(double) * 3 * N) != hipSuccess ||
        hipMalloc(&b->mass, sizeof(double) * ntypes) != hipSuccess ||
        hipMalloc(&b->imass, sizeof(double) * ntypes) != hipSuccess ||
        hipMalloc(&b->cell_ke, sizeof(double) * C) != hipSuccess || hipMalloc(&b->bke, sizeof(double) * B) != hipSuccess ||
        hipMalloc(&b->esteps, sizeof(double) * 4 * (size_t)nsteps) != hipSuccess ||
        hipMalloc(&b->id2, sizeof(int) * N) != hipSuccess || hipMalloc(&b->type2, sizeof(int) * N) != hipSuccess ||
        hipMalloc(&b->slot, sizeof(int) * N) != hipSuccess || hipMalloc(&b->qsrc, sizeof(int) * N) != hipSuccess ||
        hipMalloc(&b->qdest, sizeof(int) * N) != hipSuccess || hipMalloc(&b->stay, sizeof(int) * C) != hipSuccess ||
        hipMalloc(&b->ndep, sizeof(int) * C) != hipSuccess || hipMalloc(&b->narr, sizeof(int) * C) != hipSuccess ||
        hipMalloc(&b->bsum, sizeof(int) * B) != hipSuccess || hipMalloc(&b->bmax, sizeof(int) * B) != hipSuccess ||
        hipMalloc(&b->st, sizeof(int) * 2) != hipSuccess || hipMalloc(&b->count2, sizeof(int) * C) != hipSuccess ||
        hipMalloc(&b->offset2, sizeof(int) * (C + 1)) != hipSuccess ||
        hipMalloc(&b->oldx, sizeof(double) * 3 * N) != hipSuccess ||
        hipMalloc(&b->lmax2, sizeof(double) * C) != hipSuccess || hipMalloc(&b->bmx, sizeof(double) * B) != hipSuccess ||
        hipMalloc(&b->vmax, sizeof(double)) != hipSuccess ||
        hipMemset(b->narr, 0, sizeof(int) * C) != hipSuccess) {
        step_free(b);
        return -1;
    }
    return 0;
}

/* The step buffers of g for N particle slots, C cells, T types and S steps:
   those of an earlier call when they are large enough (a call then spends
   no time in device allocation), else new ones.  The rigid-group buffers
   stay per call.  The arrival counts start at zero.  A call swaps layout
   buffers between these and g's own, so it hands its set back at the end
   (step_put). */
struct step_cache {
    struct step_bufs b;
    size_t N;
    int C, T, S;
};

static int step_get(struct gpu_state *g, struct step_bufs *b, size_t N, int C, int T, int S) {
    struct step_cache *sc = (struct step_cache *)g->steps;
    /* the particle and cell buffers trade places with g's, so they keep
       exactly g's sizes */
    if (sc == NULL || sc->N != N || sc->C != C || sc->T < T || sc->S < S) {
        if (sc == NULL) {
            if ((sc = (struct step_cache *)calloc(1, sizeof(*sc))) == NULL) return -1;
            g->steps = sc;
        } else {
            /* no pending work may still use the old buffers */
            if (hipStreamSynchronize(g->stream) != hipSuccess) return -1;
            step_free(&sc->b);
        }
        sc->N = 0;
        if (step_alloc(&sc->b, N, C, T, S) < 0) return -1;
        sc->N = N;
        sc->C = C;
        sc->T = T;
        sc->S = S;
    }
    *b = sc->b;
    if (hipMemsetAsync(b->narr, 0, sizeof(int) * (size_t)C, g->stream) != hipSuccess) return -1;
    return 0;
}

static void step_put(struct gpu_state *g, const struct step_bufs *b) {
    struct step_cache *sc = (struct step_cache *)g->steps;
    if (sc != NULL) sc->b = *b;
}

static void step_free_rigid(struct step_bufs *b) {
    void *r[] = {b->rg_np, b->rg_nc, b->rg_parts, b->rg_ij, b->rg_d2, b->rg_a, b->sid, b->fcorr};
    for (size_t k = 0; k < sizeof(r) / sizeof(r[0]); k++) (void)hipFree(r[k]);
    b->rg_np = b->rg_nc = b->rg_parts = b->rg_ij = b->sid = NULL;
    b->rg_d2 = b->fcorr = NULL;
    b->rg_a = NULL;
    b->nrig = 0;
}

void gpu_steps_free(struct gpu_state *g) {
    struct step_cache *sc = (struct step_cache *)g->steps;
    if (sc == NULL) return;
    step_free(&sc->b);
    free(sc);
    g->steps = NULL;
}

/* Shuffle (es == NULL: no kick) or advance (kick-drift, then migration, the
   step's energies into es[0..3]): departures, the new layout, stayers and
   arrivals copied into it, then the layouts swapped.  Asynchronous; the
   largest new cell count and the outside flag land in b->st. */
static int step_migrate(struct gpu_state *g, struct step_bufs *b, const struct step_geom &G, double dt,
                        double *es) {
    const int n = g->nparts, nc = g->ncells, kick = es != NULL;
    const int nbp = n > 0 ? (n + 255) / 256 : 1;
    hipLaunchKernelGGL(k_migrate, dim3(nc), dim3(64), 0, g->stream, G, kick, dt, b->mass, b->imass, g->d_fhi,
                       g->d_flo, g->d_count, g->d_offset, g->d_x, b->v, g->d_type, g->d_cell_of, b->slot, b->qsrc,
                       b->qdest, b->stay, b->ndep, b->narr, b->cell_ke);
    hipLaunchKernelGGL(k_scan_a, dim3(b->nblocks), dim3(256), 0, g->stream, nc, b->stay, b->narr,
                       kick ? b->cell_ke : (const double *)NULL, b->bsum, b->bmax, b->bke);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(64), 0, g->stream, b->nblocks, nc, b->bsum, b->bmax, b->bke,
                       b->offset2, b->st, g->d_tot, es);
    hipLaunchKernelGGL(k_scan_c, dim3(b->nblocks), dim3(256), 0, g->stream, nc, b->stay, b->narr, b->bsum,
                       b->count2, b->offset2, g->d_cnt_off);
    hipLaunchKernelGGL(k_stayers, dim3(nbp), dim3(256), 0, g->stream, n, g->d_cell_of, b->slot, b->offset2, g->d_x,
                       b->v, g->d_id, g->d_type, b->x2, b->v2, b->id2, b->type2);
    hipLaunchKernelGGL(k_arrive, dim3((nc + 127) / 128), dim3(128), 0, g->stream, G, g->d_offset, b->ndep, b->qsrc,
                       b->qdest, b->offset2, b->stay, g->d_x, b->v, g->d_id, g->d_type, b->x2, b->v2, b->id2,
                       b->type2, b->st);
    HIPCHECK(hipGetLastError());
    double *t;
    int *u;
    t = g->d_x, g->d_x = b->x2, b->x2 = t;
    t = b->v, b->v = b->v2, b->v2 = t;
    u = g->d_id, g->d_id = b->id2, b->id2 = u;
    u = g->d_type, g->d_type = b->type2, b->type2 = u;
    u = g->d_count, g->d_count = b->count2, b->count2 = u;
    u = g->d_offset, g->d_offset = b->offset2, b->offset2 = u;
    return engine_err_ok;
}

/* Wait for the last migration; its largest cell count and whether an
   arrival is still outside its cell. */
static int step_status(struct gpu_state *g, struct step_bufs *b, int *maxc, int *outside) {
    int st[2];
    if (hipMemcpyAsync(st, b->st, sizeof(st), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess)
        return gerror_(engine_err_gpu, "migration failed");
    *maxc = st[0];
    *outside = st[1];
    if (st[0] > g->maxcell) return gerror_(engine_err_maxparts, "more particles in a cell than the packed record holds");
    return engine_err_ok;
}

/* Back to the host cells: layout, positions, velocities, the last pass's
   forces by particle id, and the energies engine_step leaves behind. */
static int step_download(struct engine *e, struct gpu_state *g, struct step_bufs *b, int nsteps, double *epot,
                         double *ekin, int verlet, double smaxdx) {
    struct space *s = &e->s;
    const int n = g->nparts, nc = g->ncells;
    int rc = engine_err_ok;
    /* pinned staging (allocated with the particle buffers) for the bulk
       downloads; the cell counts get their own array, since h_count and
       h_offset describe the layout of the last upload */
    double *x = g->h_x, *v = g->h_v;
    double *es = (double *)malloc(sizeof(double) * 4 * (size_t)nsteps);
    int *id = g->h_id, *type = g->h_type, *ioi = g->h_ioi;
    int *count = (int *)malloc(sizeof(int) * (size_t)nc);
    double *ox = verlet ? (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1)) : NULL;
    double *cm = verlet ? (double *)malloc(sizeof(double) * (size_t)nc) : NULL;
    if (x == NULL || v == NULL || es == NULL || id == NULL || type == NULL || count == NULL || ioi == NULL ||
        (verlet && (ox == NULL || cm == NULL))) {
        rc = gerror_(engine_err_malloc, "malloc failed");
        goto done;
    }
    if (hipMemcpyAsync(x, g->d_x, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(v, b->v, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(id, g->d_id, sizeof(int) * n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(type, g->d_type, sizeof(int) * n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(count, g->d_count, sizeof(int) * nc, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(ioi, g->d_index_of_id, sizeof(int) * g->nids, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(es, b->esteps, sizeof(double) * 4 * nsteps, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(g->h_fhi, g->d_fhi, sizeof(float4) * n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(g->h_flo, g->d_flo, sizeof(float4) * n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(g->h_tot, g->d_tot, sizeof(double) * 3, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemcpyAsync(g->h_tot_cnt, g->d_tot_cnt, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost,
                       g->stream) != hipSuccess ||
        (verlet && (hipMemcpyAsync(ox, b->oldx, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                    hipMemcpyAsync(cm, g->d_cell_maxdx, sizeof(double) * nc, hipMemcpyDeviceToHost, g->stream) !=
                        hipSuccess)) ||
        hipStreamSynchronize(g->stream) != hipSuccess) {
        rc = gerror_(engine_err_gpu, "download failed");
        goto done;
    }
    if (space_setcells(s, count, x, v, id, type) < 0) {
        rc = gerror_(engine_err_space, "rebuilding the cells failed");
        goto done;
    }
    if (verlet) {
        /* reference positions per slot, each cell's maxdx, s->maxdx; the
           last sort lists belong to this layout (no migration since) */
        for (int cid = 0, a = 0; cid < nc; cid++) {
            struct cell *c = &s->cells[cid];
            if (c->oldx == NULL || c->oldx_size < c->count) {
                free(c->oldx);
                c->oldx_size = c->size + 20;
                if ((c->oldx = (FPTYPE *)malloc(sizeof(FPTYPE) * 3 * c->oldx_size)) == NULL) {
                    c->oldx_size = 0;
                    rc = gerror_(engine_err_malloc, "malloc failed");
                    goto done;
                }
            }
            memcpy(c->oldx, &ox[3 * (size_t)a], sizeof(double) * 3 * c->count);
            c->maxdx = cm[cid];
            g->h_sort_count[cid] = c->count;
            a += c->count;
        }
        s->maxdx = smaxdx;
        s->verlet_rebuild = 0;
    }
    space_prepare(s);
    if (space_add_forces(s, ioi, n, (const float *)g->h_fhi, (const float *)g->h_flo) < 0) {
        rc = gerror_(engine_err_range, "particle missing from the last pass");
        goto done;
    }
    if (b->nrig > 0) {
        /* P-SHAKE's coupling matrices and the constraint forces of the last
           step, added as rigid_eval_pshake adds them */
        const int shake = (e->flags & engine_flag_shake) != 0;
        float *ra = (float *)malloc(sizeof(float) * (size_t)b->nrig * MD_RIG_NC * MD_RIG_NC);
        double *fc = shake ? NULL : (double *)malloc(sizeof(double) * 3 * (size_t)g->nids);
        if (ra == NULL || (!shake && fc == NULL) ||
            hipMemcpy(ra, b->rg_a, sizeof(float) * (size_t)b->nrig * MD_RIG_NC * MD_RIG_NC, hipMemcpyDeviceToHost) !=
                hipSuccess ||
            (!shake && hipMemcpy(fc, b->fcorr, sizeof(double) * 3 * g->nids, hipMemcpyDeviceToHost) != hipSuccess)) {
            free(ra);
            free(fc);
            rc = gerror_(engine_err_gpu, "rigid download failed");
            goto done;
        }
        for (int r = 0; r < b->nrig; r++)
            memcpy(e->rigids[r].a, &ra[(size_t)r * MD_RIG_NC * MD_RIG_NC],
                   sizeof(float) * e->rigids[r].nr_constr * e->rigids[r].nr_constr);
        if (!shake)
            for (int cid = 0; cid < nc; cid++) {
                struct cell *c = &s->cells[cid];
                for (int k = 0; k < c->count; k++) {
                    const int id = c->parts[k].id;
                    if (id < e->part2rigid_size && e->part2rigid[id] >= 0)
                        for (int j = 0; j < 3; j++) c->parts[k].f[j] += fc[3 * id + j];
                }
            }
        free(ra);
        free(fc);
    }
    s->epot_nonbond += g->h_tot[0];
    s->epot_exclusion += g->h_tot[1];
    s->epot += g->h_tot[0] + g->h_tot[1];
    s->virial += g->h_tot[2];
    s->nr_pairs += (long long)g->h_tot_cnt[0];
    s->nr_excl_evaluated += (long long)g->h_tot_cnt[1];
    for (int k = 0; k < nsteps; k++) {
        if (epot) epot[k] = es[4 * k] + es[4 * k + 1];
        if (ekin) ekin[k] = es[4 * k + 3];
    }
done:
    free(ox);
    free(cm);
    free(es);
    free(count);
    return rc;
}

/* Rigid groups to the device (engine_gpu_steps handles groups of up to
   MD_RIG_NP particles and MD_RIG_NC constraints, e.g. rigid water). */
static int rigid_upload(struct engine *e, struct gpu_state *g, struct step_bufs *b) {
    const int n = e->nr_rigids;
    int rc = engine_err_ok;
    b->nrig = n;
    if (n == 0) return engine_err_ok;
    for (int r = 0; r < n; r++) {
        if (e->rigids[r].nr_parts > MD_RIG_NP || e->rigids[r].nr_constr > MD_RIG_NC)
            return gerror_(engine_err_range, "rigid groups of more than 4 particles or 6 constraints run through engine_step");
        for (int k = 0; k < e->rigids[r].nr_parts; k++)
            if (e->rigids[r].parts[k] < 0 || e->rigids[r].parts[k] >= g->nids)
                return gerror_(engine_err_range, "rigid constraint on a particle id unknown to the device");
    }
    int *np = (int *)calloc(n, sizeof(int)), *nc = (int *)calloc(n, sizeof(int));
    int *parts = (int *)calloc((size_t)n * MD_RIG_NP, sizeof(int));
    int *ij = (int *)calloc((size_t)n * MD_RIG_NC * 2, sizeof(int));
    double *d2 = (double *)calloc((size_t)n * MD_RIG_NC, sizeof(double));
    float *a = (float *)calloc((size_t)n * MD_RIG_NC * MD_RIG_NC, sizeof(float));
    if (np == NULL || nc == NULL || parts == NULL || ij == NULL || d2 == NULL || a == NULL) {
        rc = gerror_(engine_err_malloc, "malloc failed");
        goto done;
    }
    for (int r = 0; r < n; r++) {
        const struct rigid *R = &e->rigids[r];
        np[r] = R->nr_parts;
        nc[r] = R->nr_constr;
        for (int k = 0; k < R->nr_parts; k++) parts[r * MD_RIG_NP + k] = R->parts[k];
        for (int k = 0; k < R->nr_constr; k++) {
            ij[(r * MD_RIG_NC + k) * 2] = R->constr[k].i;
            ij[(r * MD_RIG_NC + k) * 2 + 1] = R->constr[k].j;
            d2[r * MD_RIG_NC + k] = R->constr[k].d2;
        }
        memcpy(&a[(size_t)r * MD_RIG_NC * MD_RIG_NC], R->a, sizeof(float) * R->nr_constr * R->nr_constr);
    }
    if (hipMalloc(&b->rg_np, sizeof(int) * n) != hipSuccess || hipMalloc(&b->rg_nc, sizeof(int) * n) != hipSuccess ||
        hipMalloc(&b->rg_parts, sizeof(int) * n * MD_RIG_NP) != hipSuccess ||
        hipMalloc(&b->rg_ij, sizeof(int) * n * MD_RIG_NC * 2) != hipSuccess ||
        hipMalloc(&b->rg_d2, sizeof(double) * n * MD_RIG_NC) != hipSuccess ||
        hipMalloc(&b->rg_a, sizeof(float) * n * MD_RIG_NC * MD_RIG_NC) != hipSuccess ||
        hipMalloc(&b->sid, sizeof(int) * g->nids) != hipSuccess ||
        hipMalloc(&b->fcorr, sizeof(double) * 3 * g->nids) != hipSuccess ||
        hipMemcpy(b->rg_np, np, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->rg_nc, nc, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->rg_parts, parts, sizeof(int) * n * MD_RIG_NP, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->rg_ij, ij, sizeof(int) * n * MD_RIG_NC * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->rg_d2, d2, sizeof(double) * n * MD_RIG_NC, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->rg_a, a, sizeof(float) * n * MD_RIG_NC * MD_RIG_NC, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->fcorr, 0, sizeof(double) * 3 * g->nids) != hipSuccess ||
        hipMemset(b->sid, 0xff, sizeof(int) * g->nids) != hipSuccess)
        rc = gerror_(engine_err_gpu, "rigid upload failed");
done:
    free(np);
    free(nc);
    free(parts);
    free(ij);
    free(d2);
    free(a);
    return rc;
}

/* After an advance: the constraints (engine.c:2196-2232), then the kinetic
   energy of the corrected velocities and, with oldx, the displacements, and
   the step's energies into es. */
static int rigid_step(struct engine *e, struct gpu_state *g, struct step_bufs *b, const struct step_geom &G,
                      const double *oldx, double *es) {
    const int n = g->nparts, nc = G.ncells;
    const int nbl = ((n > nc ? n : nc) + 255) / 256;
    struct rigid_dev R;
    R.n = b->nrig;
    R.np = b->rg_np;
    R.nc = b->rg_nc;
    R.parts = b->rg_parts;
    R.ij = b->rg_ij;
    R.d2 = b->rg_d2;
    R.a = b->rg_a;
    hipLaunchKernelGGL(k_layout, dim3(nbl > 0 ? nbl : 1), dim3(256), 0, g->stream, n, nc, g->d_id, g->d_offset, b->sid,
                       g->d_cell_of);
    hipLaunchKernelGGL(k_rigid, dim3((R.n + 127) / 128), dim3(128), 0, g->stream, R,
                       (e->flags & engine_flag_shake) != 0, (e->time + 1) < engine_pshake_steps, e->dt, e->tol_rigid,
                       G, b->mass, b->sid, g->d_cell_of, g->d_type, g->d_x, b->v, b->fcorr, b->st);
    hipLaunchKernelGGL(k_advance_verlet, dim3(nc), dim3(64), 0, g->stream, 0, 0.0, b->mass, b->imass, g->d_fhi,
                       g->d_flo, g->d_count, g->d_offset, g->d_x, b->v, g->d_type, oldx, b->lmax2, b->cell_ke);
    hipLaunchKernelGGL(k_verlet_red_a, dim3(b->nblocks), dim3(256), 0, g->stream, nc, b->cell_ke, b->lmax2, b->bke,
                       b->bmx);
    hipLaunchKernelGGL(k_verlet_red_b, dim3(1), dim3(64), 0, g->stream, b->nblocks, b->bke, b->bmx, g->d_tot, es,
                       b->vmax);
    if (hipGetLastError() != hipSuccess) return gerror_(engine_err_gpu, "constraints failed");
    return engine_err_ok;
}

/* Largest squared displacement of the last k_advance_verlet. */
static int verlet_status(struct gpu_state *g, struct step_bufs *b, double *vmax) {
    if (hipMemcpyAsync(vmax, b->vmax, sizeof(double), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess)
        return gerror_(engine_err_gpu, "advance failed");
    return engine_err_ok;
}

/* The steps with pseudo-Verlet sort reuse: engine_verlet_update
   (engine.c:193-315) at the start of each step, the pass sorting only on a
   rebuild, the kick-only advance.  Positions, velocities and forces are
   resident as in the plain loop; the rebuild positions (oldx, per slot) and
   each cell's maxdx are too.  *smaxdx carries s->maxdx. */
static int steps_verlet(struct engine *e, struct gpu_state *g, struct step_bufs *b, const struct step_geom &G,
                        int nsteps, double *smaxdx) {
    struct space *s = &e->s;
    const double skin = fmin(s->h[0], fmin(s->h[1], s->h[2])) - s->cutoff;
    const int nc = G.ncells, nbc = (nc + 255) / 256, mlp = (e->flags & engine_flag_mlp) != 0;
    int rebuild = s->verlet_rebuild, rc = engine_err_ok, maxc = g->maxcount, outside = 0;
    double vmax = 0.0;
    g->P.maxdx = g->d_cell_maxdx;
    {
        /* each cell's maxdx, and the rebuild positions unless this step rebuilds */
        double *cm = (double *)malloc(sizeof(double) * (size_t)nc);
        double *ox = rebuild ? NULL : (double *)malloc(sizeof(double) * 3 * (size_t)(g->nparts > 0 ? g->nparts : 1));
        if (cm == NULL || (!rebuild && ox == NULL)) rc = gerror_(engine_err_malloc, "malloc failed");
        for (int cid = 0; cid < nc && rc == engine_err_ok; cid++) {
            const struct cell *c = &s->cells[cid];
            cm[cid] = c->maxdx;
            if (rebuild) continue;
            if (c->count > 0 && (c->oldx == NULL || c->oldx_size < c->count))
                rc = gerror_(engine_err_range, "no rebuild positions (engine_start sets them)");
            else if (c->count > 0)
                memcpy(&ox[3 * (size_t)g->h_offset[cid]], c->oldx, sizeof(double) * 3 * c->count);
        }
        if (rc == engine_err_ok &&
            (hipMemcpy(g->d_cell_maxdx, cm, sizeof(double) * nc, hipMemcpyHostToDevice) != hipSuccess ||
             (!rebuild && hipMemcpy(b->oldx, ox, sizeof(double) * 3 * g->nparts, hipMemcpyHostToDevice) != hipSuccess)))
            rc = gerror_(engine_err_gpu, "upload failed");
        free(cm);
        free(ox);
        if (rc < 0) return rc;
    }
    if (!rebuild) {
        /* displacements of the uploaded positions */
        hipLaunchKernelGGL(k_advance_verlet, dim3(nc), dim3(64), 0, g->stream, 0, 0.0, b->mass, b->imass, g->d_fhi,
                           g->d_flo, g->d_count, g->d_offset, g->d_x, b->v, g->d_type, b->oldx, b->lmax2, b->cell_ke);
        hipLaunchKernelGGL(k_verlet_red_a, dim3(b->nblocks), dim3(256), 0, g->stream, nc, b->cell_ke, b->lmax2, b->bke,
                           b->bmx);
        hipLaunchKernelGGL(k_verlet_red_b, dim3(1), dim3(64), 0, g->stream, b->nblocks, b->bke, b->bmx, g->d_tot,
                           (double *)NULL, b->vmax);
        if ((rc = verlet_status(g, b, &vmax)) < 0) return rc;
    }
    for (int step = 0; step < nsteps; step++) {
        if (!rebuild) {
            /* engine_verlet_update without a pending rebuild */
            hipLaunchKernelGGL(k_cell_maxdx, dim3(nbc), dim3(256), 0, g->stream, nc, b->lmax2, g->d_cell_maxdx);
            const double maxdx = sqrt(vmax);
            rebuild = (2.0 * maxdx > skin) || (maxdx > s->cutoff * engine_maxdxratio);
            if (!rebuild) *smaxdx = maxdx;
        }
        if (rebuild) {
            /* engine_shuffle, then the rebuild positions */
            if ((rc = step_migrate(g, b, G, 0.0, NULL)) < 0 || (rc = step_status(g, b, &maxc, &outside)) < 0) break;
            if (hipMemcpyAsync(b->oldx, g->d_x, sizeof(double) * 3 * g->nparts, hipMemcpyDeviceToDevice, g->stream) !=
                hipSuccess) {
                rc = gerror_(engine_err_gpu, "rebuild failed");
                break;
            }
            *smaxdx = 0.0;
            g->maxcount = maxc;
            g->P.maxn = maxc > 0 ? maxc : 1;
        }
        if ((rc = gpu_launch(g, 0, mlp, rebuild)) < 0) break;
        rebuild = 0;
        hipLaunchKernelGGL(k_advance_verlet, dim3(nc), dim3(64), 0, g->stream, 1, e->dt, b->mass, b->imass, g->d_fhi,
                           g->d_flo, g->d_count, g->d_offset, g->d_x, b->v, g->d_type, b->oldx, b->lmax2, b->cell_ke);
        hipLaunchKernelGGL(k_verlet_red_a, dim3(b->nblocks), dim3(256), 0, g->stream, nc, b->cell_ke, b->lmax2, b->bke,
                           b->bmx);
        hipLaunchKernelGGL(k_verlet_red_b, dim3(1), dim3(64), 0, g->stream, b->nblocks, b->bke, b->bmx, g->d_tot,
                           b->esteps + 4 * step, b->vmax);
        if (hipGetLastError() != hipSuccess) {
            rc = gerror_(engine_err_gpu, "advance failed");
            break;
        }
        if (b->nrig > 0 && (rc = rigid_step(e, g, b, G, b->oldx, b->esteps + 4 * step)) < 0) break;
        if ((rc = verlet_status(g, b, &vmax)) < 0) break;
        e->time += 1;
    }
    return rc;
}

extern "C" int engine_gpu_steps(struct engine *e, int nsteps, double *epot, double *ekin) {
    struct gpu_state *g;
    struct space *s;
    struct step_bufs b;
    struct step_geom G;
    int rc = engine_err_ok, maxc = 0, outside = 0, flag = 0;
    if (e == NULL) return gerror_(engine_err_null, "NULL engine");
    g = (struct gpu_state *)e->gpu;
    if (g == NULL) return gerror_(engine_err_nogpu, "engine_gpu_load was not called");
    if (nsteps < 0) return gerror_(engine_err_range, "negative step count");
    if (nsteps == 0) return engine_err_ok;
    s = &e->s;
    HIPCHECK(hipSetDevice(g->device));
    if (gpu_upload(e) < 0) return engine_err_gpu;
    for (int k = 0; k < 3; k++) {
        G.h[k] = s->h[k];
        G.cdim[k] = s->cdim[k];
    }
    G.ncells = s->nr_cells;
    g->P.maxdx = NULL;
    if (step_get(g, &b, (size_t)g->nparts_cap, g->ncells, g->max_type, nsteps) < 0)
        return gerror_(engine_err_gpu, "step buffers");
    {
        double *mass = (double *)malloc(sizeof(double) * g->max_type);
        double *imass = (double *)malloc(sizeof(double) * g->max_type);
        double *v = g->h_v; /* pinned */
        if (mass == NULL || imass == NULL) {
            rc = gerror_(engine_err_malloc, "malloc failed");
        } else {
            for (int k = 0; k < g->max_type; k++) {
                mass[k] = e->types[k].mass;
                imass[k] = e->types[k].imass;
            }
            space_pack_v(s, g->h_offset, v);
            if (hipMemcpy(b.mass, mass, sizeof(double) * g->max_type, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b.imass, imass, sizeof(double) * g->max_type, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpyAsync(b.v, v, sizeof(double) * 3 * g->nparts, hipMemcpyHostToDevice, g->stream) != hipSuccess ||
                hipMemsetAsync(g->d_queue + MD_NQUEUE, 0, sizeof(int), g->stream) != hipSuccess)
                rc = gerror_(engine_err_gpu, "upload failed");
        }
        free(mass);
        free(imass);
    }
    const int verlet = (e->flags & engine_flag_verlet) != 0;
    double smaxdx = s->maxdx;
    if (rc == engine_err_ok) rc = rigid_upload(e, g, &b);
    if (verlet && rc == engine_err_ok)
        rc = steps_verlet(e, g, &b, G, nsteps, &smaxdx);
    /* engine_step: shuffle, pass, advance.  After an advance that left every
       particle inside its cell, the next shuffle moves nothing and is skipped. */
    outside = 1;
    for (int step = 0; step < nsteps && rc == engine_err_ok && !verlet; step++) {
        if (outside) {
            if ((rc = step_migrate(g, &b, G, 0.0, NULL)) < 0 || (rc = step_status(g, &b, &maxc, &outside)) < 0)
                break;
        }
        g->maxcount = maxc;
        g->P.maxn = maxc > 0 ? maxc : 1;
        if ((rc = gpu_launch(g, 0, (e->flags & engine_flag_mlp) != 0, 1)) < 0) break;
        if ((rc = step_migrate(g, &b, G, e->dt, b.esteps + 4 * step)) < 0 ||
            (b.nrig > 0 && (rc = rigid_step(e, g, &b, G, NULL, b.esteps + 4 * step)) < 0) ||
            (rc = step_status(g, &b, &maxc, &outside)) < 0)
            break;
        e->time += 1;
    }
    if (rc == engine_err_ok &&
        (hipMemcpyAsync(&flag, g->d_queue + MD_NQUEUE, sizeof(int), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
         hipStreamSynchronize(g->stream) != hipSuccess))
        rc = gerror_(engine_err_gpu, "step failed");
    if (rc == engine_err_ok && flag) rc = gerror_(engine_err_range, "sort stack overflow");
    if (rc == engine_err_ok) rc = step_download(e, g, &b, nsteps, epot, ekin, verlet, smaxdx);
    (void)hipStreamSynchronize(g->stream);
    step_free_rigid(&b);
    step_put(g, &b);
    /* without pseudo-Verlet the last sort lists belong to the layout before
       the last migration */
    if (!verlet || rc != engine_err_ok) g->sort_valid = 0;
    return rc;
}
