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
 *     its departures and the cell's last particle takes its slot
 *     (k_migrate, one thread per cell, the loop of engine_advance);
 *   - each destination appends its arrivals ordered by source cell and
 *     departure order (k_arrive, one thread per cell), after its stayers;
 *   - new counts and offsets come from one fixed-order scan (k_scan).
 * No reference counterpart as a device path: the reference's CUDA engine
 * returns to the host every step (runner_cuda.cu:2405-2430, 2514-2521).
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gpu_state.h"

struct step_geom {
    double h[3];
    int cdim[3];
    int ncells;
};

/* v += f dt / m, x += dt v, in the host's FP64 order (engine.c, engine_advance). */
__global__ __launch_bounds__(256) void k_kick(int n, double dt, const double *__restrict__ imass,
                                              const int *__restrict__ type, const float4 *__restrict__ fhi,
                                              const float4 *__restrict__ flo, double *__restrict__ x,
                                              double *__restrict__ v) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const double w = dt * imass[type[a]];
    const float4 hi = fhi[a], lo = flo[a];
    /* the host adds the pass's force into a zeroed FP64 force */
    const double f[3] = {0.0 + ((double)hi.x + (double)lo.x), 0.0 + ((double)hi.y + (double)lo.y),
                         0.0 + ((double)hi.z + (double)lo.z)};
    for (int k = 0; k < 3; k++) {
        v[3 * a + k] += f[k] * w;
        x[3 * a + k] += dt * v[3 * a + k];
    }
}

/* Per source cell: departures in the host's order (space_shuffle_local,
   engine_advance).  Writes the stayers compacted in place, the departures
   into the cell's segment of the departure arrays, and slot -> cell of the
   old layout. */
__global__ __launch_bounds__(256) void k_migrate(struct step_geom G, const int *__restrict__ count,
                                                 const int *__restrict__ offset, double *__restrict__ x,
                                                 double *__restrict__ v, int *__restrict__ id,
                                                 int *__restrict__ type, int *__restrict__ stay,
                                                 int *__restrict__ ndep, double *__restrict__ qx,
                                                 double *__restrict__ qv, int *__restrict__ qid,
                                                 int *__restrict__ qtype, int *__restrict__ qdest,
                                                 int *__restrict__ cell_of) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G.ncells) return;
    const int off = offset[c], n0 = count[c];
    const int loc[3] = {c / (G.cdim[1] * G.cdim[2]), (c / G.cdim[2]) % G.cdim[1], c % G.cdim[2]};
    for (int k = 0; k < n0; k++) cell_of[off + k] = c;
    int cnt = n0, nd = 0, pid = 0;
    while (pid < cnt) {
        const int a = off + pid;
        int d[3];
        for (int k = 0; k < 3; k++) d[k] = (x[3 * a + k] >= G.h[k]) - (x[3 * a + k] < 0.0);
        if (d[0] || d[1] || d[2]) {
            const int b = off + nd;
            for (int k = 0; k < 3; k++) {
                qx[3 * b + k] = x[3 * a + k] - d[k] * G.h[k];
                qv[3 * b + k] = v[3 * a + k];
            }
            qid[b] = id[a];
            qtype[b] = type[a];
            qdest[b] = (((loc[0] + d[0] + G.cdim[0]) % G.cdim[0]) * G.cdim[1] +
                        (loc[1] + d[1] + G.cdim[1]) % G.cdim[1]) * G.cdim[2] +
                       (loc[2] + d[2] + G.cdim[2]) % G.cdim[2];
            nd++;
            cnt--;
            if (pid < cnt) {
                const int e = off + cnt;
                for (int k = 0; k < 3; k++) {
                    x[3 * a + k] = x[3 * e + k];
                    v[3 * a + k] = v[3 * e + k];
                }
                id[a] = id[e];
                type[a] = type[e];
            }
        } else {
            pid++;
        }
    }
    stay[c] = cnt;
    ndep[c] = nd;
}

/* The up to 27 cells a particle can arrive from (itself included when a
   dimension has fewer than 3 cells), ascending and without repeats. */
__device__ __forceinline__ int source_cells(const struct step_geom &G, int c, int *src) {
    const int loc[3] = {c / (G.cdim[1] * G.cdim[2]), (c / G.cdim[2]) % G.cdim[1], c % G.cdim[2]};
    int n = 0;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                const int s = (((loc[0] + i + G.cdim[0]) % G.cdim[0]) * G.cdim[1] +
                               (loc[1] + j + G.cdim[1]) % G.cdim[1]) * G.cdim[2] +
                              (loc[2] + k + G.cdim[2]) % G.cdim[2];
                int p = n;
                while (p > 0 && src[p - 1] > s) {
                    src[p] = src[p - 1];
                    p--;
                }
                if (p > 0 && src[p - 1] == s) {
                    for (int q = p; q < n; q++) src[q] = src[q + 1];
                    continue;
                }
                src[p] = s;
                n++;
            }
    return n;
}

/* Per destination cell: arrivals by source cell then departure order.
   Without out, only counts them. */
__global__ __launch_bounds__(128) void k_arrive(struct step_geom G, const int *__restrict__ offset,
                                                const int *__restrict__ ndep, const int *__restrict__ qdest,
                                                const double *__restrict__ qx, const double *__restrict__ qv,
                                                const int *__restrict__ qid, const int *__restrict__ qtype,
                                                int *__restrict__ narr, const int *__restrict__ new_off,
                                                const int *__restrict__ stay, double *__restrict__ x2,
                                                double *__restrict__ v2, int *__restrict__ id2,
                                                int *__restrict__ type2, int write) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G.ncells) return;
    int src[27];
    const int ns = source_cells(G, c, src);
    int j = 0;
    for (int q = 0; q < ns; q++) {
        const int s = src[q], off = offset[s], nd = ndep[s];
        for (int k = 0; k < nd; k++) {
            const int b = off + k;
            if (qdest[b] != c) continue;
            if (write) {
                const int a = new_off[c] + stay[c] + j;
                for (int m = 0; m < 3; m++) {
                    x2[3 * a + m] = qx[3 * b + m];
                    v2[3 * a + m] = qv[3 * b + m];
                }
                id2[a] = qid[b];
                type2[a] = qtype[b];
            }
            j++;
        }
    }
    if (!write) narr[c] = j;
}

/* New counts and offsets: one block, fixed-order exclusive scan; the
   largest count goes to *maxc. */
__global__ __launch_bounds__(1024) void k_scan(int ncells, const int *__restrict__ stay,
                                               const int *__restrict__ narr, int *__restrict__ count,
                                               int *__restrict__ offset, int2 *__restrict__ cnt_off,
                                               int *__restrict__ maxc) {
    __shared__ int part[1024];
    __shared__ int pmax[1024];
    const int t = threadIdx.x;
    const int per = (ncells + 1023) / 1024;
    const int c0 = t * per, c1 = c0 + per < ncells ? c0 + per : ncells;
    int sum = 0, mx = 0;
    for (int c = c0; c < c1; c++) {
        const int n = stay[c] + narr[c];
        sum += n;
        mx = n > mx ? n : mx;
    }
    part[t] = sum;
    pmax[t] = mx;
    __syncthreads();
    if (t == 0) {
        int run = 0, m = 0;
        for (int k = 0; k < 1024; k++) {
            const int v = part[k];
            part[k] = run;
            run += v;
            m = pmax[k] > m ? pmax[k] : m;
        }
        offset[ncells] = run;
        *maxc = m;
    }
    __syncthreads();
    int run = part[t];
    for (int c = c0; c < c1; c++) {
        const int n = stay[c] + narr[c];
        count[c] = n;
        offset[c] = run;
        cnt_off[c] = make_int2(n, run);
        run += n;
    }
}

/* Stayers: old slot a of cell c, if still in the cell, to its new slot. */
__global__ __launch_bounds__(256) void k_stayers(int n, const int *__restrict__ cell_of,
                                                 const int *__restrict__ old_off, const int *__restrict__ stay,
                                                 const int *__restrict__ new_off, const double *__restrict__ x,
                                                 const double *__restrict__ v, const int *__restrict__ id,
                                                 const int *__restrict__ type, double *__restrict__ x2,
                                                 double *__restrict__ v2, int *__restrict__ id2,
                                                 int *__restrict__ type2) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int c = cell_of[a], k = a - old_off[c];
    if (k >= stay[c]) return;
    const int b = new_off[c] + k;
    for (int m = 0; m < 3; m++) {
        x2[3 * b + m] = x[3 * a + m];
        v2[3 * b + m] = v[3 * a + m];
    }
    id2[b] = id[a];
    type2[b] = type[a];
}

/* Kinetic energy 1/2 m v^2 per cell in storage order, then in cell order. */
__global__ __launch_bounds__(256) void k_kin_cells(int ncells, const int *__restrict__ count,
                                                   const int *__restrict__ offset, const double *__restrict__ mass,
                                                   const int *__restrict__ type, const double *__restrict__ v,
                                                   double *__restrict__ cell_ke) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    double s = 0.0;
    for (int k = offset[c]; k < offset[c] + count[c]; k++) {
        const double m = mass[type[k]];
        s += 0.5 * m * (v[3 * k] * v[3 * k] + v[3 * k + 1] * v[3 * k + 1] + v[3 * k + 2] * v[3 * k + 2]);
    }
    cell_ke[c] = s;
}

__global__ __launch_bounds__(1024) void k_sum1(int n, const double *__restrict__ a, double *__restrict__ out) {
    __shared__ double sd[1024];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int k = t; k < n; k += 1024) s += a[k];
    sd[t] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) sd[t] += sd[t + w];
        __syncthreads();
    }
    if (t == 0) *out = sd[0];
}

/* ------------------------------------------------------------ host side */

struct step_bufs {
    double *v, *x2, *v2, *qx, *qv, *mass, *imass, *cell_ke, *esteps;
    int *id2, *type2, *qid, *qtype, *qdest, *stay, *ndep, *narr, *maxc, *count2, *offset2;
};

static void step_free(struct step_bufs *b) {
    double *d[] = {b->v, b->x2, b->v2, b->qx, b->qv, b->mass, b->imass, b->cell_ke, b->esteps};
    int *i[] = {b->id2, b->type2, b->qid, b->qtype, b->qdest, b->stay, b->ndep, b->narr, b->maxc, b->count2, b->offset2};
    for (size_t k = 0; k < sizeof(d) / sizeof(d[0]); k++) (void)hipFree(d[k]);
    for (size_t k = 0; k < sizeof(i) / sizeof(i[0]); k++) (void)hipFree(i[k]);
    memset(b, 0, sizeof(*b));
}

static int step_alloc(struct step_bufs *b, size_t N, int ncells, int ntypes, int nsteps) {
    memset(b, 0, sizeof(*b));
    const size_t C = (size_t)ncells;
    if (hipMalloc(&b->v, sizeof(double) * 3 * N) != hipSuccess || hipMalloc(&b->x2, sizeof(double) * 3 * N) != hipSuccess ||
        hipMalloc(&b->v2, sizeof(double) * 3 * N) != hipSuccess || hipMalloc(&b->qx, sizeof(double) * 3 * N) != hipSuccess ||
        hipMalloc(&b->qv, sizeof(double) * 3 * N) != hipSuccess || hipMalloc(&b->mass, sizeof(double) * ntypes) != hipSuccess ||
        hipMalloc(&b->imass, sizeof(double) * ntypes) != hipSuccess || hipMalloc(&b->cell_ke, sizeof(double) * C) != hipSuccess ||
        hipMalloc(&b->esteps, sizeof(double) * 4 * (size_t)nsteps) != hipSuccess ||
        hipMalloc(&b->id2, sizeof(int) * N) != hipSuccess || hipMalloc(&b->type2, sizeof(int) * N) != hipSuccess ||
        hipMalloc(&b->qid, sizeof(int) * N) != hipSuccess || hipMalloc(&b->qtype, sizeof(int) * N) != hipSuccess ||
        hipMalloc(&b->qdest, sizeof(int) * N) != hipSuccess || hipMalloc(&b->stay, sizeof(int) * C) != hipSuccess ||
        hipMalloc(&b->ndep, sizeof(int) * C) != hipSuccess || hipMalloc(&b->narr, sizeof(int) * C) != hipSuccess ||
        hipMalloc(&b->maxc, sizeof(int)) != hipSuccess || hipMalloc(&b->count2, sizeof(int) * C) != hipSuccess ||
        hipMalloc(&b->offset2, sizeof(int) * (C + 1)) != hipSuccess) {
        step_free(b);
        return -1;
    }
    return 0;
}

/* Shuffle (kick == 0) or advance: kick-drift, migrate, rebuild the cell
   layout and swap it in.  Asynchronous; the largest new cell count lands in
   b->maxc. */
static int step_migrate(struct gpu_state *g, struct step_bufs *b, const struct step_geom &G, int kick, double dt) {
    const int n = g->nparts, nc = g->ncells;
    const int nbp = n > 0 ? (n + 255) / 256 : 1, nbc = (nc + 255) / 256, nba = (nc + 127) / 128;
    if (kick)
        hipLaunchKernelGGL(k_kick, dim3(nbp), dim3(256), 0, g->stream, n, dt, b->imass, g->d_type, g->d_fhi,
                           g->d_flo, g->d_x, b->v);
    hipLaunchKernelGGL(k_migrate, dim3(nbc), dim3(256), 0, g->stream, G, g->d_count, g->d_offset, g->d_x, b->v,
                       g->d_id, g->d_type, b->stay, b->ndep, b->qx, b->qv, b->qid, b->qtype, b->qdest, g->d_cell_of);
    hipLaunchKernelGGL(k_arrive, dim3(nba), dim3(128), 0, g->stream, G, g->d_offset, b->ndep, b->qdest, b->qx, b->qv,
                       b->qid, b->qtype, b->narr, (const int *)NULL, (const int *)NULL, (double *)NULL, (double *)NULL,
                       (int *)NULL, (int *)NULL, 0);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, g->stream, nc, b->stay, b->narr, b->count2, b->offset2,
                       g->d_cnt_off, b->maxc);
    hipLaunchKernelGGL(k_stayers, dim3(nbp), dim3(256), 0, g->stream, n, g->d_cell_of, g->d_offset, b->stay,
                       b->offset2, g->d_x, b->v, g->d_id, g->d_type, b->x2, b->v2, b->id2, b->type2);
    hipLaunchKernelGGL(k_arrive, dim3(nba), dim3(128), 0, g->stream, G, g->d_offset, b->ndep, b->qdest, b->qx, b->qv,
                       b->qid, b->qtype, b->narr, b->offset2, b->stay, b->x2, b->v2, b->id2, b->type2, 1);
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

/* Back to the host cells: layout, positions, velocities, the last pass's
   forces by particle id, and the energies engine_step leaves behind. */
static int step_download(struct engine *e, struct gpu_state *g, struct step_bufs *b, int nsteps, double *epot,
                         double *ekin) {
    struct space *s = &e->s;
    const int n = g->nparts, nc = g->ncells;
    const size_t M = (size_t)(n > 0 ? n : 1);
    int rc = engine_err_ok;
    double *x = (double *)malloc(sizeof(double) * 3 * M), *v = (double *)malloc(sizeof(double) * 3 * M);
    double *es = (double *)malloc(sizeof(double) * 4 * (size_t)nsteps);
    int *id = (int *)malloc(sizeof(int) * M), *type = (int *)malloc(sizeof(int) * M);
    int *count = (int *)malloc(sizeof(int) * (size_t)nc), *ioi = (int *)malloc(sizeof(int) * (size_t)g->nids);
    if (x == NULL || v == NULL || es == NULL || id == NULL || type == NULL || count == NULL || ioi == NULL) {
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
        hipStreamSynchronize(g->stream) != hipSuccess) {
        rc = gerror_(engine_err_gpu, "download failed");
        goto done;
    }
    if (space_setcells(s, count, x, v, id, type) < 0) {
        rc = gerror_(engine_err_space, "rebuilding the cells failed");
        goto done;
    }
    space_prepare(s);
    for (int cid = 0; cid < nc; cid++) {
        struct cell *c = &s->cells[cid];
        for (int k = 0; k < c->count; k++) {
            const int a = ioi[c->parts[k].id];
            if (a < 0 || a >= n) {
                rc = gerror_(engine_err_range, "particle missing from the last pass");
                goto done;
            }
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
    for (int k = 0; k < nsteps; k++) {
        if (epot) epot[k] = es[4 * k] + es[4 * k + 1];
        if (ekin) ekin[k] = es[4 * k + 3];
    }
done:
    free(x);
    free(v);
    free(es);
    free(id);
    free(type);
    free(count);
    free(ioi);
    return rc;
}

extern "C" int engine_gpu_steps(struct engine *e, int nsteps, double *epot, double *ekin) {
    struct gpu_state *g;
    struct space *s;
    struct step_bufs b;
    struct step_geom G;
    int rc = engine_err_ok, maxc = 0, flag = 0;
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
    if (step_alloc(&b, (size_t)g->nparts_cap, g->ncells, g->max_type, nsteps) < 0)
        return gerror_(engine_err_gpu, "step buffers");
    {
        double *mass = (double *)malloc(sizeof(double) * g->max_type);
        double *imass = (double *)malloc(sizeof(double) * g->max_type);
        double *v = (double *)malloc(sizeof(double) * 3 * (size_t)(g->nparts > 0 ? g->nparts : 1));
        if (mass == NULL || imass == NULL || v == NULL) {
            rc = gerror_(engine_err_malloc, "malloc failed");
        } else {
            for (int k = 0; k < g->max_type; k++) {
                mass[k] = e->types[k].mass;
                imass[k] = e->types[k].imass;
            }
            for (int cid = 0; cid < s->nr_cells; cid++)
                for (int k = 0; k < s->cells[cid].count; k++)
                    for (int m = 0; m < 3; m++) v[3 * (g->h_offset[cid] + k) + m] = s->cells[cid].parts[k].v[m];
            if (hipMemcpy(b.mass, mass, sizeof(double) * g->max_type, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b.imass, imass, sizeof(double) * g->max_type, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b.v, v, sizeof(double) * 3 * g->nparts, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemsetAsync(g->d_queue + MD_NQUEUE, 0, sizeof(int), g->stream) != hipSuccess)
                rc = gerror_(engine_err_gpu, "upload failed");
        }
        free(mass);
        free(imass);
        free(v);
    }
    for (int step = 0; step < nsteps && rc == engine_err_ok; step++) {
        /* engine_step: shuffle, then the pass on the new layout */
        if (step_migrate(g, &b, G, 0, 0.0) < 0 ||
            hipMemcpyAsync(&maxc, b.maxc, sizeof(int), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
            hipStreamSynchronize(g->stream) != hipSuccess) {
            rc = gerror_(engine_err_gpu, "shuffle failed");
            break;
        }
        if (maxc > MD_MAX_CELL) {
            rc = gerror_(engine_err_maxparts, "more than 1024 particles in a cell");
            break;
        }
        g->maxcount = maxc;
        g->P.maxn = maxc > 0 ? maxc : 1;
        if (gpu_launch(g, 0, (e->flags & engine_flag_mlp) != 0) < 0) {
            rc = engine_err_gpu;
            break;
        }
        /* the pass's totals, then kick-drift and migration (engine_advance) */
        if (hipMemcpyAsync(b.esteps + 4 * step, g->d_tot, sizeof(double) * 3, hipMemcpyDeviceToDevice, g->stream) !=
                hipSuccess ||
            step_migrate(g, &b, G, 1, e->dt) < 0) {
            rc = gerror_(engine_err_gpu, "advance failed");
            break;
        }
        hipLaunchKernelGGL(k_kin_cells, dim3((G.ncells + 255) / 256), dim3(256), 0, g->stream, G.ncells, g->d_count,
                           g->d_offset, b.mass, g->d_type, b.v, b.cell_ke);
        hipLaunchKernelGGL(k_sum1, dim3(1), dim3(1024), 0, g->stream, G.ncells, b.cell_ke, b.esteps + 4 * step + 3);
        e->time += 1;
    }
    if (rc == engine_err_ok &&
        (hipMemcpyAsync(&flag, g->d_queue + MD_NQUEUE, sizeof(int), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
         hipStreamSynchronize(g->stream) != hipSuccess))
        rc = gerror_(engine_err_gpu, "step failed");
    if (rc == engine_err_ok && flag) rc = gerror_(engine_err_range, "sort stack overflow");
    if (rc == engine_err_ok) rc = step_download(e, g, &b, nsteps, epot, ekin);
    (void)hipStreamSynchronize(g->stream);
    step_free(&b);
    return rc;
}
