/*
 * ORACLE — test infrastructure only (see oracle.h).  FP64 CPU restatement of
 * mdcore's nonbonded path, compiled with -ffp-contract=off so that every
 * floating-point expression rounds like the reference's double build.
 *
 * Followed reference code (file:line under the reference's src/):
 *   grid, cell origins, span, task loops + box-distance test + dedup
 *                                        space.c:530-752
 *   pair orientation (minimum image, sid, flip)   space.c:80-114, cell.c:56-101
 *   particle insertion, cell-relative x  space.c:382-434 (engine_load, engine.c:1345-1392)
 *   one shuffle before the pass          space.c:307-364, engine.c:117-156
 *   sort keys and quicksort              runner_dosort.c:97-161, runner.c:110-174
 *   sorted pair loop                     runner_dopair.c:397-537, force sign as
 *                                        runner_doself (823-826): the double
 *                                        build's own pair loop (497-501) adds
 *                                        the force with the opposite sign
 *   self loop                            runner_dopair.c:753-855
 *   r^2                                  fptype.h:275-278 (3 components)
 *   table evaluation                     potential_eval.h:67-104
 *   exclusions                           exclusion.c:81-282 (scalar branch)
 *   energy sum over cells                engine.c:1963-1985
 *   kick-drift (or_advance)              engine.c:1945-2052, threads over cells
 * The scalar virial sum(dx . F_i) = -sum(f r^2) and the projected window
 * (OR_WINDOW_PROJ) have no reference counterpart.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

const double or_axis[39] = {
    5.773502691896258e-01,  5.773502691896258e-01,  5.773502691896258e-01,
    7.071067811865475e-01,  7.071067811865475e-01,  0.0,
    5.773502691896258e-01,  5.773502691896258e-01,  -5.773502691896258e-01,
    7.071067811865475e-01,  0.0,                    7.071067811865475e-01,
    1.0,                    0.0,                    0.0,
    7.071067811865475e-01,  0.0,                    -7.071067811865475e-01,
    5.773502691896258e-01,  -5.773502691896258e-01, 5.773502691896258e-01,
    7.071067811865475e-01,  -7.071067811865475e-01, 0.0,
    5.773502691896258e-01,  -5.773502691896258e-01, -5.773502691896258e-01,
    0.0,                    7.071067811865475e-01,  7.071067811865475e-01,
    0.0,                    1.0,                    0.0,
    0.0,                    7.071067811865475e-01,  -7.071067811865475e-01,
    0.0,                    0.0,                    1.0,
};

/* Sign pattern (base 3, x-major) -> axis, and whether to swap the cells. */
static const int or_sid_of[27] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 0,
                                  12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0};
static const int or_flip[27] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

enum { T_SELF = 1, T_PAIR = 2 };

struct or_system {
    double origin[3], dim[3], h[3], ih[3], cutoff, cutoff2, bias, dscale;
    int cdim[3], span[3], ncells, N;
    double *corg;
    int *cloc;
    int *cnt, *cap;
    int **cid;          /* particle ids per cell, in storage order */
    int **cty;          /* particle types per cell, in storage order */
    double **cx;        /* cell-relative positions, 3 per particle */
    unsigned int **sl;  /* 13 lists of cnt words per cell */
    unsigned int *sflags;
    int *type;          /* by id */
    int *cell_of, *slot_of;
    int ntask, *tt, *ti, *tj, *tsid;
    /* tables */
    int max_type, ntab, *pind, *tn, *toff;
    double *talpha, *tc;
    int nexcl, *excl;
    /* pseudo-Verlet state: per-cell maxdx widening the pair windows
       (runner_dopair.c:428) and whether or_nonbond keeps the sort lists of
       the last sort (runner.c:773: cells are re-sorted only on a rebuild) */
    double *maxdx;
    int keep_sort;
    /* pairs evaluated by each task in the last or_nonbond */
    long long *tcount;
    /* or_advance: velocities by id, inverse mass by type, and the cell
       offsets of the last or_nonbond's force buffer */
    double *v, *imass;
    int *foff;
    /* per-thread scratch of or_nonbond, kept between calls */
    int nbuf;
    double *fbuf, *ebuf;
};

/* ------------------------------------------------------------ primitives */

void or_eval(int n, const double *alpha, const double *c, double r2, double *e, double *f) {
    double r = sqrt(r2), x, ee, eff;
    int ind = (int)fmax(0.0, alpha[0] + r * (alpha[1] + r * alpha[2]));
    const double *cc;
    (void)n;
    cc = &c[ind * 8];
    x = (r - cc[0]) * cc[1];
    ee = cc[2] * x + cc[3];
    eff = cc[2];
    for (int k = 4; k < 8; k++) {
        eff = eff * x + ee;
        ee = ee * x + cc[k];
    }
    *e = ee;
    *f = eff * cc[1] / r;
}

unsigned int or_key(const double *x, const double *axis, double dscale, double bias) {
    return (unsigned int)(dscale * (bias + x[0] * axis[0] + x[1] * axis[1] + x[2] * axis[2]));
}

/* Planted-mistake hooks.  The tests compile this file once more with one of
   these defined and check that the hand-traced tie orders then fail; the
   oracle itself is always built without them.
     OR_PLANT_PIVOT_UP  pivot at (lo+hi+1)/2 instead of (lo+hi)/2
     OR_PLANT_CUT16     selection sort below 16 elements instead of 15
     OR_PLANT_LAST_MAX  selection sort keeps the last maximum ('>=')
     OR_PLANT_GE_SCANS  partition scans pass keys equal to the pivot
                        (bounded to the segment so the variant stays in it) */
#ifdef OR_PLANT_PIVOT_UP
#define OR_QS_MID(lo, hi) (((lo) + (hi) + 1) / 2)
#else
#define OR_QS_MID(lo, hi) (((lo) + (hi)) / 2)
#endif
#ifdef OR_PLANT_CUT16
#define OR_QS_CUT 16
#else
#define OR_QS_CUT 15
#endif
#ifdef OR_PLANT_LAST_MAX
#define OR_QS_BETTER(kj, km) ((kj) >= (km))
#else
#define OR_QS_BETTER(kj, km) ((kj) > (km))
#endif
#ifdef OR_PLANT_GE_SCANS
#define OR_QS_SCAN_I (i < hi && (a[i] & 0xffff) >= pivot)
#define OR_QS_SCAN_J (j > lo && (a[j] & 0xffff) <= pivot)
#else
#define OR_QS_SCAN_I ((a[i] & 0xffff) > pivot)
#define OR_QS_SCAN_J ((a[j] & 0xffff) < pivot)
#endif

void or_sort_desc(unsigned int *a, int N) {
    struct {
        short lo, hi;
    } st[10];
    int qpos = 0, i, j, lo, hi, imax;
    unsigned int pivot, t;
    if (N <= 0) return;
    st[0].lo = 0;
    st[0].hi = (short)(N - 1);
    while (qpos >= 0) {
        lo = st[qpos].lo;
        hi = st[qpos].hi;
        qpos--;
        if (hi - lo < OR_QS_CUT) {
            for (i = lo; i < hi; i++) {
                imax = i;
                for (j = i + 1; j <= hi; j++)
                    if (OR_QS_BETTER(a[j] & 0xffff, a[imax] & 0xffff)) imax = j;
                if (imax != i) {
                    t = a[imax];
                    a[imax] = a[i];
                    a[i] = t;
                }
            }
            continue;
        }
        pivot = a[OR_QS_MID(lo, hi)] & 0xffff;
        i = lo;
        j = hi;
        while (i <= j) {
            while (OR_QS_SCAN_I) i++;
            while (OR_QS_SCAN_J) j--;
            if (i <= j) {
                if (i < j) {
                    t = a[i];
                    a[i] = a[j];
                    a[j] = t;
                }
                i++;
                j--;
            }
        }
        if (j > (lo + hi) / 2) {
            if (lo < j) { qpos++; st[qpos].lo = lo; st[qpos].hi = j; }
            if (i < hi) { qpos++; st[qpos].lo = i; st[qpos].hi = hi; }
        } else {
            if (i < hi) { qpos++; st[qpos].lo = i; st[qpos].hi = hi; }
            if (lo < j) { qpos++; st[qpos].lo = lo; st[qpos].hi = j; }
        }
    }
}

/* Minimum-image shift from cell a to cell b, the pattern id and the flip. */
static int or_orient(const or_system *s, int a, int b, double *sh) {
    int k, pat = 0;
    for (k = 0; k < 3; k++) {
        sh[k] = s->corg[3 * b + k] - s->corg[3 * a + k];
        if (sh[k] * 2 > s->dim[k])
            sh[k] -= s->dim[k];
        else if (sh[k] * 2 < -s->dim[k])
            sh[k] += s->dim[k];
        pat = 3 * pat + ((sh[k] < 0) ? 0 : ((sh[k] > 0) ? 2 : 1));
    }
    return pat;
}

/* ------------------------------------------------------------ construction */

static int cell_index(const or_system *s, int i, int j, int k) { return (i * s->cdim[1] + j) * s->cdim[2] + k; }

static void cell_push(or_system *s, int c, int id, const double *x) {
    if (s->cnt[c] == s->cap[c]) {
        s->cap[c] = s->cap[c] ? 2 * s->cap[c] : 16;
        s->cid[c] = (int *)realloc(s->cid[c], sizeof(int) * s->cap[c]);
        s->cx[c] = (double *)realloc(s->cx[c], sizeof(double) * 3 * s->cap[c]);
    }
    s->cid[c][s->cnt[c]] = id;
    memcpy(&s->cx[c][3 * s->cnt[c]], x, sizeof(double) * 3);
    s->cnt[c]++;
}

static int wrapc(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

or_system *or_create(const double *origin, const double *dim, const double *L, double cutoff, int N,
                     const double *x, const int *type, const or_tables *tab, int nexcl, const int *excl) {
    or_system *s = (or_system *)calloc(1, sizeof(or_system));
    int i, j, k, l0, l1, l2, c, nc, t;
    if (s == NULL) return NULL;
    for (k = 0; k < 3; k++) {
        s->origin[k] = origin[k];
        s->dim[k] = dim[k];
        s->cdim[k] = (int)floor(dim[k] / L[k]);
        s->h[k] = dim[k] / s->cdim[k];
        s->ih[k] = 1.0 / s->h[k];
    }
    s->cutoff = cutoff;
    s->cutoff2 = cutoff * cutoff;
    nc = s->ncells = s->cdim[0] * s->cdim[1] * s->cdim[2];
    for (k = 0; k < 3; k++) s->span[k] = (int)ceil(cutoff * s->ih[k]);
    s->bias = sqrt(s->h[0] * s->h[0] + s->h[1] * s->h[1] + s->h[2] * s->h[2]);
    s->dscale = (double)32767 / (2 * s->bias);
    s->corg = (double *)malloc(sizeof(double) * 3 * nc);
    s->cloc = (int *)malloc(sizeof(int) * 3 * nc);
    for (i = 0; i < s->cdim[0]; i++)
        for (j = 0; j < s->cdim[1]; j++)
            for (k = 0; k < s->cdim[2]; k++) {
                c = cell_index(s, i, j, k);
                s->corg[3 * c] = origin[0] + i * s->h[0];
                s->corg[3 * c + 1] = origin[1] + j * s->h[1];
                s->corg[3 * c + 2] = origin[2] + k * s->h[2];
                s->cloc[3 * c] = i;
                s->cloc[3 * c + 1] = j;
                s->cloc[3 * c + 2] = k;
            }
    s->cnt = (int *)calloc(nc, sizeof(int));
    s->cap = (int *)calloc(nc, sizeof(int));
    s->cid = (int **)calloc(nc, sizeof(int *));
    s->cx = (double **)calloc(nc, sizeof(double *));
    s->sl = (unsigned int **)calloc(nc, sizeof(unsigned int *));
    s->sflags = (unsigned int *)calloc(nc, sizeof(unsigned int));
    s->maxdx = (double *)calloc(nc, sizeof(double));

    /* Tasks in the reference's order (periodic in all directions). */
    s->ntask = 0;
    {
        int cap = nc * ((2 * s->span[0] + 1) * (2 * s->span[1] + 1) * (2 * s->span[2] + 1) + 2);
        s->tt = (int *)malloc(sizeof(int) * cap);
        s->ti = (int *)malloc(sizeof(int) * cap);
        s->tj = (int *)malloc(sizeof(int) * cap);
        s->tsid = (int *)malloc(sizeof(int) * cap);
    }
    for (i = 0; i < s->cdim[0]; i++)
        for (j = 0; j < s->cdim[1]; j++)
            for (k = 0; k < s->cdim[2]; k++) {
                int id1 = cell_index(s, i, j, k);
                for (l0 = -s->span[0]; l0 <= s->span[0]; l0++)
                    for (l1 = -s->span[1]; l1 <= s->span[1]; l1++)
                        for (l2 = -s->span[2]; l2 <= s->span[2]; l2++) {
                            double lh0 = s->h[0] * fmax(abs(l0) - 1, 0);
                            double lh1 = s->h[1] * fmax(abs(l1) - 1, 0);
                            double lh2 = s->h[2] * fmax(abs(l2) - 1, 0);
                            int id2, pat, a, b;
                            double sh[3];
                            if (lh0 * lh0 + lh1 * lh1 + lh2 * lh2 > s->cutoff2) continue;
                            id2 = cell_index(s, wrapc(i + l0, s->cdim[0]), wrapc(j + l1, s->cdim[1]),
                                             wrapc(k + l2, s->cdim[2]));
                            if (!(id1 < id2 || (id1 == id2 && l0 == 0 && l1 == 0 && l2 == 0))) continue;
                            pat = or_orient(s, id1, id2, sh);
                            a = id1;
                            b = id2;
                            if (or_flip[pat]) {
                                a = id2;
                                b = id1;
                            }
                            t = s->ntask++;
                            s->tt[t] = (id1 == id2) ? T_SELF : T_PAIR;
                            s->ti[t] = a;
                            s->tj[t] = b;
                            s->tsid[t] = or_sid_of[pat];
                            if (id1 != id2) {
                                s->sflags[a] |= 1u << s->tsid[t];
                                s->sflags[b] |= 1u << s->tsid[t];
                            }
                        }
            }
    s->tcount = (long long *)calloc(s->ntask + 1, sizeof(long long));

    /* Tables (copied). */
    s->max_type = tab->max_type;
    s->ntab = tab->ntab;
    s->pind = (int *)malloc(sizeof(int) * tab->max_type * tab->max_type);
    memcpy(s->pind, tab->pind, sizeof(int) * tab->max_type * tab->max_type);
    s->tn = (int *)malloc(sizeof(int) * (tab->ntab + 1));
    s->toff = (int *)malloc(sizeof(int) * (tab->ntab + 1));
    s->talpha = (double *)malloc(sizeof(double) * 4 * (tab->ntab + 1));
    {
        int ncoef = 0;
        for (t = 0; t < tab->ntab; t++) {
            s->tn[t] = tab->n[t];
            s->toff[t] = tab->off[t];
            memcpy(&s->talpha[4 * t], &tab->alpha[4 * t], sizeof(double) * 4);
            if (tab->off[t] + (tab->n[t] + 1) * 8 > ncoef) ncoef = tab->off[t] + (tab->n[t] + 1) * 8;
        }
        s->tc = (double *)malloc(sizeof(double) * (ncoef + 8));
        memcpy(s->tc, tab->c, sizeof(double) * ncoef);
    }

    /* Particles in input order (engine_load), positions made cell-relative. */
    s->N = N;
    s->type = (int *)malloc(sizeof(int) * N);
    s->cell_of = (int *)malloc(sizeof(int) * N);
    s->slot_of = (int *)malloc(sizeof(int) * N);
    for (i = 0; i < N; i++) {
        int ind[3];
        double xl[3];
        for (k = 0; k < 3; k++) ind[k] = (int)((x[3 * i + k] - origin[k]) * s->ih[k]);
        for (k = 0; k < 3; k++)
            if (ind[k] < 0 || ind[k] >= s->cdim[k]) {
                or_free(s);
                return NULL;
            }
        c = cell_index(s, ind[0], ind[1], ind[2]);
        for (k = 0; k < 3; k++) xl[k] = x[3 * i + k] - s->corg[3 * c + k];
        s->type[i] = type[i];
        cell_push(s, c, i, xl);
    }

    /* One shuffle: particles outside [0,h) move to the neighbour cell; the
       hole is filled with the cell's last particle; arrivals are appended
       per destination in the order they were sent. */
    {
        int *icnt = (int *)calloc(nc, sizeof(int)), *icap = (int *)calloc(nc, sizeof(int));
        int **iid = (int **)calloc(nc, sizeof(int *));
        double **ix = (double **)calloc(nc, sizeof(double *));
        for (c = 0; c < nc; c++) {
            int p = 0;
            while (p < s->cnt[c]) {
                double *xp = &s->cx[c][3 * p];
                int d[3];
                for (k = 0; k < 3; k++) d[k] = (xp[k] >= s->h[k]) - (xp[k] < 0.0);
                if (d[0] || d[1] || d[2]) {
                    int dest;
                    for (k = 0; k < 3; k++) xp[k] -= d[k] * s->h[k];
                    dest = cell_index(s, (s->cloc[3 * c] + d[0] + s->cdim[0]) % s->cdim[0],
                                      (s->cloc[3 * c + 1] + d[1] + s->cdim[1]) % s->cdim[1],
                                      (s->cloc[3 * c + 2] + d[2] + s->cdim[2]) % s->cdim[2]);
                    if (icnt[dest] == icap[dest]) {
                        icap[dest] = icap[dest] ? 2 * icap[dest] : 4;
                        iid[dest] = (int *)realloc(iid[dest], sizeof(int) * icap[dest]);
                        ix[dest] = (double *)realloc(ix[dest], sizeof(double) * 3 * icap[dest]);
                    }
                    iid[dest][icnt[dest]] = s->cid[c][p];
                    memcpy(&ix[dest][3 * icnt[dest]], xp, sizeof(double) * 3);
                    icnt[dest]++;
                    s->cnt[c]--;
                    if (p < s->cnt[c]) {
                        s->cid[c][p] = s->cid[c][s->cnt[c]];
                        memcpy(&s->cx[c][3 * p], &s->cx[c][3 * s->cnt[c]], sizeof(double) * 3);
                    }
                } else
                    p++;
            }
        }
        for (c = 0; c < nc; c++) {
            for (k = 0; k < icnt[c]; k++) cell_push(s, c, iid[c][k], &ix[c][3 * k]);
            free(iid[c]);
            free(ix[c]);
        }
        free(icnt);
        free(icap);
        free(iid);
        free(ix);
    }
    s->cty = (int **)calloc(nc, sizeof(int *));
    for (c = 0; c < nc; c++) {
        s->cty[c] = (int *)malloc(sizeof(int) * (s->cnt[c] + 1));
        for (k = 0; k < s->cnt[c]; k++) {
            s->cell_of[s->cid[c][k]] = c;
            s->slot_of[s->cid[c][k]] = k;
            s->cty[c][k] = s->type[s->cid[c][k]];
        }
        s->sl[c] = (unsigned int *)malloc(sizeof(unsigned int) * 13 * (s->cnt[c] + 1));
    }

    /* Exclusions: i <= j, sorted, unique (engine_bonded.c:1301-1434). */
    s->excl = (int *)malloc(sizeof(int) * 2 * (nexcl + 1));
    s->nexcl = 0;
    for (k = 0; k < nexcl; k++) {
        int a = excl[2 * k], b = excl[2 * k + 1];
        s->excl[2 * s->nexcl] = a <= b ? a : b;
        s->excl[2 * s->nexcl + 1] = a <= b ? b : a;
        s->nexcl++;
    }
    /* insertion sort is fine for test sizes; use a simple merge-free qsort */
    {
        int n = s->nexcl;
        long long *keys = (long long *)malloc(sizeof(long long) * (n + 1));
        for (k = 0; k < n; k++) keys[k] = ((long long)s->excl[2 * k] << 32) | (unsigned int)s->excl[2 * k + 1];
        /* heap sort on 64-bit keys */
        for (int start = n / 2 - 1; start >= 0; start--) {
            int root = start;
            while (2 * root + 1 < n) {
                int ch = 2 * root + 1;
                if (ch + 1 < n && keys[ch] < keys[ch + 1]) ch++;
                if (keys[root] < keys[ch]) {
                    long long tmp = keys[root];
                    keys[root] = keys[ch];
                    keys[ch] = tmp;
                    root = ch;
                } else
                    break;
            }
        }
        for (int end = n - 1; end > 0; end--) {
            long long tmp = keys[0];
            keys[0] = keys[end];
            keys[end] = tmp;
            int root = 0;
            while (2 * root + 1 < end) {
                int ch = 2 * root + 1;
                if (ch + 1 < end && keys[ch] < keys[ch + 1]) ch++;
                if (keys[root] < keys[ch]) {
                    tmp = keys[root];
                    keys[root] = keys[ch];
                    keys[ch] = tmp;
                    root = ch;
                } else
                    break;
            }
        }
        int m = 0;
        for (k = 0; k < n; k++) {
            if (m > 0 && keys[k] == keys[m - 1]) continue;
            keys[m++] = keys[k];
        }
        for (k = 0; k < m; k++) {
            s->excl[2 * k] = (int)(keys[k] >> 32);
            s->excl[2 * k + 1] = (int)(keys[k] & 0xffffffff);
        }
        s->nexcl = m;
        free(keys);
    }
    return s;
}

void or_free(or_system *s) {
    int c;
    if (s == NULL) return;
    for (c = 0; c < s->ncells; c++) {
        if (s->cid) free(s->cid[c]);
        if (s->cx) free(s->cx[c]);
        if (s->sl) free(s->sl[c]);
        if (s->cty) free(s->cty[c]);
    }
    free(s->cty);
    free(s->cid);
    free(s->cx);
    free(s->sl);
    free(s->cnt);
    free(s->cap);
    free(s->sflags);
    free(s->maxdx);
    free(s->corg);
    free(s->cloc);
    free(s->type);
    free(s->cell_of);
    free(s->slot_of);
    free(s->tt);
    free(s->ti);
    free(s->tj);
    free(s->tsid);
    free(s->tcount);
    free(s->v);
    free(s->imass);
    free(s->foff);
    free(s->pind);
    free(s->tn);
    free(s->toff);
    free(s->talpha);
    free(s->tc);
    free(s->excl);
    free(s->fbuf);
    free(s->ebuf);
    free(s);
}

/* ------------------------------------------------------------ sort */

static void sort_cell(or_system *s, int c) {
    int n = s->cnt[c], sid, i;
    if (n == 0) return;
    for (sid = 0; sid < 13; sid++) {
        unsigned int *lst;
        if (!(s->sflags[c] & (1u << sid))) continue;
        lst = &s->sl[c][n * sid];
        for (i = 0; i < n; i++)
            lst[i] = ((unsigned int)i << 16) | or_key(&s->cx[c][3 * i], &or_axis[3 * sid], s->dscale, s->bias);
        or_sort_desc(lst, n);
    }
}

int or_sort(or_system *s) {
    for (int c = 0; c < s->ncells; c++) sort_cell(s, c);
    return 0;
}

/* ------------------------------------------------------------ pair work */

typedef struct {
    double *f;      /* 3 per (cell, slot) flattened with foff */
    double epot_c;  /* energy of the current task */
    double vir;
    long long np;
} or_acc;

static void do_self(const or_system *s, int c, double *f, const int *foff, or_acc *A) {
    int n = s->cnt[c], i, j, k;
    const double *x = s->cx[c];
    double *fc = &f[3 * foff[c]];
    const int *ty = s->cty[c];
    for (i = 1; i < n; i++) {
        double pix[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        const int *prow = &s->pind[ty[i] * s->max_type];
        for (j = 0; j < i; j++) {
            double dx[3], r2, e, ff, w;
            int tb;
            for (k = 0; k < 3; k++) dx[k] = pix[k] - x[3 * j + k];
            r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (r2 > s->cutoff2) continue;
            tb = prow[ty[j]];
            if (tb < 0) continue;
            or_eval(s->tn[tb], &s->talpha[4 * tb], &s->tc[s->toff[tb]], r2, &e, &ff);
            for (k = 0; k < 3; k++) {
                w = ff * dx[k];
                fc[3 * i + k] -= w;
                fc[3 * j + k] += w;
            }
            A->epot_c += e;
            A->vir -= ff * r2;
            A->np++;
        }
    }
}

static void do_pair(const or_system *s, int ci, int cj, int sid, int mode, double *f, const int *foff,
                    or_acc *A) {
    int ni = s->cnt[ci], nj = s->cnt[cj], i, j, k, dmaxdist, dnshift;
    double sh[3];
    const unsigned int *ip, *jp;
    const double *xi = s->cx[ci], *xj = s->cx[cj];
    double *fi = &f[3 * foff[ci]], *fj = &f[3 * foff[cj]];
    if (ni == 0 || nj == 0) return;
    /* runner_dopair.c:428, with the task's cells in task order */
    dmaxdist = (int)(2 + s->dscale * (s->cutoff + s->maxdx[ci] + s->maxdx[cj]));
    (void)or_orient(s, ci, cj, sh);
    if (mode == OR_WINDOW_REF)
        dnshift = (int)(s->dscale * sqrt(sh[0] * sh[0] + sh[1] * sh[1] + sh[2] * sh[2]));
    else
        dnshift = (int)(s->dscale * (sh[0] * or_axis[3 * sid] + sh[1] * or_axis[3 * sid + 1] +
                                     sh[2] * or_axis[3 * sid + 2]));
    ip = &s->sl[ci][ni * sid];
    jp = &s->sl[cj][nj * sid];
    for (i = 0; i < ni; i++) {
        int pi = (int)(ip[i] >> 16), ti;
        double pix[3];
        if (mode == OR_WINDOW_REF) {
            if ((jp[nj - 1] & 0xffffu) + (unsigned int)dnshift - (ip[i] & 0xffffu) > (unsigned int)dmaxdist) break;
        } else {
            if ((int)(jp[nj - 1] & 0xffff) + dnshift - (int)(ip[i] & 0xffff) > dmaxdist) break;
        }
        for (k = 0; k < 3; k++) pix[k] = xi[3 * pi + k] - sh[k];
        ti = s->cty[ci][pi];
        const int *prow = &s->pind[ti * s->max_type], *tyj = s->cty[cj];
        for (j = nj - 1; j >= 0; j--) {
            int pj, tb;
            double dx[3], r2, e, ff, w;
            if (mode == OR_WINDOW_REF) {
                if (!((jp[j] & 0xffffu) + (unsigned int)dnshift - (ip[i] & 0xffffu) < (unsigned int)dmaxdist)) break;
            } else {
                if (!((int)(jp[j] & 0xffff) + dnshift - (int)(ip[i] & 0xffff) < dmaxdist)) break;
            }
            pj = (int)(jp[j] >> 16);
            tb = prow[tyj[pj]];
            if (tb < 0) continue;
            for (k = 0; k < 3; k++) dx[k] = pix[k] - xj[3 * pj + k];
            r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (r2 > s->cutoff2) continue;
            or_eval(s->tn[tb], &s->talpha[4 * tb], &s->tc[s->toff[tb]], r2, &e, &ff);
            for (k = 0; k < 3; k++) {
                w = ff * dx[k];
                fi[3 * pi + k] -= w;
                fj[3 * pj + k] += w;
            }
            A->epot_c += e;
            A->vir -= ff * r2;
            A->np++;
        }
    }
}

typedef struct {
    or_system *s;
    int mode, *next, *foff, tid, nthreads;
    double *f, *cepot, vir;
    long long np;
    int phase;
} or_worker;

static void *worker_main(void *arg) {
    or_worker *w = (or_worker *)arg;
    or_system *s = w->s;
    or_acc A;
    int t;
    size_t k, lo, hi, n3 = (size_t)3 * s->N;
    A.f = w->f;
    A.vir = 0.0;
    A.np = 0;
    if (w->phase == 0) {
        memset(w->f, 0, sizeof(double) * n3);
        memset(w->cepot, 0, sizeof(double) * s->ncells);
        if (!s->keep_sort)
            while ((t = __atomic_fetch_add(w->next, 1, __ATOMIC_RELAXED)) < s->ncells) sort_cell(s, t);
        return NULL;
    }
    if (w->phase == 2) {
        /* Sum the private buffers into buffer 0 over this thread's slice. */
        lo = n3 * w->tid / w->nthreads;
        hi = n3 * (w->tid + 1) / w->nthreads;
        for (int b = 1; b < w->nthreads; b++) {
            const double *src = s->fbuf + n3 * b;
            for (k = lo; k < hi; k++) s->fbuf[k] += src[k];
        }
        return NULL;
    }
    while ((t = __atomic_fetch_add(w->next, 1, __ATOMIC_RELAXED)) < s->ntask) {
        long long np0 = A.np;
        A.epot_c = 0.0;
        if (s->tt[t] == T_SELF)
            do_self(s, s->ti[t], w->f, w->foff, &A);
        else
            do_pair(s, s->ti[t], s->tj[t], s->tsid[t], w->mode, w->f, w->foff, &A);
        w->cepot[s->ti[t]] += A.epot_c;
        s->tcount[t] = A.np - np0;
    }
    w->vir = A.vir;
    w->np = A.np;
    return NULL;
}

static double exclusions(const or_system *s, double *f, const int *foff, long long *count) {
    double epot = 0.0;
    long long n = 0;
    int b, k;
    for (b = 0; b < s->nexcl; b++) {
        int i = s->excl[2 * b], j = s->excl[2 * b + 1], ci, cj, tb;
        double dx[3], r2 = 0.0, e, ff, w;
        const double *xi, *xj;
        if (i < 0 || j < 0 || i >= s->N || j >= s->N) continue;
        tb = s->pind[s->type[j] * s->max_type + s->type[i]];
        if (tb < 0) continue;
        ci = s->cell_of[i];
        cj = s->cell_of[j];
        xi = &s->cx[ci][3 * s->slot_of[i]];
        xj = &s->cx[cj][3 * s->slot_of[j]];
        for (k = 0; k < 3; k++) {
            int sh = s->cloc[3 * ci + k] - s->cloc[3 * cj + k];
            if (sh > 1)
                sh = -1;
            else if (sh < -1)
                sh = 1;
            dx[k] = xi[k] - xj[k] + s->h[k] * sh;
            r2 += dx[k] * dx[k];
        }
        if (r2 > s->cutoff2) continue;
        or_eval(s->tn[tb], &s->talpha[4 * tb], &s->tc[s->toff[tb]], r2, &e, &ff);
        for (k = 0; k < 3; k++) {
            w = ff * dx[k];
            f[3 * (foff[ci] + s->slot_of[i]) + k] += w;
            f[3 * (foff[cj] + s->slot_of[j]) + k] -= w;
        }
        epot += e;
        n++;
    }
    *count = n;
    return -epot;
}

int or_nonbond(or_system *s, int mode, int nthreads, double *f, double *energies, long long *counts) {
    int nc = s->ncells, c, k, t, id;
    int *foff = (int *)malloc(sizeof(int) * (nc + 1));
    double *fl, epot = 0.0, vir = 0.0, eexcl;
    long long np = 0, nx = 0;
    size_t n3 = (size_t)3 * s->N;
    pthread_t th[256];
    or_worker w[256];
    int next;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    foff[0] = 0;
    for (c = 0; c < nc; c++) foff[c + 1] = foff[c] + s->cnt[c];
    if (s->nbuf < nthreads) {
        free(s->fbuf);
        free(s->ebuf);
        s->fbuf = (double *)malloc(sizeof(double) * n3 * nthreads + 8);
        s->ebuf = (double *)malloc(sizeof(double) * (size_t)nc * nthreads + 8);
        s->nbuf = nthreads;
    }
    for (int phase = 0; phase < 3; phase++) {
        if (phase == 2 && nthreads == 1) break;
        next = 0;
        for (t = 0; t < nthreads; t++) {
            w[t].s = s;
            w[t].mode = mode;
            w[t].next = &next;
            w[t].foff = foff;
            w[t].tid = t;
            w[t].nthreads = nthreads;
            w[t].f = s->fbuf + n3 * t;
            w[t].cepot = s->ebuf + (size_t)nc * t;
            if (phase == 1) {
                w[t].vir = 0.0;
                w[t].np = 0;
            }
            w[t].phase = phase;
        }
        if (nthreads == 1)
            worker_main(&w[0]);
        else {
            for (t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, worker_main, &w[t]);
            for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        }
        if (phase == 1)
            for (t = 0; t < nthreads; t++) {
                vir += w[t].vir;
                np += w[t].np;
            }
    }
    fl = s->fbuf;
    for (c = 0; c < nc; c++)
        for (t = 0; t < nthreads; t++) epot += s->ebuf[(size_t)nc * t + c];
    eexcl = exclusions(s, fl, foff, &nx);
    if (f)
        for (id = 0; id < s->N; id++)
            for (k = 0; k < 3; k++) f[3 * id + k] = fl[3 * (foff[s->cell_of[id]] + s->slot_of[id]) + k];
    if (energies) {
        energies[0] = epot;
        energies[1] = eexcl;
        energies[2] = vir;
    }
    if (counts) {
        counts[0] = np;
        counts[1] = nx;
    }
    free(s->foff);
    s->foff = foff;
    return 0;
}

/* ------------------------------------------------------------ advance */

int or_set_motion(or_system *s, const double *v, const double *imass_by_type) {
    if (s->v == NULL) s->v = (double *)calloc((size_t)3 * s->N + 1, sizeof(double));
    if (s->imass == NULL) s->imass = (double *)calloc(s->max_type + 1, sizeof(double));
    if (s->v == NULL || s->imass == NULL) return -1;
    if (v) memcpy(s->v, v, sizeof(double) * 3 * s->N);
    if (imass_by_type) memcpy(s->imass, imass_by_type, sizeof(double) * s->max_type);
    return 0;
}

typedef struct {
    or_system *s;
    double dt;
    int c0, c1, out;
} or_adv_job;

static void *advance_main(void *arg) {
    or_adv_job *J = (or_adv_job *)arg;
    or_system *s = J->s;
    const double dt = J->dt;
    for (int c = J->c0; c < J->c1; c++) {
        double *x = s->cx[c];
        const double *f = &s->fbuf[3 * (size_t)s->foff[c]];
        for (int k = 0; k < s->cnt[c]; k++) {
            double *v = &s->v[3 * (size_t)s->cid[c][k]];
            const double w = dt * s->imass[s->cty[c][k]];
            int out = 0;
            for (int m = 0; m < 3; m++) {
                v[m] += f[3 * k + m] * w;
                x[3 * k + m] += dt * v[m];
                out |= (x[3 * k + m] >= s->h[m]) | (x[3 * k + m] < 0.0);
            }
            J->out += out;
        }
    }
    return NULL;
}

int or_advance(or_system *s, double dt, int nthreads) {
    pthread_t th[256];
    or_adv_job J[256];
    int t, out = 0;
    if (s->foff == NULL || s->v == NULL || s->imass == NULL) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    for (t = 0; t < nthreads; t++) {
        J[t].s = s;
        J[t].dt = dt;
        J[t].c0 = (int)((long long)s->ncells * t / nthreads);
        J[t].c1 = (int)((long long)s->ncells * (t + 1) / nthreads);
        J[t].out = 0;
    }
    if (nthreads == 1)
        advance_main(&J[0]);
    else {
        for (t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, advance_main, &J[t]);
        for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    for (t = 0; t < nthreads; t++) out += J[t].out;
    return out;
}

/* ------------------------------------------------------------ brute force */

/* Binary search in the sorted unique exclusion list (a < b). */
static int is_excluded(const or_system *s, int a, int b) {
    int lo = 0, hi = s->nexcl - 1;
    while (lo <= hi) {
        int m = (lo + hi) / 2, i = s->excl[2 * m], j = s->excl[2 * m + 1];
        if (i == a && j == b) return 1;
        if (i < a || (i == a && j < b))
            lo = m + 1;
        else
            hi = m - 1;
    }
    return 0;
}

int or_bruteforce(or_system *s, double *f, double *energies, long long *count) {
    int nc = s->ncells, a, k, o0, o1, o2;
    double epot = 0.0, vir = 0.0;
    long long n = 0;
    for (k = 0; k < 3; k++)
        if (s->cdim[k] < 2 * s->span[k] + 1) return -1;
    if (f) memset(f, 0, sizeof(double) * 3 * s->N);
    for (a = 0; a < s->N; a++) {
        int ca = s->cell_of[a];
        for (o0 = -s->span[0]; o0 <= s->span[0]; o0++)
            for (o1 = -s->span[1]; o1 <= s->span[1]; o1++)
                for (o2 = -s->span[2]; o2 <= s->span[2]; o2++) {
                    int cb = cell_index(s, wrapc(s->cloc[3 * ca] + o0, s->cdim[0]),
                                        wrapc(s->cloc[3 * ca + 1] + o1, s->cdim[1]),
                                        wrapc(s->cloc[3 * ca + 2] + o2, s->cdim[2]));
                    int q;
                    (void)nc;
                    for (q = 0; q < s->cnt[cb]; q++) {
                        int b = s->cid[cb][q], tb, ia, ib, cia, cib;
                        double sh[3], dx[3], r2, e, ff;
                        if (b <= a) continue;
                        if (is_excluded(s, a, b)) continue;
                        tb = s->pind[s->type[a] * s->max_type + s->type[b]];
                        if (tb < 0) continue;
                        if (ca == cb) {
                            ia = a;
                            ib = b;
                            cia = cib = ca;
                            sh[0] = sh[1] = sh[2] = 0.0;
                        } else {
                            int pat = or_orient(s, ca, cb, sh);
                            ia = a;
                            ib = b;
                            cia = ca;
                            cib = cb;
                            if (or_flip[pat]) {
                                ia = b;
                                ib = a;
                                cia = cb;
                                cib = ca;
                                for (k = 0; k < 3; k++) sh[k] = -sh[k];
                            }
                        }
                        for (k = 0; k < 3; k++)
                            dx[k] = (s->cx[cia][3 * s->slot_of[ia] + k] - sh[k]) - s->cx[cib][3 * s->slot_of[ib] + k];
                        r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
                        if (r2 > s->cutoff2) continue;
                        or_eval(s->tn[tb], &s->talpha[4 * tb], &s->tc[s->toff[tb]], r2, &e, &ff);
                        if (f)
                            for (k = 0; k < 3; k++) {
                                f[3 * ia + k] -= ff * dx[k];
                                f[3 * ib + k] += ff * dx[k];
                            }
                        epot += e;
                        vir -= ff * r2;
                        n++;
                    }
                }
    }
    if (energies) {
        energies[0] = epot;
        energies[1] = 0.0;
        energies[2] = vir;
    }
    if (count) *count = n;
    return 0;
}

/* ------------------------------------------------------------ introspection */

int or_ncells(const or_system *s) { return s->ncells; }

int or_cell(const or_system *s, int c, int *count, int *ids, unsigned int *sortflags) {
    if (c < 0 || c >= s->ncells) return -1;
    if (count) *count = s->cnt[c];
    if (ids) memcpy(ids, s->cid[c], sizeof(int) * s->cnt[c]);
    if (sortflags) *sortflags = s->sflags[c];
    return s->cnt[c];
}

int or_set_verlet(or_system *s, const double *xrel, const double *cell_maxdx, int keep_sort) {
    int id, k, c;
    if (s == NULL) return -1;
    if (xrel)
        for (id = 0; id < s->N; id++)
            for (k = 0; k < 3; k++) s->cx[s->cell_of[id]][3 * s->slot_of[id] + k] = xrel[3 * id + k];
    if (cell_maxdx)
        for (c = 0; c < s->ncells; c++) s->maxdx[c] = cell_maxdx[c];
    s->keep_sort = keep_sort;
    return 0;
}

int or_sortlist(const or_system *s, int c, int sid, unsigned int *out) {
    if (c < 0 || c >= s->ncells || sid < 0 || sid > 12) return -1;
    memcpy(out, &s->sl[c][s->cnt[c] * sid], sizeof(unsigned int) * s->cnt[c]);
    return s->cnt[c];
}

int or_ntasks(const or_system *s, int *npair, int *nself) {
    int np = 0, ns = 0;
    for (int t = 0; t < s->ntask; t++) {
        if (s->tt[t] == T_PAIR) np++;
        if (s->tt[t] == T_SELF) ns++;
    }
    if (npair) *npair = np;
    if (nself) *nself = ns;
    return s->ntask;
}

int or_task(const or_system *s, int k, int *type, int *i, int *j, int *sid) {
    if (k < 0 || k >= s->ntask) return -1;
    if (type) *type = s->tt[k];
    if (i) *i = s->ti[k];
    if (j) *j = s->tj[k];
    if (sid) *sid = s->tsid[k];
    return 0;
}

double or_dscale(const or_system *s) { return s->dscale; }

int or_task_counts(const or_system *s, long long *out) {
    memcpy(out, s->tcount, sizeof(long long) * s->ntask);
    return s->ntask;
}

int or_cell_x(const or_system *s, int c, double *x) {
    if (c < 0 || c >= s->ncells) return -1;
    memcpy(x, s->cx[c], sizeof(double) * 3 * s->cnt[c]);
    return s->cnt[c];
}

long long or_sortlists(const or_system *s, unsigned int *out) {
    long long off = 0;
    for (int c = 0; c < s->ncells; c++) {
        if (out) memcpy(&out[off], s->sl[c], sizeof(unsigned int) * 13 * s->cnt[c]);
        off += 13LL * s->cnt[c];
    }
    return off;
}
