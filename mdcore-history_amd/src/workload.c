/*
 * Synthetic input systems for the nonbonded path (seeded, reproducible):
 *   md_gen_fcc        fcc lattice (argon fluid start)
 *   md_gen_maxwell    Maxwell-Boltzmann velocities with zero total momentum
 *   md_gen_water      3-site water molecules on a lattice, random orientation
 *   md_gen_rsa        random sequential addition with a minimum separation
 *   md_gen_types      uniform random species labels
 *   md_sample_pairs   in-cutoff pair distances of a configuration (training
 *                     samples for the MLP surrogate)
 * These are drivers' helpers (the reference builds its systems inside the
 * example programs, e.g. examples/argon/argon.c:160-200); they are not part
 * of the nonbonded path itself.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "workload.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* splitmix64: small, fast, reproducible across platforms. */
static uint64_t sm_next(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double md_uniform(uint64_t *s) { return (sm_next(s) >> 11) * (1.0 / 9007199254740992.0); }

static double gauss(uint64_t *s) {
    double u1 = md_uniform(s), u2 = md_uniform(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

int md_gen_fcc(int nx, int ny, int nz, double a, double offset, double *x) {
    static const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
    int i, j, k, b, n = 0;
    if (x == NULL || nx < 1 || ny < 1 || nz < 1) return -1;
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++)
            for (k = 0; k < nz; k++)
                for (b = 0; b < 4; b++) {
                    x[3 * n] = offset + (i + basis[b][0]) * a;
                    x[3 * n + 1] = offset + (j + basis[b][1]) * a;
                    x[3 * n + 2] = offset + (k + basis[b][2]) * a;
                    n++;
                }
    return n;
}

int md_gen_maxwell(int N, const double *mass, double kT, uint64_t seed, double *v) {
    double p[3] = {0, 0, 0}, mtot = 0.0;
    uint64_t s = seed;
    int i, k;
    if (v == NULL || mass == NULL || N < 1) return -1;
    for (i = 0; i < N; i++) {
        double sig = sqrt(kT / mass[i]);
        for (k = 0; k < 3; k++) {
            v[3 * i + k] = sig * gauss(&s);
            p[k] += mass[i] * v[3 * i + k];
        }
        mtot += mass[i];
    }
    for (i = 0; i < N; i++)
        for (k = 0; k < 3; k++) v[3 * i + k] -= p[k] / mtot;
    return N;
}

/* Wrap into [0, L). */
static double wrap(double x, double L) {
    x = fmod(x, L);
    if (x < 0) x += L;
    if (x >= L) x -= L;
    return x;
}

int md_gen_water(int nmol, const double *box, int nlx, int nly, int nlz, double roh, double theta, double jitter,
                 uint64_t seed, double *x) {
    uint64_t s = seed;
    int m, k;
    double hx, hy, hz;
    if (x == NULL || box == NULL || (long)nlx * nly * nlz < nmol) return -1;
    hx = box[0] / nlx;
    hy = box[1] / nly;
    hz = box[2] / nlz;
    for (m = 0; m < nmol; m++) {
        int i = m / (nly * nlz), j = (m / nlz) % nly, l = m % nlz;
        double o[3], q[4], R[9], h1[3], h2[3], nrm = 0.0;
        o[0] = (i + 0.5) * hx + jitter * (md_uniform(&s) - 0.5);
        o[1] = (j + 0.5) * hy + jitter * (md_uniform(&s) - 0.5);
        o[2] = (l + 0.5) * hz + jitter * (md_uniform(&s) - 0.5);
        /* Uniform random rotation from a normalised 4-d Gaussian. */
        for (k = 0; k < 4; k++) {
            q[k] = gauss(&s);
            nrm += q[k] * q[k];
        }
        nrm = sqrt(nrm);
        for (k = 0; k < 4; k++) q[k] /= nrm;
        R[0] = 1 - 2 * (q[2] * q[2] + q[3] * q[3]);
        R[1] = 2 * (q[1] * q[2] - q[0] * q[3]);
        R[2] = 2 * (q[1] * q[3] + q[0] * q[2]);
        R[3] = 2 * (q[1] * q[2] + q[0] * q[3]);
        R[4] = 1 - 2 * (q[1] * q[1] + q[3] * q[3]);
        R[5] = 2 * (q[2] * q[3] - q[0] * q[1]);
        R[6] = 2 * (q[1] * q[3] - q[0] * q[2]);
        R[7] = 2 * (q[2] * q[3] + q[0] * q[1]);
        R[8] = 1 - 2 * (q[1] * q[1] + q[2] * q[2]);
        /* Body frame: H at (+-sin(theta/2), cos(theta/2), 0) * roh. */
        double b1[3] = {roh * sin(0.5 * theta), roh * cos(0.5 * theta), 0.0};
        double b2[3] = {-roh * sin(0.5 * theta), roh * cos(0.5 * theta), 0.0};
        for (k = 0; k < 3; k++) {
            h1[k] = o[k] + R[3 * k] * b1[0] + R[3 * k + 1] * b1[1] + R[3 * k + 2] * b1[2];
            h2[k] = o[k] + R[3 * k] * b2[0] + R[3 * k + 1] * b2[1] + R[3 * k + 2] * b2[2];
        }
        for (k = 0; k < 3; k++) {
            x[9 * m + k] = wrap(o[k], box[k]);
            x[9 * m + 3 + k] = wrap(h1[k], box[k]);
            x[9 * m + 6 + k] = wrap(h2[k], box[k]);
        }
    }
    return 3 * nmol;
}

/* Random sequential addition: uniform trial points, rejected if closer than
   rmin (minimum image) to an accepted point; a cell grid of width >= rmin
   keeps each test to 27 cells. */
int md_gen_rsa(int N, const double *box, double rmin, uint64_t seed, long long max_trials, double *x) {
    uint64_t s = seed;
    int g[3], k, n = 0;
    double hg[3], rmin2 = rmin * rmin;
    int *head, *next;
    long long trials = 0;
    if (x == NULL || box == NULL || N < 1) return -1;
    for (k = 0; k < 3; k++) {
        g[k] = (int)floor(box[k] / rmin);
        if (g[k] < 3) g[k] = 3;
        hg[k] = box[k] / g[k];
    }
    head = (int *)malloc(sizeof(int) * (size_t)g[0] * g[1] * g[2]);
    next = (int *)malloc(sizeof(int) * (size_t)N);
    if (head == NULL || next == NULL) {
        free(head);
        free(next);
        return -2;
    }
    for (k = 0; k < g[0] * g[1] * g[2]; k++) head[k] = -1;
    while (n < N && trials < max_trials) {
        double p[3];
        int c[3], ok = 1, a, b, d;
        trials++;
        for (k = 0; k < 3; k++) {
            p[k] = md_uniform(&s) * box[k];
            c[k] = (int)(p[k] / hg[k]);
            if (c[k] >= g[k]) c[k] = g[k] - 1;
        }
        for (a = -1; a <= 1 && ok; a++)
            for (b = -1; b <= 1 && ok; b++)
                for (d = -1; d <= 1 && ok; d++) {
                    int cc = (((c[0] + a + g[0]) % g[0]) * g[1] + (c[1] + b + g[1]) % g[1]) * g[2] +
                             (c[2] + d + g[2]) % g[2];
                    for (int q = head[cc]; q >= 0; q = next[q]) {
                        double r2 = 0.0;
                        for (k = 0; k < 3; k++) {
                            double dx = p[k] - x[3 * q + k];
                            if (dx > 0.5 * box[k]) dx -= box[k];
                            if (dx < -0.5 * box[k]) dx += box[k];
                            r2 += dx * dx;
                        }
                        if (r2 < rmin2) {
                            ok = 0;
                            break;
                        }
                    }
                }
        if (!ok) continue;
        for (k = 0; k < 3; k++) x[3 * n + k] = p[k];
        {
            int cc = (c[0] * g[1] + c[1]) * g[2] + c[2];
            next[n] = head[cc];
            head[cc] = n;
        }
        n++;
    }
    free(head);
    free(next);
    return n;
}

int md_gen_types(int N, int ntypes, int first, uint64_t seed, int *type) {
    uint64_t s = seed;
    int i;
    if (type == NULL || ntypes < 1) return -1;
    for (i = 0; i < N; i++) {
        int t = (int)(md_uniform(&s) * ntypes);
        if (t >= ntypes) t = ntypes - 1;
        type[i] = first + t;
    }
    return N;
}

int md_gen_uniform(int n, double lo, double hi, uint64_t seed, double *out) {
    uint64_t s = seed;
    int i;
    if (out == NULL) return -1;
    for (i = 0; i < n; i++) out[i] = lo + (hi - lo) * md_uniform(&s);
    return n;
}

/* In-cutoff pairs of a periodic configuration: atoms i are visited in a
   seeded random order and every j != i within the cutoff (minimum image) is
   emitted as (r, type[i], type[j]) until nmax pairs are written.  The
   distances therefore follow the configuration's own pair distribution.
   Returns the number written, or -1 on bad input or allocation failure. */
long long md_sample_pairs(const double *x, const int *type, int N, const double *box, double cutoff,
                          long long nmax, uint64_t seed, double *r, int *ti, int *tj) {
    int nc[3], k, ncell;
    int *head, *next, *order;
    long long w = 0;
    uint64_t s = seed;
    if (x == NULL || type == NULL || box == NULL || r == NULL || ti == NULL || tj == NULL || N < 2 || !(cutoff > 0))
        return -1;
    for (k = 0; k < 3; k++) {
        nc[k] = (int)floor(box[k] / cutoff);
        if (nc[k] < 3) return -1;
    }
    ncell = nc[0] * nc[1] * nc[2];
    head = (int *)malloc(sizeof(int) * ncell);
    next = (int *)malloc(sizeof(int) * N);
    order = (int *)malloc(sizeof(int) * N);
    if (head == NULL || next == NULL || order == NULL) {
        free(head);
        free(next);
        free(order);
        return -1;
    }
    for (k = 0; k < ncell; k++) head[k] = -1;
    for (int i = 0; i < N; i++) {
        int c[3];
        for (k = 0; k < 3; k++) {
            double u = x[3 * i + k] - floor(x[3 * i + k] / box[k]) * box[k];
            c[k] = (int)(u / box[k] * nc[k]);
            if (c[k] >= nc[k]) c[k] = nc[k] - 1;
        }
        const int cid = (c[0] * nc[1] + c[1]) * nc[2] + c[2];
        next[i] = head[cid];
        head[cid] = i;
        order[i] = i;
    }
    for (int i = N - 1; i > 0; i--) {
        const int j = (int)(sm_next(&s) % (uint64_t)(i + 1));
        const int t = order[i];
        order[i] = order[j];
        order[j] = t;
    }
    for (int q = 0; q < N && w < nmax; q++) {
        const int i = order[q];
        int c[3];
        for (k = 0; k < 3; k++) {
            double u = x[3 * i + k] - floor(x[3 * i + k] / box[k]) * box[k];
            c[k] = (int)(u / box[k] * nc[k]);
            if (c[k] >= nc[k]) c[k] = nc[k] - 1;
        }
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    const int cx = (c[0] + dx + nc[0]) % nc[0], cy = (c[1] + dy + nc[1]) % nc[1],
                              cz = (c[2] + dz + nc[2]) % nc[2];
                    for (int j = head[(cx * nc[1] + cy) * nc[2] + cz]; j >= 0 && w < nmax; j = next[j]) {
                        double d2 = 0.0;
                        if (j == i) continue;
                        for (k = 0; k < 3; k++) {
                            double d = x[3 * i + k] - x[3 * j + k];
                            d -= box[k] * nearbyint(d / box[k]);
                            d2 += d * d;
                        }
                        if (d2 > cutoff * cutoff) continue;
                        r[w] = sqrt(d2);
                        ti[w] = type[i];
                        tj[w] = type[j];
                        w++;
                    }
                }
    }
    free(head);
    free(next);
    free(order);
    return w;
}
