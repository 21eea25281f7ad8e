/* Host packing between cells and cell-ordered arrays (space_setcells in
   space.c, space_pack / space_add_forces in hostpack.c), run with OpenMP:
   20,000 particles are given a new layout (random cell counts, ids in a
   random order); every field, the id -> particle / cell maps, the packed
   arrays and the force write-back are checked.  Prints "ok" on success. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "hostpack.h"

#define N 20000
#define FAIL(...) do { printf(__VA_ARGS__); printf("\n"); return 1; } while (0)

static unsigned long long rng = 88172645463325252ull;
static unsigned int rnd(void) {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (unsigned int)(rng >> 11);
}

int main(void) {
    static struct space s;
    double origin[3] = {0, 0, 0}, dim[3] = {6.0, 6.0, 6.0}, L[3] = {1.0, 1.0, 1.0};
    if (space_init(&s, origin, dim, L, 1.0, space_periodic_full) < 0) FAIL("space_init");
    for (int i = 0; i < N; i++) {
        struct part p;
        memset(&p, 0, sizeof(p));
        double x[3];
        for (int k = 0; k < 3; k++) x[k] = dim[k] * (rnd() % 1000000) / 1000000.0;
        p.id = i;
        p.vid = 7 * i + 1;
        p.q = 0.001f * (float)i;
        p.flags = (unsigned short)(i % 3);
        p.type = (short)(1 + i % 4);
        if (space_addpart(&s, &p, x) < 0) FAIL("space_addpart %d", i);
    }
    const int nc = s.nr_cells;
    /* new layout: random counts, a random permutation of the ids */
    int *count = calloc(nc, sizeof(int)), *id = malloc(sizeof(int) * N), *type = malloc(sizeof(int) * N);
    double *x = malloc(sizeof(double) * 3 * N), *v = malloc(sizeof(double) * 3 * N);
    for (int i = 0; i < N; i++) count[rnd() % nc]++;
    for (int i = 0; i < N; i++) id[i] = i;
    for (int i = N - 1; i > 0; i--) {
        int j = rnd() % (i + 1), t = id[i];
        id[i] = id[j];
        id[j] = t;
    }
    for (int a = 0; a < N; a++) {
        type[a] = 1 + (id[a] + 1) % 4;
        for (int k = 0; k < 3; k++) {
            x[3 * a + k] = 0.001 * a + k;
            v[3 * a + k] = -0.002 * a + k;
        }
    }
    if (space_setcells(&s, count, x, v, id, type) < 0) FAIL("space_setcells");
    for (int cid = 0, a = 0; cid < nc; cid++) {
        struct cell *c = &s.cells[cid];
        if (c->count != count[cid]) FAIL("count of cell %d", cid);
        for (int k = 0; k < c->count; k++, a++) {
            struct part *p = &c->parts[k];
            const int i = id[a];
            if (p->id != i || p->type != type[a]) FAIL("id/type at slot %d", a);
            if (p->vid != 7 * i + 1 || p->q != 0.001f * (float)i || p->flags != i % 3) FAIL("q/vid/flags of id %d", i);
            for (int m = 0; m < 3; m++)
                if (p->x[m] != x[3 * a + m] || p->v[m] != v[3 * a + m] || p->f[m] != 0.0) FAIL("x/v/f at slot %d", a);
            if (s.partlist[i] != p || s.celllist[i] != c) FAIL("partlist/celllist of id %d", i);
        }
    }
    /* pack back */
    int *cnt2 = malloc(sizeof(int) * nc), *off = malloc(sizeof(int) * (nc + 1)), *co = malloc(sizeof(int) * 2 * nc);
    int *id2 = malloc(sizeof(int) * N), *type2 = malloc(sizeof(int) * N), maxc = 0, maxid = 0;
    double *x2 = malloc(sizeof(double) * 3 * N);
    if (space_pack(&s, 5, cnt2, off, co, x2, type2, id2, &maxc, &maxid) != N) FAIL("space_pack");
    int mc = 0;
    for (int cid = 0; cid < nc; cid++) {
        if (cnt2[cid] != count[cid] || co[2 * cid] != count[cid] || co[2 * cid + 1] != off[cid]) FAIL("pack counts");
        if (count[cid] > mc) mc = count[cid];
    }
    if (off[nc] != N || maxc != mc || maxid != N - 1) FAIL("pack totals");
    for (int a = 0; a < N; a++) {
        if (id2[a] != id[a] || type2[a] != type[a]) FAIL("pack id/type at %d", a);
        for (int m = 0; m < 3; m++)
            if (x2[3 * a + m] != x[3 * a + m]) FAIL("pack x at %d", a);
    }
    if (space_pack(&s, 2, cnt2, off, co, x2, type2, id2, &maxc, &maxid) >= 0) FAIL("type range not checked");
    /* forces by slot of another layout: ioi[id] = slot */
    int *ioi = malloc(sizeof(int) * N);
    float *fhi = malloc(sizeof(float) * 4 * N), *flo = malloc(sizeof(float) * 4 * N);
    for (int i = 0; i < N; i++) ioi[i] = (i * 7919) % N;
    for (int a = 0; a < N; a++)
        for (int m = 0; m < 4; m++) {
            fhi[4 * a + m] = (float)(a + m);
            flo[4 * a + m] = 1e-3f * (float)m;
        }
    if (space_add_forces(&s, ioi, N, fhi, flo) < 0) FAIL("space_add_forces");
    for (int i = 0; i < N; i++) {
        const struct part *p = s.partlist[i];
        const int a = ioi[i];
        for (int m = 0; m < 3; m++)
            if (p->f[m] != (double)fhi[4 * a + m] + (double)flo[4 * a + m]) FAIL("force of id %d", i);
    }
    ioi[5] = N;
    if (space_add_forces(&s, ioi, N, fhi, flo) >= 0) FAIL("missing slot not reported");
    printf("ok\n");
    return 0;
}
