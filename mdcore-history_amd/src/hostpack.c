/* Host packing between cells and the device's cell-ordered arrays; see
   hostpack.h.  Each loop runs over cells in parallel; every cell writes its
   own slot range, so the results do not depend on the thread count. */
#include "hostpack.h"

int space_pack(const struct space *s, int max_type, int *count, int *offset, int *cnt_off, double *x, int *type,
               int *id, int *maxc, int *maxid) {
    const int nc = s->nr_cells;
    int n = 0, mc = 0, mid = -1, bad = 0;
    for (int cid = 0; cid < nc; cid++) {
        const int c = s->cells[cid].count;
        count[cid] = c;
        offset[cid] = n;
        cnt_off[2 * cid] = c;
        cnt_off[2 * cid + 1] = n;
        if (c > mc) mc = c;
        n += c;
    }
    offset[nc] = n;
#pragma omp parallel for schedule(static) reduction(max : mid) reduction(| : bad)
    for (int cid = 0; cid < nc; cid++) {
        const struct cell *c = &s->cells[cid];
        for (int k = 0; k < c->count; k++) {
            const int a = offset[cid] + k;
            const struct part *p = &c->parts[k];
            if (p->type < 0 || p->type >= max_type) bad = 1;
            x[3 * a] = p->x[0];
            x[3 * a + 1] = p->x[1];
            x[3 * a + 2] = p->x[2];
            type[a] = p->type;
            id[a] = p->id;
            if (p->id > mid) mid = p->id;
        }
    }
    *maxc = mc;
    *maxid = mid;
    return bad ? -1 : n;
}

void space_pack_v(const struct space *s, const int *offset, double *v) {
#pragma omp parallel for schedule(static)
    for (int cid = 0; cid < s->nr_cells; cid++) {
        const struct cell *c = &s->cells[cid];
        for (int k = 0; k < c->count; k++)
            for (int m = 0; m < 3; m++) v[3 * (offset[cid] + k) + m] = c->parts[k].v[m];
    }
}

int space_add_forces(struct space *s, const int *ioi, int n, const float *fhi, const float *flo) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int cid = 0; cid < s->nr_cells; cid++) {
        struct cell *c = &s->cells[cid];
        for (int k = 0; k < c->count; k++) {
            const int a = ioi[c->parts[k].id];
            if (a < 0 || a >= n) {
                bad = 1;
                continue;
            }
            for (int m = 0; m < 3; m++)
                c->parts[k].f[m] += (double)fhi[4 * (size_t)a + m] + (double)flo[4 * (size_t)a + m];
        }
    }
    return bad ? -1 : 0;
}

void space_add_forces_slots(struct space *s, const int *offset, const float *fhi, const float *flo) {
#pragma omp parallel for schedule(static)
    for (int cid = 0; cid < s->nr_cells; cid++) {
        struct cell *c = &s->cells[cid];
        for (int k = 0; k < c->count; k++) {
            const size_t a = (size_t)(offset[cid] + k);
            for (int m = 0; m < 3; m++) c->parts[k].f[m] += (double)fhi[4 * a + m] + (double)flo[4 * a + m];
        }
    }
}
