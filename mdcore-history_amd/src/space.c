/*
 * Cell grid, particle storage and the cell-pair task list.
 *
 * Data model of the reference (space.c, cell.c): the box is cut into
 * cdim = floor(dim/L) cells per axis of width h = dim/cdim; each cell stores
 * its particles as an array of struct part with positions relative to the
 * cell origin.  The task list holds one self task per cell and one pair task
 * per pair of cells whose box distance is within the cutoff, each tagged
 * with the sort direction (sid) of the pair; every cell gets a sort task
 * whose direction mask is the OR of its pairs' sids.
 *   space_getsid    space.c:80-114   (minimum-image shift, sign-pattern sid,
 *                                     flip so that the pair uses one of 13 axes)
 *   space_init      space.c:530-783  (grid, span = ceil(rc/h), task loops,
 *                                     id1 < id2 dedup, sort masks)
 *   space_addpart   space.c:382-434
 *   space_prepare   space.c:186-219
 *   space_shuffle_local, cell_welcome  space.c:307-364, cell.c
 * The direction tables below are the reference's (cell.c:56-101): they are
 * part of the sortlist format, so they must be the same values.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mdcore.h"

int space_err = space_err_ok;

static const char *space_err_msg[9] = {
    "Nothing bad happened.",
    "An unexpected NULL pointer was encountered.",
    "A call to malloc failed, probably due to insufficient memory.",
    "An error occured when calling a cell function.",
    "A call to a pthread routine failed.",
    "One or more values were outside of the allowed range.",
    "Too many pairs associated with a single particle in Verlet list.",
    "Task list too short.",
    "An error occured when calling a task function."};

#define serror_(id) (space_err = errs_register((id), space_err_msg[-(id)], __LINE__, __func__, __FILE__))

/* Offset (dx,dy,dz) with components in {-1,0,1}, index 9(dx+1)+3(dy+1)+(dz+1),
   to the sort-list id of its axis. */
const char cell_sortlistID[27] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 0,
                                  12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0};

/* The 13 unit axes (sign patterns normalised), as decimal literals. */
const FPTYPE cell_shift[13 * 3] = {
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

/* Offsets with a negative leading non-zero component are flipped onto their
   mirror image so that only 13 axes are needed. */
const char cell_flip[27] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0,
                            0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

/* ------------------------------------------------------------------ cells */

int cell_init(struct cell *c, int *loc, double *origin, double *dim) {
    int k;
    if (c == NULL || loc == NULL || origin == NULL || dim == NULL) return -1;
    memset(c, 0, sizeof(*c));
    for (k = 0; k < 3; k++) {
        c->loc[k] = loc[k];
        c->origin[k] = origin[k];
        c->dim[k] = dim[k];
    }
    c->size = cell_incr;
    if (posix_memalign((void **)&c->parts, cell_partalign, sizeof(struct part) * c->size) != 0) return -2;
    if ((c->sortlist = (unsigned int *)malloc(sizeof(unsigned int) * 13 * c->size)) == NULL) return -2;
    c->incomming_size = cell_incr;
    if (posix_memalign((void **)&c->incomming, cell_partalign, sizeof(struct part) * c->incomming_size) != 0)
        return -2;
    return 0;
}

static int cell_grow(struct cell *c, int need) {
    struct part *tmp;
    unsigned int *sl;
    int size = c->size + cell_incr;
    if (size < need) size = need;
    if (posix_memalign((void **)&tmp, cell_partalign, sizeof(struct part) * size) != 0) return -1;
    if ((sl = (unsigned int *)malloc(sizeof(unsigned int) * 13 * size)) == NULL) {
        free(tmp);
        return -1;
    }
    memcpy(tmp, c->parts, sizeof(struct part) * c->count);
    free(c->parts);
    free(c->sortlist);
    c->parts = tmp;
    c->sortlist = sl;
    c->size = size;
    return 0;
}

struct part *cell_add(struct cell *c, struct part *p, struct part **partlist) {
    int k;
    if (c == NULL || p == NULL) return NULL;
    if (c->count == c->size) {
        if (cell_grow(c, c->count + 1) < 0) return NULL;
        if (partlist != NULL)
            for (k = 0; k < c->count; k++) partlist[c->parts[k].id] = &c->parts[k];
    }
    c->parts[c->count] = *p;
    return &c->parts[c->count++];
}

struct part *cell_add_incomming(struct cell *c, struct part *p) {
    struct part *tmp;
    if (c == NULL || p == NULL) return NULL;
    if (c->incomming_count == c->incomming_size) {
        int size = c->incomming_size + cell_incr;
        if (posix_memalign((void **)&tmp, cell_partalign, sizeof(struct part) * size) != 0) return NULL;
        memcpy(tmp, c->incomming, sizeof(struct part) * c->incomming_count);
        free(c->incomming);
        c->incomming = tmp;
        c->incomming_size = size;
    }
    c->incomming[c->incomming_count] = *p;
    return &c->incomming[c->incomming_count++];
}

/* Append the incoming particles at the end of the cell, in arrival order. */
int cell_welcome(struct cell *c, struct part **partlist) {
    int k;
    if (c == NULL) return -1;
    if (c->count + c->incomming_count > c->size) {
        if (cell_grow(c, c->count + c->incomming_count) < 0) return -2;
        if (partlist != NULL)
            for (k = 0; k < c->count; k++) partlist[c->parts[k].id] = &c->parts[k];
    }
    memcpy(&c->parts[c->count], c->incomming, sizeof(struct part) * c->incomming_count);
    if (partlist != NULL)
        for (k = c->count; k < c->count + c->incomming_count; k++) partlist[c->parts[k].id] = &c->parts[k];
    c->count += c->incomming_count;
    c->incomming_count = 0;
    return 0;
}

int cell_flush(struct cell *c, struct part **partlist, struct cell **celllist) {
    int k;
    if (c == NULL) return -1;
    for (k = 0; k < c->count; k++) {
        if (partlist) partlist[c->parts[k].id] = NULL;
        if (celllist) celllist[c->parts[k].id] = NULL;
    }
    c->count = 0;
    return 0;
}

/* ------------------------------------------------------------------ space */

int space_getsid(struct space *s, struct cell **ci, struct cell **cj, FPTYPE *shift) {
    int k, sid;
    struct cell *tmp;
    FPTYPE lshift[3];
    if (shift == NULL) shift = lshift;
    for (k = 0; k < 3; k++) {
        shift[k] = (*cj)->origin[k] - (*ci)->origin[k];
        if (shift[k] * 2 > s->dim[k])
            shift[k] -= s->dim[k];
        else if (shift[k] * 2 < -s->dim[k])
            shift[k] += s->dim[k];
    }
    for (sid = 0, k = 0; k < 3; k++) sid = 3 * sid + ((shift[k] < 0) ? 0 : ((shift[k] > 0) ? 2 : 1));
    if (cell_flip[sid]) {
        tmp = *ci;
        *ci = *cj;
        *cj = tmp;
        shift[0] = -shift[0];
        shift[1] = -shift[1];
        shift[2] = -shift[2];
    }
    return cell_sortlistID[sid];
}

static int space_addtask(struct space *s, int type, int flags, int i, int j) {
    struct task *t;
    if (s->nr_tasks >= s->tasks_size) return serror_(space_err_nrtasks);
    t = &s->tasks[s->nr_tasks++];
    t->type = type;
    t->subtype = task_subtype_none;
    t->flags = flags;
    t->i = i;
    t->j = j;
    t->wait = 0;
    return s->nr_tasks - 1;
}

/* Wrap a neighbour coordinate; returns -1 if it leaves a non-periodic box. */
static int wrap_coord(int c, int n, int periodic) {
    if (c < 0) return periodic ? c + n : -1;
    if (c >= n) return periodic ? c - n : -1;
    return c;
}

int space_init(struct space *s, const double *origin, const double *dim, double *L, double cutoff,
               unsigned int period) {
    int i, j, k, l[3], ii, jj, kk, id1, id2, sid, ti;
    double o[3], lh[3];
    struct cell *ci, *cj;

    if (s == NULL || origin == NULL || dim == NULL || L == NULL) return serror_(space_err_null);
    memset(s, 0, sizeof(*s));
    for (i = 0; i < 3; i++) {
        s->origin[i] = origin[i];
        s->dim[i] = dim[i];
        s->cdim[i] = floor(dim[i] / L[i]);
        if (s->cdim[i] < 1) return serror_(space_err_range);
    }
    s->cutoff = cutoff;
    s->cutoff2 = cutoff * cutoff;
    s->period = period;
    s->nr_cells = s->cdim[0] * s->cdim[1] * s->cdim[2];
    if ((s->cells = (struct cell *)calloc(s->nr_cells, sizeof(struct cell))) == NULL)
        return serror_(space_err_malloc);
    for (i = 0; i < 3; i++) {
        s->h[i] = s->dim[i] / s->cdim[i];
        s->ih[i] = 1.0 / s->h[i];
    }
    for (l[0] = 0; l[0] < s->cdim[0]; l[0]++) {
        o[0] = origin[0] + l[0] * s->h[0];
        for (l[1] = 0; l[1] < s->cdim[1]; l[1]++) {
            o[1] = origin[1] + l[1] * s->h[1];
            for (l[2] = 0; l[2] < s->cdim[2]; l[2]++) {
                o[2] = origin[2] + l[2] * s->h[2];
                if (cell_init(&s->cells[space_cellid(s, l[0], l[1], l[2])], l, o, s->h) < 0)
                    return serror_(space_err_cell);
            }
        }
    }
    if (period & space_periodic_ghost_full) return serror_(space_err_range); /* single domain only */

    if ((s->cid_real = (int *)malloc(sizeof(int) * s->nr_cells)) == NULL ||
        (s->cid_ghost = (int *)malloc(sizeof(int) * s->nr_cells)) == NULL ||
        (s->cid_marked = (int *)malloc(sizeof(int) * s->nr_cells)) == NULL)
        return serror_(space_err_malloc);
    for (k = 0; k < s->nr_cells; k++) {
        s->cells[k].flags |= cell_flag_marked;
        s->cid_marked[s->nr_marked++] = k;
        s->cells[k].id = s->nr_real;
        s->cid_real[s->nr_real++] = k;
    }

    for (k = 0; k < 3; k++) s->span[k] = ceil(cutoff * s->ih[k]);

    s->tasks_size = s->nr_cells * ((2 * s->span[0] + 1) * (2 * s->span[1] + 1) * (2 * s->span[2] + 1) + 2);
    if ((s->tasks = (struct task *)malloc(sizeof(struct task) * s->tasks_size)) == NULL)
        return serror_(space_err_malloc);

    /* Self and pair tasks, in the reference's loop order. */
    for (i = 0; i < s->cdim[0]; i++)
        for (j = 0; j < s->cdim[1]; j++)
            for (k = 0; k < s->cdim[2]; k++) {
                id1 = space_cellid(s, i, j, k);
                for (l[0] = -s->span[0]; l[0] <= s->span[0]; l[0]++) {
                    if ((ii = wrap_coord(i + l[0], s->cdim[0], period & space_periodic_x)) < 0) continue;
                    for (l[1] = -s->span[1]; l[1] <= s->span[1]; l[1]++) {
                        if ((jj = wrap_coord(j + l[1], s->cdim[1], period & space_periodic_y)) < 0) continue;
                        for (l[2] = -s->span[2]; l[2] <= s->span[2]; l[2]++) {
                            /* Closest distance between the two cells' boxes. */
                            lh[0] = s->h[0] * fmax(abs(l[0]) - 1, 0);
                            lh[1] = s->h[1] * fmax(abs(l[1]) - 1, 0);
                            lh[2] = s->h[2] * fmax(abs(l[2]) - 1, 0);
                            if (lh[0] * lh[0] + lh[1] * lh[1] + lh[2] * lh[2] > s->cutoff2) continue;
                            if ((kk = wrap_coord(k + l[2], s->cdim[2], period & space_periodic_z)) < 0) continue;
                            id2 = space_cellid(s, ii, jj, kk);
                            ci = &s->cells[id1];
                            cj = &s->cells[id2];
                            sid = space_getsid(s, &ci, &cj, NULL);
                            if (id1 < id2 || (id1 == id2 && l[0] == 0 && l[1] == 0 && l[2] == 0)) {
                                if (space_addtask(s, (id1 == id2) ? task_type_self : task_type_pair, sid,
                                                  (int)(ci - s->cells), (int)(cj - s->cells)) < 0)
                                    return serror_(space_err);
                            }
                        }
                    }
                }
            }

    /* One sort task per cell; its mask collects the sids of its pairs. */
    for (k = 0; k < s->nr_cells; k++) s->cells[k].sortflags = 0;
    for (ti = 0; ti < s->nr_tasks; ti++)
        if (s->tasks[ti].type == task_type_pair) {
            s->cells[s->tasks[ti].i].sortflags |= 1u << s->tasks[ti].flags;
            s->cells[s->tasks[ti].j].sortflags |= 1u << s->tasks[ti].flags;
        }
    for (k = 0; k < s->nr_cells; k++)
        if (space_addtask(s, task_type_sort, (int)s->cells[k].sortflags, k, -1) < 0) return serror_(space_err);

    if ((s->partlist = (struct part **)malloc(sizeof(struct part *) * space_partlist_incr)) == NULL ||
        (s->celllist = (struct cell **)malloc(sizeof(struct cell *) * space_partlist_incr)) == NULL)
        return serror_(space_err_malloc);
    s->size_parts = space_partlist_incr;
    s->nr_parts = 0;
    s->verlet_rebuild = 1;
    s->maxdx = 0.0;
    return space_err_ok;
}

int space_addpart(struct space *s, struct part *p, double *x) {
    int k, ind[3];
    struct cell *c;
    if (s == NULL || p == NULL || x == NULL) return serror_(space_err_null);
    if (p->id < 0) return serror_(space_err_range);
    /* The partlist is indexed by particle id. */
    if (p->id >= s->size_parts) {
        int size = s->size_parts;
        struct part **pl;
        struct cell **cl;
        while (size <= p->id) size += (size > space_partlist_incr ? size / 2 : space_partlist_incr);
        if ((pl = (struct part **)calloc(size, sizeof(struct part *))) == NULL ||
            (cl = (struct cell **)calloc(size, sizeof(struct cell *))) == NULL)
            return serror_(space_err_malloc);
        memcpy(pl, s->partlist, sizeof(struct part *) * s->size_parts);
        memcpy(cl, s->celllist, sizeof(struct cell *) * s->size_parts);
        free(s->partlist);
        free(s->celllist);
        s->partlist = pl;
        s->celllist = cl;
        s->size_parts = size;
    }
    for (k = 0; k < 3; k++) ind[k] = (x[k] - s->origin[k]) * s->ih[k];
    for (k = 0; k < 3; k++)
        if (ind[k] < 0 || ind[k] >= s->cdim[k]) return serror_(space_err_range);
    c = &s->cells[space_cellid(s, ind[0], ind[1], ind[2])];
    for (k = 0; k < 3; k++) p->x[k] = x[k] - c->origin[k];
    if ((s->partlist[p->id] = cell_add(c, p, s->partlist)) == NULL) return serror_(space_err_cell);
    s->celllist[p->id] = c;
    s->nr_parts++;
    return space_err_ok;
}

int space_getpos(struct space *s, int id, double *x) {
    int k;
    if (s == NULL || x == NULL) return serror_(space_err_null);
    if (id < 0 || id >= s->size_parts || s->partlist[id] == NULL) return serror_(space_err_range);
    for (k = 0; k < 3; k++) x[k] = s->partlist[id]->x[k] + s->celllist[id]->origin[k];
    return space_err_ok;
}

int space_prepare(struct space *s) {
    int cid, pid, k;
    s->epot = 0.0;
    s->epot_nonbond = 0.0;
    s->epot_bond = 0.0;
    s->epot_angle = 0.0;
    s->epot_dihedral = 0.0;
    s->epot_exclusion = 0.0;
    s->virial = 0.0;
    s->nr_pairs = 0;
    s->nr_excl_evaluated = 0;
#pragma omp parallel for schedule(static) private(pid, k)
    for (cid = 0; cid < s->nr_cells; cid++) {
        struct cell *c = &s->cells[cid];
        c->epot = 0.0;
        for (pid = 0; pid < c->count; pid++)
            for (k = 0; k < 3; k++) c->parts[pid].f[k] = 0.0;
    }
    return space_err_ok;
}

/* Move particles that left their cell (x outside [0,h)) to the neighbour
   cell, filling the hole with the cell's last particle; newcomers are
   appended in cell-id order of their origin, then particle order.  The
   cells that hold a leaver are found in parallel; only they go through the
   serial move, in cell order, so the result is that of one serial pass over
   all cells.  The newcomers are then welcomed cell by cell in parallel. */
int space_shuffle_local(struct space *s) {
    int k, cid, pid, delta[3], bad = 0;
    struct cell *c, *dest;
    struct part *p;
    unsigned char *leaves = (unsigned char *)malloc(s->nr_real > 0 ? (size_t)s->nr_real : 1);
    if (leaves == NULL) return serror_(space_err_malloc);
#pragma omp parallel for schedule(static) private(pid, k)
    for (cid = 0; cid < s->nr_real; cid++) {
        const struct cell *cc = &s->cells[s->cid_real[cid]];
        unsigned char out = 0;
        for (pid = 0; pid < cc->count && !out; pid++)
            for (k = 0; k < 3; k++) out |= (cc->parts[pid].x[k] >= s->h[k]) | (cc->parts[pid].x[k] < 0.0);
        leaves[cid] = out;
    }
    for (cid = 0; cid < s->nr_real; cid++) {
        if (!leaves[cid]) continue;
        c = &s->cells[s->cid_real[cid]];
        pid = 0;
        while (pid < c->count) {
            p = &c->parts[pid];
            for (k = 0; k < 3; k++) delta[k] = (p->x[k] >= s->h[k]) - (p->x[k] < 0.0);
            if (delta[0] || delta[1] || delta[2]) {
                for (k = 0; k < 3; k++) p->x[k] -= delta[k] * s->h[k];
                dest = &s->cells[space_cellid(s, (c->loc[0] + delta[0] + s->cdim[0]) % s->cdim[0],
                                              (c->loc[1] + delta[1] + s->cdim[1]) % s->cdim[1],
                                              (c->loc[2] + delta[2] + s->cdim[2]) % s->cdim[2])];
                if (cell_add_incomming(dest, p) == NULL) {
                    free(leaves);
                    return serror_(space_err_cell);
                }
                s->celllist[p->id] = dest;
                c->count -= 1;
                if (pid < c->count) {
                    c->parts[pid] = c->parts[c->count];
                    s->partlist[c->parts[pid].id] = &c->parts[pid];
                }
            } else
                pid += 1;
        }
    }
    free(leaves);
    /* each cell appends its own newcomers and rewrites its own particles'
       partlist entries */
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (cid = 0; cid < s->nr_marked; cid++)
        bad |= cell_welcome(&s->cells[s->cid_marked[cid]], s->partlist) < 0;
    if (bad) return serror_(space_err_cell);
    return space_err_ok;
}

int space_flush(struct space *s) {
    int cid;
    if (s == NULL) return serror_(space_err_null);
    for (cid = 0; cid < s->nr_cells; cid++) cell_flush(&s->cells[cid], s->partlist, s->celllist);
    s->nr_parts = 0;
    return space_err_ok;
}

int space_setcells(struct space *s, const int *count, const double *x, const double *v, const int *id,
                   const int *type) {
    /* The fields the layout does not carry (charge, vid, flags), gathered
       by id into a scratch array kept between calls; then every cell is
       rewritten in slot order, forces zeroed.  Cells run in parallel: each
       writes its own parts and its own particles' partlist/celllist
       entries. */
    struct meta {
        float q;
        int vid;
        unsigned short flags;
    } *meta;
    int cid, a, n = 0, bad = 0;
    int *start;
    if (s == NULL || count == NULL || x == NULL || v == NULL || id == NULL || type == NULL)
        return serror_(space_err_null);
    for (cid = 0; cid < s->nr_cells; cid++) {
        if (count[cid] < 0) return serror_(space_err_range);
        n += count[cid];
    }
    if (n != s->nr_parts) return serror_(space_err_range);
    const size_t need = (sizeof(struct meta) + sizeof(int)) * (size_t)(n > 0 ? n : 1) + sizeof(int) * (s->nr_cells + 1);
    if (s->setcells_scratch_size < need) {
        free(s->setcells_scratch);
        s->setcells_scratch_size = 0;
        if ((s->setcells_scratch = malloc(need)) == NULL) return serror_(space_err_malloc);
        s->setcells_scratch_size = need;
    }
    meta = (struct meta *)s->setcells_scratch;
    start = (int *)(meta + (n > 0 ? n : 1));
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (a = 0; a < n; a++) {
        if (id[a] < 0 || id[a] >= s->size_parts || s->partlist[id[a]] == NULL) {
            bad = 1;
            continue;
        }
        const struct part *p = s->partlist[id[a]];
        meta[a].q = p->q;
        meta[a].vid = p->vid;
        meta[a].flags = p->flags;
    }
    if (bad) return serror_(space_err_range);
    for (a = 0, cid = 0; cid < s->nr_cells; cid++) {
        struct cell *c = &s->cells[cid];
        start[cid] = a;
        a += count[cid];
        if (count[cid] > c->size && cell_grow(c, count[cid]) < 0) return serror_(space_err_malloc);
    }
#pragma omp parallel for schedule(static)
    for (cid = 0; cid < s->nr_cells; cid++) {
        struct cell *c = &s->cells[cid];
        for (int k = 0; k < count[cid]; k++) {
            const int b = start[cid] + k;
            struct part *p = &c->parts[k];
            for (int m = 0; m < 3; m++) {
                p->x[m] = x[3 * b + m];
                p->v[m] = v[3 * b + m];
                p->f[m] = 0;
            }
            p->x[3] = p->v[3] = p->f[3] = 0;
            p->q = meta[b].q;
            p->id = id[b];
            p->vid = meta[b].vid;
            p->type = (short)type[b];
            p->flags = meta[b].flags;
            s->partlist[p->id] = p;
            s->celllist[p->id] = c;
        }
        c->count = count[cid];
    }
    return space_err_ok;
}

void space_free(struct space *s) {
    int cid;
    if (s == NULL) return;
    for (cid = 0; cid < s->nr_cells; cid++) {
        free(s->cells[cid].parts);
        free(s->cells[cid].sortlist);
        free(s->cells[cid].incomming);
        free(s->cells[cid].oldx);
    }
    free(s->cells);
    free(s->cid_real);
    free(s->cid_ghost);
    free(s->cid_marked);
    free(s->tasks);
    free(s->partlist);
    free(s->celllist);
    free(s->setcells_scratch);
    memset(s, 0, sizeof(*s));
}
