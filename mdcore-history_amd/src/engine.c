/*
 * Engine: the public entry points of the nonbonded path.
 *
 * Same calling contract as the reference's engine.c (engine_init 2439-2591,
 * engine_addtype 1543-1568, engine_addpot 1584-1600, engine_load 1345-1392,
 * engine_unload 1076-1145, engine_start 1674-1826, engine_step 2070-2240,
 * engine_advance 1945-2052, engine_shuffle 117-156) and engine_bonded.c
 * (engine_exclusion_add 1400-1434, engine_exclusion_shrink 1301-1387).
 * The nonbonded pass runs only on the GPU: engine_start requires
 * engine_flag_gpu (bit 4, the reference's engine_flag_cuda) and engine_step
 * calls engine_nonbond_gpu; there is no CPU runner in this library.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "engine_gpu.h"
#include "mdcore.h"

int engine_err = engine_err_ok;

const char *engine_err_msg[31] = {
    "Nothing bad happened.",
    "An unexpected NULL pointer was encountered.",
    "A call to malloc failed, probably due to insufficient memory.",
    "An error occured when calling a space function.",
    "A call to a pthread routine failed.",
    "An error occured when calling a runner function.",
    "One or more values were outside of the allowed range.",
    "An error occured while calling a cell function.",
    "The computational domain is too small for the requested operation.",
    "mdcore was not compiled with MPI.",
    "An error occured while calling an MPI function.",
    "An error occured when calling a bond function.",
    "An error occured when calling an angle function.",
    "An error occured when calling a reader function.",
    "An error occured while interpreting the PSF file.",
    "An error occured while interpreting the PDB file.",
    "An error occured while interpreting the CPF file.",
    "An error occured when calling a potential function.",
    "An error occured when calling an exclusion function.",
    "An error occured while computing the bonded sets.",
    "An error occured when calling a dihedral funtion.",
    "A HIP runtime call failed.",
    "No GPU support: set engine_flag_gpu and build the HIP library.",
    "Double-precision device arithmetic is not supported.",
    "Max. number of parts per cell exceeded.",
    "An error occured when calling a queue funtion.",
    "An error occured when evaluating a rigid constraint.",
    "Cell cutoff size doesn't work with METIS",
    "METIS library undefined",
    "Not all task dependencies can be satified.",
    "Not yet implemented."};

#define eerror_(id) (engine_err = errs_register((id), engine_err_msg[-(id)], __LINE__, __func__, __FILE__))

static double wall_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

unsigned long engine_sizeof(void) { return sizeof(struct engine); }

int engine_setflags(struct engine *e, unsigned int set, unsigned int clear) {
    if (e == NULL) return eerror_(engine_err_null);
    if (((set | clear) & engine_flag_gpu) && e->gpu != NULL) return eerror_(engine_err_range);
    e->flags = (e->flags | set) & ~clear;
    return engine_err_ok;
}

int engine_setdt(struct engine *e, double dt) {
    if (e == NULL) return eerror_(engine_err_null);
    if (!(dt >= 0.0)) return eerror_(engine_err_range);
    e->dt = dt;
    return engine_err_ok;
}

struct engine *engine_alloc(void) {
    struct engine *e;
    if (posix_memalign((void **)&e, 64, sizeof(struct engine)) != 0) return NULL;
    memset(e, 0, sizeof(*e));
    return e;
}

void engine_free(struct engine *e) {
    if (e == NULL) return;
    engine_finalize(e);
    free(e);
}

int engine_init(struct engine *e, const double *origin, const double *dim, double *L, double cutoff,
                unsigned int period, int max_type, unsigned int flags) {
    if (e == NULL || origin == NULL || dim == NULL || L == NULL) return eerror_(engine_err_null);
    if (max_type < 1) return eerror_(engine_err_range);
    memset(e, 0, sizeof(*e));
    if (space_init(&e->s, origin, dim, L, cutoff, period) < 0) return eerror_(engine_err_space);
    /* As in the reference, the GPU flag implies a reserved null type 0. */
    if (flags & engine_flag_verlet_pseudo) flags |= engine_flag_verlet_pairwise;
    if (flags & engine_flag_verlet_pairwise) flags |= engine_flag_verlet;
    if (flags & engine_flag_gpu) flags |= engine_flag_nullpart;
    e->flags = flags;
    e->time = 0;
    e->dt = 0.0;
    e->tol_rigid = 1e-6;
    e->exclusions_size = 100;
    if ((e->exclusions = (struct exclusion *)malloc(sizeof(struct exclusion) * e->exclusions_size)) == NULL)
        return eerror_(engine_err_malloc);
    e->nr_exclusions = 0;
    if (flags & engine_flag_nullpart) max_type += 1;
    e->max_type = max_type;
    e->nr_types = 0;
    if ((e->types = (struct part_type *)calloc(max_type, sizeof(struct part_type))) == NULL)
        return eerror_(engine_err_malloc);
    if (flags & engine_flag_nullpart) {
        e->types[0].id = 0;
        strcpy(e->types[0].name, "NULL");
        strcpy(e->types[0].name2, "NULL");
        e->nr_types = 1;
    }
    if ((e->p = (struct potential **)calloc((size_t)max_type * max_type, sizeof(struct potential *))) == NULL)
        return eerror_(engine_err_malloc);
    e->gpu = NULL;
    return engine_err_ok;
}

int engine_addtype(struct engine *e, double mass, double charge, char *name, char *name2) {
    struct part_type *t;
    if (e == NULL) return eerror_(engine_err_null);
    if (e->nr_types >= e->max_type) return eerror_(engine_err_range);
    t = &e->types[e->nr_types];
    t->id = e->nr_types;
    t->mass = mass;
    t->imass = 1.0 / mass;
    t->charge = charge;
    snprintf(t->name, sizeof(t->name), "%s", name ? name : "X");
    snprintf(t->name2, sizeof(t->name2), "%s", name2 ? name2 : "X");
    return e->nr_types++;
}

int engine_addpot(struct engine *e, struct potential *p, int i, int j) {
    if (e == NULL) return eerror_(engine_err_null);
    if (i < 0 || i >= e->max_type || j < 0 || j >= e->max_type) return eerror_(engine_err_range);
    e->p[i * e->max_type + j] = p;
    if (i != j) e->p[j * e->max_type + i] = p;
    return engine_err_ok;
}

int engine_load(struct engine *e, double *x, double *v, int *type, int *pid, int *vid, double *q,
                unsigned int *flags, int N) {
    struct part p;
    int j, k;
    if (e == NULL || x == NULL || type == NULL) return eerror_(engine_err_null);
    memset(&p, 0, sizeof(p));
    for (j = 0; j < N; j++) {
        if (type[j] < 0 || type[j] >= e->max_type) return eerror_(engine_err_range);
        p.type = type[j];
        p.id = pid ? pid[j] : j;
        p.vid = vid ? vid[j] : 0;
        p.flags = flags ? flags[j] : part_flag_none;
        for (k = 0; k < 3; k++) {
            p.v[k] = v ? v[3 * j + k] : 0.0;
            p.f[k] = 0.0;
        }
        p.x[3] = p.v[3] = p.f[3] = 0.0;
        p.q = q ? q[j] : 0.0;
        if (space_addpart(&e->s, &p, &x[3 * j]) < 0) return eerror_(engine_err_space);
    }
    return engine_err_ok;
}

int engine_unload(struct engine *e, double *x, double *v, int *type, int *pid, int *vid, double *q,
                  unsigned int *flags, double *epot, int N) {
    int cid, k, j, count = 0;
    double eacc = 0.0;
    if (e == NULL) return eerror_(engine_err_null);
    if (e->s.nr_parts > N) return eerror_(engine_err_range);
    for (cid = 0; cid < e->s.nr_cells; cid++) {
        struct cell *c = &e->s.cells[cid];
        eacc += c->epot;
        for (k = 0; k < c->count; k++, count++) {
            struct part *p = &c->parts[k];
            if (x)
                for (j = 0; j < 3; j++) x[3 * count + j] = c->origin[j] + p->x[j];
            if (v)
                for (j = 0; j < 3; j++) v[3 * count + j] = p->v[j];
            if (type) type[count] = p->type;
            if (pid) pid[count] = p->id;
            if (vid) vid[count] = p->vid;
            if (q) q[count] = p->q;
            if (flags) flags[count] = p->flags;
        }
    }
    if (epot) *epot += eacc;
    return count;
}

int engine_prepare(struct engine *e) {
    if (e == NULL) return eerror_(engine_err_null);
    if (space_prepare(&e->s) != space_err_ok) return eerror_(engine_err_space);
    return engine_err_ok;
}

int engine_shuffle(struct engine *e) {
    if (space_shuffle_local(&e->s) < 0) return eerror_(engine_err_space);
    return engine_err_ok;
}

static int verlet_store(struct space *s);

int engine_start(struct engine *e, int nr_runners, int nr_queues) {
    (void)nr_queues;
    if (e == NULL) return eerror_(engine_err_null);
    if (!(e->flags & engine_flag_gpu)) return eerror_(engine_err_nogpu);
    e->nr_runners = nr_runners;
    /* Pseudo-Verlet: shuffle, reference positions, rebuild on the first
       step (engine.c:1712-1737). */
    if (e->flags & engine_flag_verlet) {
        if (engine_shuffle(e) < 0) return eerror_(engine_err);
        if (verlet_store(&e->s) < 0) return eerror_(engine_err);
        e->s.verlet_rebuild = 1;
    }
    if (engine_gpu_load(e) < 0) return eerror_(engine_err_gpu);
    return engine_err_ok;
}

int engine_nonbond_eval(struct engine *e) {
    if (e == NULL) return eerror_(engine_err_null);
    if (!(e->flags & engine_flag_gpu) || e->gpu == NULL) return eerror_(engine_err_nogpu);
    if (engine_nonbond_gpu(e) < 0) return eerror_(engine_err_gpu);
    return engine_err_ok;
}

/* Store the current positions as the reference of the displacement test
   (engine.c:286-297, 1717-1730). */
static int verlet_store(struct space *s) {
    int cid, pid, k;
    for (cid = 0; cid < s->nr_real; cid++) {
        struct cell *c = &s->cells[s->cid_real[cid]];
        if (c->oldx == NULL || c->oldx_size < c->count) {
            free(c->oldx);
            c->oldx_size = c->size + 20;
            if ((c->oldx = (FPTYPE *)malloc(sizeof(FPTYPE) * 3 * c->oldx_size)) == NULL) {
                c->oldx_size = 0;
                return eerror_(engine_err_malloc);
            }
        }
        for (pid = 0; pid < c->count; pid++)
            for (k = 0; k < 3; k++) c->oldx[3 * pid + k] = c->parts[pid].x[k];
    }
    return engine_err_ok;
}

/* engine.c:193-315 (the OpenMP branch, which sets each cell's maxdx): the
   largest displacement since the last rebuild, per cell and overall; a
   rebuild (shuffle, new reference positions, re-sort) once 2 maxdx exceeds
   the skin min(h) - rc or maxdx exceeds rc / 4.  As in the reference, a
   rebuild leaves the cells' maxdx as they were and zeroes only s->maxdx. */
int engine_verlet_update(struct engine *e) {
    struct space *s;
    int cid, pid, k;
    double dx, w, maxdx = 0.0, skin;
    if (e == NULL) return eerror_(engine_err_null);
    if (!(e->flags & engine_flag_verlet)) return engine_err_ok;
    s = &e->s;
    skin = fmin(s->h[0], fmin(s->h[1], s->h[2])) - s->cutoff;
    if (!s->verlet_rebuild) {
        for (cid = 0; cid < s->nr_real; cid++) {
            struct cell *c = &s->cells[s->cid_real[cid]];
            double lmaxdx = 0.0;
            for (pid = 0; pid < c->count; pid++) {
                for (dx = 0.0, k = 0; k < 3; k++) {
                    w = c->parts[pid].x[k] - c->oldx[3 * pid + k];
                    dx += w * w;
                }
                lmaxdx = fmax(dx, lmaxdx);
            }
            c->maxdx = sqrt(lmaxdx);
            maxdx = fmax(lmaxdx, maxdx);
        }
        maxdx = sqrt(maxdx);
        s->verlet_rebuild = (2.0 * maxdx > skin) || (maxdx > s->cutoff * engine_maxdxratio);
    }
    if (s->verlet_rebuild) {
        if (engine_shuffle(e) < 0) return eerror_(engine_err);
        if (verlet_store(s) < 0) return eerror_(engine_err);
        s->maxdx = 0.0;
    } else
        s->maxdx = maxdx;
    return engine_err_ok;
}

int engine_getmaxdx(struct engine *e, double *maxdx, double *cell_maxdx) {
    int cid;
    if (e == NULL) return eerror_(engine_err_null);
    if (maxdx) *maxdx = e->s.maxdx;
    if (cell_maxdx)
        for (cid = 0; cid < e->s.nr_cells; cid++) cell_maxdx[cid] = e->s.cells[cid].maxdx;
    return e->s.verlet_rebuild;
}

int engine_getlocal(struct engine *e, double *x, int *cell, int N) {
    int cid, k, m, n = 0;
    if (e == NULL || x == NULL) return eerror_(engine_err_null);
    for (cid = 0; cid < e->s.nr_cells; cid++) {
        struct cell *c = &e->s.cells[cid];
        for (k = 0; k < c->count; k++) {
            const int id = c->parts[k].id;
            if (id < 0 || id >= N) return eerror_(engine_err_range);
            for (m = 0; m < 3; m++) x[3 * id + m] = c->parts[k].x[m];
            if (cell) cell[id] = cid;
            n++;
        }
    }
    return n;
}

/* Kick-drift, then migration (engine.c:1945-2052).  With the pseudo-Verlet
   flag the particles stay in their cells until the next rebuild, so only
   the kick-drift is done (the Verlet branch, 1961-1985). */
int engine_advance(struct engine *e) {
    struct space *s = &e->s;
    int cid, pid, k;
    double epot = 0.0, dt = e->dt, w;
    /* the energy sum in cell order */
    for (cid = 0; cid < s->nr_real; cid++) epot += s->cells[s->cid_real[cid]].epot;
    /* the kick-drift of every particle, each its own: in parallel */
#pragma omp parallel for schedule(static) private(pid, k, w)
    for (cid = 0; cid < s->nr_real; cid++) {
        struct cell *c = &s->cells[s->cid_real[cid]];
        for (pid = 0; pid < c->count; pid++) {
            struct part *p = &c->parts[pid];
            w = dt * e->types[p->type].imass;
            for (k = 0; k < 3; k++) {
                p->v[k] += p->f[k] * w;
                p->x[k] += dt * p->v[k];
            }
        }
    }
    /* the migration of the particles that left their cell: the same cell
       contents and order as the reference's loop, which moves each particle
       right after its kick-drift (a particle's move depends only on its own
       position) */
    if (!(e->flags & engine_flag_verlet) && space_shuffle_local(s) < 0) return eerror_(engine_err_cell);
    s->epot_nonbond += epot;
    s->epot += epot;
    return engine_err_ok;
}

int engine_step(struct engine *e) {
    double t0 = wall_seconds(), t;
    if (e == NULL) return eerror_(engine_err_null);
    e->time += 1;
    t = wall_seconds();
    if (e->flags & engine_flag_verlet) {
        if (engine_verlet_update(e) < 0) return eerror_(engine_err);
    } else if (engine_shuffle(e) < 0)
        return eerror_(engine_err_space);
    e->timers[engine_timer_shuffle] = wall_seconds() - t;
    t = wall_seconds();
    if (space_prepare(&e->s) != space_err_ok) return eerror_(engine_err_space);
    e->timers[engine_timer_prepare] = wall_seconds() - t;
    t = wall_seconds();
    if (engine_nonbond_eval(e) < 0) return eerror_(engine_err);
    e->timers[engine_timer_nonbond] = wall_seconds() - t;
    if ((e->flags & engine_flag_verlet) && e->s.verlet_rebuild) e->s.verlet_rebuild = 0;
    t = wall_seconds();
    if (engine_advance(e) < 0) return eerror_(engine_err);
    /* constraints after the advance (engine.c:2196-2232) */
    if (e->nr_rigids > 0 && engine_rigid_eval(e) < 0) return eerror_(engine_err_rigid);
    e->timers[engine_timer_advance] = wall_seconds() - t;
    e->timers[engine_timer_step] = wall_seconds() - t0;
    return engine_err_ok;
}

int engine_exclusion_add(struct engine *e, int i, int j) {
    struct exclusion *tmp;
    if (e == NULL) return eerror_(engine_err_null);
    if (e->nr_exclusions == e->exclusions_size) {
        int size = e->exclusions_size * 1.414;
        if (size <= e->exclusions_size) size = e->exclusions_size + 16;
        if ((tmp = (struct exclusion *)malloc(sizeof(struct exclusion) * size)) == NULL)
            return eerror_(engine_err_malloc);
        memcpy(tmp, e->exclusions, sizeof(struct exclusion) * e->nr_exclusions);
        free(e->exclusions);
        e->exclusions = tmp;
        e->exclusions_size = size;
    }
    e->exclusions[e->nr_exclusions].i = (i <= j) ? i : j;
    e->exclusions[e->nr_exclusions].j = (i <= j) ? j : i;
    e->nr_exclusions += 1;
    return engine_err_ok;
}

static int excl_cmp(const void *a, const void *b) {
    const struct exclusion *x = (const struct exclusion *)a, *y = (const struct exclusion *)b;
    if (x->i != y->i) return (x->i < y->i) ? -1 : 1;
    if (x->j != y->j) return (x->j < y->j) ? -1 : 1;
    return 0;
}

/* Sort by (i, j) and drop duplicates.  The result is a sorted unique list,
   which is all the reference's hand-written quicksort guarantees too. */
int engine_exclusion_shrink(struct engine *e) {
    int j, k;
    if (e == NULL) return eerror_(engine_err_null);
    if (e->nr_exclusions == 0) return engine_err_ok;
    qsort(e->exclusions, e->nr_exclusions, sizeof(struct exclusion), excl_cmp);
    for (j = 0, k = 1; k < e->nr_exclusions; k++)
        if (e->exclusions[k].j != e->exclusions[j].j || e->exclusions[k].i != e->exclusions[j].i)
            e->exclusions[++j] = e->exclusions[k];
    e->nr_exclusions = j + 1;
    return engine_err_ok;
}

int engine_finalize(struct engine *e) {
    if (e == NULL) return eerror_(engine_err_null);
    if (e->gpu != NULL) engine_gpu_finalize(e);
    space_free(&e->s);
    free(e->types);
    free(e->p);
    free(e->exclusions);
    free(e->rigids);
    free(e->part2rigid);
    free(e->bonds);
    free(e->angles);
    free(e->dihedrals);
    e->bonds = NULL;
    e->angles = NULL;
    e->dihedrals = NULL;
    e->nr_bonds = e->bonds_size = e->nr_angles = e->angles_size = e->nr_dihedrals = e->dihedrals_size = 0;
    e->rigids = NULL;
    e->part2rigid = NULL;
    e->nr_rigids = e->rigids_size = e->part2rigid_size = 0;
    e->types = NULL;
    e->p = NULL;
    e->exclusions = NULL;
    return engine_err_ok;
}

/* ------------------------------------------------------------ accessors */

int engine_getforces(struct engine *e, double *f, int N) {
    int id, k;
    if (e == NULL || f == NULL) return eerror_(engine_err_null);
    for (id = 0; id < N; id++) {
        struct part *p = (id < e->s.size_parts) ? e->s.partlist[id] : NULL;
        for (k = 0; k < 3; k++) f[3 * id + k] = p ? p->f[k] : 0.0;
    }
    return engine_err_ok;
}

int engine_getpositions(struct engine *e, double *x, int N) {
    int id;
    if (e == NULL || x == NULL) return eerror_(engine_err_null);
    for (id = 0; id < N; id++)
        if (space_getpos(&e->s, id, &x[3 * id]) < 0) return eerror_(engine_err_space);
    return engine_err_ok;
}

int engine_getvelocities(struct engine *e, double *v, int N) {
    int id, k;
    if (e == NULL || v == NULL) return eerror_(engine_err_null);
    for (id = 0; id < N; id++) {
        const struct part *p = (id < e->s.size_parts) ? e->s.partlist[id] : NULL;
        if (p == NULL) return eerror_(engine_err_range);
        for (k = 0; k < 3; k++) v[3 * id + k] = p->v[k];
    }
    return engine_err_ok;
}

int engine_setvelocities(struct engine *e, const double *v, int N) {
    int id, k;
    if (e == NULL || v == NULL) return eerror_(engine_err_null);
    for (id = 0; id < N; id++) {
        struct part *p = (id < e->s.size_parts) ? e->s.partlist[id] : NULL;
        if (p == NULL) return eerror_(engine_err_range);
        for (k = 0; k < 3; k++) p->v[k] = (FPTYPE)v[3 * id + k];
    }
    return engine_err_ok;
}

int engine_getenergies(struct engine *e, double *energies, long long *counts) {
    if (e == NULL) return eerror_(engine_err_null);
    if (energies) {
        energies[0] = e->s.epot;
        energies[1] = e->s.epot_nonbond;
        energies[2] = e->s.epot_exclusion;
        energies[3] = e->s.virial;
    }
    if (counts) {
        counts[0] = e->s.nr_pairs;
        counts[1] = e->s.nr_excl_evaluated;
    }
    return engine_err_ok;
}

int engine_getsortlist(struct engine *e, int cid, int sid, unsigned int *out) {
    struct cell *c;
    if (e == NULL || out == NULL) return eerror_(engine_err_null);
    if (cid < 0 || cid >= e->s.nr_cells || sid < 0 || sid >= 13) return eerror_(engine_err_range);
    c = &e->s.cells[cid];
    memcpy(out, &c->sortlist[c->count * sid], sizeof(unsigned int) * c->count);
    return c->count;
}

long long engine_getsortlists(struct engine *e, unsigned int *out) {
    long long off = 0;
    int cid;
    if (e == NULL) return eerror_(engine_err_null);
    for (cid = 0; cid < e->s.nr_cells; cid++) {
        struct cell *c = &e->s.cells[cid];
        if (out != NULL && c->count > 0) memcpy(&out[off], c->sortlist, sizeof(unsigned int) * 13 * c->count);
        off += 13LL * c->count;
    }
    return off;
}

int engine_gettasks(struct engine *e, int *out, int max) {
    int k;
    if (e == NULL) return eerror_(engine_err_null);
    if (out != NULL)
        for (k = 0; k < e->s.nr_tasks && k < max; k++) {
            out[4 * k] = e->s.tasks[k].type;
            out[4 * k + 1] = e->s.tasks[k].i;
            out[4 * k + 2] = e->s.tasks[k].j;
            out[4 * k + 3] = e->s.tasks[k].flags;
        }
    return e->s.nr_tasks;
}

int engine_getcells(struct engine *e, int *counts, int *ids) {
    int cid, k, n = 0;
    if (e == NULL) return eerror_(engine_err_null);
    for (cid = 0; cid < e->s.nr_cells; cid++) {
        struct cell *c = &e->s.cells[cid];
        if (counts) counts[cid] = c->count;
        if (ids)
            for (k = 0; k < c->count; k++) ids[n + k] = c->parts[k].id;
        n += c->count;
    }
    return n;
}
