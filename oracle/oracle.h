/*
 * ORACLE — test infrastructure only.  A CPU, FP64 restatement of mdcore's
 * nonbonded path used to check the GPU engine; nothing in the product links
 * or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 */
#ifndef MD_ORACLE_H
#define MD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Window modes of the sorted pair loop. */
#define OR_WINDOW_REF  0 /* reference: |shift|, unsigned arithmetic (runner_dopair.c:451-475) */
#define OR_WINDOW_PROJ 1 /* projected shift.n, signed arithmetic */

typedef struct or_system or_system;

/* Tables: ntab tables; table t has n[t] intervals, alpha[4t..4t+3] and
   coefficients c[off[t] .. off[t] + (n[t]+1)*8).  pind[ti*max_type+tj] is the
   table of a type pair or -1. */
typedef struct {
    int max_type, ntab;
    const int *pind;
    const int *n;
    const double *alpha;
    const double *c;
    const int *off;
} or_tables;

/* Build the cell grid and task list (space_init), insert the particles in
   input order (engine_load) and apply one engine_shuffle.  x are absolute
   positions, 3 per particle; particle ids are 0..N-1. */
or_system *or_create(const double *origin, const double *dim, const double *L, double cutoff, int N,
                     const double *x, const int *type, const or_tables *tab, int nexcl, const int *excl);
void or_free(or_system *s);

/* Sort every cell along its flagged directions (runner_dosort). */
int or_sort(or_system *s);

/* One nonbonded pass: sort, pair/self tasks, exclusions.  f[3*id+k] gets the
   forces by particle id.  energies = {epot_nonbond, epot_exclusion, virial};
   counts = {pairs evaluated, exclusions subtracted}.  nthreads > 1 runs the
   tasks on pthreads with private force buffers (summation order then
   depends on scheduling). */
int or_nonbond(or_system *s, int mode, int nthreads, double *f, double *energies, long long *counts);

/* All-pairs check with the reference's r^2 arithmetic but no window and no
   excluded pairs:
   counts every pair with r^2 <= rc^2 and a potential, and sums FP64 forces. */
int or_bruteforce(or_system *s, double *f, double *energies, long long *count);

/* Pseudo-Verlet steps (engine_flag_verlet_pseudo): move the particles
   without changing their cells (xrel: cell-relative positions by id, may lie
   outside the cell; NULL keeps them), set each cell's maxdx (NULL keeps
   them), and with keep_sort != 0 make or_nonbond reuse the sort lists of the
   last sort instead of sorting (runner.c:773, runner_dopair.c:428). */
int or_set_verlet(or_system *s, const double *xrel, const double *cell_maxdx, int keep_sort);

/* Kick-drift of engine_advance (engine.c:1945-2052) with the forces of the
   last or_nonbond, over cells on nthreads threads: v += dt f / m, x += dt v.
   Velocities by id and inverse masses by type come from or_set_motion (v
   NULL keeps them, zero at first).  Particles are not moved between cells:
   or_advance returns how many now lie outside their cell (none at dt = 0,
   the setting of the CPU baseline timings), or -1 before any or_nonbond. */
int or_set_motion(or_system *s, const double *v, const double *imass_by_type);
int or_advance(or_system *s, double dt, int nthreads);

/* Introspection. */
int or_ncells(const or_system *s);
int or_cell(const or_system *s, int cid, int *count, int *ids, unsigned int *sortflags);
int or_sortlist(const or_system *s, int cid, int sid, unsigned int *out);
int or_ntasks(const or_system *s, int *npair, int *nself);
int or_task(const or_system *s, int k, int *type, int *i, int *j, int *sid);
double or_dscale(const or_system *s);
/* Pairs evaluated by each task (or_task order) in the last or_nonbond. */
int or_task_counts(const or_system *s, long long *out);
/* Cell-relative positions of cell c in storage order (3 per particle). */
int or_cell_x(const or_system *s, int c, double *x);
/* Every cell's 13 sort lists, cells in id order, each cell's block laid out
   as cell.c:443 does (direction sid at sid*count); directions not flagged
   hold whatever the last sort left.  out == NULL returns the word count. */
long long or_sortlists(const or_system *s, unsigned int *out);

/* Building blocks, exported for unit tests. */
void or_sort_desc(unsigned int *a, int N);
unsigned int or_key(const double *x, const double *axis, double dscale, double bias);
void or_eval(int n, const double *alpha, const double *c, double r2, double *e, double *f);
extern const double or_axis[39];

#ifdef __cplusplus
}
#endif

#endif
