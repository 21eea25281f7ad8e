/*
 * mdcore host API for the MI355X nonbonded engine.
 *
 * This header keeps the C API of mdcore's nonbonded path so that drivers
 * written against the reference (examples/argon/argon.c, examples/bulk/bulk.c)
 * compile against it: engine_init / engine_addtype / engine_addpot /
 * engine_load / engine_start / engine_step / engine_unload, the
 * potential_create_* piecewise-quintic tables, and the space/cell/task data
 * model.  The host build is double precision (FPTYPE = double); the GPU
 * computes pair arithmetic in FP32 (see engine_gpu.h).
 *
 * Reference interfaces mirrored here (file:line in the reference tree):
 *   struct part            src/part.h:46-69
 *   struct part_type       src/part.h:72-89
 *   struct potential       src/potential.h:57-74, constants 31-40
 *   struct cell            src/cell.h:58-118, flags 34-44
 *   struct task            src/task.h:24-88
 *   struct space           src/space.h:65-142, constants 31-51
 *   struct engine          src/engine.h:179-300, error codes 21-52, flags 56-72
 *   errs_register/dump     src/errs.h, src/errs.c:61-111
 */
#ifndef MDCORE_H
#define MDCORE_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

/* Host floating-point type: the double build of the reference (fptype.h:21-40). */
typedef double FPTYPE;
#define FPTYPE_DOUBLE 1

/* ------------------------------------------------------------------ errs */

/* Error stack (errs.c:61-111): every failing call pushes one entry. */
int errs_register(int id, const char *msg, int line, const char *func, const char *file);
int errs_dump(FILE *out);
int errs_clear(void);
/* Copy the error stack as text into buf (for callers without a FILE*). */
int errs_gettext(char *buf, int size);

/* ------------------------------------------------------------------ part */

#define part_flag_none   0
#define part_flag_frozen 1
#define part_flag_ghost  2

/* Particle record (part.h:46-69): positions are relative to the owning cell. */
struct part {
    FPTYPE x[4] __attribute__((aligned(16)));
    FPTYPE v[4] __attribute__((aligned(16)));
    FPTYPE f[4] __attribute__((aligned(16)));
    float q;
    int id, vid;
    short int type;
    unsigned short int flags;
} __attribute__((aligned(64)));

struct part_type {
    int id;
    double mass, imass, charge;
    double eps, rmin;
    char name[64], name2[64];
};

/* ------------------------------------------------------------------ potential */

#define potential_err_ok        0
#define potential_err_null      -1
#define potential_err_malloc    -2
#define potential_err_bounds    -3
#define potential_err_nyi       -4
#define potential_err_ivalsmax  -5

#define potential_degree   5
#define potential_chunk    (potential_degree + 3)
#define potential_ivalsa   1
#define potential_ivalsb   10
#define potential_N        100
#define potential_align    64
#define potential_ivalsmax 320

/* Electrostatic prefactor 1/(4 pi) of the reference (potential.h:40): the
   reference's Coulomb tables are in units where eps0 = 1. */
#define potential_escale (0.079577471545947667882)

#define potential_flag_none    0
#define potential_flag_LJ126   1
#define potential_flag_Ewald   2
#define potential_flag_Coulomb 4

extern int potential_err;

/* A piecewise-quintic table on [a,b] with n intervals.  c holds (n+1) chunks
   of potential_chunk coefficients {centre, 1/halfwidth, x^5..x^0}; chunk 0 is
   the linear extension below a (potential.c:1167-1185). */
struct potential {
    FPTYPE alpha[4];
    FPTYPE *c;
    double a, b;
    unsigned int flags;
    int n;
};

extern struct potential potential_null;

int potential_init(struct potential *p, double (*f)(double), double (*fp)(double),
                   double (*f6p)(double), FPTYPE a, FPTYPE b, FPTYPE tol);
int potential_getcoeffs(double (*f)(double), double (*fp)(double), FPTYPE *xi, int n,
                        FPTYPE *c, FPTYPE *err);
double potential_getalpha(double (*f6p)(double), double a, double b);
void potential_clear(struct potential *p);

struct potential *potential_create_LJ126(double a, double b, double A, double B, double tol);
struct potential *potential_create_LJ126_switch(double a, double b, double A, double B, double s, double tol);
struct potential *potential_create_LJ126_Ewald(double a, double b, double A, double B, double q, double kappa, double tol);
struct potential *potential_create_LJ126_Ewald_switch(double a, double b, double A, double B, double q, double kappa, double s, double tol);
struct potential *potential_create_LJ126_Coulomb(double a, double b, double A, double B, double q, double tol);
struct potential *potential_create_Ewald(double a, double b, double q, double kappa, double tol);
struct potential *potential_create_Coulomb(double a, double b, double q, double tol);
struct potential *potential_create_harmonic(double a, double b, double K, double r0, double tol);

/* New (no reference counterpart; built through the reference's extension point
   potential_init(f, fp, f6p, ...), potential.h:83): Lennard-Jones 12-6 plus
   reaction-field Coulomb, shifted so that E(rc) = 0:
     E(r) = A/r^12 - B/r^6 + ke*q*(1/r + krf*r^2 - crf),
     krf = (eps_rf - 1)/((2 eps_rf + 1) rc^3), crf = 1/rc + krf rc^2.
   ke is the Coulomb constant in the caller's units (138.935458 kJ mol^-1 nm e^-2
   for GROMACS units); eps_rf <= 0 means eps_rf = infinity. */
struct potential *potential_create_LJ126_RF(double a, double b, double A, double B, double q,
                                            double ke, double eps_rf, double tol);
/* New, like potential_create_LJ126_RF, with the Lennard-Jones part switched off
   between s and b by the reference's potential_switch (potential.c:60-96):
     E(r) = (A/r^12 - B/r^6) S(r) + ke*q*(1/r + krf*r^2 - crf).
   With eps_rf <= 0 (conducting boundary) E and dE/dr both vanish at b, so
   the table is force-continuous at the cutoff. */
struct potential *potential_create_LJ126_switch_RF(double a, double b, double A, double B, double s, double q,
                                                   double ke, double eps_rf, double tol);

/* Host FP64 evaluation of a table (potential_eval.h:67-104): e and f=(dE/dr)/r. */
void potential_eval(const struct potential *p, FPTYPE r2, FPTYPE *e, FPTYPE *f);
/* New: e[k] and dE/dr[k] at distances r[k], k < n. */
int potential_eval_r(const struct potential *p, const double *r, long long n, double *e, double *dedr);

/* Analytic helpers (potential.c:66-267). */
double potential_LJ126(double r, double A, double B);
double potential_LJ126_p(double r, double A, double B);
double potential_LJ126_6p(double r, double A, double B);
double potential_Coulomb(double r);
double potential_Coulomb_p(double r);
double potential_Coulomb_6p(double r);
double potential_Ewald(double r, double kappa);
double potential_Ewald_p(double r, double kappa);
double potential_Ewald_6p(double r, double kappa);
double potential_switch(double r, double A, double B);
double potential_switch_p(double r, double A, double B);

/* ------------------------------------------------------------------ cell */

#define cell_err_ok      0
#define cell_err_null    -1
#define cell_err_malloc  -2
#define cell_err_pthread -3

#define cell_incr      10
#define cell_partalign 64

#define cell_flag_none   0
#define cell_flag_ghost  1
#define cell_flag_wait   2
#define cell_flag_waited 4
#define cell_flag_marked 8

/* Direction tables (cell.c:56-101): sort-list id of each of the 27 neighbour
   offsets, the 13 unit axes, and whether a pair must be flipped. */
extern const char cell_sortlistID[27];
extern const FPTYPE cell_shift[13 * 3];
extern const char cell_flip[27];

struct cell {
    unsigned int flags;
    int id;
    int loc[3];
    double origin[3];
    double dim[3];
    int size, count;
    struct part *parts;
    double epot;
    struct part *incomming;
    int incomming_size, incomming_count;
    FPTYPE maxdx;
    /* Positions at the last pseudo-Verlet rebuild, 3 per particle slot
       (the reference's c->oldx, 4 per slot, engine.c:286-297). */
    FPTYPE *oldx;
    int oldx_size;
    /* 13 sorted lists of (index << 16 | key), count entries each (cell.c:443). */
    unsigned int *sortlist;
    /* Bitmask of the sort directions this cell's pair tasks need (space.c:745-752). */
    unsigned int sortflags;
};

int cell_init(struct cell *c, int *loc, double *origin, double *dim);
struct part *cell_add(struct cell *c, struct part *p, struct part **partlist);
struct part *cell_add_incomming(struct cell *c, struct part *p);
int cell_welcome(struct cell *c, struct part **partlist);
int cell_flush(struct cell *c, struct part **partlist, struct cell **celllist);

/* ------------------------------------------------------------------ task */

#define task_type_none 0
#define task_type_self 1
#define task_type_pair 2
#define task_type_sort 3

#define task_subtype_none 0

/* Cell-pair / self / sort task (task.h:24-88).  For pairs, flags = sid. */
struct task {
    int type, subtype, flags;
    int i, j;
    int wait;
};

/* ------------------------------------------------------------------ space */

#define space_err_ok       0
#define space_err_null     -1
#define space_err_malloc   -2
#define space_err_cell     -3
#define space_err_pthread  -4
#define space_err_range    -5
#define space_err_maxpairs -6
#define space_err_nrtasks  -7
#define space_err_task     -8

#define space_periodic_none       0
#define space_periodic_x          1
#define space_periodic_y          2
#define space_periodic_z          4
#define space_periodic_full       7
#define space_periodic_ghost_x    8
#define space_periodic_ghost_y    16
#define space_periodic_ghost_z    32
#define space_periodic_ghost_full 56

#define space_partlist_incr 100

#define space_cellid(s, i, j, k) (((i) * (s)->cdim[1] + (j)) * (s)->cdim[2] + (k))

extern int space_err;

struct space {
    double dim[3];
    double origin[3];
    int cdim[3];
    int span[3];
    double h[3], ih[3];
    double cutoff, cutoff2;
    unsigned int period;
    int nr_cells;
    int *cid_real, *cid_ghost, *cid_marked;
    int nr_real, nr_ghost, nr_marked;
    struct cell *cells;
    int nr_tasks, tasks_size;
    struct task *tasks;
    struct part **partlist;
    struct cell **celllist;
    int nr_parts, size_parts;
    int verlet_rebuild;
    FPTYPE maxdx;
    double epot, epot_nonbond, epot_bond, epot_angle, epot_dihedral, epot_exclusion;
    /* Scalar virial sum_pairs dx.F of the last nonbonded pass (new). */
    double virial;
    /* Pair interactions evaluated by the last nonbonded pass (new). */
    long long nr_pairs;
    /* Exclusions subtracted by the last pass (new). */
    long long nr_excl_evaluated;
    /* Scratch of space_setcells, kept between calls (new). */
    void *setcells_scratch;
    size_t setcells_scratch_size;
};

int space_init(struct space *s, const double *origin, const double *dim, double *L, double cutoff,
               unsigned int period);
int space_getsid(struct space *s, struct cell **ci, struct cell **cj, FPTYPE *shift);
int space_addpart(struct space *s, struct part *p, double *x);
int space_prepare(struct space *s);
int space_shuffle_local(struct space *s);
int space_getpos(struct space *s, int id, double *x);
int space_flush(struct space *s);
/* New helper: replace the cell contents by a layout built elsewhere (the
   device step).  count[cid] particles per cell in storage order, given by
   id with their new cell-relative x, v and type; charge, vid and flags
   follow the id.  Forces are set to zero. */
int space_setcells(struct space *s, const int *count, const double *x, const double *v, const int *id,
                   const int *type);
void space_free(struct space *s);

/* ------------------------------------------------------------------ exclusion */

struct exclusion {
    int i, j;
};

/* ------------------------------------------------------------------ engine */

#define engine_err_ok        0
#define engine_err_null      -1
#define engine_err_malloc    -2
#define engine_err_space     -3
#define engine_err_pthread   -4
#define engine_err_runner    -5
#define engine_err_range     -6
#define engine_err_cell      -7
#define engine_err_domain    -8
#define engine_err_nompi     -9
#define engine_err_mpi       -10
#define engine_err_bond      -11
#define engine_err_angle     -12
#define engine_err_reader    -13
#define engine_err_psf       -14
#define engine_err_pdb       -15
#define engine_err_cpf       -16
#define engine_err_potential -17
#define engine_err_exclusion -18
#define engine_err_sets      -19
#define engine_err_dihedral  -20
/* engine_err_cuda (-21) of the reference: any HIP runtime failure. */
#define engine_err_gpu       -21
#define engine_err_nogpu     -22
#define engine_err_cudasp    -23
#define engine_err_maxparts  -24
#define engine_err_queue     -25
#define engine_err_rigid     -26
#define engine_err_cutoff    -27
#define engine_err_nometis   -28
#define engine_err_deps      -29
#define engine_err_nyi       -30

#define engine_flag_none          0
#define engine_flag_static        1
#define engine_flag_localparts    2
/* Bit 4 is the reference's engine_flag_cuda: here it selects the GPU path. */
#define engine_flag_gpu           4
#define engine_flag_cuda          engine_flag_gpu
#define engine_flag_explepot      8
#define engine_flag_verlet        16
#define engine_flag_verlet_pairwise 32
#define engine_flag_affinity      64
#define engine_flag_prefetch      128
#define engine_flag_verlet_pseudo 256
#define engine_flag_unsorted      512
#define engine_flag_shake         1024
#define engine_flag_mpi           2048
#define engine_flag_parbonded     4096
#define engine_flag_async         8192
#define engine_flag_dpd           16384
#define engine_flag_nullpart      32768
/* New: evaluate the sorted window with the projected shift shift.n instead of
   the reference's |shift| (runner_dopair.c:451-452).  The projected window
   never drops an in-cutoff pair; the reference window does for cell offsets
   with more than one non-zero component and span > 1. */
#define engine_flag_projwindow    65536

/* Pseudo-Verlet rebuild threshold: maxdx > cutoff * engine_maxdxratio
   (engine.h:83). */
#define engine_maxdxratio 0.25

extern int engine_err;
extern const char *engine_err_msg[];

/* ---------------------------------------------------------------- rigid */

struct engine;

/* Rigid constraint groups (rigid.h): up to rigid_maxparts particles and
   rigid_maxconstr distance constraints, solved by SHAKE or P-SHAKE after
   each advance. */
#define rigid_err_ok          0
#define rigid_err_null        -1
#define rigid_err_malloc      -2
#define rigid_maxparts        10
#define rigid_maxconstr       (3 * rigid_maxparts)
#define rigid_maxiter         100
#define rigid_pshake_refine   10
#define rigid_pshake_maxalpha 0.1f

struct rigid {
    int nr_parts;
    int parts[rigid_maxparts]; /* particle ids */
    int nr_constr;
    struct {
        int i, j; /* indices into parts */
        double d2;
    } constr[rigid_maxconstr];
    /* P-SHAKE constraint coupling matrix, nr_constr x nr_constr used. */
    float a[rigid_maxconstr * rigid_maxconstr];
};

/* Bonded topology as the structure readers store it (bond.h, angle.h,
   dihedral.h).  This library evaluates no bonded terms: the lists serve to
   build exclusions and rigid constraints, and pid stays -1 (or -2 for an
   improper dihedral, as engine_read_psf marks them). */
struct bond {
    int i, j, pid;
};
struct angle {
    int i, j, k, pid;
};
struct dihedral {
    int i, j, k, l, pid;
};

/* Inner radius of the tables made by the parameter readers: where the pair
   potential reaches engine_maxKcutoff times its value at r0 (engine.h:82);
   reader buffer size (engine.h:79). */
#define engine_maxKcutoff 2
#define engine_readbuff   16384

/* P-SHAKE refines its coupling matrices on the first engine_pshake_steps
   steps (engine.h:81). */
#define engine_pshake_steps 20

int rigid_eval_shake(struct rigid *rs, int N, struct engine *e);
int rigid_eval_pshake(struct rigid *rs, int N, struct engine *e, int a_update);

struct engine {
    unsigned int flags;
    struct space s;
    int time;
    double dt;
    int max_type;
    int nr_types;
    struct part_type *types;
    /* Type-pair potentials, max_type x max_type, NULL = no interaction (engine.h:192). */
    struct potential **p;
    struct exclusion *exclusions;
    int nr_exclusions, exclusions_size;
    int nr_runners;
    /* Opaque device state owned by the GPU layer (engine_gpu.h). */
    void *gpu;
    /* Wall-clock seconds of the last engine_step phases (new). */
    double timers[8];
    /* Rigid constraints (engine.h:239-250). */
    struct rigid *rigids;
    int nr_rigids, rigids_size, nr_constr;
    int *part2rigid, part2rigid_size;
    double tol_rigid;
    /* Bonded topology (engine.h:213-233). */
    struct bond *bonds;
    int nr_bonds, bonds_size;
    struct angle *angles;
    int nr_angles, angles_size;
    struct dihedral *dihedrals;
    int nr_dihedrals, dihedrals_size;
};

enum {
    engine_timer_step = 0,
    engine_timer_prepare,
    engine_timer_nonbond,
    engine_timer_advance,
    engine_timer_shuffle,
    engine_timer_gpu_load,
    engine_timer_gpu_kernels,
    engine_timer_gpu_unload
};

int engine_init(struct engine *e, const double *origin, const double *dim, double *L, double cutoff,
                unsigned int period, int max_type, unsigned int flags);
int engine_addtype(struct engine *e, double mass, double charge, char *name, char *name2);
int engine_addpot(struct engine *e, struct potential *p, int i, int j);
int engine_load(struct engine *e, double *x, double *v, int *type, int *pid, int *vid, double *charge,
                unsigned int *flags, int N);
int engine_unload(struct engine *e, double *x, double *v, int *type, int *pid, int *vid, double *charge,
                  unsigned int *flags, double *epot, int N);
int engine_start(struct engine *e, int nr_runners, int nr_queues);
int engine_step(struct engine *e);
int engine_nonbond_eval(struct engine *e);
int engine_advance(struct engine *e);
/* Pseudo-Verlet bookkeeping at the start of a step (engine.c:193-315). */
int engine_verlet_update(struct engine *e);
/* Rigid constraints (engine_rigid.c:78-230, 352-583): add a distance
   constraint d between particles pid and pjd (merging groups), and solve
   all groups (SHAKE with engine_flag_shake, P-SHAKE otherwise). */
int engine_rigid_add(struct engine *e, int pid, int pjd, double d);
int engine_rigid_eval(struct engine *e);
/* New helper: group rid's particle ids (parts[0] = count, then the ids),
   constraint index pairs (ij, 2 per constraint) and d^2; returns the
   number of constraints (rid == -1: the number of groups). */
int engine_getrigid(struct engine *e, int rid, int *parts, int *ij, double *d2);

/* Topology (engine_bonded.c): append a bond, angle or dihedral between
   particle ids; pid is the (unused here) potential id. */
int engine_bond_add(struct engine *e, int i, int j);
int engine_angle_add(struct engine *e, int i, int j, int k, int pid);
int engine_dihedral_add(struct engine *e, int i, int j, int k, int l, int pid);
/* Type lookup by name (engine.c:1496-1540): the first type whose name
   (engine_gettype) or second name (engine_gettype2) matches, or
   engine_err_range. */
int engine_gettype(struct engine *e, char *name);
int engine_gettype2(struct engine *e, char *name2);

/* Structure and parameter readers (engine_io.c), on open file descriptors:
   - engine_read_psf (1234-1618): types (one per atom type and charge),
     particles from the PDB (Angstrom, wrapped into the box), bonds, angles,
     dihedrals and impropers into the topology lists;
   - engine_read_cpf (630-1232, CHARMM parameter file) and engine_read_xplor
     (82-628): the NONBONDED parameters (eps in kcal/mol, Rmin/2 in A) become
     Lennard-Jones tables for every type pair, with Coulomb (kappa == 0),
     real-space Ewald (kappa > 0) or none (kappa < 0) added for charged
     pairs; cpf uses the switched forms over switchlen.  With rigidH, bonds
     to hydrogen types become rigid constraints and leave the bond list.
     Bonded parameters are read and not used. */
/* New helpers for inspection: the table of a type pair (NULL if none); a
   type's mass, charge, eps and rmin (out[4]) and names (at least 64 chars
   each, may be NULL); the topology lists (which: 0 bonds as i j, 1 angles
   as i j k, 2 dihedrals as i j k l pid; out may be NULL), returning the
   count. */
struct potential *engine_getpot(struct engine *e, int i, int j);
int engine_gettypeinfo(struct engine *e, int k, double *out, char *name, char *name2);
int engine_gettopology(struct engine *e, int which, int *out);
int engine_read_psf(struct engine *e, int psf, int pdb);
int engine_read_cpf(struct engine *e, int cpf, double kappa, double tol, int rigidH, double switchlen);
int engine_read_xplor(struct engine *e, int xplor, double kappa, double tol, int rigidH);
/* New helper: s->maxdx and the per-cell maxdx (cell_maxdx may be NULL). */
int engine_getmaxdx(struct engine *e, double *maxdx, double *cell_maxdx);
/* New helper: by particle id, the cell-relative position (x, 3 per id) and
   the cell id (cell, may be NULL), as stored; returns the count. */
int engine_getlocal(struct engine *e, double *x, int *cell, int N);
int engine_shuffle(struct engine *e);
/* space_prepare(&e->s): zero forces and energies before a pass (new helper). */
int engine_prepare(struct engine *e);
int engine_exclusion_add(struct engine *e, int i, int j);
int engine_exclusion_shrink(struct engine *e);
int engine_finalize(struct engine *e);

/* Helpers for callers that cannot reach into struct engine (ctypes, other
   FFIs).  New; the reference's drivers read e.s.cells[].parts[] directly. */
unsigned long engine_sizeof(void);
struct engine *engine_alloc(void);
void engine_free(struct engine *e);
/* Set the time step used by engine_step's kick-drift (the reference's drivers
   assign e->dt directly). */
int engine_setdt(struct engine *e, double dt);
/* Set and clear bits of e->flags (the reference's drivers assign the field);
   engine_flag_gpu cannot be changed after engine_start. */
int engine_setflags(struct engine *e, unsigned int set, unsigned int clear);
/* Per-particle forces f[3*id+k] (host copy), by particle id. */
int engine_getforces(struct engine *e, double *f, int N);
/* Absolute positions x[3*id+k], by particle id. */
int engine_getpositions(struct engine *e, double *x, int N);
/* Velocities v[3*id+k], by particle id. */
int engine_getvelocities(struct engine *e, double *v, int N);
/* Set the velocities from v[3*id+k], by particle id (the reference's drivers
   assign p->v directly, e.g. to rescale to a temperature); engine_gpu_steps
   takes them at its next call. */
int engine_setvelocities(struct engine *e, const double *v, int N);
/* energies[0..3] = epot, epot_nonbond, epot_exclusion, virial; counts[0..1] = pairs, exclusions. */
int engine_getenergies(struct engine *e, double *energies, long long *counts);
/* Copy the host sortlist of cell cid, direction sid, into out (count entries). */
int engine_getsortlist(struct engine *e, int cid, int sid, unsigned int *out);
/* Every cell's host sortlist block (13*count words, cell.c:443 layout), cells
   in id order; out == NULL returns the word count. */
long long engine_getsortlists(struct engine *e, unsigned int *out);
/* The task list as rows (type, i, j, flags); returns the number of tasks. */
int engine_gettasks(struct engine *e, int *out, int max);
/* Per-cell geometry: counts[cid], and the particle ids in cell order (N entries). */
int engine_getcells(struct engine *e, int *counts, int *ids);
void potential_free(struct potential *p);
/* Read back a table: n, a, b, alpha[4], c[(n+1)*8]. */
int potential_getdata(const struct potential *p, int *n, double *ab, double *alpha, double *c, int csize);

#ifdef __cplusplus
}
#endif

#endif /* MDCORE_H */
