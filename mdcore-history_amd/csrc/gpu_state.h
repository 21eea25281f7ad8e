/*
 * Device state of the gfx950 nonbonded engine, shared by the pair kernels
 * (engine_gpu.hip) and the MLP surrogate (mlp_gpu.hip).  Internal header.
 */
#ifndef GPU_STATE_H
#define GPU_STATE_H

#include <hip/hip_runtime.h>

#include "engine_gpu.h"
#include "mlp_tile.h"

#define MD_NQUEUE 8
#define MD_MAX_CELL 1024  /* particles per cell: the packed record's slot field is 10 bits at most */
#define MD_MAX_TYPES 1024 /* particle types: 6, 8 or 10 type bits, leaving 10, 8 or 6 slot bits */

#define gerror_(id, msg) errs_register((id), (msg), __LINE__, __func__, __FILE__)

#define HIPCHECK(call)                                                                            \
    do {                                                                                          \
        hipError_t err_ = (call);                                                                 \
        if (err_ != hipSuccess) {                                                                 \
            errs_register(engine_err_gpu, hipGetErrorString(err_), __LINE__, __func__, __FILE__); \
            return engine_err_gpu;                                                                \
        }                                                                                         \
    } while (0)

/* ------------------------------------------------------------ parameters */

struct nb_params {
    int ncells, nparts, nids, max_type, proj;
    int dmaxdist;
    double dscale, bias, cutoff2;
    double dim[3], h[3];
    float rc2f, band;
    double cutoff;
    int tshift, imask; /* packed record: type = w >> tshift, slot = (w >> 16) & imask */
    /* Pseudo-Verlet: per-cell maxdx widening each pair window
       (runner_dopair.c:428), or NULL (no widening: dmaxdist). */
    const double *maxdx;
    int maxn, maxnbr;
    int ptab_bytes, coef_bytes; /* sizes of the table headers and coefficients */
    /* dynamic LDS layout (byte offsets) */
    int lds_own_key, lds_kfl, lds_stage, lds_hit, lds_dense, lds_jf;
    int nsplit; /* neighbour groups per owner cell (work items per cell) */
    const int *order; /* owner cell of each owner position in the work order */
};

/* Neighbour-list entry of an owner cell: x = other cell, y = sid | orient<<4
   | self<<5 (orient 0: owner is the task's i cell), z = dnshift with the
   reference window |shift|, w = dnshift with the projected shift.n.
   Type-pair table header ptab[ti*mt+tj] = (alpha0, alpha1, alpha2,
   bits(n | chunk0 << 9)) or w = bits(-1) when the pair has no potential. */

struct gpu_state {
    int device;
    hipStream_t stream;
    hipEvent_t ev[8];
    int ncells, ntab, max_type, nparts_cap, nids_cap, maxcount, maxnbr;
    int maxcell; /* particles per cell the packed record holds */
    int nparts, nids;
    struct nb_params P;
    /* [0] table mode, [1] surrogate mode, [2] table mode with the 5-wave
       budget, [3] / [4] the half-shell table kernels with 4 / 5 waves, [5]
       the half-shell surrogate kernel, [6] / [7] the paired table kernels
       with 4 / 5 waves */
    int grid_nb[8], ncu;
    size_t lds_nb[8];
    /* static */
    float4 *d_ptab;
    float *d_coeffs;
    int4 *d_nbr;
    float4 *d_nbs; /* per neighbour entry: FP32 minimum-image shift owner -> neighbour */
    int *d_nbr_off;
    /* the same entries with those of opposite offsets side by side: the
       first of a pair (its owner is the task's i cell) has bit 6 of .y set,
       its partner bit 7 */
    int4 *d_nbr_p;
    float4 *d_nbs_p;
    int paired; /* set unless engine_flag_unpaired */
    /* half shell: each task once at its i cell; the self entry of each cell;
       the entries that visit each cell as a task's j cell; per entry the
       forces on the visited atoms (jcap slots) */
    int4 *d_nbr_h;
    float4 *d_nbs_h;
    int *d_nbr_off_h, *d_self_entry, *d_inv_off, *d_inv_list;
    int n_entries_h, jcap, half, mlp_half;
    float4 *d_jpart;
    /* split owners: nsplit neighbour groups per cell; per group the owner
       forces of every atom (pcap float4 allocated in each) */
    int nsplit;
    size_t pcap;
    float4 *d_phi, *d_plo;
    double *d_origin;
    int *d_loc;
    unsigned int *d_sortflags;
    int *d_excl_off, *d_excl_list;
    /* per pass */
    int2 *d_cnt_off;
    int *d_count, *d_offset;
    double *d_x;
    int *d_type, *d_id, *d_index_of_id, *d_cell_of;
    float4 *d_xf, *d_fhi, *d_flo;
    unsigned int *d_sortlist;
    unsigned short *d_keyof;
    double *d_cell_acc, *d_tot;
    unsigned long long *d_cell_cnt, *d_tot_cnt;
    double *d_red_acc; /* per-block partial sums of the totals */
    unsigned long long *d_red_cnt;
    int *d_queue;
    int *d_order; /* owner cells in blocked work order */
    /* pseudo-Verlet: per-cell maxdx on the device; whether the sort lists
       are those of the last sort of the current layout, and its counts */
    double *d_cell_maxdx, *h_cell_maxdx;
    int sort_valid;
    int *h_sort_count;
    /* pinned host staging */
    double *h_x;
    int *h_type, *h_id, *h_count, *h_offset;
    int2 *h_cnt_off;
    float4 *h_fhi, *h_flo;
    double *h_v; /* pinned: velocities to and from the resident steps */
    int *h_ioi;  /* pinned: id -> slot of the last pass, downloaded */
    double *h_tot;
    unsigned long long *h_tot_cnt;
    /* species pairs = distinct potentials, in table order */
    int *h_tabof;      /* [mt*mt] type pair -> pair id or -1 */
    int *h_pair_ti, *h_pair_tj;
    int *d_ptype;      /* [mt*mt] type pair -> pair id or -1 */
    /* MLP surrogate (mlp_gpu.hip); mlp_np == 0 until engine_mlp_init */
    int mlp_np, mlp_nparams;
    long long mlp_step;
    float *d_theta, *d_grad, *d_adam_m, *d_adam_v, *d_gb1e, *d_b1e, *d_loss, *d_wpart;
    float *d_tpart, *d_err; /* per-tile gradient rows; bad-pair flag */
    mlp_bf16x8 *d_w2f, *d_w2b;
    float4 *d_msc, *h_msc;  /* per pair (r_lo, r_hi, sE, r_switch) */
    float *d_mrlo2;         /* [mt*mt] per type pair: r_lo^2 of its species pair (FP32 product), +inf without one */
    float *d_msh, *d_mshy;  /* per pair energy shift; y at r_hi (scratch of the shift) */
    float *d_sr, *d_sE, *d_sF;
    int *d_sp, *d_tiles;
    long long d_samples_cap;
    /* engine_gpu_steps' device buffers (step_gpu.hip), kept between calls
       and grown as needed; NULL until the first call */
    void *steps;
};

/* Device view of the published network of g. */
static inline struct mlp_net gpu_mlp_net(const struct gpu_state *g) {
    struct mlp_net N;
    N.w2f = g->d_w2f;
    N.w2b = g->d_w2b;
    N.w1 = g->d_theta + MLP_OFF_W1;
    N.b1e = g->d_b1e;
    N.b2 = g->d_theta + MLP_OFF_B2;
    N.w3 = g->d_theta + MLP_OFF_W3(g->mlp_np);
    N.b3 = g->d_theta + MLP_OFF_B3(g->mlp_np);
    N.sc = g->d_msc;
    N.rlo2 = g->d_mrlo2;
    N.shift = g->d_msh;
    return N;
}

void gpu_mlp_free(struct gpu_state *g);
void gpu_steps_free(struct gpu_state *g);

/* engine_gpu.hip: pack the host cells and upload them (no kernels); launch
   one pass on the resident particles (mlp: surrogate mode; sort == 0 keeps
   the sort lists of the last sort), with events around each kernel when
   timed. */
int gpu_upload(struct engine *e);
int gpu_launch(struct gpu_state *g, int timed, int mlp, int sort);

#endif /* GPU_STATE_H */


