/*
 * C-ABI boundary between the mdcore host library and the gfx950 nonbonded
 * engine.  These entry points replace the reference's CUDA boundary
 * (src/engine.h:381-398, implemented in src/runner_cuda.cu and
 * src/engine_cuda.cu) one for one; the reference dispatches to them from
 * engine_start (engine.c:1747-1761) and engine_step (engine.c:2164-2172)
 * when engine_flag_cuda (bit 4) is set, and so does this host library.
 *
 * Conventions (same as the reference boundary):
 *   - every call returns engine_err_ok (0) or a negative engine_err_* code
 *     after pushing a message on the errs stack; HIP runtime failures map to
 *     engine_err_gpu with the HIP error string;
 *   - all calls come from the thread that calls engine_step;
 *   - forces are ADDED into the host part->f (zeroed by space_prepare) and the
 *     energy is ADDED into e->s.epot / e->s.epot_nonbond.
 * Differences: device state lives in the opaque e->gpu (freed by
 * engine_gpu_finalize); the double-precision host build drives the FP32
 * device arithmetic; exclusion subtraction runs on the device in the same
 * launch (the reference runs exclusion_eval on the host, engine_bonded.c:1618).
 */
#ifndef ENGINE_GPU_H
#define ENGINE_GPU_H

#include "mdcore.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces engine_cuda_setdevice (engine.h:396, engine_cuda.cu:121-140):
   select the HIP device used by this engine. */
int engine_gpu_setdevice(struct engine *e, int id);

/* Replaces engine_cuda_load (engine.h:392, runner_cuda.cu:2931-3545):
   upload the type-pair tables (rounded to FP32), the per-cell neighbour
   lists derived from the task list, the exclusion lists and the work
   queues; allocate device buffers. */
int engine_gpu_load(struct engine *e);

/* Replaces engine_nonbond_cuda (engine.h:391, runner_cuda.cu:2287-2628):
   one nonbonded pass.  Uploads the cell-ordered positions, runs the sort,
   pair/self and exclusion kernels, reduces the energy in FP64 and adds
   forces and energy into the host particles. */
int engine_nonbond_gpu(struct engine *e);

/* Replaces engine_cuda_load_parts / engine_cuda_unload_parts
   (engine.h:393-394, runner_cuda.cu:2641-2825): explicit host<->device
   synchronisation.  what is a bitmask of engine_gpu_sync_*. */
#define engine_gpu_sync_parts_up     1  /* host positions -> device          */
#define engine_gpu_sync_forces_down  2  /* device forces/energy -> host      */
#define engine_gpu_sync_sort_down    4  /* device sortlists -> cell sortlist */
int engine_gpu_sync(struct engine *e, unsigned int what);

/* No reference counterpart (the reference never frees device memory):
   release everything engine_gpu_load allocated. */
int engine_gpu_finalize(struct engine *e);

/* Run only the device part of a nonbonded pass on resident data, nsteps
   times after nwarm untimed passes, and report the mean duration of each
   kernel measured with HIP events on the stream the kernels run on:
   ms[0] = sort (incl. FP32 packing), ms[1] = pair/self/exclusion kernel,
   ms[2] = energy reduction, ms[3] = whole pass.  Inputs must have been
   uploaded by engine_gpu_sync(e, engine_gpu_sync_parts_up).  New; the
   reference times with cudaEvents around load/run/unload
   (runner_cuda.cu:2307-2316, 2606-2622). */
int engine_gpu_time(struct engine *e, int nsteps, int nwarm, double *ms);

/* Device results of the last pass, in FP64 (new): out[0] = epot (pairs),
   out[1] = epot_exclusion, out[2] = virial; counts[0] = pairs evaluated,
   counts[1] = exclusions subtracted, counts[2] = pairs evaluated that the
   reference window also admits (equal to counts[0] unless
   engine_flag_projwindow is set). */
int engine_gpu_results(struct engine *e, double *out, long long *counts);

/* nsteps calls of engine_step (engine.c:2070-2240: engine_shuffle at 2101,
   the nonbonded pass, engine_advance at 2191 and 1945-2052 with
   space_shuffle_local, space.c:307-364) with the particles resident on the
   device.  Positions, velocities, the per-cell particle order, the last
   pass's forces and the energies left in e->s are bitwise those of the
   host-driven loop; e->time advances by nsteps.  epot[k] (may be NULL)
   receives the pass energy of step k (pairs + exclusions, what e->s.epot
   holds after that engine_step) and ekin[k] (may be NULL) the kinetic
   energy sum 1/2 m v^2 after its kick, summed per cell in FP64.  No
   reference counterpart: the reference's CUDA path returns to the host
   every step (runner_cuda.cu:2405-2430, 2514-2521).  As in engine_step of
   this library, the nonbonded forces are the only forces integrated. */
int engine_gpu_steps(struct engine *e, int nsteps, double *epot, double *ekin);

/* ------------------------------------------------------------------------
   MLP surrogate of the pair potential.  No reference counterpart: the
   reference evaluates only its piecewise-polynomial tables
   (potential_eval.h:67-104), which remain the measure of its error and the fallback.

   A species pair is a distinct potential of the engine; pair ids number
   them in the order engine_gpu_load found them.  Per pair p the network
   covers r in [r_lo[p], r_hi[p]] with one output y(r), from inputs
   (r, 1/r) and an 8-dim learned pair embedding: 64 -> 64 hidden units
   (SiLU), bf16 weights on MFMA with fp32 accumulation for the 64x64 layer,
   and a 64 -> 1 head per pair.  Its r-derivative yd is carried through
   both layers alongside (forward mode), and
     E     = S(r) (escale[p] sinh(y) - shift[p]),
     dE/dr = S'(r) (escale[p] sinh(y) - shift[p]) + S(r) escale[p] cosh(y) yd:
   the force is the exact derivative of the energy, so the dynamics conserve
   it.  S is the reference's switch (potential.c:69-100) from rswitch[p] to
   r_hi[p] (S = 1 when rswitch[p] >= r_hi[p]): with a switch the pair energy
   and force go to zero at r_hi.  shift[p] is set after every change of the
   parameters (engine_mlp_init, _train, _params) so that
   escale sinh(y(r_hi)) - shift equals the FP64 table's energy at r_hi.
   Parameter vector: fp32, engine_mlp_init returns its length.  With
   engine_flag_mlp set, engine_nonbond_gpu evaluates the pairs with
   r_lo[p] <= r <= cutoff through the network and the others through the
   tables. */
#define engine_flag_mlp 131072
/* Pair-kernel shell, chosen at every particle upload (each pass, each
   engine_gpu_steps call; no reference counterpart):
   full shell    every pair evaluated by both of its atoms' cells, forces
                 kept in registers;
   half shell    every pair task visited once, the opposite forces added in
                 LDS per visit and folded per atom in a fixed order.
   Without either flag the half shell is taken when the cells hold
   engine_gpu_half_mean atoms or more on average and are fewer than the
   device's compute units (or engine_flag_unpaired is set): it saves work
   there.  Otherwise the paired full shell is faster (100k water, 1000 cells
   of 100 atoms: 0.69 against 0.75 ms), and with few atoms per cell the half
   shell's per-visit cost outweighs its halved pair work.
   The surrogate pass (engine_flag_mlp) takes the half shell unless
   engine_flag_fullshell is set: its network evaluation dominates a visit. */
#define engine_flag_fullshell 262144
#define engine_flag_halfshell 524288
#define engine_gpu_half_mean 48
/* Owner split (no reference counterpart): with fewer than 32 cells per
   compute unit and engine_gpu_half_mean or more atoms per cell on average,
   engine_gpu_load splits each owner cell's neighbour list into up to 8
   groups, each a work item of the pair kernel with its own partial forces
   and sums, folded per atom in group order.  engine_flag_nosplit keeps one
   item per cell. */
#define engine_flag_nosplit 1048576
/* Paired visits (no reference counterpart): the full-shell table kernel
   visits the neighbour cells at opposite offsets +d and -d of an owner cell
   together, so that each lane walks its atom's window into both (long on one
   side where it is short on the other); with the reference window only
   (engine_flag_projwindow passes visit one neighbour at a time).
   engine_flag_unpaired visits them one at a time. */
#define engine_flag_unpaired 2097152

/* Species pairs: ti[p], tj[p] (ti <= tj) for p < returned count. */
int engine_mlp_pairs(struct engine *e, int *ti, int *tj, int max);

/* Create the network for every species pair with the given domains, energy
   scales and switch starts (rswitch[p] > r_lo[p]; >= r_hi[p] for no switch)
   and seeded random weights; returns the parameter count. */
int engine_mlp_init(struct engine *e, const double *r_lo, const double *r_hi, const double *escale,
                    const double *rswitch, unsigned long long seed);

/* Copy the fp32 parameters out (get != 0) or in (get == 0); theta == NULL
   returns the parameter count. */
int engine_mlp_params(struct engine *e, float *theta, int get);

/* Train on n samples (r, pair id, target E, target dE/dr) on the GPU: steps
   Adam steps over shuffled tiles of 16 same-pair samples, batch_tiles tiles
   per step, learning rate cosine-decayed from lr0 to lr1; loss per sample
   wE (y - a)^2 + wF (yd - ad)^2 / (ad^2 + 1/c^2) with a = asinh(E/escale),
   ad = F/sqrt(escale^2 + E^2) (its r-derivative for the target's dE/dr = F)
   and c = (r_hi - r_lo)/2: the output against a, and its r-derivative
   against ad, relatively where |ad| is large (the repulsive wall).  The
   targets are the unswitched potential; the switch is applied exactly at
   evaluation.  loss_hist[steps] (may be NULL) receives the mean loss of
   each step's batch. */
int engine_mlp_train(struct engine *e, const float *r, const int *pair, const float *E, const float *F,
                     long long n, int steps, int batch_tiles, double lr0, double lr1, double wE, double wF,
                     unsigned long long seed, float *loss_hist);

/* Gradient of the mean loss over n samples (n a multiple of 16, each run of
   16 samples one pair id) with respect to the parameters, without an
   update; grad has engine_mlp_params(e, NULL, 0) entries. */
int engine_mlp_grad(struct engine *e, const float *r, const int *pair, const float *E, const float *F,
                    long long n, double wE, double wF, float *grad, double *loss);

/* No reference counterpart; for callers that cannot allocate device memory
   (code under HIP graph capture, tight training loops, harnesses that forbid
   hipMalloc).  engine_mlp_grad on DEVICE buffers r, pair, E, F (n samples)
   and outputs grad (engine_mlp_params(e, NULL, 0) floats) and loss (one
   float), with no hipMalloc, hipFree or other allocation: its scratch is a
   static per-device array of 6.8 MB set up by engine_mlp_init.  Limits:
   16 <= n <= 65536, n a multiple of 16, at most 1024 species pairs; beyond
   them it returns engine_err_range.  Asynchronous on the engine's stream:
   the outputs are valid once the device is synchronized.  A pair id outside
   the network's pairs makes *loss NaN.  Calls that overlap on one device
   race on the shared scratch. */
int engine_mlp_grad_noalloc(struct engine *e, const float *r, const int *pair, const float *E, const float *F,
                            long long n, double wE, double wF, float *grad, float *loss);

/* Evaluate the network on n samples: E[k] and dE/dr F[k], switch and shift
   included. */
int engine_mlp_eval(struct engine *e, const float *r, const int *pair, long long n, float *E, float *F);

#ifdef __cplusplus
}
#endif

#endif /* ENGINE_GPU_H */
