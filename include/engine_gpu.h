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

#ifdef __cplusplus
}
#endif

#endif /* ENGINE_GPU_H */
