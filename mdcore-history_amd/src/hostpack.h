/* Host-side packing between the cells (struct space) and the device's
   cell-ordered arrays, parallel over cells (OpenMP).  Internal to this
   library: the GPU host code calls these instead of looping itself. */
#ifndef MD_HOSTPACK_H
#define MD_HOSTPACK_H

#include "mdcore.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Cell-ordered slots: count[c], offset[c] (offset[nr_cells] = total),
   cnt_off[2c..2c+1] = (count, offset); x (3 per slot), type, id.  *maxc =
   largest count, *maxid = largest id.  Returns -1 if a type is outside
   [0, max_type), else the number of particles. */
int space_pack(const struct space *s, int max_type, int *count, int *offset, int *cnt_off, double *x, int *type,
               int *id, int *maxc, int *maxid);

/* Velocities in the same slot order (offset from space_pack). */
void space_pack_v(const struct space *s, const int *offset, double *v);

/* p->f += fhi + flo of slot ioi[p->id] (float4 pairs, x y z w) for every
   particle.  Returns -1 if a particle has no slot in [0, n). */
int space_add_forces(struct space *s, const int *ioi, int n, const float *fhi, const float *flo);

/* p->f += fhi + flo of slot offset[c] + k for the k-th particle of cell c
   (the layout space_pack made, unchanged since). */
void space_add_forces_slots(struct space *s, const int *offset, const float *fhi, const float *flo);

#ifdef __cplusplus
}
#endif

#endif
