/*
 * pylith_amd_debug.h -- test hooks of libpylith_amd.so (no counterpart in the reference).
 *
 * plx_debug_emulate_apply runs the tile plan and the element math of the gfx950 residual /
 * Jacobian-action kernels on the CPU, bounds-checking every LDS, halo and vector index, so the
 * kernels' indexing can be verified on a machine without a GPU.  Returns 0, 1 (bad
 * arguments) or 3 (an index out of bounds, or two cells of one colour touching one node).
 *   lam_mu [num_cells][2]: Lame lambda and shear modulus per cell (original cell order);
 *   algorithm: plx_algorithm (the element path to emulate).
 */
#ifndef PYLITH_AMD_DEBUG_H
#define PYLITH_AMD_DEBUG_H

#include <stdint.h>

#include "pylith_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tile-plan statistics of a key: tiles, halo slots (shared-node partial sums per apply) and
 * whether the planner took the cells in Morton order of their centroids (input order tiled with
 * more than 1.5 halo slots per cell). */
int plx_debug_plan_info(plx_dm dm, int label_value, int part, int *num_tiles, int64_t *halo_slots,
                        int *reordered);

int plx_debug_emulate_apply(int cell_type, int num_cells, int num_vertices, int num_nodes,
                            const int32_t *closure, const int32_t *cone, const double *coords,
                            const int32_t *labels, int label_value, const double *lam_mu,
                            int scalar_type, const void *u, void *r, int mode, int algorithm);

#ifdef __cplusplus
}
#endif
#endif
