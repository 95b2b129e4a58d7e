/* kelp_grid_hash.h -- grid-hash set-up helpers of the MI355X particle
 * library (the C ABI of kelp_hip.h; this header adds the per-sample
 * statistics of the reference's occupancy cell rule).
 *
 * The reference picks the grid-hash cell of a mesh by measuring a ladder of
 * candidate cells on sampled element centres (choose_grid_hash_cell,
 * src/kelp/core/mesh.py:464, applied by apply_grid_hash_cell_rule :619 from
 * run_tracking, src/kelp/tracking.py:3378-3386); the per-sample numbers
 * come from _gh_sample_stats (core/mesh.py:386).  This entry point computes
 * those numbers on the device for one candidate grid; the ladder and the
 * choice stay on the host.
 */
#ifndef KELP_GRID_HASH_H
#define KELP_GRID_HASH_H

#include "kelp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* For n sampled points (lon, lat: device doubles) with their exact nearest
 * node nn (device int32, mesh numbering), against the candidate grid
 * (cell, lon_min, lat_min, nx, ny, wrap) whose node counts are given as the
 * 2-D prefix sum cum ((ny+1) x (nx+1) int32, device; cum[j][i] = nodes in
 * rows < j, columns < i): the 3x3 window's node count (occ, int64), the
 * exact search's scan-cost proxy (cost: node distance evaluations + cell
 * visits + 20 per certificate evaluation), whether the search must widen
 * beyond the 3x3 window (unprov) and whether the true nearest node lies
 * outside it (approx).  Outputs are device arrays of n.  Replaces
 * core/mesh.py:386 _gh_sample_stats. */
int kelp_gh_sample_stats(kelp_ctx *ctx, int64_t n, const double *lon, const double *lat, const int32_t *nn,
                         const int32_t *cum, double cell, double lon_min, double lat_min, int64_t nx,
                         int64_t ny, int wrap, int64_t *occ, double *cost, int8_t *unprov, int8_t *approx);

#ifdef __cplusplus
}
#endif
#endif
