/* kelp_oracle.h -- CPU oracle for the KELP particle time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C, scalar restatement of the
 * reference's Numba kernels, used by tests/, the smoke check and the
 * benchmark's CPU-baseline leg as the checker.  The product (kelp_amd)
 * never links, loads or calls it.
 *
 * Two builds of one source:
 *   liboracle_glibc.so  transcendental functions from the host glibc libm,
 *                       as the reference's Numba JIT lowers them;
 *   liboracle_kmath.so  the same code with kelp_amd/csrc/kmath.h in place
 *                       of libm: bit-for-bit the arithmetic of the device
 *                       kernels, so device results can be compared exactly.
 */
#ifndef KELP_ORACLE_H
#define KELP_ORACLE_H

#include <stdint.h>

/* field storage dtypes */
#define OR_F64 0
#define OR_F32 1

typedef struct {
    int32_t n_nodes, n_elems, max_v, max_conn;
    const double *nodes_lon, *nodes_lat, *depth;
    const int32_t *elem_nodes;   /* n_elems * max_v, -1 padded */
    const int32_t *n2e;          /* n_nodes * max_conn, -1 padded */
    const int32_t *n2e_count;    /* n_nodes */
    const double *node_xyz;      /* n_nodes * 3 */
    const int32_t *gh_offsets;   /* nx*ny + 1 */
    const int32_t *gh_node_ids;  /* n_nodes */
    double gh_cell, gh_lon_min, gh_lat_min;
    int32_t gh_nx, gh_ny, gh_wrap, gh_far_rows;
    double r_earth;
    /* vertical (3-D): interfaces z_levels[K], positive down */
    int32_t n_levels;
    const double *z_levels;
    const uint8_t *elem_coastal;  /* (E,) coastal-element flags or NULL */
    const double *elem_center_lon, *elem_center_lat;  /* (E,) element centres */
    int32_t vertical_coord;      /* 0 z-level; 1 sigma: z_levels holds sigma_lev (1 .. 0) */
    const int32_t *elem_neighbors;  /* (E, max_v) core/mesh.py:2824, or NULL */
} or_mesh_t;

/* Snapshot triple of one step: bracket pair (a, b) blended by alpha0 and
 * the stage target t; each array holds n_layers planes of n_nodes values
 * (2-D fields: one plane). */
typedef struct {
    const void *u_a, *v_a, *u_b, *v_b, *u_t, *v_t;
    const void *w_a, *w_b, *w_t;   /* 3-D: (K, N) interface W, or NULL */
    int32_t dtype;
    double alpha0;
    int32_t uv_layout;             /* 0: u/v at nodes, 1: u/v at element centres */
    const double *eta_a, *eta_b;   /* (N,) surface elevation of the bracket pair, or NULL */
} or_fields_t;

/* One step's configuration (the keyword arguments of step_fn,
 * tracking.py:2643-2655, plus the beaching hook's per-run constants). */
typedef struct {
    int32_t order;     /* 1 Euler, 2 RK2, 4 RK4 */
    int32_t is3d;
    int32_t w_method;  /* 0 nearest (linear at the anchor), 1 leastsq, 2 barycentric, 3 idw */
    int32_t exact;     /* mesh.exact_nearest */
    double dt, time_dir;
    double alphas[4];  /* stage fractions towards the target snapshot */
    /* split uniform-Kh kick (integrators.py:673) */
    int32_t kick;
    int32_t beach;
    double kh_x, kh_y;
    uint64_t seed, real_id, step_i;
    /* beaching (numerics/beaching.py:95, tracking.py:199) */
    int32_t use_depth, use_coastal, use_stochastic, on_failure;
    int32_t resuspension, min_wet_steps;
    double depth_min, zone_depth, t_beach, rewet_hysteresis, t_resuspend;
    uint64_t beach_seed;
    double beach_dt;
    /* vertical Az kick (integrators.py:842): the constant source, or the
     * step payload's Az field (K, N) at the interfaces (vel.az, already
     * blended over the bracket) when az_field is not NULL */
    int32_t az;
    double az_const;
    const double *az_field;
    int32_t use_eta;   /* beaching water depth = depth + eta (beaching.py:121) */
    /* periodic fold windows [x0, x1) / [y0, y1) (tracking.py:1089) */
    int32_t periodic_lon, periodic_lat;
    double px0, px1, py0, py1;
    /* sigma meshes: the surface elevation of stages 0..3 and of the step
     * end, (N,) each (integrators.py:1381-1441), the column floor and the
     * bed distance */
    const double *sig_eta[5];
    double Dmin, z0b_min;
    int32_t walk;      /* elem_search walk: containing elements by the edge walk */
    /* order 5 = rk45 (integrators.py:6448): error tolerance (m) per substep,
     * smallest substep (s), 1 / dt_to_next (0: steady), substep cap */
    double rk45_tol_m, rk45_h_min, rk45_inv_dtn;
    int64_t rk45_max_sub;
} or_step_args_t;

int or_set_threads(int n);

void or_gh_sample_stats(const or_mesh_t *m, int64_t n, const double *lon, const double *lat, const int32_t *nn,
                        const int32_t *cum, int64_t *occ_out, double *cost_out, int8_t *unprov_out,
                        int8_t *approx_out);

#endif
