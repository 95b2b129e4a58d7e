/* kelp_hip.h -- C ABI of the MI355X particle time-stepping library.
 *
 * The reference (CliDyn/kelp) is Python + Numba; its step boundary is the
 * integrator registry (src/kelp/numerics/integrators.py:6635-6660) called
 * once per step by run_single_realization (src/kelp/tracking.py:2643-2655)
 * as step_fn(state, vel, mesh, dt, time_dir, ...).  A ctypes binding of the
 * entry points below replaces that call: the mesh and snapshots live in HBM
 * (uploaded once), the particle state lives in caller-owned device buffers,
 * and kelp_step advances every particle by one Euler/RK2/RK4 step in place.
 *
 * Conventions: every function returns KELP_OK (0) or a negative error code;
 * kelp_last_error() describes the last failure of the calling thread.  All
 * pointers named host_* are host memory; pointers in kelp_particles are
 * device memory.  A context owns one HIP stream (default priority) and is
 * not re-entrant: calls on one context must be serialised by the caller.
 */
#ifndef KELP_HIP_H
#define KELP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KELP_OK 0
#define KELP_EINVAL (-1)  /* invalid argument / shape */
#define KELP_EHIP (-2)    /* HIP runtime error */
#define KELP_ENOMEM (-3)  /* device allocation failed */
#define KELP_ESTATE (-4)  /* call out of order (no mesh, empty slot, ...) */

#define KELP_DT_F64 0
#define KELP_DT_F32 1
/* bfloat16 storage of node-layout z-level snapshots (kelp_field_upload*: the
 * host planes are float32, rounded to nearest even on the device); read
 * back exactly as float32 values.  No counterpart in the reference, which
 * keeps float32 or float64 snapshots: halves the records' HBM footprint. */
#define KELP_DT_BF16 2
/* float16 storage with a power-of-two scale per level (node-layout z-level
 * snapshots; host planes float32): every u/v layer and every W interface
 * gets the scale 2^(q - 15) of its largest finite |value| m = f 2^q (u and v
 * of a layer share the larger), the values are stored as half(value /
 * scale) rounded to nearest even and read back exactly as half * scale.
 * No counterpart in the reference (float32 / float64 snapshots). */
#define KELP_DT_F16S 3

/* kelp_step_args.sig_start_a / sig_end_a: the surface given by
 * kelp_sigma_surface_upload instead of a blend of slot surfaces. */
#define KELP_SURFACE_GIVEN (-2)

typedef struct kelp_ctx kelp_ctx;

/* Mesh contract of the reference's Mesh dataclass (src/kelp/core/mesh.py:41)
 * plus its grid hash (build_grid_hash, core/mesh.py:844). */
typedef struct {
    int64_t n_nodes, n_elems, max_v, max_conn;
    const double *host_nodes_lon, *host_nodes_lat; /* (N,) degrees */
    const double *host_depth;                      /* (N,) m, positive down; may be NULL (2-D) */
    const int32_t *host_elem_nodes;                /* (E, max_v), -1 padded; max_v 3 or 4 */
    const int32_t *host_node_to_elems;             /* (N, max_conn), -1 padded */
    const int32_t *host_node_to_elems_count;       /* (N,) */
    const int32_t *host_gh_offsets;                /* (nx*ny + 1,) CSR over cells */
    const int32_t *host_gh_node_ids;               /* (N,) */
    double gh_cell, gh_lon_min, gh_lat_min;
    int64_t gh_nx, gh_ny, gh_wrap, gh_far_rows;
    double r_earth;
    int64_t n_levels;                              /* K interfaces (0 for 2-D) */
    const double *host_z_levels;                   /* (K,) m, positive down */
    const uint8_t *host_elem_coastal;              /* (E,) coastal flags (core/mesh.py:2765) or NULL */
    const double *host_elem_center_lon;            /* (E,) element centres (Mesh.elem_center_lon/lat); */
    const double *host_elem_center_lat;            /*   needed only for element-centre u/v */
    int64_t vertical_coord;                        /* 0 z-level (host_z_levels = interface
                                                      depths, non-decreasing); 1 sigma
                                                      (Mesh.vertical_coord == "sigma",
                                                      host_z_levels = sigma_lev, 1 at the
                                                      surface .. 0 at the bottom,
                                                      non-increasing) */
    const int32_t *host_elem_neighbors;            /* (E, max_v) core/mesh.py:2824
                                                      build_elem_neighbors (mesh ids, -1
                                                      boundary) or NULL; needed by the
                                                      edge walk (elem_search: walk) */
} kelp_mesh_desc;

/* Particle state, device memory, structure of arrays (the fields of the
 * reference's ParticleState, core/particles.py:153).  Updated in place. */
typedef struct {
    int64_t n;
    double *lon, *lat, *z;
    int32_t *status;   /* ParticleStatus: 1 alive, 2 beached, 3 outside, 4 not released */
    int32_t *nn;       /* nearest node, internal numbering (see kelp_nodes_to_mesh_ids) */
    int8_t *fr;        /* FailureReason (sticky; written on the step a particle dies) */
    double *u, *v, *w; /* diagnostic velocity at the step start (w may be NULL in 2-D) */
    int64_t *pid;      /* global particle id = RNG key; needed with the kick, beaching, sorting */
    int32_t *wet_streak; /* beaching wet streak (tracking loop state); needed with beaching */
    uint64_t *alive_count; /* optional: += particles alive at the step start */
    int32_t *elem;     /* carried containing element (internal numbering, -1 unknown), with
                          elem_search walk (ParticleState.elem); NULL otherwise */
} kelp_particles;

/* One step (the arguments of step_fn at tracking.py:2643-2655 and the
 * beaching hook's per-run constants, tracking.py:199-359).  A field slot
 * holding W makes the step 3-D (z-level, mesh n_levels >= 2). */
typedef struct {
    int32_t order;                 /* 1 Euler, 2 RK2 (midpoint), 4 RK4 */
    int32_t slot_a, slot_b, slot_t; /* bracket pair (a, b) and stage target t; b/t may be -1 */
    double alpha0;                 /* bracket fraction of a->b at the step start */
    double alphas[4];              /* per-stage fraction towards t (c_k*dt/dt_to_next) */
    double dt;                     /* positive seconds */
    double time_dir;               /* +1 forward, -1 backward */
    int32_t exact_nearest;         /* anchor: 1 exact (mesh.exact_nearest, the default);
                                      0 guard off, fused kernels' widened window;
                                      2 guard off, batch 3x3 window healed exactly */
    int32_t w_method;              /* 3-D W: 0 linear at the anchor ("nearest"), 1 leastsq,
                                      2 barycentric, 3 idw */
    int32_t kick;                  /* split uniform-Kh random walk (integrators.py:673) */
    int32_t beach;                 /* beaching hook */
    double kh_x, kh_y;             /* m^2/s */
    uint64_t seed, real_id, step_i; /* counter-RNG key: (seed, real_id, pid, step_i, comp) */
    int32_t use_depth, use_coastal, use_stochastic, on_failure; /* beaching criteria */
    int32_t resuspension;          /* 0 never, 1 auto, 2 stochastic, 3 auto_or_stochastic */
    int32_t min_wet_steps;
    double depth_min, zone_depth, t_beach, rewet_hysteresis, t_resuspend;
    double beach_dt;               /* step length seen by the hook (abs() taken) */
    uint64_t beach_seed;
    int32_t az;                    /* vertical Az kick after the horizontal one
                                      (integrators.py:842), constant source, 3-D */
    double az_const;               /* m^2/s; 0 = no kick */
    int32_t use_eta;               /* beaching: water depth = depth + eta of the step's
                                      field (beaching.py:121; needs kelp_field_upload_eta) */
    int32_t periodic_lon, periodic_lat; /* fold ALIVE particles back into [px0, px1) /
                                      [py0, py1) after the kicks, before beaching
                                      (tracking.py:1089 _apply_periodic_wrap) */
    double px0, px1, py0, py1;
    int32_t reuse_anchor;          /* the caller vouches that every ALIVE particle's nn >= 0
                                      is the anchor of its current position under this
                                      exact_nearest mode (true after a kelp_step in the same
                                      mode with the positions left alone; set nn = -1 for
                                      particles placed or released since): the step skips
                                      that search.  Ignored in mode 0. */
    /* Sigma meshes (integrators.py:1310-1441: the step integrates the sigma
     * fraction s = z / H_p with H_p = max(depth + eta, Dmin) interpolated like
     * W; vertical_sigma.py).  The surface elevation at the step start is
     * the time blend of the eta of slots (a, b, t) with fractions
     * (alpha0, alpha) -- the nested blend of the velocities; b/t -1 when
     * unused -- and likewise at the step end (sig_end_a = -1: the start
     * surface).  Stage surfaces follow the stages' velocity blend, or on a
     * steady step (slot_t = -1) the blend of the start and end surfaces. */
    int32_t sig_start_a, sig_start_b, sig_start_t;
    double sig_start_alpha0, sig_start_alpha;
    int32_t sig_end_a, sig_end_b, sig_end_t;
    double sig_end_alpha0, sig_end_alpha;
    double Dmin;                   /* minimum column height (m) */
    double z0b_min;                /* distance kept above the bed (m) */
    int32_t walk;                  /* elem_search walk (core/mesh.py:2489): every site's
                                      containing element by the edge walk from the carried
                                      one, the domain verdict too; node-layout u/v, z-level */
    /* Az from a field (integrators.py:842 with vel.az, numerics/diffusion.py:427
     * _loop_az_kick_zlevel): with `az` and az_field != 0 the kick takes Az and
     * dAz/dz on the nearest node's column of slot az_slot_a's Az, blended with
     * az_slot_b's by az_alpha0 when az_slot_b >= 0 (the step payload's bracket,
     * core/io.py:68), instead of az_const; z-level 3-D meshes, both slots filled
     * by kelp_field_upload_az */
    int32_t az_field;
    int32_t az_slot_a, az_slot_b;
    double az_alpha0;
    /* order 5 = rk45 (integrators.py:6448 rk45_step): adaptive Dormand-Prince
     * 5(4) per particle over exactly dt, 2-D, u/v at element centres (least
     * squares), steady or blended towards slot_t; the embedded error in
     * metres per substep against rk45_tol_m (<= 0: 1 m), substeps never below
     * rk45_h_min seconds (<= 0: dt 1e-4); rk45_inv_dtn = 1 / dt_to_next (0:
     * steady), stage fractions (t / dt_to_next) take the place of alphas */
    double rk45_tol_m, rk45_h_min, rk45_inv_dtn;
    int32_t block_size;            /* threads per workgroup of the batched step (kelp_batch_*):
                                      64, 128 or 256; 0 = 256.  kelp_step runs 256. */
} kelp_step_args;

/* ---------------------------------------------------------------------------
 * Batched clouds: many independent releases advanced together, one kernel
 * per step.  No single counterpart in the reference: it runs each release
 * (a realization, utils/split_seeds.py:89, tracking.py:1805
 * run_single_realization) as its own loop; here every cloud is such a run,
 * and its particles' results equal that run's (same integrator, dt, fields
 * and draws).  A particle belongs to cloud pid >> KELP_CLOUD_SHIFT and its
 * random draws are keyed by the cloud-local part pid & (2^40 - 1) with the
 * cloud's seed and the cloud-local step, so a cloud's results do not depend
 * on which other clouds share the batch.
 * ------------------------------------------------------------------------- */
#define KELP_CLOUD_SHIFT 40
#define KELP_CLOUD_VALID 1      /* the entry describes a cloud */
#define KELP_CLOUD_3D 2         /* 3-D cloud (the 3-D snapshot axis), else 2-D */
#define KELP_CLOUD_KICK 4       /* split uniform-Kh random walk */
#define KELP_MAX_CLOUDS 65536

typedef struct {
    int32_t flags;       /* KELP_CLOUD_* */
    int32_t n_steps;     /* steps the cloud runs */
    int64_t start_step;  /* batch step of its first step */
    double time_dir;     /* +1 forward, -1 backward */
    double t0;           /* time of its first step on the snapshot axis (s) */
    double kh_x, kh_y;   /* Kh of its random walk (m^2/s) */
    uint64_t seed, real_id; /* counter-RNG key of its kicks and beaching */
} kelp_cloud;

/* The snapshot axes the clouds read: slots slot0 .. slot0 + n_snap - 1 hold
 * snapshots at 0, snap_dt, 2 snap_dt, ... cycled (time.on_exhausted:
 * cycle); n_snap 1 = a steady field.  A cloud at time t steps with the
 * bracket of t (the step payload of tracking.py:3136 _build_vel_payload:
 * alpha0 snapped to 0 / 1 within 1e-12, the stage target b forward and a
 * backward, stage fractions c_k dt / dt_to_next).  2-D clouds read the 2-D
 * axis, 3-D clouds the 3-D one (unused axis: n_snap 0). */
typedef struct {
    int32_t slot0_3d, n_snap_3d;
    int32_t slot0_2d, n_snap_2d;
    double snap_dt;
} kelp_batch_axis;

/* Write cloud entries [first, first + n) of the context's cloud table
 * (grown as needed, at most KELP_MAX_CLOUDS; entries past the old end are
 * cleared) and set the axes.  Takes effect at the next batched step. */
int kelp_batch_set_clouds(kelp_ctx *ctx, int64_t first, int64_t n, const kelp_cloud *host_clouds,
                          const kelp_batch_axis *axis);
/* Start a new table of n_clouds empty entries (no cloud runs) and set the
 * step counter to 0. */
int kelp_batch_reset(kelp_ctx *ctx, int64_t n_clouds);
/* Set the batch step counter (the step the next kelp_batch_step runs). */
int kelp_batch_set_step(kelp_ctx *ctx, int64_t step);
/* Read the batch step counter (waits for the context stream). */
int kelp_batch_get_step(kelp_ctx *ctx, int64_t *step);
/* One batched step: every cloud with start_step <= step < start_step +
 * n_steps advances its particles by one step (the base arguments give the
 * integrator, dt, anchoring, W method, beaching and, on sigma meshes, the
 * surfaces; slot and time fractions come from each cloud's bracket, and
 * time_dir, kick, seed, real_id and step_i from the cloud); particles of
 * other clouds, and entries whose cloud is not in the table (free pool
 * entries), are left alone.  Then the step counter advances.  base->slot_a
 * names a slot of the axes (dtype and shape); on sigma meshes the 3-D axis
 * must be steady.  Node-layout u/v. */
int kelp_batch_step(kelp_ctx *ctx, const kelp_particles *p, const kelp_step_args *base);
/* n_steps batched steps as one HIP graph launch on the context stream: the
 * graph is captured at the first call and reused while the particle
 * arrays, the base arguments, n_steps and the cloud table's allocation stay
 * the same (the table's contents may change between launches). */
int kelp_batch_run_graph(kelp_ctx *ctx, const kelp_particles *p, const kelp_step_args *base, int n_steps);
/* Copy the particles of cloud `cloud` into host-visible or device arrays
 * indexed by their cloud-local id (pid & (2^40 - 1) < n_out): lon, lat, z
 * (may be NULL), status, fr; then, when `release` is set, mark them free
 * (status 0, pid of no cloud), so a re-sort moves them to the end. */
int kelp_batch_collect(kelp_ctx *ctx, const kelp_particles *p, int64_t cloud, int64_t n_out, double *lon,
                       double *lat, double *z, int32_t *status, int8_t *fr, int release);

const char *kelp_last_error(void);
int kelp_version(void);

int kelp_ctx_create(int device, kelp_ctx **out);
int kelp_ctx_destroy(kelp_ctx *ctx);
int kelp_synchronize(kelp_ctx *ctx);
/* The context's HIP stream (hipStream_t) as an opaque pointer. */
void *kelp_ctx_stream(kelp_ctx *ctx);

/* Upload the mesh (replaces any earlier mesh).  Builds the device layout
 * and the node unit-vector table (core/mesh.py:713 _node_xyz_fill). */
int kelp_mesh_upload(kelp_ctx *ctx, const kelp_mesh_desc *desc);
/* Device copy of the node unit-vector table, node order, (N, 3) doubles. */
int kelp_mesh_node_xyz(kelp_ctx *ctx, double *host_out);
/* The library numbers nodes internally (grid-hash scan order) for locality;
 * kelp_particles.nn holds internal ids.  Translate a device array of node ids
 * in place between the mesh's numbering and the internal one (-1 kept).  No
 * reference counterpart (the reference works in mesh numbering throughout). */
int kelp_nodes_to_mesh_ids(kelp_ctx *ctx, int64_t n, int32_t *ids);
/* Element ids (device int32, n) between the mesh's and the internal
 * numbering (kelp_particles.elem), in place. */
int kelp_elems_to_mesh_ids(kelp_ctx *ctx, int64_t n, int32_t *ids);
int kelp_elems_from_mesh_ids(kelp_ctx *ctx, int64_t n, int32_t *ids);
int kelp_nodes_from_mesh_ids(kelp_ctx *ctx, int64_t n, int32_t *ids);

/* Field snapshot slot (the reference's VelocityField, core/io.py:180).
 * u, v: n_planes x N node values in mesh node order (2-D: n_planes = 1; 3-D:
 * K-1 layers); w: K x N interface values positive down, or NULL (2-D).
 * dtype KELP_DT_*.  Stored packed per (layer, node). */
int kelp_field_upload(kelp_ctx *ctx, int slot, int dtype, int64_t n_planes,
                      const void *host_u, const void *host_v,
                      const void *host_w, int64_t n_w_planes);
/* The same with u, v at the element centres (the reference's default
 * uv_layout "element", VelocityField.u of shape (E,) / (K-1, E)): the step
 * interpolates them by least squares over the anchor's element ring
 * (numerics/interpolate.py:52).  W stays at the nodes. */
int kelp_field_upload_elem(kelp_ctx *ctx, int slot, int dtype, int64_t n_planes,
                           const void *host_u, const void *host_v,
                           const void *host_w, int64_t n_w_planes);
/* Attach the snapshot's surface elevation (VelocityField.eta, (N,) doubles
 * in mesh node order) to an uploaded slot; read by beaching with use_eta. */
int kelp_field_upload_eta(kelp_ctx *ctx, int slot, const double *host_eta);
int kelp_field_clear(kelp_ctx *ctx, int slot);

/* Sigma meshes: the surface elevation a step starts (which 0) or ends
 * (which 1) on, given as an array -- the sigma_eta_start / sigma_eta_end
 * keywords of the reference's step functions (numerics/integrators.py:
 * 5642-5643, used at :1381 _sigma_end_eta and :1423 _sigma_given).  (N,)
 * doubles in mesh node order; NULL forgets the surface.  A step selects it
 * with sig_start_a / sig_end_a = KELP_SURFACE_GIVEN. */
int kelp_sigma_surface_upload(kelp_ctx *ctx, int which, const double *host_eta);

/* Node scalar `var` (0..7) of a snapshot slot for record-point sampling
 * (VelocityField.scalars, numerics/sampling.py:727): n_planes x N doubles in
 * mesh node order, n_planes 1 (2-D), K-1 (mid-layer) or K (interface). */
int kelp_scalar_upload(kelp_ctx *ctx, int slot, int var, int64_t n_planes, const double *host_values);

/* Sample scalar `var` at every particle (sample_field): the field of slot_a,
 * or slot_a blended towards slot_b by alpha; method 0 nearest, 1 leastsq,
 * 2 barycentric, 3 idw.  out: device doubles (n), NaN where the particle is
 * not alive or has no anchor.  Particle order as in `p`. */
int kelp_sample(kelp_ctx *ctx, const kelp_particles *p, int var, int slot_a, int slot_b, double alpha,
                int method, double *out);

/* Device rows of one staged record (kelp_record_stage): n_rows entries each
 * in output particle order; w may be NULL (2-D); samples: n_vars x n_rows
 * doubles, or NULL without sampled scalars. */
typedef struct {
    double *lon, *lat, *z, *u, *v, *w;
    int32_t *status;
    int8_t *fr;
    double *samples;
} kelp_record_rows;

/* Record staging: the record append of tracking.py:2757-2781 on the
 * device, for a tracking loop that keeps its records in HBM and copies
 * them out in batches.  Every particle's record goes to row pid - pid_offset
 * of `out`, masked as core/release.py:132 mask_dormant_for_record (lon,
 * lat, z NaN where NOT_RELEASED) and :170 mask_inactive_velocity (u, v, w
 * NaN unless ALIVE both at the step start -- prev_status, device int32 per
 * particle slot, from kelp_activate -- and at its end), the velocities
 * negated when time_dir < 0 (tracking.py:2772-2777); status and failure
 * reason as they are; the n_vars sampled scalars (device doubles, n_vars x
 * n in particle-slot order, as kelp_sample_vars writes them) scattered the
 * same way.  Particles whose pid falls outside [pid_offset, pid_offset +
 * n_rows) are skipped.  Enqueued on the context stream; no wait. */
int kelp_record_stage(kelp_ctx *ctx, const kelp_particles *p, const int32_t *prev_status, int64_t pid_offset,
                      int64_t n_rows, double time_dir, int n_vars, const double *sampled,
                      const kelp_record_rows *out);

/* Release (core/release.py:112 activate_released): status NOT_RELEASED (4)
 * becomes ALIVE where step >= release_step[pid - pid_offset] (device int64,
 * n_release entries; NULL = no release schedule); a released particle's nn
 * becomes -1 (its first step anchors it).  prev_status (device,
 * n int32, or NULL) receives every particle's status after activation,
 * the step-start status the record masks of core/release.py:132 use. */
int kelp_activate(kelp_ctx *ctx, const kelp_particles *p, const int64_t *release_step, int64_t pid_offset,
                  int64_t n_release, int64_t step, int32_t *prev_status);

/* Refill a slot asynchronously (the snapshot ring: the next snapshot
 * streams in while steps run).  The slot must already hold a snapshot of
 * the same dtype, layout and shape (one kelp_field_upload fills it first);
 * host_eta (N doubles, or NULL) also refills its surface.  The copies and
 * the packing run on the context's copy stream after every call already
 * enqueued on the compute stream (a step still reading the old snapshot
 * finishes first); steps and sampling that read the slot wait for the
 * refill on the device.  The host arrays must stay valid until the refill
 * has run (kelp_field_upload_wait); page-locked host memory makes the
 * copies truly asynchronous.  No device allocation after the first refill
 * of a slot.  Replaces the reference's _VelCache prefetch
 * (tracking.py:584-735), which reads ahead on a host thread. */
int kelp_field_upload_async(kelp_ctx *ctx, int slot, int dtype, int layout, int64_t n_planes,
                            const void *host_u, const void *host_v, const void *host_w, int64_t n_w_planes,
                            const double *host_eta);
/* 4-snapshot PCHIP temporal reconstruction (numerics/temporal_reconstruct.py:
 * 218 reconstruct_pchip, the per-value arithmetic of :184
 * _pchip_value_scalar): slot `dst` receives the monotone cubic of slots
 * src[0..3] (same dtype, layout and shape) at times[0..3] (strictly
 * increasing), evaluated at tq in [times[1], times[2]] (within 1e-9 of the
 * interval), every u, v and W value, stored as float64 (allocated at the
 * first call, or when the shape changes).  A step reads a stage field from
 * it like any snapshot.  Runs on the context stream. */
int kelp_field_reconstruct(kelp_ctx *ctx, int dst, const int32_t *src, const double *times, double tq);

/* The vertical diffusivity of a snapshot (readers/fesom2.py:1596-1636: the
 * `var_az` record, float64, NaN as 0; core/io.py:246 VelocityField.az): K
 * interface planes of N values in mesh node order, stored one column per
 * node.  Re-uploading the slot's snapshot (kelp_field_upload*,
 * kelp_field_reconstruct) drops it. */
int kelp_field_upload_az(kelp_ctx *ctx, int slot, int64_t n_planes, const double *host_az);

/* Block the host until the last refill of `slot` has run. */
int kelp_field_upload_wait(kelp_ctx *ctx, int slot);

/* Several variables at once (vars[0..n_vars), host array): row k of `out`
 * (device, n_vars x n doubles) is variable vars[k] as kelp_sample computes
 * it; least-squares pairs of 2-D/mid-layer variables share one pass. */
int kelp_sample_vars(kelp_ctx *ctx, const kelp_particles *p, int n_vars, const int32_t *vars,
                     int slot_a, int slot_b, double alpha, int method, double *out);

/* Advance every particle by one step, in place, on the context stream.
 * Each workgroup stages the grid-hash cells around its particles in LDS
 * (unless KELP_GH_TILE=0 was set when the context was created); same
 * results either way. */
int kelp_step(kelp_ctx *ctx, const kelp_particles *p, const kelp_step_args *a);

/* Reorder every particle array in place by the grid-hash cell of each
 * particle (alive first, then beached, then the rest) so that the step's
 * mesh and field gathers of neighbouring lanes share cache lines.  Results
 * do not depend on the order (draws are keyed by pid); restore the release
 * order from pid.  No reference counterpart: the reference's threads work
 * on release order (numerics/integrators.py prange loops). */
int kelp_sort_particles(kelp_ctx *ctx, const kelp_particles *p);

/* kelp_sort_particles without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  No counterpart in the reference.  The keys, permutation,
 * staging and the radix sort's temporary storage live in a static device
 * array (48 MiB per GPU, held until the process ends); at most 1,048,576
 * particles per call (KELP_EINVAL above that).  All calls on one GPU share
 * that array: calls that overlap on one GPU (several contexts or threads)
 * race and must be serialised by the caller.  Same results as
 * kelp_sort_particles. */
int kelp_sort_particles_noalloc(kelp_ctx *ctx, const kelp_particles *p);

/* Batch queries (core/mesh.py:1896 find_nearest_nodes and :2289
 * find_containing_elements) on device arrays. */
int kelp_nearest(kelp_ctx *ctx, int64_t n, const double *lon, const double *lat,
                 int exact, int32_t *out);
int kelp_containing(kelp_ctx *ctx, int64_t n, const double *lon, const double *lat,
                    const int32_t *nn, int32_t *out);

/* Device memory helpers for callers without their own allocator. */
int kelp_malloc(kelp_ctx *ctx, int64_t bytes, void **out);
int kelp_free(kelp_ctx *ctx, void *ptr);
int kelp_memcpy_h2d(kelp_ctx *ctx, void *dst, const void *src, int64_t bytes);
int kelp_memcpy_d2h(kelp_ctx *ctx, void *dst, const void *src, int64_t bytes);

/* Timing on the context stream: record event `slot` (0..1023) now; elapsed
 * milliseconds between two recorded slots (synchronises on `b`). */
int kelp_event_record(kelp_ctx *ctx, int slot);
int kelp_event_elapsed(kelp_ctx *ctx, int a, int b, float *ms);

/* Diagnostics: evaluate the device math routines elementwise on device
 * arrays (fn: 0 sin, 1 cos, 2 asin, 3 acos, 4 atan2(x, y), 5 log, 6 exp,
 * 7 Python x % 360, 8 sqrt, 9 Python x % y, 10 reflective fold of x into
 * [0, y], 11 / 12 the first / second normal of the random walk's pair for
 * key (seed 20240501, realization 0, particle x, step y, component 0),
 * 13 x^y).  No reference counterpart; used to check the device compiler
 * against a host build of the same routines. */
int kelp_math_probe(kelp_ctx *ctx, int fn, int64_t n, const double *x, const double *y,
                    double *out);

#ifdef __cplusplus
}
#endif
#endif
