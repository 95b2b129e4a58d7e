// kelp_step.cuh -- the per-step configuration shared by the C ABI (host) and
// the step kernels, and the declaration of the kernel launchers.
//
// The step kernel is instantiated for ORDER x MAXV x 3-D x field dtype x u/v
// layout.  The instantiations are spread over several translation units
// (kelp_step_variant.hip compiled once per (3-D, dtype, layout)), so the
// library builds in parallel; this header only declares the launchers.
#pragma once

#include "kelp_device.cuh"

namespace kelp {

// One step's configuration (step_fn keywords, tracking.py:2643-2655, plus the
// beaching hook's per-run constants, tracking.py:199).
struct StepCfg {
    // rk45 (order 5): tolerance (m) per substep, smallest substep (s),
    // 1 / dt_to_next (0: steady), substep cap
    double rk45_tol, rk45_h_min, rk45_inv_dtn;
    int64_t rk45_max_sub;
    double alpha0;
    double alphas[4];
    double dt, time_dir;
    int exact, w_method, kick, beach;
    double kh_x, kh_y;
    uint64_t seed, real_id, step_i;
    int use_depth, use_coastal, use_stochastic, on_failure;
    int resuspension, min_wet_steps;
    double depth_min, zone_depth, t_beach, rewet_hysteresis, t_resuspend, beach_dt;
    uint64_t beach_seed;
    int az;           // vertical Az kick
    double az_const;  // the constant source when az_a is null
    const double* az_a;  // Az columns (n_levels per node, internal order) of the
    const double* az_b;  // step payload's bracket, blended by az_alpha0 when az_b
    double az_alpha0;    // is not null
    int use_eta;      // beaching water depth = depth + eta of the step's field
    const double* eta_a;  // (N,) eta of the bracket pair (internal node order)
    const double* eta_b;
    int periodic_lon, periodic_lat;  // post-step fold windows [x0, x1), [y0, y1)
    double px0, px1, py0, py1;
    int reuse_anchor;  // an ALIVE particle's nn >= 0 is the anchor of its position
    int walk;          // elem_search = walk: containing elements by the edge walk (PState.elem carried)
    // sigma meshes: surface elevation per node of stages 0..3 and of the step
    // end (internal node order), column floor and bed distance
    const double* sig_eta[5];
    double Dmin, z0b_min;
};

struct PState {
    double* __restrict__ lon;
    double* __restrict__ lat;
    double* __restrict__ z;
    int32_t* __restrict__ status;
    int32_t* __restrict__ nn;
    int8_t* __restrict__ fr;
    double* __restrict__ u;
    double* __restrict__ v;
    double* __restrict__ w;
    const int64_t* __restrict__ pid;
    int32_t* __restrict__ ws;
    unsigned long long* __restrict__ alive_count;
    int64_t n;
    int32_t* __restrict__ elem;  // carried containing element (internal numbering) or null
};

// Launch step_kernel<order, MAXV, IS3D, T, ELEM> over all particles on
// `stream`.  `f`: the node snapshot records (u, v[, w pair]); `fe`: the
// element-centre u/v records when ELEM (u/v at element centres, least
// squares), else unused.  m.tile_on: stage grid-hash tiles in LDS.
template <int MAXV, bool IS3D, typename T, bool ELEM>
void launch_step_v(hipStream_t stream, int order, const DMesh& m, const Snap<T, IS3D>& f,
                   const Snap<T, false>& fe, const StepCfg& cfg, const PState& ps);

// The adaptive Dormand-Prince 5(4) step (2-D, u/v at element centres):
// launch rk45_kernel<MAXV, T>.
template <int MAXV, typename T>
void launch_rk45_v(hipStream_t stream, const DMesh& m, const Snap<T, false>& fe, const StepCfg& cfg,
                   const PState& ps);

// The sigma-mesh step (3-D): launch sigma_step_kernel<order, MAXV, T, ELEM>.
template <int MAXV, typename T, bool ELEM>
void launch_sigma_step(hipStream_t stream, int order, const DMesh& m, const Snap<T, true>& f,
                       const Snap<T, false>& fe, const StepCfg& cfg, const PState& ps);

// Surfaces of one sigma step into out[0..4] (stages 0..3, end), N each:
// start/end/stage sources are nested time blends of slot eta arrays.
struct EtaSrc {
    const double* a;
    const double* b;
    const double* t;
    double alpha0, alpha;
};
void launch_sigma_eta(hipStream_t stream, int64_t n_nodes, int order, int steady, EtaSrc start, EtaSrc end,
                      int end_is_start, EtaSrc stages, const double* stage_alphas, double* out);

}  // namespace kelp
