// kelp_step_impl.cuh -- the step kernel: one particle-step per thread.
//
// Included only by kelp_step_variant.hip, which instantiates the launchers
// of one (3-D, dtype, layout) combination per object file.
#pragma once

#include "kelp_sigma.cuh"
#include "kelp_step.cuh"

namespace kelp {

__device__ __forceinline__ double wrap_lon(double x) { return km_pymod360(x + 180.0) - 180.0; }

__device__ __forceinline__ double clamp_lat(double y)
{
    if (y > 90.0) return 90.0;
    if (y < -90.0) return -90.0;
    return y;
}

__device__ __forceinline__ double z0_of(double lat)
{
    double z0 = km_cos(lat * kDeg2Rad) * kSLM;
    if (z0 < 1.0) z0 = 1.0;
    return z0;
}

__device__ __forceinline__ double clamp_z(const DMesh& m, double z, int nn)
{
    if (z < m.zlev[0]) z = m.zlev[0];
    double bottom = m.depth[nn];
    if (z > bottom) z = bottom;
    return z;
}

// interpolate.py:1341/:1444 _interp_barycentric_core_tinterp(_3d):
// containment around the anchor, vertex weights, blended vertex values of
// the layer plane starting at `off`.
template <int MAXV, typename T, bool IS3D>
__device__ inline bool interp_uv(const DMesh& m, const Snap<T, IS3D>& f, double alpha0, double alpha,
                                 double time_dir, double xp, double yp, int nn, int64_t off, double& U,
                                 double& V, int ehint = -1)
{
    U = 0.0;
    V = 0.0;
    // a containing element validated for this position (the edge walk,
    // interpolate.py:1266) skips the search
    int e = ehint >= 0 ? ehint : find_containing<MAXV>(m, xp, yp, nn);
    if (e < 0) return false;
    Weights W;
    if (!vertex_weights<MAXV>(m, e, xp, yp, W)) return false;
    double su = 0.0, sv = 0.0;
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < W.n) {
            double un, vn;
            uv_blend(f, off + W.v[k], alpha0, alpha, un, vn);
            su += W.w[k] * un;
            sv += W.w[k] * vn;
        }
    }
    U = time_dir * su;
    V = time_dir * sv;
    return true;
}

// One evaluation site: anchor, layer at the anchor (vertical.py:40), u/v on
// that layer (node layout: barycentric in the containing element; element
// layout: least squares over the anchor's element centres), W by the
// configured method (integrators.py:1260 _compute_w).
template <int MAXV, bool IS3D, typename T, bool ELEM>
__device__ inline bool eval_site(const DMesh& m, const Snap<T, IS3D>& f, const Snap<T, false>& fe,
                                 const StepCfg& c, double alpha, double lon, double lat, double zp, int& nn,
                                 double& U, double& V, double& W, int ehint = -1)
{
    // nn >= 0 on entry: the anchor of (lon, lat) under this mode is known
    if (nn < 0) nn = site_anchor(m, lon, lat, c.exact);
    int64_t layer = 0;
    if (IS3D) layer = find_layer(m, zp, m.depth[nn]);
    bool ok;
    if constexpr (ELEM)
        ok = interp_lsq(m, fe, c.alpha0, alpha, c.time_dir, lon, lat, nn, layer * m.n_elems, U, V);
    else
        ok = interp_uv<MAXV>(m, f, c.alpha0, alpha, c.time_dir, lon, lat, nn, layer * m.n_nodes, U, V, ehint);
    W = 0.0;
    if constexpr (IS3D) {
        if (c.w_method == 0)
            W = w_linear(m, f, c.alpha0, alpha, zp, nn, c.time_dir);
        else if (c.w_method == 1)
            W = w_lsq<MAXV>(m, f, c.alpha0, alpha, lon, lat, zp, nn, c.time_dir);
        else if (c.w_method == 3)
            W = w_idw<MAXV>(m, f, c.alpha0, alpha, lon, lat, zp, nn, c.time_dir);
        else
            W = w_bary<MAXV>(m, f, c.alpha0, alpha, lon, lat, zp, nn, c.time_dir);
    }
    return ok;
}

// The deterministic step of one alive particle, in place.  Euler follows
// The walk kept out of line: it runs only with elem_search = walk.
template <int MAXV>
__device__ __noinline__ __attribute__((cold)) int walk_at(const DMesh& m, double lon, double lat, int start)
{
    return walk_locate<MAXV>(m, lon, lat, start);
}

// integrators.py:4997 _euler_step_unfused, RK2 the fused barycentric kernels
// :2522 (2-D) / :3648 (3-D), RK4 :2642 (2-D) and :5760 _rk4_step_unfused
// (3-D).  Returns the failure code (0 = alive).
//
// With elem_search = walk (c.walk, node-layout u/v; the reference's
// multi-pass steps with carried elements, integrators.py:5020-5022,
// :5415-5417, :5792-5794) every site's containing element comes from the
// edge walk started at the previous site's (the carried element at site 1)
// and the end-of-step domain verdict from the walk too; `elem` is carried.
template <int ORDER, int MAXV, bool IS3D, typename T, bool ELEM>
__device__ inline int det_step(const DMesh& m, const Snap<T, IS3D>& f, const Snap<T, false>& fe, const StepCfg& c,
                               double& lon, double& lat, double& z, int& nn, double& Uo, double& Vo, double& Wo,
                               int& elem)
{
    const double dt = c.dt;
    const double lon0 = lon, lat0 = lat, zz0 = z;
    int node_i;
    double U1, V1, W1;
    // site 1 is the particle's position: its anchor is the one the previous
    // step (or kick, or fold) stored, when the caller vouches for it
    node_i = (c.reuse_anchor && nn >= 0) ? nn : -1;
    const int e1 = c.walk ? walk_at<MAXV>(m, lon0, lat0, elem) : -1;
    int e_last = e1;
    bool ok = eval_site<MAXV, IS3D, T, ELEM>(m, f, fe, c, c.alphas[0], lon0, lat0, zz0, node_i, U1, V1, W1, e1);
    Uo = U1;
    Vo = V1;
    Wo = W1;
    double lon_new = lon0, lat_new = lat0, z_new = zz0;
    if (ok) {
        if (ORDER == 1) {
            lon_new = wrap_lon(lon0 + dt * U1 / z0_of(lat0));
            lat_new = clamp_lat(lat0 + dt * V1 / kSLM);
            z_new = zz0 + W1 * dt;
        } else if (ORDER == 2) {
            double z0 = z0_of(lat0);
            double lon_mid = wrap_lon(lon0 + 0.5 * dt * U1 / z0);
            double lat_mid = clamp_lat(lat0 + 0.5 * dt * V1 / kSLM);
            double z_mid = zz0 + W1 * 0.5 * dt;
            int nn_mid = -1;
            double U2, V2, W2;
            if (c.walk) e_last = walk_at<MAXV>(m, lon_mid, lat_mid, e_last);
            ok = eval_site<MAXV, IS3D, T, ELEM>(m, f, fe, c, c.alphas[1], lon_mid, lat_mid, z_mid, nn_mid, U2, V2, W2,
                                                e_last);
            if (ok) {
                double z0m = z0_of(lat_mid);
                lon_new = wrap_lon(lon0 + dt * U2 / z0m);
                lat_new = clamp_lat(lat0 + dt * V2 / kSLM);
                z_new = zz0 + W2 * dt;
            }
        } else {
            double dlon1 = dt * U1 / z0_of(lat0);
            double dlat1 = dt * V1 / kSLM;
            double dz1 = dt * W1;
            double lon2 = wrap_lon(lon0 + 0.5 * dlon1);
            double lat2 = clamp_lat(lat0 + 0.5 * dlat1);
            double z2 = zz0 + 0.5 * dz1;
            int nn2 = -1, nn3 = -1, nn4 = -1;
            double U2, V2, W2, U3, V3, W3, U4, V4, W4;
            if (c.walk) e_last = walk_at<MAXV>(m, lon2, lat2, e_last);
            ok = eval_site<MAXV, IS3D, T, ELEM>(m, f, fe, c, c.alphas[1], lon2, lat2, z2, nn2, U2, V2, W2, e_last);
            if (ok) {
                double dlon2 = dt * U2 / z0_of(lat2);
                double dlat2 = dt * V2 / kSLM;
                double dz2 = dt * W2;
                double lon3 = wrap_lon(lon0 + 0.5 * dlon2);
                double lat3 = clamp_lat(lat0 + 0.5 * dlat2);
                double z3 = zz0 + 0.5 * dz2;
                if (c.walk) e_last = walk_at<MAXV>(m, lon3, lat3, e_last);
                ok = eval_site<MAXV, IS3D, T, ELEM>(m, f, fe, c, c.alphas[2], lon3, lat3, z3, nn3, U3, V3, W3,
                                                    e_last);
                if (ok) {
                    double dlon3 = dt * U3 / z0_of(lat3);
                    double dlat3 = dt * V3 / kSLM;
                    double dz3 = dt * W3;
                    double lon4 = wrap_lon(lon0 + dlon3);
                    double lat4 = clamp_lat(lat0 + dlat3);
                    double z4 = zz0 + dz3;
                    if (c.walk) e_last = walk_at<MAXV>(m, lon4, lat4, e_last);
                    ok = eval_site<MAXV, IS3D, T, ELEM>(m, f, fe, c, c.alphas[3], lon4, lat4, z4, nn4, U4, V4,
                                                        W4, e_last);
                    if (ok) {
                        double dlon4 = dt * U4 / z0_of(lat4);
                        double dlat4 = dt * V4 / kSLM;
                        double dz4 = dt * W4;
                        double dlon = (dlon1 + 2.0 * dlon2 + 2.0 * dlon3 + dlon4) / 6.0;
                        double dlat = (dlat1 + 2.0 * dlat2 + 2.0 * dlat3 + dlat4) / 6.0;
                        double dz = (dz1 + 2.0 * dz2 + 2.0 * dz3 + dz4) / 6.0;
                        lon_new = wrap_lon(lon0 + dlon);
                        lat_new = clamp_lat(lat0 + dlat);
                        z_new = zz0 + dz;
                    }
                }
            }
        }
    }
    if (!ok) {
        nn = node_i;
        if (c.walk) elem = e1;
        return kFrInterp;
    }
    int nn_new = site_anchor(m, lon_new, lat_new, c.exact);
    int e = c.walk ? walk_at<MAXV>(m, lon_new, lat_new, e_last) : find_containing<MAXV>(m, lon_new, lat_new, nn_new);
    if (e < 0) {
        nn = node_i;
        if (c.walk) elem = e1;
        return kFrOOB;
    }
    lon = lon_new;
    lat = lat_new;
    if (IS3D) z = clamp_z(m, z_new, nn_new);
    nn = nn_new;
    if (c.walk) elem = e;
    return 0;
}

// integrators.py:673 _apply_split_diffusion_kick for one particle alive after
// the deterministic step (uniform Kh, counter RNG keyed by the particle id).
// SIGMA: the particle keeps its sigma fraction in the end-of-step surface,
// z / H_p(post-step position) * H_p(kicked position), clamped in the new
// column (integrators.py:784-801).
template <int MAXV, bool IS3D, bool SIGMA = false>
__device__ inline void kick_one(const DMesh& m, const StepCfg& c, uint64_t pid, double pre_lon,
                                      double pre_lat, double pre_z, int pre_nn, double& lon, double& lat,
                                      double& z, int32_t& st, int& nn, int8_t& fr)
{
    double g1, g2;
    normal_pair(c.seed, c.real_id, pid, c.step_i, 0, g1, g2);
    double abs_dt = c.dt >= 0.0 ? c.dt : -c.dt;
    double sqrt_abs_dt = km_sqrt(abs_dt);
    double sigma_x = km_sqrt(2.0 * c.kh_x);
    double sigma_y = km_sqrt(2.0 * c.kh_y);
    double dx = sigma_x * sqrt_abs_dt * g1;
    double dy = sigma_y * sqrt_abs_dt * g2;
    if (c.kh_x != c.kh_y) {
        double la = lat;
        if (la < -89.9) la = -89.9;
        if (la > 89.9) la = 89.9;
        double corr = (c.kh_x - c.kh_y) * (km_sin(la * kDeg2Rad) / km_cos(la * kDeg2Rad)) / kRMetric *
                      km_fabs(c.dt);
        dy = dy + corr;
    }
    double lon_k, lat_k;
    sphere_displace(lon, lat, dx, dy, lon_k, lat_k);
    int nn_k = site_anchor(m, lon_k, lat_k, c.exact == 1 ? 1 : 2);  // batch find_nearest_nodes
    int e = find_containing<MAXV>(m, lon_k, lat_k, nn_k);
    if (e < 0) {
        lon = pre_lon;
        lat = pre_lat;
        z = pre_z;
        nn = pre_nn;
        st = kOutside;
        fr = kFrOOB;
        return;
    }
    if constexpr (SIGMA) {
        const double Hp = column_height<MAXV>(m, c.sig_eta[4], c.Dmin, c.w_method, lon, lat, nn);
        const double Hk = column_height<MAXV>(m, c.sig_eta[4], c.Dmin, c.w_method, lon_k, lat_k, nn_k);
        z = clamp_sigma(z / Hp * Hk, Hk, c.Dmin, c.z0b_min);
    } else if (IS3D) {
        z = clamp_z(m, z, nn_k);
    }
    lon = lon_k;
    lat = lat_k;
    nn = nn_k;
}

// numerics/diffusion.py:378 az_column_value_gradient on the column of `node`
// of the step payload's Az (the bracket pair blended per value,
// core/io.py:68): the depth clamped to [z_levels[0], min(depth, last
// level)], the value linear between interfaces (:357), the gradient the
// slope of the segment the layer search picks.
__device__ inline void az_column(const DMesh& m, const double* __restrict__ az_a, const double* __restrict__ az_b,
                                 double alpha0, int node, double z, double& val, double& grad)
{
    const int K = m.n_levels;
    const double* zl = m.zlev;
    const double* ca = az_a + (int64_t)node * K;
    const double* cb = az_b ? az_b + (int64_t)node * K : nullptr;
    const double beta = 1.0 - alpha0;
    auto A = [&](int k) { return cb ? beta * ca[k] + alpha0 * cb[k] : ca[k]; };
    const double top = zl[0];
    const double dn = m.depth[node];
    double bot = dn < zl[K - 1] ? dn : zl[K - 1];
    if (bot < top) bot = top;
    double ze = z;
    if (ze < top) ze = top;
    if (ze > bot) ze = bot;
    if (K < 2) {
        val = A(0);
        grad = 0.0;
        return;
    }
    // the first segment whose lower interface is at or below ze (the last
    // one when none is): the gradient's segment, and the value's inside
    int k = K - 2;
    for (int kk = 0; kk < K - 1; ++kk)
        if (ze <= zl[kk + 1]) {
            k = kk;
            break;
        }
    if (ze <= zl[0]) {
        val = A(0);
    } else if (ze >= zl[K - 1]) {
        val = A(K - 1);
    } else {
        const double dzl = zl[k + 1] - zl[k];
        if (dzl <= 0.0) {
            val = A(k);
        } else {
            const double frac = (ze - zl[k]) / dzl;
            val = (1.0 - frac) * A(k) + frac * A(k + 1);
        }
    }
    const double dz = zl[k + 1] - zl[k];
    grad = dz <= 0.0 ? 0.0 : (A(k + 1) - A(k)) / dz;
}

// integrators.py:842 _apply_split_az_kick on a z-level mesh
// (numerics/diffusion.py:427 _loop_az_kick_zlevel and :302
// diffusion_az_step_meters): the constant source (gradient 0) or the
// payload's Az field, one particle alive after the horizontal kick, folded
// back into [z_levels[0], depth[nn]].
__device__ inline void az_one(const DMesh& m, const StepCfg& c, uint64_t pid, int node, double& z)
{
    double az = c.az_const, g = 0.0;
    if (c.az_a) {
        az_column(m, c.az_a, c.az_b, c.az_alpha0, node, z, az, g);
        if (az < 0.0) {  // numerics/diffusion.py:448: a negative value kills the kick
            az = 0.0;
            g = 0.0;
        }
    }
    double xi, unused;
    normal_pair(c.seed, c.real_id, pid, c.step_i, 1, xi, unused);
    const double sdt = c.time_dir * c.dt;
    const double abs_dt = sdt >= 0.0 ? sdt : -sdt;
    const double dz = g * abs_dt + km_sqrt(2.0 * az * abs_dt) * xi;
    z = km_reflect_depth(z + dz, m.zlev[0], m.depth[node]);
}

// integrators.py:842 _apply_split_az_kick on a sigma mesh (constant source,
// numerics/diffusion.py:459 _loop_az_kick_sigma): folded into
// [0, max(H_p, Dmin) - z0b_min] of the end-of-step column.
template <int MAXV>
__device__ inline void az_sigma_one(const DMesh& m, const StepCfg& c, uint64_t pid, double lon, double lat, int node,
                                    double& z)
{
    double H = column_height<MAXV>(m, c.sig_eta[4], c.Dmin, c.w_method, lon, lat, node);
    if (H < c.Dmin) H = c.Dmin;
    double bot = H - c.z0b_min;
    if (bot < 0.0) bot = 0.0;
    double xi, unused;
    normal_pair(c.seed, c.real_id, pid, c.step_i, 1, xi, unused);
    const double sdt = c.time_dir * c.dt;
    const double abs_dt = sdt >= 0.0 ? sdt : -sdt;
    const double g = 0.0;
    const double dz = g * abs_dt + km_sqrt(2.0 * c.az_const * abs_dt) * xi;
    z = km_reflect_depth(z + dz, 0.0, bot);
}

// tracking.py:1089 _apply_periodic_wrap for one ALIVE particle: fold a
// coordinate outside [x0, x1) back with numpy's mod (Python % semantics),
// strictly inside the window, and re-anchor a folded particle with the batch
// search (find_nearest_nodes).
__device__ inline void periodic_one(const DMesh& m, const StepCfg& c, double& lon, double& lat, int& nn,
                                    int& elem)
{
    bool moved = false;
    if (c.periodic_lon && (lon < c.px0 || lon >= c.px1)) {
        double fo = c.px0 + km_pymod_pos(lon - c.px0, c.px1 - c.px0);
        if (fo >= c.px1) fo = c.px0;
        lon = fo;
        moved = true;
    }
    if (c.periodic_lat && (lat < c.py0 || lat >= c.py1)) {
        double fo = c.py0 + km_pymod_pos(lat - c.py0, c.py1 - c.py0);
        if (fo >= c.py1) fo = c.py0;
        lat = fo;
        moved = true;
    }
    if (moved) {
        nn = site_anchor(m, lon, lat, c.exact == 1 ? 1 : 2);
        elem = -1;   // tracking.py:1143: a folded particle's carried element is dropped
    }
}

// tracking.py:199 _apply_beaching_hook + numerics/beaching.py:95
// evaluate_beaching, one particle (prev = status at the step start).
template <int MAXV>
__device__ inline void beach_one(const DMesh& m, const StepCfg& c, uint64_t pid, int32_t prev, double lon,
                                       double lat, int32_t& st, int& nn, int8_t& fr, int32_t& ws)
{
    const bool need_topo = c.use_coastal || c.use_stochastic;
    const bool newly_failed = (st == kOutside) && (prev == kAlive);
    if (newly_failed && nn < 0) nn = site_anchor(m, lon, lat, c.exact == 1 ? 1 : 2);
    const int n = nn;
    int containing = -1;
    if (need_topo && (st == kAlive || st == kBeached || newly_failed) && n >= 0)
        containing = find_containing<MAXV>(m, lon, lat, n);
    const bool coastal = need_topo && containing >= 0 && m.coastal != nullptr && m.coastal[containing];
    double h_raw = n >= 0 ? m.depth[n] : __builtin_inf();
    // + vel.eta[nn] of the step's field: the bracket pair blended by alpha0
    // (core/io.py:386 LazyLerpField), or the plain snapshot
    if (c.use_eta && n >= 0) {
        double eta = c.eta_a[n];
        if (c.alpha0 != 0.0) eta = (1.0 - c.alpha0) * c.eta_a[n] + c.alpha0 * c.eta_b[n];
        h_raw = h_raw + eta;
    }
    if (c.on_failure && newly_failed) {
        if (h_raw < c.zone_depth || coastal) {
            st = kBeached;
            fr = 0;
        }
    }
    const bool alive = st == kAlive, beached = st == kBeached;
    bool to_beach = false;
    if (c.use_depth) to_beach |= h_raw < c.depth_min;
    if (c.use_coastal) to_beach |= coastal;
    if (c.use_stochastic && c.t_beach > 0) {
        bool in_zone = h_raw < c.zone_depth || coastal;
        double p = 1.0 - km_exp(-km_fabs(c.beach_dt) / c.t_beach);
        if (alive && in_zone && u01(c.beach_seed, c.real_id, pid, c.step_i, 10) < p) to_beach = true;
    }
    to_beach = to_beach && alive;
    bool to_release = false;
    int32_t w = ws;
    if (c.resuspension != 0 && beached) {
        bool rewet = h_raw > (c.depth_min + c.rewet_hysteresis) && n >= 0;
        w = rewet ? w + 1 : 0;
        if (c.resuspension == 1 || c.resuspension == 3) to_release |= w >= c.min_wet_steps;
        if ((c.resuspension == 2 || c.resuspension == 3) && c.t_resuspend > 0) {
            double p_r = 1.0 - km_exp(-km_fabs(c.beach_dt) / c.t_resuspend);
            if (rewet && u01(c.beach_seed, c.real_id, pid, c.step_i, 11) < p_r) to_release = true;
        }
    }
    if (to_release) w = 0;
    if (!beached) w = 0;
    ws = w;
    if (to_beach) st = kBeached;
    if (to_release) {
        st = kAlive;
        fr = 0;
    }
}

// One full step per particle: deterministic step, split Kh kick, beaching.
// Waves per SIMD the compiler budgets registers for (0: its own choice).
#ifndef KELP_STEP_WAVES
#define KELP_STEP_WAVES 3
#endif
#if KELP_STEP_WAVES > 0
#define KELP_STEP_ATTR __attribute__((amdgpu_waves_per_eu(KELP_STEP_WAVES)))
#else
#define KELP_STEP_ATTR
#endif

// One particle's full step (deterministic step, kicks, periodic fold,
// beaching), the state read and written at slot i.  `prev` is its status at
// the step start; its random draws are keyed by pid[i] & key_mask (the
// whole pid in a plain step; the cloud-local part in a batched one).
template <int ORDER, int MAXV, bool IS3D, typename T, bool ELEM>
__device__ __forceinline__ void step_particle(const DMesh& m, const Snap<T, IS3D>& f, const Snap<T, false>& fe,
                                              const StepCfg& c, const PState& s, int64_t i, int32_t prev,
                                              uint64_t key_mask)
{
    int32_t st = prev;
    double lon = s.lon[i], lat = s.lat[i];
    double z = IS3D ? s.z[i] : 0.0;
    int nn = s.nn[i];
    int8_t fr = s.fr[i];
    int elem = s.elem ? s.elem[i] : -1;
    double U = 0.0, V = 0.0, W = 0.0;
    if (prev == kAlive) {
        const double pre_lon = lon, pre_lat = lat, pre_z = z;
        const int pre_nn = nn;
        int code = det_step<ORDER, MAXV, IS3D, T, ELEM>(m, f, fe, c, lon, lat, z, nn, U, V, W, elem);
        if (code) {
            st = kOutside;
            fr = (int8_t)code;
        } else if (c.kick && !(c.kh_x == 0.0 && c.kh_y == 0.0)) {
            kick_one<MAXV, IS3D>(m, c, (uint64_t)s.pid[i] & key_mask, pre_lon, pre_lat, pre_z, pre_nn, lon, lat, z,
                                 st, nn, fr);
        }
        if (IS3D && c.az && st == kAlive) az_one(m, c, (uint64_t)s.pid[i] & key_mask, nn, z);
        if ((c.periodic_lon || c.periodic_lat) && st == kAlive) periodic_one(m, c, lon, lat, nn, elem);
    }
    if (c.beach) beach_one<MAXV>(m, c, (uint64_t)s.pid[i] & key_mask, prev, lon, lat, st, nn, fr, s.ws[i]);
    if (s.elem) s.elem[i] = elem;
    s.lon[i] = lon;
    s.lat[i] = lat;
    if (IS3D) s.z[i] = z;
    s.status[i] = st;
    s.nn[i] = nn;
    s.fr[i] = fr;
    s.u[i] = U;
    s.v[i] = V;
    if (IS3D && s.w) s.w[i] = W;
}

// particles alive at the step start: one atomic per wavefront (the lanes
// that reach the call)
__device__ __forceinline__ void count_alive(unsigned long long* count, bool alive_lane)
{
    if (!count) return;
    unsigned long long alive = __ballot(alive_lane);
    unsigned long long active = __ballot(1);
    if ((int)(threadIdx.x & 63) == __ffsll((long long)active) - 1 && alive)
        atomicAdd(count, (unsigned long long)__popcll(alive));
}

template <int ORDER, int MAXV, bool IS3D, typename T, bool ELEM>
__global__ __launch_bounds__(256) KELP_STEP_ATTR void step_kernel(DMesh m, Snap<T, IS3D> f, Snap<T, false> fe,
                                                                  StepCfg c, PState s)
{
    int64_t i = xcd_thread_index();
    const bool in = i < s.n;
    const int32_t prev = in ? s.status[i] : 0;
    // every thread of the workgroup takes part in staging its grid-hash
    // tile (workgroup barriers) before any returns
    if (m.tile_on) stage_gh_tile(m, prev == kAlive, in ? s.lon[i] : 0.0, in ? s.lat[i] : 0.0);
    if (!in) return;
    count_alive(s.alive_count, prev == kAlive);
    step_particle<ORDER, MAXV, IS3D, T, ELEM>(m, f, fe, c, s, i, prev, ~0ull);
}

// ---------------------------------------------------------------------------
// rk45: integrators.py:6233 _rk45_kernel_2d, per-particle adaptive
// Dormand-Prince 5(4) over exactly dt (2-D, u/v at element centres by least
// squares, steady or time-blended), then rk45_step's re-validation
// (:6586-6611).  Tableau of Dormand & Prince (1980), error weights b5 - b4
// as exact fractions.

// a trial position: lon wrapped only when it left [-180, 180), lat clamped
// (integrators.py:6182 _norm_lonlat)
__device__ __forceinline__ void norm_lonlat(double& lon, double& lat)
{
    if (lon >= 180.0 || lon < -180.0) lon = km_pymod360(lon + 180.0) - 180.0;
    if (lat > 90.0)
        lat = 90.0;
    else if (lat < -90.0)
        lat = -90.0;
}

// integrators.py:6197 _rk45_eval: velocity and degree rates at (lon, lat)
template <int MAXV, typename T>
__device__ inline bool rk45_eval(const DMesh& m, const Snap<T, false>& fe, const StepCfg& c, double lon, double lat,
                                 double alpha, double& U, double& V, double& kx, double& ky)
{
    U = V = kx = ky = 0.0;
    const int nn = site_anchor(m, lon, lat, c.exact == 1 ? 1 : 0);
    if (nn < 0) return false;
    double u, v;
    if (!interp_lsq(m, fe, c.alpha0, alpha, c.time_dir, lon, lat, nn, 0, u, v)) return false;
    const double z0 = z0_of(lat);
    U = u;
    V = v;
    kx = u / z0;
    ky = v / kSLM;
    return true;
}

__device__ __forceinline__ double rk45_alpha(double t, double inv_dtn)
{
    const double al = t * inv_dtn;
    return al > 1.0 ? 1.0 : al;
}

template <int MAXV, typename T>
__device__ __noinline__ void rk45_det(const DMesh& m, const Snap<T, false>& fe, const StepCfg& c, double& lon,
                                      double& lat, int32_t& st, int& nn, int8_t& fr, double& Uo, double& Vo)
{
    const double dt = c.dt, tol = c.rk45_tol, h_min = c.rk45_h_min, inv = c.rk45_inv_dtn;
    double y0 = lon, y1 = lat, t = 0.0, h = dt;
    double Uv, Vv, k1x, k1y;
    bool ok_out = true;
    Uo = Vo = 0.0;
    if (!rk45_eval<MAXV, T>(m, fe, c, y0, y1, rk45_alpha(t, inv), Uv, Vv, k1x, k1y)) {
        ok_out = false;
    } else {
        int64_t nsub = 0;
        double k2x = 0, k2y = 0, k3x = 0, k3y = 0, k4x = 0, k4y = 0, k5x = 0, k5y = 0, k6x = 0, k6y = 0, k7x = 0,
               k7y = 0, U7 = 0, V7 = 0, y5x = y0, y5y = y1, du, dv;
        while (t < dt * (1.0 - 1e-12)) {
            if (h > dt - t) h = dt - t;
            bool clamped = false;
            if (h < h_min) {
                h = h_min;
                if (h > dt - t) h = dt - t;
                clamped = true;
            }
            double px, py;
            px = y0 + h * (1.0 / 5.0) * k1x;
            py = y1 + h * (1.0 / 5.0) * k1y;
            norm_lonlat(px, py);
            bool okst = rk45_eval<MAXV, T>(m, fe, c, px, py, rk45_alpha(t + (1.0 / 5.0) * h, inv), du, dv, k2x, k2y);
            if (okst) {
                px = y0 + h * ((3.0 / 40.0) * k1x + (9.0 / 40.0) * k2x);
                py = y1 + h * ((3.0 / 40.0) * k1y + (9.0 / 40.0) * k2y);
                norm_lonlat(px, py);
                okst = rk45_eval<MAXV, T>(m, fe, c, px, py, rk45_alpha(t + (3.0 / 10.0) * h, inv), du, dv, k3x, k3y);
            }
            if (okst) {
                px = y0 + h * ((44.0 / 45.0) * k1x + (-56.0 / 15.0) * k2x + (32.0 / 9.0) * k3x);
                py = y1 + h * ((44.0 / 45.0) * k1y + (-56.0 / 15.0) * k2y + (32.0 / 9.0) * k3y);
                norm_lonlat(px, py);
                okst = rk45_eval<MAXV, T>(m, fe, c, px, py, rk45_alpha(t + (4.0 / 5.0) * h, inv), du, dv, k4x, k4y);
            }
            if (okst) {
                px = y0 + h * ((19372.0 / 6561.0) * k1x + (-25360.0 / 2187.0) * k2x + (64448.0 / 6561.0) * k3x +
                               (-212.0 / 729.0) * k4x);
                py = y1 + h * ((19372.0 / 6561.0) * k1y + (-25360.0 / 2187.0) * k2y + (64448.0 / 6561.0) * k3y +
                               (-212.0 / 729.0) * k4y);
                norm_lonlat(px, py);
                okst = rk45_eval<MAXV, T>(m, fe, c, px, py, rk45_alpha(t + (8.0 / 9.0) * h, inv), du, dv, k5x, k5y);
            }
            const double a6 = rk45_alpha(t + h, inv);
            if (okst) {
                px = y0 + h * ((9017.0 / 3168.0) * k1x + (-355.0 / 33.0) * k2x + (46732.0 / 5247.0) * k3x +
                               (49.0 / 176.0) * k4x + (-5103.0 / 18656.0) * k5x);
                py = y1 + h * ((9017.0 / 3168.0) * k1y + (-355.0 / 33.0) * k2y + (46732.0 / 5247.0) * k3y +
                               (49.0 / 176.0) * k4y + (-5103.0 / 18656.0) * k5y);
                norm_lonlat(px, py);
                okst = rk45_eval<MAXV, T>(m, fe, c, px, py, a6, du, dv, k6x, k6y);
            }
            if (okst) {
                y5x = y0 + h * ((35.0 / 384.0) * k1x + (500.0 / 1113.0) * k3x + (125.0 / 192.0) * k4x +
                                (-2187.0 / 6784.0) * k5x + (11.0 / 84.0) * k6x);
                y5y = y1 + h * ((35.0 / 384.0) * k1y + (500.0 / 1113.0) * k3y + (125.0 / 192.0) * k4y +
                                (-2187.0 / 6784.0) * k5y + (11.0 / 84.0) * k6y);
                norm_lonlat(y5x, y5y);
                okst = rk45_eval<MAXV, T>(m, fe, c, y5x, y5y, a6, U7, V7, k7x, k7y);
            }
            nsub += 1;
            if (!okst) {
                // a failed stage is a rejected attempt, terminal only at the
                // h_min floor or the cap
                if (nsub > c.rk45_max_sub || h <= h_min * (1.0 + 1e-12)) {
                    ok_out = false;
                    break;
                }
                h = h * 0.2;
                if (h < h_min) h = h_min;
                continue;
            }
            const double errx = h * ((71.0 / 57600.0) * k1x + (-71.0 / 16695.0) * k3x + (71.0 / 1920.0) * k4x +
                                     (-17253.0 / 339200.0) * k5x + (22.0 / 525.0) * k6x + (-1.0 / 40.0) * k7x);
            const double erry = h * ((71.0 / 57600.0) * k1y + (-71.0 / 16695.0) * k3y + (71.0 / 1920.0) * k4y +
                                     (-17253.0 / 339200.0) * k5y + (22.0 / 525.0) * k6y + (-1.0 / 40.0) * k7y);
            const double z0m = z0_of(y1);
            const double ex = errx * z0m, ey = erry * kSLM;
            const double err = km_sqrt(ex * ex + ey * ey);
            if (nsub > c.rk45_max_sub) {
                ok_out = false;
                break;
            }
            if (err <= tol || clamped) {
                t += h;
                y0 = y5x;
                y1 = y5y;
                k1x = k7x;
                k1y = k7y;
                Uv = U7;
                Vv = V7;
            }
            double fac;
            if (err > 0.0) {
                fac = 0.9 * km_pow(tol / err, 0.2);
                if (fac < 0.2)
                    fac = 0.2;
                else if (fac > 5.0)
                    fac = 5.0;
            } else {
                fac = 5.0;
            }
            h = h * fac;
        }
    }
    if (!ok_out) {
        st = kOutside;
        fr = 1;
        return;
    }
    // rk45_step's re-validation: the batch nearest node, then the domain
    const int nn_new = site_anchor(m, y0, y1, c.exact == 1 ? 1 : 2);
    if (find_containing<MAXV>(m, y0, y1, nn_new) < 0) {
        st = kOutside;
        fr = 2;
        return;
    }
    lon = y0;
    lat = y1;
    nn = nn_new;
    Uo = Uv;
    Vo = Vv;
}

template <int MAXV, typename T>
__global__ __launch_bounds__(256) void rk45_kernel(DMesh m, Snap<T, false> fe, StepCfg c, PState s)
{
    const int64_t i = xcd_thread_index();
    if (i >= s.n) return;
    const int32_t prev = s.status[i];
    if (s.alive_count) {
        unsigned long long alive = __ballot(prev == kAlive);
        unsigned long long active = __ballot(1);
        if ((int)(threadIdx.x & 63) == __ffsll((long long)active) - 1 && alive)
            atomicAdd(s.alive_count, (unsigned long long)__popcll(alive));
    }
    int32_t st = prev;
    double lon = s.lon[i], lat = s.lat[i];
    int nn = s.nn[i];
    int8_t fr = s.fr[i];
    int elem = -1;
    double U = 0.0, V = 0.0, z = 0.0;
    if (prev == kAlive) {
        const double pre_lon = lon, pre_lat = lat;
        const int pre_nn = nn;
        rk45_det<MAXV, T>(m, fe, c, lon, lat, st, nn, fr, U, V);
        if (st == kAlive && c.kick && !(c.kh_x == 0.0 && c.kh_y == 0.0))
            kick_one<MAXV, false>(m, c, (uint64_t)s.pid[i], pre_lon, pre_lat, 0.0, pre_nn, lon, lat, z, st, nn, fr);
        if ((c.periodic_lon || c.periodic_lat) && st == kAlive) periodic_one(m, c, lon, lat, nn, elem);
    }
    if (c.beach) beach_one<MAXV>(m, c, (uint64_t)s.pid[i], prev, lon, lat, st, nn, fr, s.ws[i]);
    s.lon[i] = lon;
    s.lat[i] = lat;
    s.status[i] = st;
    s.nn[i] = nn;
    s.fr[i] = fr;
    s.u[i] = U;
    s.v[i] = V;
}

template <int MAXV, typename T>
void launch_rk45_v(hipStream_t stream, const DMesh& m, const Snap<T, false>& fe, const StepCfg& cfg,
                   const PState& ps)
{
    rk45_kernel<MAXV, T><<<(unsigned)((ps.n + 255) / 256), 256, 0, stream>>>(m, fe, cfg, ps);
}

// Never launched.  A second caller of the exact nearest search that carries
// no occupancy attribute: with the step kernels as its only callers the
// compiler narrows the out-of-line search to the step kernels' register
// budget and the step kernels then spill about eight times more (41 -> 329
// VGPRs for RK4 3-D at 3 waves/SIMD, hipcc -Rpass-analysis=kernel-resource-usage).
namespace {
__global__ void anchor_probe_kernel(DMesh m, int64_t n, const double* __restrict__ lon,
                                    const double* __restrict__ lat, int mode, int32_t* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = site_anchor(m, lon[i], lat[i], mode);
}
}  // namespace

template <int MAXV, bool IS3D, typename T, bool ELEM>
void launch_step_v(hipStream_t stream, int order, const DMesh& m, const Snap<T, IS3D>& f,
                   const Snap<T, false>& fe, const StepCfg& cfg, const PState& ps)
{
    // m.tile_on: the kernel stages its workgroup's grid-hash tile in
    // dynamic LDS
    const unsigned g = (unsigned)((ps.n + 255) / 256);
    const size_t lds = m.tile_on ? sizeof(GhTile) : 0;
    if (order == 1)
        step_kernel<1, MAXV, IS3D, T, ELEM><<<g, 256, lds, stream>>>(m, f, fe, cfg, ps);
    else if (order == 2)
        step_kernel<2, MAXV, IS3D, T, ELEM><<<g, 256, lds, stream>>>(m, f, fe, cfg, ps);
    else
        step_kernel<4, MAXV, IS3D, T, ELEM><<<g, 256, lds, stream>>>(m, f, fe, cfg, ps);
}

}  // namespace kelp
