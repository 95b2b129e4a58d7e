// kelp_hip.hip -- kernels and C ABI of the particle time-stepping library.
//
// Build (see kelp_amd/build.py):
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared
//         -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops
//         -o kelp_amd/libkelp_hip.so kelp_amd/csrc/kelp_hip.hip
// -ffp-contract=off keeps every a*b+c as two rounded operations, the
// arithmetic of the reference's Numba kernels (no fastmath there).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include <algorithm>

#include "kelp_internal.h"

using namespace kelp;

#define KELP_LIB_VERSION 1

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local char g_err[512] = "";

int kelp_fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define fail kelp_fail

int kelp_reserve(DevBuf& b, size_t bytes)
{
    if (b.bytes >= bytes && b.p) return KELP_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    HIP_TRY(hipMalloc(&b.p, bytes > 0 ? bytes : 16));
    b.bytes = bytes;
    return KELP_OK;
}

void kelp_release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b = DevBuf{};
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

static void free_mesh(kelp_ctx* c)
{
    for (void* p : c->mesh_bufs) (void)hipFree(p);
    c->mesh_bufs.clear();
    c->has_mesh = false;
    c->m = DMesh{};
    c->node_old_of_new = c->node_new_of_old = c->elem_old_of_new = c->elem_new_of_old = nullptr;
}

static void free_slot(FieldSlot& s)
{
    if (s.rec) (void)hipFree(s.rec);
    s = FieldSlot{};
}

template <typename T>
static int dev_alloc(kelp_ctx* c, T** out, size_t count)
{
    void* p = nullptr;
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 16;
    HIP_TRY(hipMalloc(&p, bytes));
    c->mesh_bufs.push_back(p);
    *out = static_cast<T*>(p);
    return KELP_OK;
}

// ---------------------------------------------------------------------------
// mesh build kernels
// ---------------------------------------------------------------------------

// core/mesh.py:713 _node_xyz_fill.  Internal node k is the node in slot k of
// the hash's CSR order, so the table is already in scan order.
__global__ void node_xyz_kernel(int64_t n, const double2* __restrict__ lonlat, double4* __restrict__ out)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double2 ll = lonlat[k];
    double lat_r = ll.y * kDeg2Rad, lon_r = ll.x * kDeg2Rad;
    double c = km_cos(lat_r);
    double4 r;
    r.x = c * km_cos(lon_r);
    r.y = c * km_sin(lon_r);
    r.z = km_sin(lat_r);
    r.w = __longlong_as_double((long long)k);
    out[k] = r;
}

// node unit vectors back in mesh (original) node order, (N, 3)
__global__ void gather_xyz_kernel(int64_t n, const double4* __restrict__ xyz, const int32_t* __restrict__ old_of_new,
                                  double* __restrict__ out)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 p = xyz[k];
    int64_t node = old_of_new[k];
    out[3 * node + 0] = p.x;
    out[3 * node + 1] = p.y;
    out[3 * node + 2] = p.z;
}

// pack a snapshot's planes (mesh node order) into records (internal order)
template <typename T>
__global__ void pack3d_kernel(int64_t N, int64_t n_layers, const int32_t* __restrict__ old_of_new,
                              const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
                              typename RecOf<T, true>::type* __restrict__ rec)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * n_layers) return;
    int64_t L = i / N, k = i - L * N;
    int64_t o = old_of_new[k];
    typename RecOf<T, true>::type r;
    r.x = u[L * N + o];
    r.y = v[L * N + o];
    r.z = w[L * N + o];
    r.w = w[(L + 1) * N + o];
    rec[i] = r;
}

template <typename T>
__global__ void pack2d_kernel(int64_t N, const int32_t* __restrict__ old_of_new, const T* __restrict__ u,
                              const T* __restrict__ v, typename RecOf<T, false>::type* __restrict__ rec)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    int64_t o = old_of_new[k];
    typename RecOf<T, false>::type r;
    r.x = u[o];
    r.y = v[o];
    rec[k] = r;
}

// in-place id translation through a table (-1 stays -1)
__global__ void map_ids_kernel(int64_t n, const int32_t* __restrict__ table, int32_t* __restrict__ ids)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t x = ids[i];
    if (x >= 0) ids[i] = table[x];
}

static unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------------------------
// step kernel
// ---------------------------------------------------------------------------

// One step's configuration (step_fn keywords, tracking.py:2643-2655, plus the
// beaching hook's per-run constants, tracking.py:199).
struct StepCfg {
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
    int az;
    double az_const;
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
};

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
                                 double& V)
{
    U = 0.0;
    V = 0.0;
    int e = find_containing<MAXV>(m, xp, yp, nn);
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
// that layer, W by the configured method (integrators.py:1260 _compute_w).
template <int MAXV, bool IS3D, typename T>
__device__ inline bool eval_site(const DMesh& m, const Snap<T, IS3D>& f, const StepCfg& c, double alpha,
                                 double lon, double lat, double zp, int& nn, double& U, double& V, double& W)
{
    nn = site_anchor(m, lon, lat, c.exact);
    int64_t off = 0;
    if (IS3D) off = (int64_t)find_layer(m, zp, m.depth[nn]) * m.n_nodes;
    bool ok = interp_uv<MAXV>(m, f, c.alpha0, alpha, c.time_dir, lon, lat, nn, off, U, V);
    W = 0.0;
    if constexpr (IS3D) {
        if (c.w_method == 0)
            W = w_linear(m, f, c.alpha0, alpha, zp, nn, c.time_dir);
        else
            W = w_bary<MAXV>(m, f, c.alpha0, alpha, lon, lat, zp, nn, c.time_dir);
    }
    return ok;
}

// The deterministic step of one alive particle, in place.  Euler follows
// integrators.py:4997 _euler_step_unfused, RK2 the fused barycentric kernels
// :2522 (2-D) / :3648 (3-D), RK4 :2642 (2-D) and :5760 _rk4_step_unfused
// (3-D).  Returns the failure code (0 = alive).
template <int ORDER, int MAXV, bool IS3D, typename T>
__device__ inline int det_step(const DMesh& m, const Snap<T, IS3D>& f, const StepCfg& c, double& lon, double& lat,
                               double& z, int& nn, double& Uo, double& Vo, double& Wo)
{
    const double dt = c.dt;
    const double lon0 = lon, lat0 = lat, zz0 = z;
    int node_i;
    double U1, V1, W1;
    bool ok = eval_site<MAXV, IS3D>(m, f, c, c.alphas[0], lon0, lat0, zz0, node_i, U1, V1, W1);
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
            int nn_mid;
            double U2, V2, W2;
            ok = eval_site<MAXV, IS3D>(m, f, c, c.alphas[1], lon_mid, lat_mid, z_mid, nn_mid, U2, V2, W2);
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
            int nn2, nn3, nn4;
            double U2, V2, W2, U3, V3, W3, U4, V4, W4;
            ok = eval_site<MAXV, IS3D>(m, f, c, c.alphas[1], lon2, lat2, z2, nn2, U2, V2, W2);
            if (ok) {
                double dlon2 = dt * U2 / z0_of(lat2);
                double dlat2 = dt * V2 / kSLM;
                double dz2 = dt * W2;
                double lon3 = wrap_lon(lon0 + 0.5 * dlon2);
                double lat3 = clamp_lat(lat0 + 0.5 * dlat2);
                double z3 = zz0 + 0.5 * dz2;
                ok = eval_site<MAXV, IS3D>(m, f, c, c.alphas[2], lon3, lat3, z3, nn3, U3, V3, W3);
                if (ok) {
                    double dlon3 = dt * U3 / z0_of(lat3);
                    double dlat3 = dt * V3 / kSLM;
                    double dz3 = dt * W3;
                    double lon4 = wrap_lon(lon0 + dlon3);
                    double lat4 = clamp_lat(lat0 + dlat3);
                    double z4 = zz0 + dz3;
                    ok = eval_site<MAXV, IS3D>(m, f, c, c.alphas[3], lon4, lat4, z4, nn4, U4, V4, W4);
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
        return kFrInterp;
    }
    int nn_new = site_anchor(m, lon_new, lat_new, c.exact);
    int e = find_containing<MAXV>(m, lon_new, lat_new, nn_new);
    if (e < 0) {
        nn = node_i;
        return kFrOOB;
    }
    lon = lon_new;
    lat = lat_new;
    if (IS3D) z = clamp_z(m, z_new, nn_new);
    nn = nn_new;
    return 0;
}

// integrators.py:673 _apply_split_diffusion_kick for one particle alive after
// the deterministic step (uniform Kh, counter RNG keyed by the particle id).
template <int MAXV, bool IS3D>
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
    lon = lon_k;
    lat = lat_k;
    nn = nn_k;
    if (IS3D) z = clamp_z(m, z, nn_k);
}

// integrators.py:842 _apply_split_az_kick, constant source on a z-level
// mesh (numerics/diffusion.py:427 _loop_az_kick_zlevel, gradient 0, and
// :302 diffusion_az_step_meters): one particle alive after the horizontal
// kick, folded back into [z_levels[0], depth[nn]].
__device__ inline void az_one(const DMesh& m, const StepCfg& c, uint64_t pid, int node, double& z)
{
    double xi, unused;
    normal_pair(c.seed, c.real_id, pid, c.step_i, 1, xi, unused);
    const double sdt = c.time_dir * c.dt;
    const double abs_dt = sdt >= 0.0 ? sdt : -sdt;
    const double g = 0.0;
    const double dz = g * abs_dt + km_sqrt(2.0 * c.az_const * abs_dt) * xi;
    z = km_reflect_depth(z + dz, m.zlev[0], m.depth[node]);
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
    const double h_raw = n >= 0 ? m.depth[n] : __builtin_inf();
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

template <int ORDER, int MAXV, bool IS3D, typename T>
__global__ __launch_bounds__(256) KELP_STEP_ATTR void step_kernel(DMesh m, Snap<T, IS3D> f, StepCfg c, PState s)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    const int32_t prev = s.status[i];
    if (s.alive_count) {
        // particles alive at the step start: one atomic per wavefront
        unsigned long long alive = __ballot(prev == kAlive);
        unsigned long long active = __ballot(1);
        if ((int)(threadIdx.x & 63) == __ffsll((long long)active) - 1 && alive)
            atomicAdd(s.alive_count, (unsigned long long)__popcll(alive));
    }
    int32_t st = prev;
    double lon = s.lon[i], lat = s.lat[i];
    double z = IS3D ? s.z[i] : 0.0;
    int nn = s.nn[i];
    int8_t fr = s.fr[i];
    double U = 0.0, V = 0.0, W = 0.0;
    if (prev == kAlive) {
        const double pre_lon = lon, pre_lat = lat, pre_z = z;
        const int pre_nn = nn;
        int code = det_step<ORDER, MAXV, IS3D>(m, f, c, lon, lat, z, nn, U, V, W);
        if (code) {
            st = kOutside;
            fr = (int8_t)code;
        } else if (c.kick && !(c.kh_x == 0.0 && c.kh_y == 0.0)) {
            kick_one<MAXV, IS3D>(m, c, (uint64_t)s.pid[i], pre_lon, pre_lat, pre_z, pre_nn, lon, lat, z, st, nn, fr);
        }
        if (IS3D && c.az && st == kAlive) az_one(m, c, (uint64_t)s.pid[i], nn, z);
    }
    if (c.beach) beach_one<MAXV>(m, c, (uint64_t)s.pid[i], prev, lon, lat, st, nn, fr, s.ws[i]);
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

// batch anchors (core/mesh.py:1896 find_nearest_nodes), mesh node ids out
__global__ void nearest_kernel(DMesh m, int64_t n, const double* __restrict__ lon,
                               const double* __restrict__ lat, int exact, const int32_t* __restrict__ old_of_new,
                               int32_t* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // guard off: the plain 3x3 window, an empty window healed exactly (the
    // reference heals it with a KD-tree query)
    int b = site_anchor(m, lon[i], lat[i], exact ? 1 : 2);
    out[i] = b >= 0 ? old_of_new[b] : -1;
}

// batch containment (core/mesh.py:2267); mesh node ids in, mesh element ids out
template <int MAXV>
__global__ void containing_kernel(DMesh m, int64_t n, const double* __restrict__ lon,
                                  const double* __restrict__ lat, const int32_t* __restrict__ nn,
                                  const int32_t* __restrict__ node_new_of_old,
                                  const int32_t* __restrict__ elem_old_of_new, int32_t* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int nd = nn[i];
    if (nd < 0 || nd >= m.n_nodes) {
        out[i] = -1;
        return;
    }
    int e = find_containing<MAXV>(m, lon[i], lat[i], node_new_of_old[nd]);
    out[i] = e >= 0 ? elem_old_of_new[e] : -1;
}

// Elementwise evaluation of the device math routines (diagnostics: lets a
// test compare them with a host build of the same kmath.h).
__global__ void math_probe_kernel(int fn, int64_t n, const double* __restrict__ x,
                                  const double* __restrict__ y, double* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = x[i], r;
    switch (fn) {
    case 0: r = km_sin(a); break;
    case 1: r = km_cos(a); break;
    case 2: r = km_asin(a); break;
    case 3: r = km_acos(a); break;
    case 4: r = km_atan2(a, y[i]); break;
    case 5: r = km_log(a); break;
    case 6: r = km_exp(a); break;
    case 7: r = km_pymod360(a); break;
    case 9: r = km_pymod_pos(a, y[i]); break;
    case 10: r = km_reflect_depth(a, 0.0, y[i]); break;
    default: r = km_sqrt(a); break;
    }
    out[i] = r;
}

template <typename T, bool IS3D>
static Snap<T, IS3D> snap_of(const FieldSlot& a, const FieldSlot* b, const FieldSlot* t)
{
    using R = typename RecOf<T, IS3D>::type;
    Snap<T, IS3D> f;
    f.a = static_cast<const R*>(a.rec);
    f.b = b ? static_cast<const R*>(b->rec) : f.a;
    f.t = t ? static_cast<const R*>(t->rec) : f.a;
    return f;
}

template <int MAXV, bool IS3D, typename T>
static void launch_step_v(kelp_ctx* c, int order, const Snap<T, IS3D>& f, const StepCfg& cfg, const PState& ps)
{
    unsigned g = grid_for(ps.n, 256);
    if (order == 1)
        step_kernel<1, MAXV, IS3D, T><<<g, 256, 0, c->stream>>>(c->m, f, cfg, ps);
    else if (order == 2)
        step_kernel<2, MAXV, IS3D, T><<<g, 256, 0, c->stream>>>(c->m, f, cfg, ps);
    else
        step_kernel<4, MAXV, IS3D, T><<<g, 256, 0, c->stream>>>(c->m, f, cfg, ps);
}

template <bool IS3D, typename T>
static void launch_step_t(kelp_ctx* c, int order, const FieldSlot& a, const FieldSlot* b, const FieldSlot* t,
                          const StepCfg& cfg, const PState& ps)
{
    Snap<T, IS3D> f = snap_of<T, IS3D>(a, b, t);
    if (c->m.max_v == 3)
        launch_step_v<3, IS3D, T>(c, order, f, cfg, ps);
    else
        launch_step_v<4, IS3D, T>(c, order, f, cfg, ps);
}

template <typename T>
static void launch_step(kelp_ctx* c, int order, bool is3d, const FieldSlot& a, const FieldSlot* b,
                        const FieldSlot* t, const StepCfg& cfg, const PState& ps)
{
    if (is3d)
        launch_step_t<true, T>(c, order, a, b, t, cfg, ps);
    else
        launch_step_t<false, T>(c, order, a, b, t, cfg, ps);
}

static const FieldSlot* opt_slot(kelp_ctx* c, int k)
{
    if (k < 0) return nullptr;
    return &c->slots[k];
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* kelp_last_error(void) { return g_err; }

int kelp_version(void) { return KELP_LIB_VERSION; }

int kelp_ctx_create(int device, kelp_ctx** out)
{
    if (!out) return fail(KELP_EINVAL, "kelp_ctx_create: out is NULL");
    *out = nullptr;
    int n_dev = 0;
    HIP_TRY(hipGetDeviceCount(&n_dev));
    if (device < 0 || device >= n_dev)
        return fail(KELP_EINVAL, "kelp_ctx_create: device %d not in [0, %d)", device, n_dev);
    HIP_TRY(hipSetDevice(device));
    kelp_ctx* c = new (std::nothrow) kelp_ctx();
    if (!c) return fail(KELP_ENOMEM, "kelp_ctx_create: host allocation failed");
    c->device = device;
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        delete c;
        return fail(KELP_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    for (int k = 0; k < kMaxEvents; ++k) {
        e = hipEventCreate(&c->events[k]);
        if (e != hipSuccess) {
            kelp_ctx_destroy(c);
            return fail(KELP_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    *out = c;
    return KELP_OK;
}

int kelp_ctx_destroy(kelp_ctx* c)
{
    if (!c) return KELP_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_mesh(c);
    for (auto& s : c->slots) free_slot(s);
    for (DevBuf* b : {&c->sort_tmp, &c->sort_keys, &c->sort_keys2, &c->sort_idx, &c->sort_perm, &c->sort_stage})
        kelp_release(*b);
    for (int k = 0; k < kMaxEvents; ++k)
        if (c->events[k]) (void)hipEventDestroy(c->events[k]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return KELP_OK;
}

int kelp_synchronize(kelp_ctx* c)
{
    if (!c) return fail(KELP_EINVAL, "kelp_synchronize: NULL context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return KELP_OK;
}

void* kelp_ctx_stream(kelp_ctx* c) { return c ? (void*)c->stream : nullptr; }

int kelp_mesh_upload(kelp_ctx* c, const kelp_mesh_desc* d)
{
    if (!c || !d) return fail(KELP_EINVAL, "kelp_mesh_upload: NULL argument");
    const int64_t N = d->n_nodes, E = d->n_elems;
    if (N <= 0 || E <= 0) return fail(KELP_EINVAL, "kelp_mesh_upload: empty mesh (N=%lld E=%lld)",
                                      (long long)N, (long long)E);
    if (N >= INT32_MAX || E >= INT32_MAX) return fail(KELP_EINVAL, "kelp_mesh_upload: int32 overflow");
    if (d->max_v != 3 && d->max_v != 4)
        return fail(KELP_EINVAL, "kelp_mesh_upload: max_v must be 3 or 4, got %lld", (long long)d->max_v);
    if (!d->host_nodes_lon || !d->host_nodes_lat || !d->host_elem_nodes || !d->host_node_to_elems ||
        !d->host_node_to_elems_count || !d->host_gh_offsets || !d->host_gh_node_ids)
        return fail(KELP_EINVAL, "kelp_mesh_upload: missing array");
    if (d->gh_nx <= 0 || d->gh_ny <= 0 || !(d->gh_cell > 0.0))
        return fail(KELP_EINVAL, "kelp_mesh_upload: bad grid hash geometry");
    const int64_t n_cells = d->gh_nx * d->gh_ny;
    if (d->host_gh_offsets[0] != 0 || d->host_gh_offsets[n_cells] != N)
        return fail(KELP_EINVAL, "kelp_mesh_upload: grid hash offsets do not span the nodes");
    if (d->n_levels < 0 || (d->n_levels > 0 && !d->host_z_levels))
        return fail(KELP_EINVAL, "kelp_mesh_upload: bad vertical levels");
    for (int64_t k = 1; k < d->n_levels; ++k)
        if (!(d->host_z_levels[k] >= d->host_z_levels[k - 1]))
            return fail(KELP_EINVAL, "kelp_mesh_upload: z levels must be non-decreasing");
    HIP_TRY(hipSetDevice(c->device));
    free_mesh(c);

    // Internal numbering.  Node k is the node in slot k of the hash's CSR
    // order (cells row-major, ascending mesh id inside a cell), so a cell's
    // nodes, a window's rows and the mesh neighbours of a region sit close in
    // memory.  Elements follow their lowest internal vertex.  Every order the
    // reference scans in is kept: the CSR slot order (nearest-node ties),
    // each node's element list in ascending mesh element id (first containing
    // element), each element's vertex slots (rings, fans).
    std::vector<int32_t> old_of_new(N), new_of_old(N, -1);
    for (int64_t k = 0; k < N; ++k) {
        int32_t o = d->host_gh_node_ids[k];
        if (o < 0 || o >= N || new_of_old[o] >= 0)
            return fail(KELP_EINVAL, "kelp_mesh_upload: grid hash node ids are not a permutation");
        old_of_new[k] = o;
        new_of_old[o] = (int32_t)k;
    }
    std::vector<int64_t> ekey(E);
    for (int64_t e = 0; e < E; ++e) {
        int64_t lo = INT64_MAX;
        for (int k = 0; k < d->max_v; ++k) {
            int x = d->host_elem_nodes[e * d->max_v + k];
            if (x < -1 || x >= N)
                return fail(KELP_EINVAL, "kelp_mesh_upload: elem %lld vertex %d out of range", (long long)e, x);
            if (x >= 0 && new_of_old[x] < lo) lo = new_of_old[x];
        }
        ekey[e] = lo;
    }
    std::vector<int32_t> e_old_of_new(E), e_new_of_old(E);
    for (int64_t e = 0; e < E; ++e) e_old_of_new[e] = (int32_t)e;
    std::stable_sort(e_old_of_new.begin(), e_old_of_new.end(),
                     [&](int32_t a, int32_t b) { return ekey[a] < ekey[b]; });
    for (int64_t e = 0; e < E; ++e) e_new_of_old[e_old_of_new[e]] = (int32_t)e;

    std::vector<double2> lonlat(N);
    std::vector<double> depth(N, 0.0);
    for (int64_t k = 0; k < N; ++k) {
        int64_t o = old_of_new[k];
        lonlat[k].x = d->host_nodes_lon[o];
        lonlat[k].y = d->host_nodes_lat[o];
        if (d->host_depth) depth[k] = d->host_depth[o];
    }
    std::vector<int4> elems(E);
    std::vector<uint8_t> coastal(d->host_elem_coastal ? E : 0);
    for (int64_t en = 0; en < E; ++en) {
        int64_t eo = e_old_of_new[en];
        int v[4] = {-1, -1, -1, -1};
        for (int k = 0; k < d->max_v; ++k) {
            int x = d->host_elem_nodes[eo * d->max_v + k];
            v[k] = x >= 0 ? new_of_old[x] : -1;
        }
        elems[en] = make_int4(v[0], v[1], v[2], v[3]);
        if (d->host_elem_coastal) coastal[en] = d->host_elem_coastal[eo] ? 1 : 0;
    }
    std::vector<int32_t> n2e_off(N + 1);
    n2e_off[0] = 0;
    for (int64_t k = 0; k < N; ++k) {
        int64_t o = old_of_new[k];
        int32_t cnt = d->host_node_to_elems_count[o];
        if (cnt < 0 || cnt > d->max_conn) return fail(KELP_EINVAL, "kelp_mesh_upload: bad node_to_elems_count");
        int32_t valid = 0;
        for (int32_t j = 0; j < cnt; ++j)
            if (d->host_node_to_elems[o * d->max_conn + j] >= 0) ++valid;
        n2e_off[k + 1] = n2e_off[k] + valid;
    }
    std::vector<int32_t> n2e(n2e_off[N] > 0 ? n2e_off[N] : 1);
    for (int64_t k = 0; k < N; ++k) {
        int64_t o = old_of_new[k];
        int32_t w = n2e_off[k];
        for (int32_t j = 0; j < d->host_node_to_elems_count[o]; ++j) {
            int32_t e = d->host_node_to_elems[o * d->max_conn + j];
            if (e >= E) return fail(KELP_EINVAL, "kelp_mesh_upload: node_to_elems entry out of range");
            if (e >= 0) n2e[w++] = e_new_of_old[e];
        }
    }

    DMesh m{};
    double2* d_lonlat;
    double4* d_xyz;
    int32_t *d_off, *d_n2e_off, *d_n2e, *d_nodeo, *d_noden, *d_elemo, *d_elemn;
    int4* d_elems;
    double *d_depth, *d_zlev;
    uint8_t* d_coastal = nullptr;
    int rc;
    if ((rc = dev_alloc(c, &d_lonlat, N)) || (rc = dev_alloc(c, &d_xyz, N)) ||
        (rc = dev_alloc(c, &d_off, n_cells + 1)) || (rc = dev_alloc(c, &d_elems, E)) ||
        (rc = dev_alloc(c, &d_n2e_off, N + 1)) || (rc = dev_alloc(c, &d_n2e, n2e.size())) ||
        (rc = dev_alloc(c, &d_depth, N)) || (rc = dev_alloc(c, &d_zlev, d->n_levels > 0 ? d->n_levels : 1)) ||
        (rc = dev_alloc(c, &d_nodeo, N)) || (rc = dev_alloc(c, &d_noden, N)) ||
        (rc = dev_alloc(c, &d_elemo, E)) || (rc = dev_alloc(c, &d_elemn, E)) ||
        (d->host_elem_coastal && (rc = dev_alloc(c, &d_coastal, E)))) {
        free_mesh(c);
        return rc;
    }
    HIP_TRY(hipMemcpy(d_lonlat, lonlat.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_off, d->host_gh_offsets, (n_cells + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_elems, elems.data(), E * sizeof(int4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_n2e_off, n2e_off.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_n2e, n2e.data(), n2e.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_depth, depth.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_nodeo, old_of_new.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_noden, new_of_old.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_elemo, e_old_of_new.data(), E * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_elemn, e_new_of_old.data(), E * sizeof(int32_t), hipMemcpyHostToDevice));
    if (d->n_levels > 0)
        HIP_TRY(hipMemcpy(d_zlev, d->host_z_levels, d->n_levels * sizeof(double), hipMemcpyHostToDevice));
    if (d_coastal) HIP_TRY(hipMemcpy(d_coastal, coastal.data(), E * sizeof(uint8_t), hipMemcpyHostToDevice));
    node_xyz_kernel<<<grid_for(N, 256), 256, 0, c->stream>>>(N, d_lonlat, d_xyz);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));

    m.lonlat = d_lonlat;
    m.gh_xyz = d_xyz;
    m.gh_offsets = d_off;
    m.elems = d_elems;
    m.n2e_off = d_n2e_off;
    m.n2e = d_n2e;
    m.depth = d_depth;
    m.zlev = d_zlev;
    m.coastal = d_coastal;
    m.cell = d->gh_cell;
    m.lon_min = d->gh_lon_min;
    m.lat_min = d->gh_lat_min;
    m.r_earth = d->r_earth;
    m.nx = (int32_t)d->gh_nx;
    m.ny = (int32_t)d->gh_ny;
    m.wrap = d->gh_wrap ? 1 : 0;
    m.far_rows = (int32_t)d->gh_far_rows;
    m.n_nodes = (int32_t)N;
    m.n_elems = (int32_t)E;
    m.max_v = (int32_t)d->max_v;
    m.n_levels = (int32_t)d->n_levels;
    c->m = m;
    c->node_old_of_new = d_nodeo;
    c->node_new_of_old = d_noden;
    c->elem_old_of_new = d_elemo;
    c->elem_new_of_old = d_elemn;
    c->has_mesh = true;
    return KELP_OK;
}

int kelp_mesh_node_xyz(kelp_ctx* c, double* host_out)
{
    if (!c || !host_out) return fail(KELP_EINVAL, "kelp_mesh_node_xyz: NULL argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_mesh_node_xyz: no mesh uploaded");
    const int64_t N = c->m.n_nodes;
    double* tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, N * 3 * sizeof(double)));
    gather_xyz_kernel<<<grid_for(N, 256), 256, 0, c->stream>>>(N, c->m.gh_xyz, c->node_old_of_new, tmp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, tmp, N * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(KELP_EHIP, "kelp_mesh_node_xyz: %s", hipGetErrorString(e));
    return KELP_OK;
}

int kelp_nodes_to_mesh_ids(kelp_ctx* c, int64_t n, int32_t* ids)
{
    if (!c || n < 0 || (n > 0 && !ids)) return fail(KELP_EINVAL, "kelp_nodes_to_mesh_ids: bad argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_nodes_to_mesh_ids: no mesh uploaded");
    if (n == 0) return KELP_OK;
    map_ids_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(n, c->node_old_of_new, ids);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

int kelp_nodes_from_mesh_ids(kelp_ctx* c, int64_t n, int32_t* ids)
{
    if (!c || n < 0 || (n > 0 && !ids)) return fail(KELP_EINVAL, "kelp_nodes_from_mesh_ids: bad argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_nodes_from_mesh_ids: no mesh uploaded");
    if (n == 0) return KELP_OK;
    map_ids_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(n, c->node_new_of_old, ids);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

int kelp_field_upload(kelp_ctx* c, int slot, int dtype, int64_t n_planes, const void* hu, const void* hv,
                      const void* hw, int64_t n_w_planes)
{
    if (!c) return fail(KELP_EINVAL, "kelp_field_upload: NULL context");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_field_upload: upload the mesh first");
    if (slot < 0 || slot >= kMaxSlots) return fail(KELP_EINVAL, "kelp_field_upload: slot %d not in [0, %d)", slot, kMaxSlots);
    if (dtype != KELP_DT_F64 && dtype != KELP_DT_F32) return fail(KELP_EINVAL, "kelp_field_upload: bad dtype %d", dtype);
    if (n_planes < 1 || !hu || !hv) return fail(KELP_EINVAL, "kelp_field_upload: u/v missing");
    const bool is3d = hw != nullptr;
    if (is3d && n_w_planes != n_planes + 1)
        return fail(KELP_EINVAL, "kelp_field_upload: W needs n_planes + 1 interface planes");
    if (!is3d && n_planes != 1) return fail(KELP_EINVAL, "kelp_field_upload: layered u/v need W (3-D)");
    HIP_TRY(hipSetDevice(c->device));
    FieldSlot& s = c->slots[slot];
    free_slot(s);
    const int64_t N = c->m.n_nodes;
    const size_t es = dtype == KELP_DT_F64 ? 8 : 4;
    const size_t uv_bytes = (size_t)n_planes * N * es;
    const size_t w_bytes = is3d ? (size_t)n_w_planes * N * es : 0;
    const size_t rec_bytes = (size_t)n_planes * N * es * (is3d ? 4 : 2);
    void *tu = nullptr, *tv = nullptr, *tw = nullptr;
    HIP_TRY(hipMalloc(&s.rec, rec_bytes));
    hipError_t e = hipMalloc(&tu, uv_bytes);
    if (e == hipSuccess) e = hipMalloc(&tv, uv_bytes);
    if (e == hipSuccess && is3d) e = hipMalloc(&tw, w_bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(tu, hu, uv_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tv, hv, uv_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && is3d) e = hipMemcpyAsync(tw, hw, w_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const int64_t n = is3d ? N * n_planes : N;
        const unsigned g = grid_for(n, 256);
        if (is3d && dtype == KELP_DT_F32)
            pack3d_kernel<float><<<g, 256, 0, c->stream>>>(N, n_planes, c->node_old_of_new, (const float*)tu,
                                                           (const float*)tv, (const float*)tw, (float4*)s.rec);
        else if (is3d)
            pack3d_kernel<double><<<g, 256, 0, c->stream>>>(N, n_planes, c->node_old_of_new, (const double*)tu,
                                                            (const double*)tv, (const double*)tw, (double4*)s.rec);
        else if (dtype == KELP_DT_F32)
            pack2d_kernel<float><<<g, 256, 0, c->stream>>>(N, c->node_old_of_new, (const float*)tu,
                                                           (const float*)tv, (float2*)s.rec);
        else
            pack2d_kernel<double><<<g, 256, 0, c->stream>>>(N, c->node_old_of_new, (const double*)tu,
                                                            (const double*)tv, (double2*)s.rec);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tu) (void)hipFree(tu);
    if (tv) (void)hipFree(tv);
    if (tw) (void)hipFree(tw);
    if (e != hipSuccess) {
        free_slot(s);
        return fail(e == hipErrorOutOfMemory ? KELP_ENOMEM : KELP_EHIP, "kelp_field_upload: %s", hipGetErrorString(e));
    }
    s.dtype = dtype;
    s.n_planes = n_planes;
    s.n_w_planes = is3d ? n_w_planes : 0;
    s.is3d = is3d;
    return KELP_OK;
}

int kelp_field_clear(kelp_ctx* c, int slot)
{
    if (!c || slot < 0 || slot >= kMaxSlots) return fail(KELP_EINVAL, "kelp_field_clear: bad argument");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_slot(c->slots[slot]);
    return KELP_OK;
}

int kelp_step(kelp_ctx* c, const kelp_particles* p, const kelp_step_args* a)
{
    if (!c || !p || !a) return fail(KELP_EINVAL, "kelp_step: NULL argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_step: no mesh uploaded");
    if (a->order != 1 && a->order != 2 && a->order != 4)
        return fail(KELP_EINVAL, "kelp_step: order must be 1, 2 or 4 (got %d)", a->order);
    if (!(a->dt > 0.0)) return fail(KELP_EINVAL, "kelp_step: dt must be positive (direction goes in time_dir)");
    if (a->time_dir != 1.0 && a->time_dir != -1.0) return fail(KELP_EINVAL, "kelp_step: time_dir must be +1 or -1");
    if (a->w_method != 0 && a->w_method != 2)
        return fail(KELP_EINVAL, "kelp_step: w_method must be 0 (nearest) or 2 (barycentric)");
    if (p->n < 0) return fail(KELP_EINVAL, "kelp_step: negative particle count");
    if (p->n == 0) return KELP_OK;
    if (!p->lon || !p->lat || !p->status || !p->nn || !p->fr || !p->u || !p->v)
        return fail(KELP_EINVAL, "kelp_step: missing particle array");
    if ((a->kick || a->beach || a->az) && !p->pid)
        return fail(KELP_EINVAL, "kelp_step: the kicks and beaching need pid");
    if (a->az && !(a->az_const >= 0.0 && a->az_const < __builtin_inf()))
        return fail(KELP_EINVAL, "kelp_step: az_const must be finite and >= 0");
    if (a->beach && !p->wet_streak) return fail(KELP_EINVAL, "kelp_step: beaching needs wet_streak");
    if (a->kick && (!(a->kh_x >= 0.0) || !(a->kh_y >= 0.0))) return fail(KELP_EINVAL, "kelp_step: Kh must be >= 0");
    if (a->beach && (a->use_coastal || a->use_stochastic) && !c->m.coastal)
        return fail(KELP_ESTATE, "kelp_step: coastal/stochastic beaching needs the coastal-element flags");
    for (int k : {a->slot_a, a->slot_b, a->slot_t})
        if (k >= kMaxSlots || k < -1) return fail(KELP_EINVAL, "kelp_step: slot %d out of range", k);
    if (a->slot_a < 0 || c->slots[a->slot_a].dtype < 0) return fail(KELP_ESTATE, "kelp_step: slot_a is empty");
    const FieldSlot& sa = c->slots[a->slot_a];
    const FieldSlot* sb = opt_slot(c, a->slot_b);
    const FieldSlot* st = opt_slot(c, a->slot_t);
    for (const FieldSlot* s : {sb, st}) {
        if (!s) continue;
        if (s->dtype < 0) return fail(KELP_ESTATE, "kelp_step: slot_b/slot_t is empty");
        if (s->dtype != sa.dtype || s->n_planes != sa.n_planes || s->n_w_planes != sa.n_w_planes)
            return fail(KELP_EINVAL, "kelp_step: snapshot slots differ in dtype or shape");
    }
    if (a->alpha0 != 0.0 && !sb) return fail(KELP_EINVAL, "kelp_step: alpha0 != 0 needs slot_b");
    for (int k = 0; k < a->order; ++k)
        if (a->alphas[k] >= 1e-12 && !st) return fail(KELP_EINVAL, "kelp_step: stage alpha > 0 needs slot_t");
    const bool is3d = sa.is3d;
    const int K = c->m.n_levels;
    if (is3d) {
        if (K < 2) return fail(KELP_EINVAL, "kelp_step: a 3-D field needs a mesh with >= 2 levels");
        if (sa.n_planes != K - 1 || sa.n_w_planes != K)
            return fail(KELP_EINVAL, "kelp_step: 3-D field shapes (%lld, %lld) do not match K=%d",
                        (long long)sa.n_planes, (long long)sa.n_w_planes, K);
        if (!p->z) return fail(KELP_EINVAL, "kelp_step: 3-D step needs z");
    } else if (a->az && a->az_const > 0.0) {
        return fail(KELP_EINVAL, "kelp_step: the Az kick needs a 3-D run (integrators.py:893)");
    } else if (sa.n_planes != 1) {
        return fail(KELP_EINVAL, "kelp_step: a layered u/v field needs W (3-D)");
    }
    HIP_TRY(hipSetDevice(c->device));
    StepCfg cfg{};
    cfg.alpha0 = a->alpha0;
    for (int k = 0; k < 4; ++k) cfg.alphas[k] = a->alphas[k];
    cfg.dt = a->dt;
    cfg.time_dir = a->time_dir;
    if (a->exact_nearest < 0 || a->exact_nearest > 2)
        return fail(KELP_EINVAL, "kelp_step: exact_nearest must be 0, 1 or 2");
    cfg.exact = a->exact_nearest;
    cfg.w_method = a->w_method;
    cfg.kick = a->kick ? 1 : 0;
    cfg.beach = a->beach ? 1 : 0;
    cfg.kh_x = a->kh_x;
    cfg.kh_y = a->kh_y;
    cfg.seed = a->seed;
    cfg.real_id = a->real_id;
    cfg.step_i = a->step_i;
    cfg.use_depth = a->use_depth;
    cfg.use_coastal = a->use_coastal;
    cfg.use_stochastic = a->use_stochastic;
    cfg.on_failure = a->on_failure;
    cfg.resuspension = a->resuspension;
    cfg.min_wet_steps = a->min_wet_steps;
    cfg.depth_min = a->depth_min;
    cfg.zone_depth = a->zone_depth;
    cfg.t_beach = a->t_beach;
    cfg.rewet_hysteresis = a->rewet_hysteresis;
    cfg.t_resuspend = a->t_resuspend;
    cfg.beach_dt = a->beach_dt;
    cfg.beach_seed = a->beach_seed;
    cfg.az = a->az && a->az_const > 0.0;
    cfg.az_const = a->az_const;
    PState ps{p->lon, p->lat, p->z, p->status, p->nn, p->fr, p->u, p->v, p->w, p->pid, p->wet_streak,
              reinterpret_cast<unsigned long long*>(p->alive_count), p->n};
    if (sa.dtype == KELP_DT_F32)
        launch_step<float>(c, a->order, is3d, sa, sb, st, cfg, ps);
    else
        launch_step<double>(c, a->order, is3d, sa, sb, st, cfg, ps);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

int kelp_nearest(kelp_ctx* c, int64_t n, const double* lon, const double* lat, int exact, int32_t* out)
{
    if (!c || (n > 0 && (!lon || !lat || !out))) return fail(KELP_EINVAL, "kelp_nearest: NULL argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_nearest: no mesh uploaded");
    if (n <= 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    nearest_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(c->m, n, lon, lat, exact, c->node_old_of_new, out);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

int kelp_containing(kelp_ctx* c, int64_t n, const double* lon, const double* lat, const int32_t* nn,
                    int32_t* out)
{
    if (!c || (n > 0 && (!lon || !lat || !nn || !out))) return fail(KELP_EINVAL, "kelp_containing: NULL argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_containing: no mesh uploaded");
    if (n <= 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    if (c->m.max_v == 3)
        containing_kernel<3><<<grid_for(n, 256), 256, 0, c->stream>>>(c->m, n, lon, lat, nn, c->node_new_of_old,
                                                                       c->elem_old_of_new, out);
    else
        containing_kernel<4><<<grid_for(n, 256), 256, 0, c->stream>>>(c->m, n, lon, lat, nn, c->node_new_of_old,
                                                                       c->elem_old_of_new, out);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

int kelp_malloc(kelp_ctx* c, int64_t bytes, void** out)
{
    if (!c || !out || bytes < 0) return fail(KELP_EINVAL, "kelp_malloc: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMalloc(out, bytes > 0 ? (size_t)bytes : 16));
    return KELP_OK;
}

int kelp_free(kelp_ctx* c, void* ptr)
{
    if (!c) return fail(KELP_EINVAL, "kelp_free: NULL context");
    if (!ptr) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(ptr));
    return KELP_OK;
}

int kelp_memcpy_h2d(kelp_ctx* c, void* dst, const void* src, int64_t bytes)
{
    if (!c || bytes < 0 || (bytes > 0 && (!dst || !src))) return fail(KELP_EINVAL, "kelp_memcpy_h2d: bad argument");
    if (bytes == 0) return KELP_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return KELP_OK;
}

int kelp_memcpy_d2h(kelp_ctx* c, void* dst, const void* src, int64_t bytes)
{
    if (!c || bytes < 0 || (bytes > 0 && (!dst || !src))) return fail(KELP_EINVAL, "kelp_memcpy_d2h: bad argument");
    if (bytes == 0) return KELP_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return KELP_OK;
}

int kelp_event_record(kelp_ctx* c, int slot)
{
    if (!c || slot < 0 || slot >= kMaxEvents) return fail(KELP_EINVAL, "kelp_event_record: bad argument");
    HIP_TRY(hipEventRecord(c->events[slot], c->stream));
    return KELP_OK;
}

int kelp_event_elapsed(kelp_ctx* c, int a, int b, float* ms)
{
    if (!c || !ms || a < 0 || b < 0 || a >= kMaxEvents || b >= kMaxEvents)
        return fail(KELP_EINVAL, "kelp_event_elapsed: bad argument");
    HIP_TRY(hipEventSynchronize(c->events[b]));
    HIP_TRY(hipEventElapsedTime(ms, c->events[a], c->events[b]));
    return KELP_OK;
}

int kelp_math_probe(kelp_ctx* c, int fn, int64_t n, const double* x, const double* y, double* out)
{
    if (!c || n < 0 || (n > 0 && (!x || !y || !out))) return fail(KELP_EINVAL, "kelp_math_probe: bad argument");
    if (n == 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    math_probe_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(fn, n, x, y, out);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

}  // extern "C"
