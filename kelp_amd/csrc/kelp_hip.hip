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

#include "../../include/kelp_hip.h"
#include "kelp_device.cuh"

using namespace kelp;

#define KELP_LIB_VERSION 1

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(e_ == hipErrorOutOfMemory ? KELP_ENOMEM : KELP_EHIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

constexpr int kMaxSlots = 16;
constexpr int kMaxEvents = 16;

struct FieldSlot {
    int dtype = -1;  // -1: empty
    int64_t n_planes = 0, n_w_planes = 0;
    void* u = nullptr;
    void* v = nullptr;
    void* w = nullptr;
};

struct kelp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t events[kMaxEvents] = {};
    bool has_mesh = false;
    DMesh m{};
    // owned device buffers of the mesh
    std::vector<void*> mesh_bufs;
    FieldSlot slots[kMaxSlots];
};

static void free_mesh(kelp_ctx* c)
{
    for (void* p : c->mesh_bufs) (void)hipFree(p);
    c->mesh_bufs.clear();
    c->has_mesh = false;
    c->m = DMesh{};
}

static void free_slot(FieldSlot& s)
{
    if (s.u) (void)hipFree(s.u);
    if (s.v) (void)hipFree(s.v);
    if (s.w) (void)hipFree(s.w);
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

// core/mesh.py:713 _node_xyz_fill, written straight into the CSR order of
// the hash (slot k holds node gh_node_ids[k]).
__global__ void node_xyz_csr_kernel(int64_t n, const double2* __restrict__ lonlat,
                                    const int32_t* __restrict__ ids, double4* __restrict__ out)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int32_t node = ids[k];
    double2 ll = lonlat[node];
    double lat_r = ll.y * kDeg2Rad, lon_r = ll.x * kDeg2Rad;
    double c = km_cos(lat_r);
    double4 r;
    r.x = c * km_cos(lon_r);
    r.y = c * km_sin(lon_r);
    r.z = km_sin(lat_r);
    r.w = __longlong_as_double((long long)node);
    out[k] = r;
}

__global__ void gather_xyz_kernel(int64_t n, const double4* __restrict__ csr, double* __restrict__ out)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 p = csr[k];
    int32_t node = node_id_of(p);
    out[3 * (int64_t)node + 0] = p.x;
    out[3 * (int64_t)node + 1] = p.y;
    out[3 * (int64_t)node + 2] = p.z;
}

static unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------------------------
// step kernels
// ---------------------------------------------------------------------------

struct StepParams {
    double alpha0;
    double alphas[4];
    double dt, time_dir;
    int exact;
};

struct PState {
    double* __restrict__ lon;
    double* __restrict__ lat;
    int32_t* __restrict__ status;
    int32_t* __restrict__ nn;
    int8_t* __restrict__ fr;
    double* __restrict__ u;
    double* __restrict__ v;
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

// interpolate.py:1341 _interp_barycentric_core_tinterp (2-D node layout):
// containment around the anchor, vertex weights, blended vertex values.
template <typename T>
__device__ inline bool interp_uv_2d(const DMesh& m, const StageFields<T>& fu, const StageFields<T>& fv,
                                    double alpha0, double alpha, double time_dir, double xp, double yp,
                                    int nn, double& U, double& V)
{
    U = 0.0;
    V = 0.0;
    int e = find_containing(m, xp, yp, nn);
    if (e < 0) return false;
    Weights W;
    if (!vertex_weights(m, e, xp, yp, W)) return false;
    double su = 0.0, sv = 0.0;
    for (int k = 0; k < W.n; ++k) {
        int64_t idx = W.v[k];
        double un = vblend(fu, idx, alpha0, alpha);
        double vn = vblend(fv, idx, alpha0, alpha);
        su += W.w[k] * un;
        sv += W.w[k] * vn;
    }
    U = time_dir * su;
    V = time_dir * sv;
    return true;
}

// One fused 2-D step per particle: integrators.py:2522
// _fused_rk2_step_2d_bary_tinterp (ORDER 2), :2642 _fused_rk4_step_2d_bary_tinterp
// (ORDER 4) and the barycentric Euler of :4997 _euler_step_unfused (ORDER 1).
template <int ORDER, typename T>
__global__ __launch_bounds__(256) void step2d_kernel(DMesh m, StageFields<T> fu, StageFields<T> fv,
                                                     StepParams sp, PState s)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    if (s.status[i] != kAlive) {
        s.u[i] = 0.0;
        s.v[i] = 0.0;
        return;
    }
    const double lon0 = s.lon[i], lat0 = s.lat[i];
    const double dt = sp.dt, td = sp.time_dir;
    const bool exact = sp.exact != 0;
    int node_i = site_anchor(m, lon0, lat0, exact);
    double U1, V1;
    bool ok = interp_uv_2d(m, fu, fv, sp.alpha0, sp.alphas[0], td, lon0, lat0, node_i, U1, V1);
    s.u[i] = U1;
    s.v[i] = V1;
    double lon_new = lon0, lat_new = lat0;
    if (ok) {
        if (ORDER == 1) {
            lon_new = wrap_lon(lon0 + dt * U1 / z0_of(lat0));
            lat_new = clamp_lat(lat0 + dt * V1 / kSLM);
        } else if (ORDER == 2) {
            double z0 = z0_of(lat0);
            double lon_mid = wrap_lon(lon0 + 0.5 * dt * U1 / z0);
            double lat_mid = clamp_lat(lat0 + 0.5 * dt * V1 / kSLM);
            int nn_mid = site_anchor(m, lon_mid, lat_mid, exact);
            double U2, V2;
            ok = interp_uv_2d(m, fu, fv, sp.alpha0, sp.alphas[1], td, lon_mid, lat_mid, nn_mid, U2, V2);
            if (ok) {
                double z0m = z0_of(lat_mid);
                lon_new = wrap_lon(lon0 + dt * U2 / z0m);
                lat_new = clamp_lat(lat0 + dt * V2 / kSLM);
            }
        } else {
            double dlon1 = dt * U1 / z0_of(lat0);
            double dlat1 = dt * V1 / kSLM;
            double lon2 = wrap_lon(lon0 + 0.5 * dlon1);
            double lat2 = clamp_lat(lat0 + 0.5 * dlat1);
            int nn2 = site_anchor(m, lon2, lat2, exact);
            double U2, V2, U3, V3, U4, V4;
            ok = interp_uv_2d(m, fu, fv, sp.alpha0, sp.alphas[1], td, lon2, lat2, nn2, U2, V2);
            if (ok) {
                double dlon2 = dt * U2 / z0_of(lat2);
                double dlat2 = dt * V2 / kSLM;
                double lon3 = wrap_lon(lon0 + 0.5 * dlon2);
                double lat3 = clamp_lat(lat0 + 0.5 * dlat2);
                int nn3 = site_anchor(m, lon3, lat3, exact);
                ok = interp_uv_2d(m, fu, fv, sp.alpha0, sp.alphas[2], td, lon3, lat3, nn3, U3, V3);
                if (ok) {
                    double dlon3 = dt * U3 / z0_of(lat3);
                    double dlat3 = dt * V3 / kSLM;
                    double lon4 = wrap_lon(lon0 + dlon3);
                    double lat4 = clamp_lat(lat0 + dlat3);
                    int nn4 = site_anchor(m, lon4, lat4, exact);
                    ok = interp_uv_2d(m, fu, fv, sp.alpha0, sp.alphas[3], td, lon4, lat4, nn4, U4, V4);
                    if (ok) {
                        double dlon4 = dt * U4 / z0_of(lat4);
                        double dlat4 = dt * V4 / kSLM;
                        double dlon = (dlon1 + 2.0 * dlon2 + 2.0 * dlon3 + dlon4) / 6.0;
                        double dlat = (dlat1 + 2.0 * dlat2 + 2.0 * dlat3 + dlat4) / 6.0;
                        lon_new = wrap_lon(lon0 + dlon);
                        lat_new = clamp_lat(lat0 + dlat);
                    }
                }
            }
        }
    }
    if (!ok) {
        s.status[i] = kOutside;
        s.fr[i] = kFrInterp;
        s.nn[i] = node_i;
        return;
    }
    int nn_new = site_anchor(m, lon_new, lat_new, exact);
    int e = find_containing(m, lon_new, lat_new, nn_new);
    if (e < 0) {
        s.status[i] = kOutside;
        s.fr[i] = kFrOOB;
        s.nn[i] = node_i;
        return;
    }
    s.lon[i] = lon_new;
    s.lat[i] = lat_new;
    s.nn[i] = nn_new;
}

__global__ void nearest_kernel(DMesh m, int64_t n, const double* __restrict__ lon,
                               const double* __restrict__ lat, int exact, int32_t* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (exact) {
        out[i] = nearest_exact(m, lon[i], lat[i]);
    } else {
        // guard off: the plain 3x3 window; an empty window is healed by the
        // exact search (the reference heals it with a KD-tree query).
        int ci, cj;
        gh_cell_of(m, lon[i], lat[i], ci, cj);
        Query q = make_query(lon[i], lat[i]);
        double d2;
        int b = nearest_3x3(m, ci, cj, q, d2);
        out[i] = b >= 0 ? b : nearest_exact(m, lon[i], lat[i]);
    }
}

__global__ void containing_kernel(DMesh m, int64_t n, const double* __restrict__ lon,
                                  const double* __restrict__ lat, const int32_t* __restrict__ nn,
                                  int32_t* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = find_containing(m, lon[i], lat[i], nn[i]);
}

template <typename T>
static StageFields<T> stage_of(const FieldSlot& a, const FieldSlot* b, const FieldSlot* t, bool u)
{
    StageFields<T> f;
    f.a = static_cast<const T*>(u ? a.u : a.v);
    f.b = b ? static_cast<const T*>(u ? b->u : b->v) : f.a;
    f.t = t ? static_cast<const T*>(u ? t->u : t->v) : f.a;
    return f;
}

template <typename T>
static void launch_step2d(kelp_ctx* c, int order, const FieldSlot& a, const FieldSlot* b, const FieldSlot* t,
                          const StepParams& sp, const PState& ps)
{
    StageFields<T> fu = stage_of<T>(a, b, t, true), fv = stage_of<T>(a, b, t, false);
    unsigned g = grid_for(ps.n, 256);
    if (order == 1)
        step2d_kernel<1, T><<<g, 256, 0, c->stream>>>(c->m, fu, fv, sp, ps);
    else if (order == 2)
        step2d_kernel<2, T><<<g, 256, 0, c->stream>>>(c->m, fu, fv, sp, ps);
    else
        step2d_kernel<4, T><<<g, 256, 0, c->stream>>>(c->m, fu, fv, sp, ps);
}

static const FieldSlot* opt_slot(kelp_ctx* c, int k)
{
    if (k < 0) return nullptr;
    return &c->slots[k];
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
    default: r = km_sqrt(a); break;
    }
    out[i] = r;
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
    HIP_TRY(hipSetDevice(c->device));
    free_mesh(c);

    // host staging in the device layout
    std::vector<double2> lonlat(N);
    for (int64_t n = 0; n < N; ++n) {
        lonlat[n].x = d->host_nodes_lon[n];
        lonlat[n].y = d->host_nodes_lat[n];
    }
    std::vector<int4> elems(E);
    for (int64_t e = 0; e < E; ++e) {
        int v[4] = {-1, -1, -1, -1};
        for (int k = 0; k < d->max_v; ++k) {
            int x = d->host_elem_nodes[e * d->max_v + k];
            if (x < -1 || x >= N) return fail(KELP_EINVAL, "kelp_mesh_upload: elem %lld vertex %d out of range",
                                              (long long)e, x);
            v[k] = x;
        }
        elems[e] = make_int4(v[0], v[1], v[2], v[3]);
    }
    std::vector<int32_t> n2e_off(N + 1);
    n2e_off[0] = 0;
    for (int64_t n = 0; n < N; ++n) {
        int32_t cnt = d->host_node_to_elems_count[n];
        if (cnt < 0 || cnt > d->max_conn) return fail(KELP_EINVAL, "kelp_mesh_upload: bad node_to_elems_count");
        int32_t valid = 0;
        for (int32_t k = 0; k < cnt; ++k)
            if (d->host_node_to_elems[n * d->max_conn + k] >= 0) ++valid;
        n2e_off[n + 1] = n2e_off[n] + valid;
    }
    std::vector<int32_t> n2e(n2e_off[N] > 0 ? n2e_off[N] : 1);
    for (int64_t n = 0; n < N; ++n) {
        int32_t w = n2e_off[n];
        for (int32_t k = 0; k < d->host_node_to_elems_count[n]; ++k) {
            int32_t e = d->host_node_to_elems[n * d->max_conn + k];
            if (e >= E) return fail(KELP_EINVAL, "kelp_mesh_upload: node_to_elems entry out of range");
            if (e >= 0) n2e[w++] = e;
        }
    }
    for (int64_t k = 0; k < N; ++k)
        if (d->host_gh_node_ids[k] < 0 || d->host_gh_node_ids[k] >= N)
            return fail(KELP_EINVAL, "kelp_mesh_upload: grid hash node id out of range");

    DMesh m{};
    double2* d_lonlat;
    double4* d_xyz;
    int32_t *d_off, *d_ids, *d_n2e_off, *d_n2e;
    int4* d_elems;
    double *d_depth, *d_zlev;
    int rc;
    if ((rc = dev_alloc(c, &d_lonlat, N)) || (rc = dev_alloc(c, &d_xyz, N)) ||
        (rc = dev_alloc(c, &d_off, n_cells + 1)) || (rc = dev_alloc(c, &d_ids, N)) ||
        (rc = dev_alloc(c, &d_elems, E)) || (rc = dev_alloc(c, &d_n2e_off, N + 1)) ||
        (rc = dev_alloc(c, &d_n2e, n2e.size())) || (rc = dev_alloc(c, &d_depth, N)) ||
        (rc = dev_alloc(c, &d_zlev, d->n_levels > 0 ? d->n_levels : 1))) {
        free_mesh(c);
        return rc;
    }
    std::vector<double> depth(N, 0.0);
    if (d->host_depth)
        for (int64_t n = 0; n < N; ++n) depth[n] = d->host_depth[n];
    HIP_TRY(hipMemcpy(d_lonlat, lonlat.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_off, d->host_gh_offsets, (n_cells + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_ids, d->host_gh_node_ids, N * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_elems, elems.data(), E * sizeof(int4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_n2e_off, n2e_off.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_n2e, n2e.data(), n2e.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_depth, depth.data(), N * sizeof(double), hipMemcpyHostToDevice));
    if (d->n_levels > 0)
        HIP_TRY(hipMemcpy(d_zlev, d->host_z_levels, d->n_levels * sizeof(double), hipMemcpyHostToDevice));
    node_xyz_csr_kernel<<<grid_for(N, 256), 256, 0, c->stream>>>(N, d_lonlat, d_ids, d_xyz);
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
    gather_xyz_kernel<<<grid_for(N, 256), 256, 0, c->stream>>>(N, c->m.gh_xyz, tmp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, tmp, N * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(KELP_EHIP, "kelp_mesh_node_xyz: %s", hipGetErrorString(e));
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
    if (hw && n_w_planes < 1) return fail(KELP_EINVAL, "kelp_field_upload: w planes");
    HIP_TRY(hipSetDevice(c->device));
    FieldSlot& s = c->slots[slot];
    free_slot(s);
    const size_t es = dtype == KELP_DT_F64 ? 8 : 4;
    const size_t uv_bytes = (size_t)n_planes * c->m.n_nodes * es;
    HIP_TRY(hipMalloc(&s.u, uv_bytes));
    HIP_TRY(hipMalloc(&s.v, uv_bytes));
    HIP_TRY(hipMemcpyAsync(s.u, hu, uv_bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(s.v, hv, uv_bytes, hipMemcpyHostToDevice, c->stream));
    if (hw) {
        const size_t w_bytes = (size_t)n_w_planes * c->m.n_nodes * es;
        HIP_TRY(hipMalloc(&s.w, w_bytes));
        HIP_TRY(hipMemcpyAsync(s.w, hw, w_bytes, hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    s.dtype = dtype;
    s.n_planes = n_planes;
    s.n_w_planes = hw ? n_w_planes : 0;
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
    if (p->n < 0) return fail(KELP_EINVAL, "kelp_step: negative particle count");
    if (p->n == 0) return KELP_OK;
    if (!p->lon || !p->lat || !p->status || !p->nn || !p->fr || !p->u || !p->v)
        return fail(KELP_EINVAL, "kelp_step: missing particle array");
    for (int k : {a->slot_a, a->slot_b, a->slot_t})
        if (k >= kMaxSlots || k < -1) return fail(KELP_EINVAL, "kelp_step: slot %d out of range", k);
    if (a->slot_a < 0 || c->slots[a->slot_a].dtype < 0) return fail(KELP_ESTATE, "kelp_step: slot_a is empty");
    const FieldSlot& sa = c->slots[a->slot_a];
    const FieldSlot* sb = opt_slot(c, a->slot_b);
    const FieldSlot* st = opt_slot(c, a->slot_t);
    for (const FieldSlot* s : {sb, st}) {
        if (!s) continue;
        if (s->dtype != sa.dtype || s->n_planes != sa.n_planes)
            return fail(KELP_EINVAL, "kelp_step: snapshot slots differ in dtype or shape");
    }
    if (a->alpha0 != 0.0 && !sb) return fail(KELP_EINVAL, "kelp_step: alpha0 != 0 needs slot_b");
    const int n_st = a->order;
    for (int k = 0; k < n_st; ++k)
        if (a->alphas[k] >= 1e-12 && !st) return fail(KELP_EINVAL, "kelp_step: stage alpha > 0 needs slot_t");
    if (sa.n_planes != 1) return fail(KELP_EINVAL, "kelp_step: 3-D fields are not supported by this entry point");
    HIP_TRY(hipSetDevice(c->device));
    StepParams sp{};
    sp.alpha0 = a->alpha0;
    for (int k = 0; k < 4; ++k) sp.alphas[k] = a->alphas[k];
    sp.dt = a->dt;
    sp.time_dir = a->time_dir;
    sp.exact = a->exact_nearest ? 1 : 0;
    PState ps{p->lon, p->lat, p->status, p->nn, p->fr, p->u, p->v, p->n};
    if (sa.dtype == KELP_DT_F32)
        launch_step2d<float>(c, a->order, sa, sb, st, sp, ps);
    else
        launch_step2d<double>(c, a->order, sa, sb, st, sp, ps);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

int kelp_nearest(kelp_ctx* c, int64_t n, const double* lon, const double* lat, int exact, int32_t* out)
{
    if (!c || (n > 0 && (!lon || !lat || !out))) return fail(KELP_EINVAL, "kelp_nearest: NULL argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_nearest: no mesh uploaded");
    if (n <= 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    nearest_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(c->m, n, lon, lat, exact, out);
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
    containing_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(c->m, n, lon, lat, nn, out);
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
