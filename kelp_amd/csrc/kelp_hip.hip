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

#include "../../include/kelp_grid_hash.h"
#include "kelp_internal.h"
#include "kelp_step.cuh"

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
    c->vertical_coord = 0;
    c->sig_eta = nullptr;
    c->sig_given[0] = c->sig_given[1] = false;
}

void kelp_free_scalar(ScalarField& f)
{
    if (f.planes) (void)hipFree(f.planes);
    if (f.wrec) (void)hipFree(f.wrec);
    f = ScalarField{};
}

static void free_slot(FieldSlot& s)
{
    if (s.rec) (void)hipFree(s.rec);
    if (s.erec) (void)hipFree(s.erec);
    if (s.eta) (void)hipFree(s.eta);
    if (s.az) (void)hipFree(s.az);
    if (s.pair) (void)hipFree(s.pair);
    if (s.scale) (void)hipFree(s.scale);
    for (auto& f : s.scalars) kelp_free_scalar(f);
    for (void* p : s.stage)
        if (p) (void)hipFree(p);
    if (s.ready) (void)hipEventDestroy(s.ready);
    s = FieldSlot{};
}

int kelp_wait_slot(kelp_ctx* c, FieldSlot& s)
{
    if (!s.pending) return KELP_OK;
    HIP_TRY(hipStreamWaitEvent(c->stream, s.ready, 0));
    s.pending = false;
    return KELP_OK;
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
__global__ void node_xyz_kernel(int64_t n, const double2* __restrict__ lonlat, const int32_t* __restrict__ n2e_off,
                                double4* __restrict__ out, double* __restrict__ coslat)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double2 ll = lonlat[k];
    double lat_r = ll.y * kDeg2Rad, lon_r = ll.x * kDeg2Rad;
    double c = km_cos(lat_r);
    coslat[k] = c;
    double4 r;
    r.x = c * km_cos(lon_r);
    r.y = c * km_sin(lon_r);
    r.z = km_sin(lat_r);
    // the node's element-list range [start, end) (n2e_range)
    r.w = __hiloint2double(n2e_off[k + 1], n2e_off[k]);
    out[k] = r;
}

// containment frames of all elements (elem_geom)
template <int MAXV>
__global__ void elem_geom_kernel(DMesh m, double4* __restrict__ eframe, uint8_t* __restrict__ eflags)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n_elems) return;
    double4 g0, g1;
    double2 xy[4];
    uint8_t fl;
    elem_geom<MAXV>(m, (int)c, g0, g1, xy, fl);
    eflags[c] = fl;
    eframe[4 * c] = g0;
    eframe[4 * c + 1] = g1;
    eframe[4 * c + 2] = make_double4(xy[0].x, xy[0].y, xy[1].x, xy[1].y);
    double4 last = make_double4(xy[2].x, xy[2].y, xy[3].x, xy[3].y);
    if (MAXV == 3) {
        // no fourth vertex: the slot carries the vertex ids (elem_nodes)
        const int4 en = m.elems[c];
        last.z = __hiloint2double(en.y, en.x);
        last.w = __hiloint2double(en.w, en.z);
    }
    eframe[4 * c + 3] = last;
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
    r.x = u ? u[L * N + o] : (T)0;
    r.y = v ? v[L * N + o] : (T)0;
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

// float32 planes into bfloat16 records (KELP_DT_BF16), 3-D and 2-D
__global__ void pack3d_bf16_kernel(int64_t N, int64_t n_layers, const int32_t* __restrict__ old_of_new,
                                   const float* __restrict__ u, const float* __restrict__ v,
                                   const float* __restrict__ w, Bf16x4* __restrict__ rec)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * n_layers) return;
    int64_t L = i / N, k = i - L * N;
    int64_t o = old_of_new[k];
    Bf16x4 r;
    r.x = to_bf16(u[L * N + o]);
    r.y = to_bf16(v[L * N + o]);
    r.z = to_bf16(w[L * N + o]);
    r.w = to_bf16(w[(L + 1) * N + o]);
    rec[i] = r;
}

__global__ void pack2d_bf16_kernel(int64_t N, const int32_t* __restrict__ old_of_new, const float* __restrict__ u,
                                   const float* __restrict__ v, Bf16x2* __restrict__ rec)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    int64_t o = old_of_new[k];
    Bf16x2 r;
    r.x = to_bf16(u[o]);
    r.y = to_bf16(v[o]);
    rec[k] = r;
}

// float32 planes into float16 records with a power-of-two scale per level
// (KELP_DT_F16S): the largest finite |value| of every u and v layer row and
// every W interface row, the scale 2^(q - 15) of a row maximum m = f 2^q
// (f in [0.5, 1), so |value| / scale < 2^15, inside the half range; 1 for
// an empty or all-zero row, u and v of a layer sharing the larger), then
// value * (1 / scale) rounded to the nearest even half.
__global__ void f16s_clear_kernel(int64_t n, uint32_t* __restrict__ mx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mx[i] = 0u;
}

// row r of blockIdx.y: u rows [0, n_u), v rows [n_u, 2 n_u), W rows after
__global__ __launch_bounds__(256) void f16s_max_kernel(int64_t N, int64_t n_u, const float* __restrict__ u,
                                                       const float* __restrict__ v, const float* __restrict__ w,
                                                       uint32_t* __restrict__ mx)
{
    __shared__ float red[256];
    const int64_t r = blockIdx.y;
    const float* row = r < n_u ? u + r * N : (r < 2 * n_u ? v + (r - n_u) * N : w + (r - 2 * n_u) * N);
    float m = 0.0f;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (int64_t)gridDim.x * blockDim.x) {
        const float a = fabsf(row[k]);
        if (a <= 3.402823466e38f && a > m) m = a;  // finite values only
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h && red[threadIdx.x + h] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&mx[r], __float_as_uint(red[0]));
}

__device__ __forceinline__ float f16s_scale_of(uint32_t mbits)
{
    const float m = __uint_as_float(mbits);
    if (!(m > 0.0f)) return 1.0f;
    int q;
    (void)frexpf(m, &q);
    int e = q - 15;
    if (e < -126) e = -126;
    if (e > 127) e = 127;
    return ldexpf(1.0f, e);
}

__global__ void f16s_scale_kernel(int64_t n_layers, int is3d, const uint32_t* __restrict__ mx,
                                  float4* __restrict__ scale)
{
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= n_layers) return;
    const uint32_t mu = mx[L], mv = mx[n_layers + L];
    const float suv = f16s_scale_of(mu > mv ? mu : mv);
    float4 sc;
    sc.x = suv;
    sc.y = suv;
    sc.z = is3d ? f16s_scale_of(mx[2 * n_layers + L]) : 0.0f;
    sc.w = is3d ? f16s_scale_of(mx[2 * n_layers + L + 1]) : 0.0f;
    scale[L] = sc;
}

__device__ __forceinline__ uint16_t to_f16s(float x, float inv_scale)
{
    return __builtin_bit_cast(uint16_t, (_Float16)(x * inv_scale));
}

__global__ void pack3d_f16s_kernel(int64_t N, int64_t n_layers, const int32_t* __restrict__ old_of_new,
                                   const float* __restrict__ u, const float* __restrict__ v,
                                   const float* __restrict__ w, const float4* __restrict__ scale,
                                   F16sx4* __restrict__ rec)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * n_layers) return;
    int64_t L = i / N, k = i - L * N;
    int64_t o = old_of_new[k];
    const float4 sc = scale[L];
    F16sx4 r;
    r.x = to_f16s(u[L * N + o], 1.0f / sc.x);
    r.y = to_f16s(v[L * N + o], 1.0f / sc.y);
    r.z = to_f16s(w[L * N + o], 1.0f / sc.z);
    r.w = to_f16s(w[(L + 1) * N + o], 1.0f / sc.w);
    rec[i] = r;
}

__global__ void pack2d_f16s_kernel(int64_t N, const int32_t* __restrict__ old_of_new, const float* __restrict__ u,
                                   const float* __restrict__ v, const float4* __restrict__ scale,
                                   F16sx2* __restrict__ rec)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    int64_t o = old_of_new[k];
    const float4 sc = scale[0];
    F16sx2 r;
    r.x = to_f16s(u[o], 1.0f / sc.x);
    r.y = to_f16s(v[o], 1.0f / sc.y);
    rec[k] = r;
}

// element-centre u/v planes (mesh element order) into (u, v) records per
// (layer, element) in internal element order
template <typename T>
__global__ void pack_elem_kernel(int64_t E, int64_t n_layers, const int32_t* __restrict__ old_of_new,
                                 const T* __restrict__ u, const T* __restrict__ v,
                                 typename RecOf<T, false>::type* __restrict__ rec)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E * n_layers) return;
    int64_t L = i / E, k = i - L * E;
    int64_t o = old_of_new[k];
    typename RecOf<T, false>::type r;
    r.x = u[L * E + o];
    r.y = v[L * E + o];
    rec[i] = r;
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
    case 13: r = km_pow(a, y[i]); break;
    case 11:
    case 12: {
        // the random walk's normal pair (numerics/diffusion.py:234) of key
        // (seed 20240501, real 0, pid x, step y, comp 0)
        double g1, g2;
        normal_pair(20240501ull, 0ull, (uint64_t)a, (uint64_t)y[i], 0ull, g1, g2);
        r = fn == 11 ? g1 : g2;
        break;
    }
    default: r = km_sqrt(a); break;
    }
    out[i] = r;
}

// core/mesh.py:357 _gh_rect_sum over the prefix sum cum ((ny+1) x (nx+1))
__device__ inline int64_t gh_rect_sum(const int32_t* __restrict__ cum, int nx, int ny, int wrap, int ci, int cj,
                                      int rx, int ry)
{
    auto C = [&](int j, int i) { return (int64_t)cum[(int64_t)j * (nx + 1) + i]; };
    int j0 = cj - ry;
    if (j0 < 0) j0 = 0;
    int j1 = cj + ry + 1;
    if (j1 > ny) j1 = ny;
    if (wrap && 2 * rx + 1 >= nx) return C(j1, nx) - C(j0, nx);
    int i0 = ci - rx, i1 = ci + rx + 1;
    if (wrap) {
        if (i0 < 0) return (C(j1, i1) - C(j0, i1)) + (C(j1, nx) - C(j0, nx)) - (C(j1, i0 + nx) - C(j0, i0 + nx));
        if (i1 > nx) return (C(j1, nx) - C(j0, nx)) - (C(j1, i0) - C(j0, i0)) + (C(j1, i1 - nx) - C(j0, i1 - nx));
    } else {
        if (i0 < 0) i0 = 0;
        if (i1 > nx) i1 = nx;
    }
    return (C(j1, i1) - C(j0, i1)) - (C(j1, i0) - C(j0, i0));
}

// core/mesh.py:386 _gh_sample_stats for the candidate grid held in g (the
// uploaded mesh with the candidate's cell, origin and dimensions); nodes in
// internal numbering
__global__ void gh_sample_stats_kernel(DMesh g, int64_t n, const double* __restrict__ lon,
                                       const double* __restrict__ lat, const int32_t* __restrict__ nn_mesh,
                                       const int32_t* __restrict__ new_of_old, const int32_t* __restrict__ cum,
                                       int64_t* __restrict__ occ_out, double* __restrict__ cost_out,
                                       int8_t* __restrict__ unprov_out, int8_t* __restrict__ approx_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nx = g.nx, ny = g.ny, wrap = g.wrap;
    const double cell = g.cell;
    const double lp = lon[i], tp = lat[i];
    int ci, cj, cin, cjn;
    gh_cell_of(g, lp, tp, ci, cj);
    const int64_t occ = gh_rect_sum(cum, nx, ny, wrap, ci, cj, 1, 1);
    occ_out[i] = occ;
    const int node = new_of_old[nn_mesh[i]];
    const double2 ll = g.lonlat[node];
    gh_cell_of(g, ll.x, ll.y, cin, cjn);
    int di = cin - ci;
    if (di < 0) di = -di;
    if (wrap && nx - di < di) di = nx - di;
    int dj = cjn - cj;
    if (dj < 0) dj = -dj;
    const int8_t approx = (di > 1 || dj > 1) ? 1 : 0;
    approx_out[i] = approx;
    const Query q = make_query(lp, tp);
    double theta, ext;
    cap_extents(g, q, lp, tp, node, theta, ext);
    bool provable = occ > 0 && approx == 0;
    if (provable) provable = anchor_provable_rect(g, lp, tp, theta, ext, 1, 1) == 0;
    if (provable) {
        cost_out[i] = (double)occ + 9.0 * 1.0 + 20.0;
        unprov_out[i] = 0;
        return;
    }
    int ry = (int)ceil(theta / cell);
    if (ry < 1) ry = 1;
    if (ry > ny) ry = ny;
    int rx;
    if (ext >= 1.0e29) {
        rx = nx;
    } else {
        rx = (int)ceil(ext / cell);
        if (rx < 1) rx = 1;
    }
    int cells, iters;
    if (rx * 2 + 1 >= nx) {
        rx = nx;
        cells = 9 + 2 * ry;
        iters = ry;
    } else {
        cells = 9 + rx * (2 * ry + 1) + 2 * ry;
        iters = rx > ry ? rx : ry;
    }
    cost_out[i] = ((double)gh_rect_sum(cum, nx, ny, wrap, ci, cj, rx, ry) + cells * 1.0) + (iters + 1) * 20.0;
    unprov_out[i] = 1;
}

template <typename T, bool IS3D>
static Snap<T, IS3D> snap_of(const void* a, const void* b, const void* t)
{
    using R = typename RecOf<T, IS3D>::type;
    Snap<T, IS3D> f;
    f.a = static_cast<const R*>(a);
    f.b = b ? static_cast<const R*>(b) : f.a;
    f.t = t ? static_cast<const R*>(t) : f.a;
    return f;
}

// A z-level step, the context's grid-hash tile setting in the mesh view
template <bool IS3D, typename T, bool ELEM>
static void launch_step_snaps(kelp_ctx* c, int order, const Snap<T, IS3D>& f, const Snap<T, false>& fe,
                              const StepCfg& cfg, const PState& ps)
{
    DMesh mt = c->m;
    mt.tile_on = c->gh_tile ? 1 : 0;
    if (c->m.max_v == 3)
        launch_step_v<3, IS3D, T, ELEM>(c->stream, order, mt, f, fe, cfg, ps);
    else
        launch_step_v<4, IS3D, T, ELEM>(c->stream, order, mt, f, fe, cfg, ps);
}

template <bool IS3D, typename T, bool ELEM>
static void launch_step_t(kelp_ctx* c, int order, const FieldSlot& a, const FieldSlot* b, const FieldSlot* t,
                          const StepCfg& cfg, const PState& ps)
{
    Snap<T, IS3D> f = snap_of<T, IS3D>(a.rec, b ? b->rec : nullptr, t ? t->rec : nullptr);
    Snap<T, false> fe = snap_of<T, false>(a.erec, b ? b->erec : nullptr, t ? t->erec : nullptr);
    launch_step_snaps<IS3D, T, ELEM>(c, order, f, fe, cfg, ps);
}

// float16 records: the Snap also carries each slot's row scales
template <bool IS3D>
static void launch_step_f16s(kelp_ctx* c, int order, const FieldSlot& a, const FieldSlot* b, const FieldSlot* t,
                             const StepCfg& cfg, const PState& ps)
{
    using R = typename RecOf<F16s, IS3D>::type;
    Snap<F16s, IS3D> f;
    f.a = static_cast<const R*>(a.rec);
    f.b = b ? static_cast<const R*>(b->rec) : f.a;
    f.t = t ? static_cast<const R*>(t->rec) : f.a;
    f.sa = a.scale;
    f.sb = b ? b->scale : a.scale;
    f.st = t ? t->scale : a.scale;
    f.row_len = (uint32_t)c->m.n_nodes;
    const Snap<F16s, false> fe{};
    launch_step_snaps<IS3D, F16s, false>(c, order, f, fe, cfg, ps);
}

template <typename T, bool ELEM>
static void launch_sigma_t(kelp_ctx* c, int order, const FieldSlot& a, const FieldSlot* b, const FieldSlot* t,
                           const StepCfg& cfg, const PState& ps)
{
    Snap<T, true> f = snap_of<T, true>(a.rec, b ? b->rec : nullptr, t ? t->rec : nullptr);
    Snap<T, false> fe = snap_of<T, false>(a.erec, b ? b->erec : nullptr, t ? t->erec : nullptr);
    if (c->m.max_v == 3)
        launch_sigma_step<3, T, ELEM>(c->stream, order, c->m, f, fe, cfg, ps);
    else
        launch_sigma_step<4, T, ELEM>(c->stream, order, c->m, f, fe, cfg, ps);
}

template <typename T>
static void launch_step(kelp_ctx* c, int order, bool is3d, const FieldSlot& a, const FieldSlot* b,
                        const FieldSlot* t, const StepCfg& cfg, const PState& ps)
{
    const bool elem = a.layout == 1;
    if (c->vertical_coord == 1) {
        if (elem)
            launch_sigma_t<T, true>(c, order, a, b, t, cfg, ps);
        else
            launch_sigma_t<T, false>(c, order, a, b, t, cfg, ps);
        return;
    }
    if (is3d && elem)
        launch_step_t<true, T, true>(c, order, a, b, t, cfg, ps);
    else if (is3d)
        launch_step_t<true, T, false>(c, order, a, b, t, cfg, ps);
    else if (elem)
        launch_step_t<false, T, true>(c, order, a, b, t, cfg, ps);
    else
        launch_step_t<false, T, false>(c, order, a, b, t, cfg, ps);
}

static const FieldSlot* opt_slot(kelp_ctx* c, int k)
{
    if (k < 0) return nullptr;
    return &c->slots[k];
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// numerics/temporal_reconstruct.py:168 _fc_slope_scalar
__device__ __forceinline__ double fc_slope(double d_prev, double d_next, double h_prev, double h_next)
{
    if (d_prev * d_next > 0.0) {
        const double w1 = 2.0 * h_next + h_prev;
        const double w2 = h_next + 2.0 * h_prev;
        return (w1 + w2) / (w1 / d_prev + w2 / d_next);
    }
    return 0.0;
}

// temporal_reconstruct.py:184 _pchip_value_scalar: the monotone cubic through
// four snapshot values at t0 < t1 < t2 < t3, evaluated at tq in [t1, t2]
__device__ __forceinline__ double pchip4(double f0, double f1, double f2, double f3, double t0, double t1, double t2,
                                         double t3, double tq)
{
    const double h0 = t1 - t0, h1 = t2 - t1, h2 = t3 - t2;
    const double d0 = (f1 - f0) / h0;
    const double d1 = (f2 - f1) / h1;
    const double d2 = (f3 - f2) / h2;
    const double m1 = fc_slope(d0, d1, h0, h1);
    const double m2 = fc_slope(d1, d2, h1, h2);
    const double s = (tq - t1) / h1;
    const double s2 = s * s;
    const double s3 = s2 * s;
    const double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
    const double h10 = s3 - 2.0 * s2 + s;
    const double h01 = -2.0 * s3 + 3.0 * s2;
    const double h11 = s3 - s2;
    return h00 * f1 + h10 * h1 * m1 + h01 * f2 + h11 * h1 * m2;
}

struct PchipTimes {
    double t0, t1, t2, t3, tq;
};

// every component of n records of width W (4: node 3-D, 2: node 2-D or
// element) of four same-shaped snapshots into f64 records
template <typename T, int W>
__global__ void pchip_records_kernel(int64_t n, const T* __restrict__ a, const T* __restrict__ b,
                                     const T* __restrict__ c, const T* __restrict__ d, PchipTimes t,
                                     double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * W) return;
    out[i] = pchip4((double)a[i], (double)b[i], (double)c[i], (double)d[i], t.t0, t.t1, t.t2, t.t3, t.tq);
}

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
    // plain steps stage their workgroups' grid-hash tiles in LDS unless
    // KELP_GH_TILE=0 (the scans then all read HBM; same results)
    const char* tile_env = getenv("KELP_GH_TILE");
    c->gh_tile = !(tile_env && tile_env[0] == '0');
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        delete c;
        return fail(KELP_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    e = hipStreamCreate(&c->copy_stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->compute_mark, hipEventDisableTiming);
    if (e != hipSuccess) {
        kelp_ctx_destroy(c);
        return fail(KELP_EHIP, "copy stream: %s", hipGetErrorString(e));
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
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    free_mesh(c);
    for (auto& s : c->slots) free_slot(s);
    for (DevBuf* b : {&c->sort_tmp, &c->sort_keys, &c->sort_keys2, &c->sort_idx, &c->sort_perm, &c->sort_stage,
                      &c->batch_clouds, &c->batch_steps, &c->batch_counter})
        kelp_release(*b);
    if (c->batch_graph) (void)hipGraphExecDestroy(c->batch_graph);
    for (int k = 0; k < kMaxEvents; ++k)
        if (c->events[k]) (void)hipEventDestroy(c->events[k]);
    if (c->compute_mark) (void)hipEventDestroy(c->compute_mark);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
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
    if (d->vertical_coord != 0 && d->vertical_coord != 1)
        return fail(KELP_EINVAL, "kelp_mesh_upload: vertical_coord must be 0 (z-level) or 1 (sigma)");
    for (int64_t k = 1; k < d->n_levels; ++k) {
        if (d->vertical_coord == 0 && !(d->host_z_levels[k] >= d->host_z_levels[k - 1]))
            return fail(KELP_EINVAL, "kelp_mesh_upload: z levels must be non-decreasing");
        if (d->vertical_coord == 1 && !(d->host_z_levels[k] <= d->host_z_levels[k - 1]))
            return fail(KELP_EINVAL, "kelp_mesh_upload: sigma levels must be non-increasing (1 at the surface)");
    }
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
    const bool has_centres = d->host_elem_center_lon && d->host_elem_center_lat;
    std::vector<double2> ecent(has_centres ? E : 0);
    std::vector<int4> enbr(d->host_elem_neighbors ? E : 0);
    for (int64_t en = 0; en < E; ++en) {
        int64_t eo = e_old_of_new[en];
        int v[4] = {-1, -1, -1, -1};
        for (int k = 0; k < d->max_v; ++k) {
            int x = d->host_elem_nodes[eo * d->max_v + k];
            v[k] = x >= 0 ? new_of_old[x] : -1;
        }
        elems[en] = make_int4(v[0], v[1], v[2], v[3]);
        if (d->host_elem_neighbors) {
            int b[4] = {-1, -1, -1, -1};
            for (int k = 0; k < d->max_v; ++k) {
                int x = d->host_elem_neighbors[eo * d->max_v + k];
                if (x < -1 || x >= E) return fail(KELP_EINVAL, "kelp_mesh_upload: element neighbour out of range");
                b[k] = x >= 0 ? e_new_of_old[x] : -1;
            }
            enbr[en] = make_int4(b[0], b[1], b[2], b[3]);
        }
        if (d->host_elem_coastal) coastal[en] = d->host_elem_coastal[eo] ? 1 : 0;
        if (has_centres) ecent[en] = make_double2(d->host_elem_center_lon[eo], d->host_elem_center_lat[eo]);
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
    double *d_depth, *d_zlev, *d_coslat;
    uint8_t* d_coastal = nullptr;
    double2* d_ecent = nullptr;
    double4* d_eframe = nullptr;
    uint8_t* d_eflags = nullptr;
    double* d_sig_eta = nullptr;
    int4* d_enbr = nullptr;
    int rc;
    if ((rc = dev_alloc(c, &d_lonlat, N)) || (rc = dev_alloc(c, &d_xyz, N)) ||
        (rc = dev_alloc(c, &d_off, n_cells + 1)) || (rc = dev_alloc(c, &d_elems, E)) ||
        (rc = dev_alloc(c, &d_n2e_off, N + 1)) || (rc = dev_alloc(c, &d_n2e, n2e.size())) ||
        (rc = dev_alloc(c, &d_depth, N)) || (rc = dev_alloc(c, &d_coslat, N)) || (rc = dev_alloc(c, &d_zlev, d->n_levels > 0 ? d->n_levels : 1)) ||
        (rc = dev_alloc(c, &d_nodeo, N)) || (rc = dev_alloc(c, &d_noden, N)) ||
        (rc = dev_alloc(c, &d_elemo, E)) || (rc = dev_alloc(c, &d_elemn, E)) ||
        (d->host_elem_coastal && (rc = dev_alloc(c, &d_coastal, E))) ||
        (has_centres && (rc = dev_alloc(c, &d_ecent, E))) || (rc = dev_alloc(c, &d_eframe, 4 * E)) ||
        (rc = dev_alloc(c, &d_eflags, E)) ||
        (d->vertical_coord == 1 && (rc = dev_alloc(c, &d_sig_eta, 7 * N))) ||
        (d->host_elem_neighbors && (rc = dev_alloc(c, &d_enbr, E)))) {
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
    if (d_ecent) HIP_TRY(hipMemcpy(d_ecent, ecent.data(), E * sizeof(double2), hipMemcpyHostToDevice));
    if (d_enbr) HIP_TRY(hipMemcpy(d_enbr, enbr.data(), E * sizeof(int4), hipMemcpyHostToDevice));
    node_xyz_kernel<<<grid_for(N, 256), 256, 0, c->stream>>>(N, d_lonlat, d_n2e_off, d_xyz, d_coslat);
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
    m.ecent = d_ecent;
    m.coslat = d_coslat;
    m.eframe = d_eframe;
    m.eflags = d_eflags;
    m.enbr = d_enbr;
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
    if (d->max_v == 3)
        elem_geom_kernel<3><<<grid_for(E, 256), 256, 0, c->stream>>>(m, d_eframe, d_eflags);
    else
        elem_geom_kernel<4><<<grid_for(E, 256), 256, 0, c->stream>>>(m, d_eframe, d_eflags);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->m = m;
    c->node_old_of_new = d_nodeo;
    c->node_new_of_old = d_noden;
    c->elem_old_of_new = d_elemo;
    c->elem_new_of_old = d_elemn;
    c->vertical_coord = (int)d->vertical_coord;
    c->sig_eta = d_sig_eta;
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

int kelp_elems_to_mesh_ids(kelp_ctx* c, int64_t n, int32_t* ids)
{
    if (!c || n < 0 || (n > 0 && !ids)) return fail(KELP_EINVAL, "kelp_elems_to_mesh_ids: bad argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_elems_to_mesh_ids: no mesh uploaded");
    if (n == 0) return KELP_OK;
    map_ids_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(n, c->elem_old_of_new, ids);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

int kelp_elems_from_mesh_ids(kelp_ctx* c, int64_t n, int32_t* ids)
{
    if (!c || n < 0 || (n > 0 && !ids)) return fail(KELP_EINVAL, "kelp_elems_from_mesh_ids: bad argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_elems_from_mesh_ids: no mesh uploaded");
    if (n == 0) return KELP_OK;
    map_ids_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(n, c->elem_new_of_old, ids);
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

// core/release.py:112 activate_released: NOT_RELEASED -> ALIVE once the
// global step reaches the particle's release step (indexed by pid -
// pid_offset, so the particle order on the device does not matter); the
// status at the step start is copied to prev_status for the record masks.
__global__ void activate_kernel(int64_t n, int32_t* __restrict__ status, int32_t* __restrict__ nn,
                                const int64_t* __restrict__ pid, const int64_t* __restrict__ release_step,
                                int64_t pid_offset, int64_t n_release, int64_t step,
                                int32_t* __restrict__ prev_status)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t s = status[i];
    if (release_step && s == 4) {
        const int64_t k = pid[i] - pid_offset;
        if (k >= 0 && k < n_release && step >= release_step[k]) {
            s = kAlive;
            status[i] = s;
            if (nn) nn[i] = -1;  // anchored afresh by its first step
        }
    }
    if (prev_status) prev_status[i] = s;
}

int kelp_activate(kelp_ctx* c, const kelp_particles* p, const int64_t* release_step, int64_t pid_offset,
                  int64_t n_release, int64_t step, int32_t* prev_status)
{
    if (!c || !p) return fail(KELP_EINVAL, "kelp_activate: NULL argument");
    if (p->n < 0 || !p->status) return fail(KELP_EINVAL, "kelp_activate: particles need status");
    if (release_step && (!p->pid || n_release < 0))
        return fail(KELP_EINVAL, "kelp_activate: release steps need pid and n_release >= 0");
    if (p->n == 0 || (!release_step && !prev_status)) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    activate_kernel<<<grid_for(p->n, 256), 256, 0, c->stream>>>(p->n, p->status, p->nn, p->pid, release_step,
                                                               pid_offset, n_release, step, prev_status);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

// One staged record: particle slot i to row pid - pid_offset, masked as the
// host record (core/release.py:132, :170; the backward flip after masking,
// so a masked velocity of a backward run is -NaN there too).
__global__ void record_stage_kernel(int64_t n, const double* __restrict__ lon, const double* __restrict__ lat,
                                    const double* __restrict__ z, const double* __restrict__ u,
                                    const double* __restrict__ v, const double* __restrict__ w,
                                    const int32_t* __restrict__ status, const int8_t* __restrict__ fr,
                                    const int64_t* __restrict__ pid, const int32_t* __restrict__ prev,
                                    int64_t pid_offset, int64_t n_rows, double time_dir, int n_vars,
                                    const double* __restrict__ sampled, kelp_record_rows out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = pid[i] - pid_offset;
    if (r < 0 || r >= n_rows) return;
    const int32_t st = status[i];
    const bool dormant = st == 4;
    const bool bad = st != 1 || prev[i] != 1;
    const double nan = __builtin_nan("");
    out.lon[r] = dormant ? nan : lon[i];
    out.lat[r] = dormant ? nan : lat[i];
    out.z[r] = dormant ? nan : (z ? z[i] : 0.0);
    double uu = (bad || dormant) ? nan : u[i];
    double vv = (bad || dormant) ? nan : v[i];
    double ww = (bad || dormant) ? nan : (w ? w[i] : 0.0);
    if (time_dir < 0.0) {
        uu = -uu;
        vv = -vv;
        ww = -ww;
    }
    out.u[r] = uu;
    out.v[r] = vv;
    if (out.w) out.w[r] = ww;
    out.status[r] = st;
    out.fr[r] = fr[i];
    for (int k = 0; k < n_vars; ++k) out.samples[k * n_rows + r] = sampled[k * n + i];
}

int kelp_record_stage(kelp_ctx* c, const kelp_particles* p, const int32_t* prev_status, int64_t pid_offset,
                      int64_t n_rows, double time_dir, int n_vars, const double* sampled,
                      const kelp_record_rows* out)
{
    if (!c || !p || !out) return fail(KELP_EINVAL, "kelp_record_stage: NULL argument");
    if (p->n < 0 || n_rows < 0 || n_vars < 0) return fail(KELP_EINVAL, "kelp_record_stage: negative size");
    if (!p->lon || !p->lat || !p->status || !p->fr || !p->pid || !p->u || !p->v || !prev_status)
        return fail(KELP_EINVAL, "kelp_record_stage: particles need lon, lat, status, fr, pid, u, v and prev_status");
    if (!out->lon || !out->lat || !out->z || !out->u || !out->v || !out->status || !out->fr)
        return fail(KELP_EINVAL, "kelp_record_stage: missing output row");
    if (out->w && !p->w) return fail(KELP_EINVAL, "kelp_record_stage: a w row needs particles with w");
    if (n_vars > 0 && (!sampled || !out->samples))
        return fail(KELP_EINVAL, "kelp_record_stage: sampled scalars need `sampled` and out->samples");
    if (p->n == 0 || n_rows == 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    record_stage_kernel<<<grid_for(p->n, 256), 256, 0, c->stream>>>(
        p->n, p->lon, p->lat, p->z, p->u, p->v, p->w, p->status, p->fr, p->pid, prev_status, pid_offset, n_rows,
        time_dir, n_vars, sampled, *out);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

// Pack one snapshot's planes (device, mesh order) into the slot's records on// Pack one snapshot's planes (device, mesh order) into the slot's records on
// `stream` (pack3d/pack2d for node records, pack_elem for element u/v).
static hipError_t launch_pack(kelp_ctx* c, hipStream_t stream, FieldSlot& s, int dtype, int layout, bool is3d,
                              int64_t n_planes, const void* tu, const void* tv, const void* tw)
{
    const int64_t N = c->m.n_nodes, E = c->m.n_elems;
    const bool f32 = dtype == KELP_DT_F32;
    // node records: u/v (node layout) and the W pair (3-D)
    const void* nu = layout == 0 ? tu : nullptr;
    const void* nv = layout == 0 ? tv : nullptr;
    if (dtype == KELP_DT_BF16) {  // node layout only (checked by the callers)
        if (is3d)
            pack3d_bf16_kernel<<<grid_for(N * n_planes, 256), 256, 0, stream>>>(
                N, n_planes, c->node_old_of_new, (const float*)tu, (const float*)tv, (const float*)tw,
                (Bf16x4*)s.rec);
        else
            pack2d_bf16_kernel<<<grid_for(N, 256), 256, 0, stream>>>(N, c->node_old_of_new, (const float*)tu,
                                                                     (const float*)tv, (Bf16x2*)s.rec);
        return hipGetLastError();
    }
    if (dtype == KELP_DT_F16S) {  // node layout only (checked by the callers)
        const int64_t n_rows = 2 * n_planes + (is3d ? n_planes + 1 : 0);
        float4* scale = s.scale;
        uint32_t* mx = reinterpret_cast<uint32_t*>(s.scale + n_planes);
        f16s_clear_kernel<<<grid_for(n_rows, 256), 256, 0, stream>>>(n_rows, mx);
        const unsigned gx = (unsigned)std::min<int64_t>((N + 255) / 256, 64);
        f16s_max_kernel<<<dim3(gx, (unsigned)n_rows), 256, 0, stream>>>(N, n_planes, (const float*)tu,
                                                                         (const float*)tv, (const float*)tw, mx);
        f16s_scale_kernel<<<grid_for(n_planes, 64), 64, 0, stream>>>(n_planes, is3d ? 1 : 0, mx, scale);
        if (is3d)
            pack3d_f16s_kernel<<<grid_for(N * n_planes, 256), 256, 0, stream>>>(
                N, n_planes, c->node_old_of_new, (const float*)tu, (const float*)tv, (const float*)tw, scale,
                (F16sx4*)s.rec);
        else
            pack2d_f16s_kernel<<<grid_for(N, 256), 256, 0, stream>>>(N, c->node_old_of_new, (const float*)tu,
                                                                     (const float*)tv, scale, (F16sx2*)s.rec);
        return hipGetLastError();
    }
    if (is3d) {
        const unsigned g = grid_for(N * n_planes, 256);
        if (f32)
            pack3d_kernel<float><<<g, 256, 0, stream>>>(N, n_planes, c->node_old_of_new, (const float*)nu,
                                                        (const float*)nv, (const float*)tw, (float4*)s.rec);
        else
            pack3d_kernel<double><<<g, 256, 0, stream>>>(N, n_planes, c->node_old_of_new, (const double*)nu,
                                                         (const double*)nv, (const double*)tw, (double4*)s.rec);
    } else if (layout == 0) {
        const unsigned g = grid_for(N, 256);
        if (f32)
            pack2d_kernel<float><<<g, 256, 0, stream>>>(N, c->node_old_of_new, (const float*)tu, (const float*)tv,
                                                        (float2*)s.rec);
        else
            pack2d_kernel<double><<<g, 256, 0, stream>>>(N, c->node_old_of_new, (const double*)tu,
                                                         (const double*)tv, (double2*)s.rec);
    }
    if (layout == 1) {
        const unsigned g = grid_for(E * n_planes, 256);
        if (f32)
            pack_elem_kernel<float><<<g, 256, 0, stream>>>(E, n_planes, c->elem_old_of_new, (const float*)tu,
                                                           (const float*)tv, (float2*)s.erec);
        else
            pack_elem_kernel<double><<<g, 256, 0, stream>>>(E, n_planes, c->elem_old_of_new, (const double*)tu,
                                                            (const double*)tv, (double2*)s.erec);
    }
    return hipGetLastError();
}

// Upload one snapshot: u/v planes at the nodes (layout 0) or at the element
// centres (layout 1), W planes at the nodes; packed on the device.
static int field_upload(kelp_ctx* c, const char* fn, int slot, int dtype, int layout, int64_t n_planes,
                        const void* hu, const void* hv, const void* hw, int64_t n_w_planes)
{
    if (!c) return fail(KELP_EINVAL, "%s: NULL context", fn);
    if (!c->has_mesh) return fail(KELP_ESTATE, "%s: upload the mesh first", fn);
    if (slot < 0 || slot >= kMaxSlots) return fail(KELP_EINVAL, "%s: slot %d not in [0, %d)", fn, slot, kMaxSlots);
    if (dtype != KELP_DT_F64 && dtype != KELP_DT_F32 && dtype != KELP_DT_BF16 && dtype != KELP_DT_F16S)
        return fail(KELP_EINVAL, "%s: bad dtype %d", fn, dtype);
    if ((dtype == KELP_DT_BF16 || dtype == KELP_DT_F16S) && (layout != 0 || c->vertical_coord == 1))
        return fail(KELP_EINVAL, "%s: bfloat16 / float16 storage is for node-layout z-level snapshots", fn);
    if (dtype == KELP_DT_F16S && (uint64_t)n_planes * (uint64_t)c->m.n_nodes >= (uint64_t(1) << 32))
        return fail(KELP_EINVAL, "%s: float16 storage holds at most 2^32 records per snapshot", fn);
    if (n_planes < 1 || !hu || !hv) return fail(KELP_EINVAL, "%s: u/v missing", fn);
    const bool is3d = hw != nullptr;
    if (is3d && n_w_planes != n_planes + 1) return fail(KELP_EINVAL, "%s: W needs n_planes + 1 interface planes", fn);
    if (!is3d && n_planes != 1) return fail(KELP_EINVAL, "%s: layered u/v need W (3-D)", fn);
    if (layout == 1 && !c->m.ecent)
        return fail(KELP_ESTATE, "%s: element-centre u/v need the mesh's element centres", fn);
    HIP_TRY(hipSetDevice(c->device));
    FieldSlot& s = c->slots[slot];
    if (s.pending) HIP_TRY(hipStreamSynchronize(c->copy_stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    free_slot(s);
    const int64_t N = c->m.n_nodes, E = c->m.n_elems;
    const int64_t M = layout == 1 ? E : N;  // u/v values per plane
    const size_t es = dtype == KELP_DT_F64 ? 8 : 4;                     // host values
    const size_t vs = dtype == KELP_DT_F64 ? 8 : dtype == KELP_DT_F32 ? 4 : 2;  // stored values
    const size_t uv_bytes = (size_t)n_planes * M * es;
    const size_t w_bytes = is3d ? (size_t)n_w_planes * N * es : 0;
    void *tu = nullptr, *tv = nullptr, *tw = nullptr;
    hipError_t e = hipSuccess;
    if (is3d || layout == 0) e = hipMalloc(&s.rec, (size_t)n_planes * N * vs * (is3d ? 4 : 2));
    if (e == hipSuccess && dtype == KELP_DT_F16S)  // row scales, then the row maxima
        e = hipMalloc(&s.scale, (size_t)n_planes * sizeof(float4) + (size_t)(3 * n_planes + 1) * sizeof(uint32_t));
    if (e == hipSuccess && layout == 1) e = hipMalloc(&s.erec, (size_t)n_planes * E * es * 2);
    if (e == hipSuccess) e = hipMalloc(&tu, uv_bytes);
    if (e == hipSuccess) e = hipMalloc(&tv, uv_bytes);
    if (e == hipSuccess && is3d) e = hipMalloc(&tw, w_bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(tu, hu, uv_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tv, hv, uv_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && is3d) e = hipMemcpyAsync(tw, hw, w_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_pack(c, c->stream, s, dtype, layout, is3d, n_planes, tu, tv, tw);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tu) (void)hipFree(tu);
    if (tv) (void)hipFree(tv);
    if (tw) (void)hipFree(tw);
    if (e != hipSuccess) {
        free_slot(s);
        return fail(e == hipErrorOutOfMemory ? KELP_ENOMEM : KELP_EHIP, "%s: %s", fn, hipGetErrorString(e));
    }
    s.dtype = dtype;
    s.layout = layout;
    s.n_planes = n_planes;
    s.n_w_planes = is3d ? n_w_planes : 0;
    s.is3d = is3d;
    return KELP_OK;
}

int kelp_field_upload(kelp_ctx* c, int slot, int dtype, int64_t n_planes, const void* hu, const void* hv,
                      const void* hw, int64_t n_w_planes)
{
    return field_upload(c, "kelp_field_upload", slot, dtype, 0, n_planes, hu, hv, hw, n_w_planes);
}

int kelp_field_upload_elem(kelp_ctx* c, int slot, int dtype, int64_t n_planes, const void* hu, const void* hv,
                           const void* hw, int64_t n_w_planes)
{
    return field_upload(c, "kelp_field_upload_elem", slot, dtype, 1, n_planes, hu, hv, hw, n_w_planes);
}

// eta of a snapshot (mesh node order) into internal order
__global__ void permute_nodes_kernel(int64_t N, const int32_t* __restrict__ old_of_new, const double* __restrict__ in,
                                     double* __restrict__ out)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    out[k] = in[old_of_new[k]];
}

int kelp_field_upload_eta(kelp_ctx* c, int slot, const double* host_eta)
{
    if (!c || !host_eta) return fail(KELP_EINVAL, "kelp_field_upload_eta: NULL argument");
    if (slot < 0 || slot >= kMaxSlots) return fail(KELP_EINVAL, "kelp_field_upload_eta: slot %d out of range", slot);
    FieldSlot& s = c->slots[slot];
    if (s.dtype < 0) return fail(KELP_ESTATE, "kelp_field_upload_eta: slot %d holds no snapshot", slot);
    HIP_TRY(hipSetDevice(c->device));
    if (s.pending) HIP_TRY(hipStreamSynchronize(c->copy_stream));
    const int64_t N = c->m.n_nodes;
    if (!s.eta) HIP_TRY(hipMalloc(&s.eta, N * sizeof(double)));
    double* tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, N * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, host_eta, N * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        permute_nodes_kernel<<<grid_for(N, 256), 256, 0, c->stream>>>(N, c->node_old_of_new, tmp, s.eta);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(KELP_EHIP, "kelp_field_upload_eta: %s", hipGetErrorString(e));
    return KELP_OK;
}

int kelp_sigma_surface_upload(kelp_ctx* c, int which, const double* host_eta)
{
    if (!c) return fail(KELP_EINVAL, "kelp_sigma_surface_upload: NULL context");
    if (which != 0 && which != 1) return fail(KELP_EINVAL, "kelp_sigma_surface_upload: which must be 0 or 1");
    if (!c->has_mesh || c->vertical_coord != 1 || !c->sig_eta)
        return fail(KELP_ESTATE, "kelp_sigma_surface_upload: no sigma mesh uploaded");
    if (!host_eta) {
        c->sig_given[which] = false;
        return KELP_OK;
    }
    HIP_TRY(hipSetDevice(c->device));
    const int64_t N = c->m.n_nodes;
    double* tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, N * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, host_eta, N * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        permute_nodes_kernel<<<grid_for(N, 256), 256, 0, c->stream>>>(N, c->node_old_of_new, tmp,
                                                                     c->sig_eta + (5 + which) * N);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(KELP_EHIP, "kelp_sigma_surface_upload: %s", hipGetErrorString(e));
    c->sig_given[which] = true;
    return KELP_OK;
}

int kelp_field_clear(kelp_ctx* c, int slot)
{
    if (!c || slot < 0 || slot >= kMaxSlots) return fail(KELP_EINVAL, "kelp_field_clear: bad argument");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->copy_stream);
    free_slot(c->slots[slot]);
    return KELP_OK;
}

// Device staging buffer k of a slot, at least `bytes` (grown with hipMalloc
// only when too small; the copy stream has finished with it by then).
static int slot_stage(kelp_ctx* c, FieldSlot& s, int k, size_t bytes, void** out)
{
    if (s.stage_bytes[k] < bytes) {
        HIP_TRY(hipStreamSynchronize(c->copy_stream));
        if (s.stage[k]) (void)hipFree(s.stage[k]);
        s.stage[k] = nullptr;
        s.stage_bytes[k] = 0;
        HIP_TRY(hipMalloc(&s.stage[k], bytes));
        s.stage_bytes[k] = bytes;
    }
    *out = s.stage[k];
    return KELP_OK;
}

int kelp_field_upload_async(kelp_ctx* c, int slot, int dtype, int layout, int64_t n_planes, const void* host_u,
                            const void* host_v, const void* host_w, int64_t n_w_planes, const double* host_eta)
{
    if (!c) return fail(KELP_EINVAL, "kelp_field_upload_async: NULL context");
    if (slot < 0 || slot >= kMaxSlots) return fail(KELP_EINVAL, "kelp_field_upload_async: slot %d out of range", slot);
    if (!host_u || !host_v) return fail(KELP_EINVAL, "kelp_field_upload_async: u/v missing");
    FieldSlot& s = c->slots[slot];
    const bool is3d = host_w != nullptr;
    if (s.dtype != dtype || s.layout != layout || s.n_planes != n_planes || s.is3d != is3d ||
        (is3d && s.n_w_planes != n_w_planes))
        return fail(KELP_ESTATE, "kelp_field_upload_async: slot %d must already hold a snapshot of this shape "
                                 "(fill it once with kelp_field_upload)", slot);
    if (host_eta && !s.eta)
        return fail(KELP_ESTATE, "kelp_field_upload_async: slot %d has no eta to refill (kelp_field_upload_eta)",
                    slot);
    HIP_TRY(hipSetDevice(c->device));
    const int64_t N = c->m.n_nodes, E = c->m.n_elems;
    const int64_t M = layout == 1 ? E : N;
    const size_t es = dtype == KELP_DT_F64 ? 8 : 4;
    const size_t uv_bytes = (size_t)n_planes * M * es;
    const size_t w_bytes = is3d ? (size_t)n_w_planes * N * es : 0;
    void *tu, *tv, *tw = nullptr, *te = nullptr;
    int rc;
    if ((rc = slot_stage(c, s, 0, uv_bytes, &tu)) || (rc = slot_stage(c, s, 1, uv_bytes, &tv)) ||
        (is3d && (rc = slot_stage(c, s, 2, w_bytes, &tw))) ||
        (host_eta && (rc = slot_stage(c, s, 3, (size_t)N * sizeof(double), &te))))
        return rc;
    if (!s.ready) HIP_TRY(hipEventCreateWithFlags(&s.ready, hipEventDisableTiming));
    // the refill overwrites the slot only after every step enqueued so far,
    // which may read it, has run
    HIP_TRY(hipEventRecord(c->compute_mark, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->compute_mark, 0));
    HIP_TRY(hipMemcpyAsync(tu, host_u, uv_bytes, hipMemcpyHostToDevice, c->copy_stream));
    HIP_TRY(hipMemcpyAsync(tv, host_v, uv_bytes, hipMemcpyHostToDevice, c->copy_stream));
    if (is3d) HIP_TRY(hipMemcpyAsync(tw, host_w, w_bytes, hipMemcpyHostToDevice, c->copy_stream));
    if (host_eta) HIP_TRY(hipMemcpyAsync(te, host_eta, N * sizeof(double), hipMemcpyHostToDevice, c->copy_stream));
    HIP_TRY(launch_pack(c, c->copy_stream, s, dtype, layout, is3d, n_planes, tu, tv, tw));
    if (host_eta) {
        permute_nodes_kernel<<<grid_for(N, 256), 256, 0, c->copy_stream>>>(N, c->node_old_of_new,
                                                                           (const double*)te, s.eta);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(s.ready, c->copy_stream));
    s.pending = true;
    s.az_k = 0;  // the new snapshot brings its own Az, if any
    return KELP_OK;
}

int kelp_field_upload_wait(kelp_ctx* c, int slot)
{
    if (!c || slot < 0 || slot >= kMaxSlots) return fail(KELP_EINVAL, "kelp_field_upload_wait: bad argument");
    FieldSlot& s = c->slots[slot];
    if (s.ready) HIP_TRY(hipEventSynchronize(s.ready));
    return KELP_OK;
}

int kelp_field_reconstruct(kelp_ctx* c, int dst, const int32_t* src, const double* times, double tq)
{
    if (!c || !src || !times) return fail(KELP_EINVAL, "kelp_field_reconstruct: NULL argument");
    if (dst < 0 || dst >= kMaxSlots) return fail(KELP_EINVAL, "kelp_field_reconstruct: bad destination slot");
    for (int k = 0; k < 4; ++k) {
        if (src[k] < 0 || src[k] >= kMaxSlots || c->slots[src[k]].dtype < 0)
            return fail(KELP_ESTATE, "kelp_field_reconstruct: source slot %d is empty", src[k]);
        if (src[k] == dst) return fail(KELP_EINVAL, "kelp_field_reconstruct: the destination is a source");
    }
    const FieldSlot& s0 = c->slots[src[0]];
    if (s0.dtype != KELP_DT_F32 && s0.dtype != KELP_DT_F64)
        return fail(KELP_EINVAL, "kelp_field_reconstruct: the snapshots must be stored as float32 or float64");
    for (int k = 1; k < 4; ++k) {
        const FieldSlot& sk = c->slots[src[k]];
        if (sk.dtype != s0.dtype || sk.layout != s0.layout || sk.n_planes != s0.n_planes ||
            sk.n_w_planes != s0.n_w_planes || sk.is3d != s0.is3d)
            return fail(KELP_EINVAL, "kelp_field_reconstruct: the four snapshots differ in dtype, layout or shape");
    }
    const double t0 = times[0], t1 = times[1], t2 = times[2], t3 = times[3];
    if (!(t0 < t1 && t1 < t2 && t2 < t3))
        return fail(KELP_EINVAL, "kelp_field_reconstruct: times must be strictly increasing");
    const double tol = 1e-9 * (t2 - t1);
    if (!(t1 - tol <= tq && tq <= t2 + tol))
        return fail(KELP_EINVAL, "kelp_field_reconstruct: query time outside the central interval");
    HIP_TRY(hipSetDevice(c->device));
    for (int k = 0; k < 4; ++k) {
        int rc = kelp_wait_slot(c, c->slots[src[k]]);
        if (rc) return rc;
    }
    FieldSlot& d = c->slots[dst];
    const int64_t N = c->m.n_nodes, E = c->m.n_elems;
    const int64_t n_node_rec = s0.is3d ? N * s0.n_planes : (s0.layout == 0 ? N : 0);
    const int wn = s0.is3d ? 4 : 2;
    const int64_t n_elem_rec = s0.layout == 1 ? E * s0.n_planes : 0;
    const bool fits = d.dtype == KELP_DT_F64 && d.layout == s0.layout && d.n_planes == s0.n_planes &&
                      d.n_w_planes == s0.n_w_planes && d.is3d == s0.is3d;
    if (!fits) {
        if (d.pending) HIP_TRY(hipStreamSynchronize(c->copy_stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        free_slot(d);
        hipError_t e = hipSuccess;
        if (n_node_rec) e = hipMalloc(&d.rec, (size_t)n_node_rec * wn * sizeof(double));
        if (e == hipSuccess && n_elem_rec) e = hipMalloc(&d.erec, (size_t)n_elem_rec * 2 * sizeof(double));
        if (e != hipSuccess) {
            free_slot(d);
            return fail(KELP_ENOMEM, "kelp_field_reconstruct: %s", hipGetErrorString(e));
        }
        d.dtype = KELP_DT_F64;
        d.layout = s0.layout;
        d.n_planes = s0.n_planes;
        d.n_w_planes = s0.n_w_planes;
        d.is3d = s0.is3d;
    }
    d.az_k = 0;
    const PchipTimes pt{t0, t1, t2, t3, tq};
    const FieldSlot *a = &c->slots[src[0]], *b = &c->slots[src[1]], *cc = &c->slots[src[2]],
                    *dd = &c->slots[src[3]];
    auto launch = [&](int64_t n, int w, const void* pa, const void* pb, const void* pc, const void* pd,
                      void* out) {
        const unsigned g = grid_for(n * w, 256);
        if (s0.dtype == KELP_DT_F32) {
            if (w == 4)
                pchip_records_kernel<float, 4><<<g, 256, 0, c->stream>>>(n, (const float*)pa, (const float*)pb,
                                                                         (const float*)pc, (const float*)pd, pt,
                                                                         (double*)out);
            else
                pchip_records_kernel<float, 2><<<g, 256, 0, c->stream>>>(n, (const float*)pa, (const float*)pb,
                                                                         (const float*)pc, (const float*)pd, pt,
                                                                         (double*)out);
        } else {
            if (w == 4)
                pchip_records_kernel<double, 4><<<g, 256, 0, c->stream>>>(n, (const double*)pa, (const double*)pb,
                                                                          (const double*)pc, (const double*)pd, pt,
                                                                          (double*)out);
            else
                pchip_records_kernel<double, 2><<<g, 256, 0, c->stream>>>(n, (const double*)pa, (const double*)pb,
                                                                          (const double*)pc, (const double*)pd, pt,
                                                                          (double*)out);
        }
    };
    if (n_node_rec) launch(n_node_rec, wn, a->rec, b->rec, cc->rec, dd->rec, d.rec);
    if (n_elem_rec) launch(n_elem_rec, 2, a->erec, b->erec, cc->erec, dd->erec, d.erec);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

// The eta of slot k (internal node order), or null when the slot or its eta
// is missing.
static const double* slot_eta(kelp_ctx* c, int k)
{
    if (k < 0 || k >= kMaxSlots || c->slots[k].dtype < 0) return nullptr;
    return c->slots[k].eta;
}

// A nested blend of slot surfaces; every source the blend reads must exist.
// sa == KELP_SURFACE_GIVEN: the surface of kelp_sigma_surface_upload
// (`which` 0 at the start, 1 at the end of the step).
static int eta_source(kelp_ctx* c, const char* what, int sa, int sb, int st, double alpha0, double alpha,
                      EtaSrc& out, int which = 0)
{
    if (sa == KELP_SURFACE_GIVEN) {
        if (!c->sig_given[which])
            return fail(KELP_ESTATE, "kelp_step: the sigma %s surface was not uploaded (kelp_sigma_surface_upload)",
                        what);
        const double* g = c->sig_eta + (int64_t)(5 + which) * c->m.n_nodes;
        out = EtaSrc{g, g, g, 0.0, 0.0};
        return KELP_OK;
    }
    out = EtaSrc{slot_eta(c, sa), slot_eta(c, sb), slot_eta(c, st), alpha0, alpha};
    const bool need_t = alpha >= 1e-12;
    const bool need_a = alpha <= 1.0 - 1e-12;
    const bool need_b = need_a && alpha0 != 0.0;
    if ((need_a && !out.a) || (need_b && !out.b) || (need_t && !out.t))
        return fail(KELP_ESTATE, "kelp_step: the sigma %s surface needs eta on slots (%d, %d, %d)", what, sa, sb, st);
    if (!out.a) out.a = out.t;
    if (!out.b) out.b = out.a;
    if (!out.t) out.t = out.a;
    return KELP_OK;
}

// Sigma meshes: check the step, build the stage surfaces of this step in
// the context's buffer and point the configuration at them.
static int sigma_prepare(kelp_ctx* c, const kelp_step_args* a, bool is3d, const FieldSlot* sb,
                         const FieldSlot* st, StepCfg& cfg)
{
    if (!is3d) return fail(KELP_EINVAL, "kelp_step: a sigma mesh runs 3-D fields (W on the sigma interfaces)");
    if (!(a->Dmin > 0.0) || !(a->z0b_min >= 0.0))
        return fail(KELP_EINVAL, "kelp_step: sigma steps need Dmin > 0 and z0b_min >= 0");
    const bool steady = a->slot_t < 0;
    EtaSrc start, end, stages;
    int rc = eta_source(c, "start", a->sig_start_a, a->sig_start_b, a->sig_start_t, a->sig_start_alpha0,
                        a->sig_start_alpha, start);
    if (rc) return rc;
    const bool end_is_start = a->sig_end_a < 0 && a->sig_end_a != KELP_SURFACE_GIVEN;
    if (!end_is_start) {
        rc = eta_source(c, "end", a->sig_end_a, a->sig_end_b, a->sig_end_t, a->sig_end_alpha0, a->sig_end_alpha,
                        end, 1);
        if (rc) return rc;
    } else {
        end = start;
    }
    stages = start;
    if (!steady) {
        // stage k's surface is the eta of stage k's velocity blend
        for (int k = 1; k < a->order; ++k) {
            rc = eta_source(c, "stage", a->slot_a, sb ? a->slot_b : -1, a->slot_t, a->alpha0, a->alphas[k], stages);
            if (rc) return rc;
        }
    }
    (void)st;
    const int64_t N = c->m.n_nodes;
    launch_sigma_eta(c->stream, N, a->order, steady ? 1 : 0, start, end, end_is_start ? 1 : 0, stages, a->alphas,
                     c->sig_eta);
    HIP_TRY(hipGetLastError());
    for (int k = 0; k < 5; ++k) cfg.sig_eta[k] = c->sig_eta + (int64_t)k * N;
    cfg.Dmin = a->Dmin;
    cfg.z0b_min = a->z0b_min;
    return KELP_OK;
}

// Validation and the kernel configuration of one step (everything kelp_step
// does before its launch); also the base of a batched step.
int kelp_step_prepare(kelp_ctx* c, const kelp_particles* p, const kelp_step_args* a, StepCfg& cfg, PState& ps)
{
    if (!c || !p || !a) return fail(KELP_EINVAL, "kelp_step: NULL argument");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_step: no mesh uploaded");
    if (a->order != 1 && a->order != 2 && a->order != 4 && a->order != 5)
        return fail(KELP_EINVAL, "kelp_step: order must be 1, 2, 4 or 5 (rk45) (got %d)", a->order);
    if (!(a->dt > 0.0)) return fail(KELP_EINVAL, "kelp_step: dt must be positive (direction goes in time_dir)");
    if (a->time_dir != 1.0 && a->time_dir != -1.0) return fail(KELP_EINVAL, "kelp_step: time_dir must be +1 or -1");
    const bool sigma = c->vertical_coord == 1;
    // rk45 reads no stage fractions (alphas): refuse what it does not take
    // before anything loops over the stages
    if (a->order == 5 && (sigma || a->walk))
        return fail(KELP_EINVAL, "kelp_step: rk45 takes 2-D u/v at element centres (no 3-D, sigma or walk)");
    const int n_stage = a->order == 5 ? 0 : a->order;
    if (a->walk) {
        if (!p->elem) return fail(KELP_EINVAL, "kelp_step: elem_search walk needs the carried elements (elem)");
        if (!c->m.enbr) return fail(KELP_ESTATE, "kelp_step: elem_search walk needs the element neighbours");
        if (sigma) return fail(KELP_EINVAL, "kelp_step: elem_search walk on sigma meshes is not implemented");
    }
    if (a->w_method < 0 || a->w_method > 3)
        return fail(KELP_EINVAL, "kelp_step: w_method must be 0 (nearest), 1 (leastsq), 2 (barycentric) or 3 (idw)");
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
        if (s->dtype != sa.dtype || s->layout != sa.layout || s->n_planes != sa.n_planes ||
            s->n_w_planes != sa.n_w_planes)
            return fail(KELP_EINVAL, "kelp_step: snapshot slots differ in dtype, layout or shape");
    }
    if (a->alpha0 != 0.0 && !sb) return fail(KELP_EINVAL, "kelp_step: alpha0 != 0 needs slot_b");
    for (int k = 0; k < n_stage; ++k)
        if (a->alphas[k] >= 1e-12 && !st) return fail(KELP_EINVAL, "kelp_step: stage alpha > 0 needs slot_t");
    const bool is3d = sa.is3d;
    const int K = c->m.n_levels;
    if (is3d) {
        if (K < 2) return fail(KELP_EINVAL, "kelp_step: a 3-D field needs a mesh with >= 2 levels");
        if (sa.n_planes != K - 1 || sa.n_w_planes != K)
            return fail(KELP_EINVAL, "kelp_step: 3-D field shapes (%lld, %lld) do not match K=%d",
                        (long long)sa.n_planes, (long long)sa.n_w_planes, K);
        if (!p->z) return fail(KELP_EINVAL, "kelp_step: 3-D step needs z");
    } else if (a->az && (a->az_const > 0.0 || a->az_field)) {
        return fail(KELP_EINVAL, "kelp_step: the Az kick needs a 3-D run (integrators.py:893)");
    } else if (sa.n_planes != 1) {
        return fail(KELP_EINVAL, "kelp_step: a layered u/v field needs W (3-D)");
    }
    const bool az_field = a->az && a->az_field;
    if (az_field) {
        // integrators.py:902-906: a field source on sigma meshes is refused
        if (sigma) return fail(KELP_EINVAL, "kelp_step: an Az field on a sigma mesh is not supported");
        for (int k : {a->az_slot_a, a->az_slot_b}) {
            if (k < -1 || k >= kMaxSlots) return fail(KELP_EINVAL, "kelp_step: Az slot %d out of range", k);
            if (k >= 0 && c->slots[k].az_k != K)
                return fail(KELP_ESTATE, "kelp_step: slot %d carries no Az (kelp_field_upload_az)", k);
        }
        if (a->az_slot_a < 0) return fail(KELP_EINVAL, "kelp_step: an Az field needs az_slot_a");
    }
    HIP_TRY(hipSetDevice(c->device));
    for (int k : {a->slot_a, a->slot_b, a->slot_t, a->sig_start_a, a->sig_start_b, a->sig_start_t, a->sig_end_a,
                  a->sig_end_b, a->sig_end_t, az_field ? a->az_slot_a : -1, az_field ? a->az_slot_b : -1})
        if (k >= 0 && k < kMaxSlots) {
            int rc = kelp_wait_slot(c, c->slots[k]);
            if (rc) return rc;
        }
    cfg = StepCfg{};
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
    cfg.az = a->az && (a->az_const > 0.0 || az_field);
    cfg.az_const = a->az_const;
    if (az_field) {
        cfg.az_a = c->slots[a->az_slot_a].az;
        cfg.az_b = a->az_slot_b >= 0 ? c->slots[a->az_slot_b].az : nullptr;
        cfg.az_alpha0 = a->az_alpha0;
    }
    if ((a->periodic_lon && !(a->px1 > a->px0)) || (a->periodic_lat && !(a->py1 > a->py0)))
        return fail(KELP_EINVAL, "kelp_step: a periodic window needs x0 < x1");
    // the kick and the fold re-anchor in mode 1 (exact) or 2, so a stored
    // anchor matches the step's own search in those modes only
    cfg.reuse_anchor = (a->reuse_anchor && a->exact_nearest != 0) ? 1 : 0;
    cfg.walk = a->walk ? 1 : 0;
    if (cfg.walk && sa.layout != 0)
        return fail(KELP_EINVAL, "kelp_step: elem_search walk with element-centre u/v is not implemented");
    cfg.periodic_lon = a->periodic_lon ? 1 : 0;
    cfg.periodic_lat = a->periodic_lat ? 1 : 0;
    cfg.px0 = a->px0;
    cfg.px1 = a->px1;
    cfg.py0 = a->py0;
    cfg.py1 = a->py1;
    if (a->beach && a->use_eta) {
        if (!sa.eta || (sb && !sb->eta))
            return fail(KELP_ESTATE, "kelp_step: beaching with use_eta needs eta on the bracket snapshots");
        cfg.use_eta = 1;
        cfg.eta_a = sa.eta;
        cfg.eta_b = sb ? sb->eta : sa.eta;
    }
    if (sigma) {
        int rc = sigma_prepare(c, a, is3d, sb, st, cfg);
        if (rc) return rc;
    }
    ps = PState{p->lon, p->lat, p->z, p->status, p->nn, p->fr, p->u, p->v, p->w, p->pid, p->wet_streak,
                reinterpret_cast<unsigned long long*>(p->alive_count), p->n, p->elem};
    return KELP_OK;
}

int kelp_step(kelp_ctx* c, const kelp_particles* p, const kelp_step_args* a)
{
    StepCfg cfg{};
    PState ps{};
    int rc = kelp_step_prepare(c, p, a, cfg, ps);
    if (rc) return rc;
    if (p->n == 0) return KELP_OK;
    const bool sigma = c->vertical_coord == 1;
    const FieldSlot& sa = c->slots[a->slot_a];
    const FieldSlot* sb = opt_slot(c, a->slot_b);
    const FieldSlot* st = opt_slot(c, a->slot_t);
    const bool is3d = sa.is3d;
    if (a->order == 5) {
        // integrators.py:6485-6513: rk45 is 2-D, element-layout least squares
        if (is3d || sa.layout != 1 || sigma || cfg.walk)
            return fail(KELP_EINVAL, "kelp_step: rk45 takes 2-D u/v at element centres (no 3-D, sigma or walk)");
        if (!(a->dt > 0.0)) return fail(KELP_EINVAL, "kelp_step: rk45 needs dt > 0");
        cfg.rk45_tol = a->rk45_tol_m > 0.0 ? a->rk45_tol_m : 1.0;
        cfg.rk45_h_min = a->rk45_h_min > 0.0 ? a->rk45_h_min : a->dt * 1e-4;
        cfg.rk45_inv_dtn = a->rk45_inv_dtn;
        if (!(cfg.rk45_inv_dtn >= 0.0) || (cfg.rk45_inv_dtn > 0.0 && !st))
            return fail(KELP_EINVAL, "kelp_step: rk45 blending towards slot_t needs slot_t and rk45_inv_dtn >= 0");
        // :6531: the cap exceeds the accepts of a whole step at h_min and their rejections
        const double nmin = a->dt / cfg.rk45_h_min;
        cfg.rk45_max_sub = nmin < 1e12 ? (int64_t)(2.0 * (double)(int64_t)nmin + 16.0) : (int64_t)1 << 40;
        if (cfg.rk45_max_sub < 10000) cfg.rk45_max_sub = 10000;
        Snap<float, false> ff{};
        Snap<double, false> fd{};
        if (sa.dtype == KELP_DT_F32) {
            ff = snap_of<float, false>(sa.erec, sb ? sb->erec : nullptr, st ? st->erec : nullptr);
            if (c->m.max_v == 3) launch_rk45_v<3, float>(c->stream, c->m, ff, cfg, ps);
            else launch_rk45_v<4, float>(c->stream, c->m, ff, cfg, ps);
        } else {
            fd = snap_of<double, false>(sa.erec, sb ? sb->erec : nullptr, st ? st->erec : nullptr);
            if (c->m.max_v == 3) launch_rk45_v<3, double>(c->stream, c->m, fd, cfg, ps);
            else launch_rk45_v<4, double>(c->stream, c->m, fd, cfg, ps);
        }
        HIP_TRY(hipGetLastError());
        return KELP_OK;
    }
    if (sa.dtype == KELP_DT_F16S) {
        // node-layout z-level only (refused at upload otherwise)
        if (sigma) return fail(KELP_EINVAL, "kelp_step: float16 snapshots on a sigma mesh");
        if (is3d)
            launch_step_f16s<true>(c, a->order, sa, sb, st, cfg, ps);
        else
            launch_step_f16s<false>(c, a->order, sa, sb, st, cfg, ps);
        HIP_TRY(hipGetLastError());
        return KELP_OK;
    }
    if (sa.dtype == KELP_DT_BF16) {
        // node-layout z-level only (refused at upload otherwise)
        if (sigma) return fail(KELP_EINVAL, "kelp_step: bfloat16 snapshots on a sigma mesh");
        if (is3d)
            launch_step_t<true, Bf16, false>(c, a->order, sa, sb, st, cfg, ps);
        else
            launch_step_t<false, Bf16, false>(c, a->order, sa, sb, st, cfg, ps);
    } else if (sa.dtype == KELP_DT_F32) {
        launch_step<float>(c, a->order, is3d, sa, sb, st, cfg, ps);
    } else {
        launch_step<double>(c, a->order, is3d, sa, sb, st, cfg, ps);
    }
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

int kelp_gh_sample_stats(kelp_ctx* c, int64_t n, const double* lon, const double* lat, const int32_t* nn,
                         const int32_t* cum, double cell, double lon_min, double lat_min, int64_t nx, int64_t ny,
                         int wrap, int64_t* occ, double* cost, int8_t* unprov, int8_t* approx)
{
    if (!c) return fail(KELP_EINVAL, "kelp_gh_sample_stats: NULL context");
    if (!c->has_mesh) return fail(KELP_ESTATE, "kelp_gh_sample_stats: no mesh uploaded");
    if (n < 0 || (n > 0 && (!lon || !lat || !nn || !cum || !occ || !cost || !unprov || !approx)))
        return fail(KELP_EINVAL, "kelp_gh_sample_stats: NULL argument");
    if (!(cell > 0.0) || nx < 1 || ny < 1 || nx * ny > 50000000)
        return fail(KELP_EINVAL, "kelp_gh_sample_stats: bad candidate grid");
    if (n == 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    DMesh g = c->m;
    g.cell = cell;
    g.lon_min = lon_min;
    g.lat_min = lat_min;
    g.nx = (int)nx;
    g.ny = (int)ny;
    g.wrap = wrap ? 1 : 0;
    gh_sample_stats_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(g, n, lon, lat, nn, c->node_new_of_old, cum,
                                                                   occ, cost, unprov, approx);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

}  // extern "C"
