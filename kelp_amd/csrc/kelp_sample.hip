// kelp_sample.hip -- record-point sampling of node scalars on the device
// (numerics/sampling.py:727 sample_field, z-level meshes).
//
// One thread per particle reads the post-step state (position, depth,
// status, anchor) and writes the sampled value, NaN where the particle is
// not alive or has no anchor.  2-D and mid-layer fields combine the unique
// 1-ring node values (sampling.py:83 _combine_horizontal: nearest, leastsq,
// barycentric, idw; non-finite values filtered as the reference does);
// interface fields go through the W interpolators of the step
// (vertical.py:1330 interp_w_spatial, time_dir +1) on a record copy of the
// field.  The record's field is snapshot a, or a blended towards b by alpha
// (core/io.py:68 _blend of the bracket's scalars).
#include "kelp_internal.h"

using namespace kelp;

namespace {

// value idx of a field is a[idx * stride] (stride 2: one variable of an
// interleaved pair, kelp_sample_vars)
struct SField {
    const double* __restrict__ a;
    const double* __restrict__ b;
    double alpha;
    int64_t n_planes;
    int64_t stride = 1;
};

__device__ __forceinline__ double sval(const SField& s, int64_t idx)
{
    const int64_t j = idx * s.stride;
    if (s.alpha == 0.0) return s.a[j];
    return (1.0 - s.alpha) * s.a[j] + s.alpha * s.b[j];
}

// The bracket of depth z among the mid-layer depths 0.5 * (z_k + z_k+1) in
// the branch structure of sampling.py:433 _interp_mid_linear_single: plane
// `k` alone (al < 0) or planes k and k+1 blended by al.  The mid-layer
// depths are non-decreasing, so the first k with z < zm[k+1] is found by
// bisection; a NaN depth ends on the last plane, as the linear scan does.
struct MidBracket {
    int k;
    double al;
};

__device__ MidBracket mid_bracket(const DMesh& m, double z)
{
    const double* zl = m.zlev;
    const int K1 = m.n_levels - 1;
    if (K1 <= 1 || z <= 0.5 * (zl[0] + zl[1])) return {0, -1.0};
    if (z >= 0.5 * (zl[K1 - 1] + zl[K1])) return {K1 - 1, -1.0};
    int lo = 0, hi = K1 - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (z < 0.5 * (zl[mid + 1] + zl[mid + 2]))
            hi = mid;
        else
            lo = mid + 1;
    }
    if (lo > K1 - 2) return {K1 - 1, -1.0};
    const double zk = 0.5 * (zl[lo] + zl[lo + 1]);
    const double zk1 = 0.5 * (zl[lo + 1] + zl[lo + 2]);
    const double dz = zk1 - zk;
    if (dz < 1e-12) return {lo, -1.0};
    return {lo, (z - zk) / dz};
}

__device__ __forceinline__ double mid_value(const DMesh& m, const SField& s, const MidBracket& b, int node)
{
    // mid-layer planes are stored per node (column layout): planes k and
    // k + 1 of one node share a cache line, and particles near one node at
    // nearby depths share its column
    const int64_t col = (int64_t)node * s.n_planes;
    if (b.al < 0.0) return sval(s, col + b.k);
    return (1.0 - b.al) * sval(s, col + b.k) + b.al * sval(s, col + b.k + 1);
}

// The value the horizontal combination uses at node nd (sampling.py:517):
// 2-D the node value; mid-layer linear in depth at min(z, depth[nd]), with
// the particle's own bracket `bz` reused for every node deeper than z.
__device__ __forceinline__ double node_value(const DMesh& m, const SField& s, double z, const MidBracket& bz,
                                             int nd)
{
    if (s.n_planes == 1) return sval(s, nd);
    const double dep = m.depth[nd];
    if (z < dep) return mid_value(m, s, bz, nd);
    return mid_value(m, s, mid_bracket(m, dep), nd);
}

// visit the unique 1-ring nodes of `node` in the order of
// vertical.py:446 _collect_neighbour_nodes (centre first, at most 64)
template <int MAXV, typename F>
__device__ inline int for_neighbours(const DMesh& m, int node, F&& fn)
{
    int k0, k1;
    n2e_range(m, node, k0, k1);
    int count = 0;
    fn(node);
    ++count;
    for (int k = k0; k < k1; ++k) {
        const int e = m.n2e[k];
        if (e < 0) continue;
        const int4 q = m.elems[e];
#pragma unroll
        for (int v = 0; v < MAXV; ++v) {
            const int nd = v == 0 ? q.x : (v == 1 ? q.y : (v == 2 ? q.z : q.w));
            if (nd < 0 || nd == node || seen_before<MAXV>(m, k0, k, v, nd)) continue;
            if (count >= 64) continue;
            fn(nd);
            ++count;
        }
    }
    return count;
}

// sampling.py:83 _combine_horizontal; method 0 nearest, 1 leastsq,
// 2 barycentric, 3 idw
template <int MAXV>
__device__ double combine(const DMesh& m, const SField& s, double z, double lon_p, double lat_p, int nn,
                          int method)
{
    const MidBracket bz = s.n_planes > 1 ? mid_bracket(m, z) : MidBracket{0, -1.0};
    const double fallback = node_value(m, s, z, bz, nn);
    if (method == 0) return fallback;
    if (method == 3) {
        double z0 = km_cos(lat_p * kDeg2Rad) * kSLM;
        if (z0 < 1.0) z0 = 1.0;
        double sw = 0.0, swv = 0.0;
        bool hit = false;
        double hit_val = 0.0;
        for_neighbours<MAXV>(m, nn, [&](int nd) {
            if (hit) return;
            const double2 ll = m.lonlat[nd];
            const double dlon = km_pymod360(ll.x - lon_p + 180.0) - 180.0;
            const double dx = dlon * z0;
            const double dy = (ll.y - lat_p) * kSLM;
            const double d2 = dx * dx + dy * dy;
            const double val = node_value(m, s, z, bz, nd);
            if (d2 < 1e-12) {
                hit = true;
                hit_val = val;
                return;
            }
            if (!__builtin_isfinite(val)) return;
            const double w = 1.0 / d2;
            sw += w;
            swv += w * val;
        });
        if (hit) return hit_val;
        if (sw < 1e-30) return fallback;
        return swv / sw;
    }
    if (method == 2) {
        const double R = 6371000.0;
        int k0, k1;
        n2e_range(m, nn, k0, k1);
        for (int k = k0; k < k1; ++k) {
            const int e = m.n2e[k];
            if (e < 0) continue;
            int verts[4] = {0, 0, 0, 0};
            const int nv = interp_ring<MAXV>(m.elems[e], verts);
            if (nv < 3) continue;
            double lo[4] = {0, 0, 0, 0}, la[4] = {0, 0, 0, 0};
#pragma unroll
            for (int v = 0; v < MAXV; ++v)
                if (v < nv) {
                    const double2 ll = m.lonlat[verts[v]];
                    lo[v] = ll.x;
                    la[v] = ll.y;
                }
            const double first = lo[0];
            double lmin = first, lmax = first, ssum = 0.0;
#pragma unroll
            for (int v = 0; v < MAXV; ++v)
                if (v < nv) {
                    ssum += lo[v];
                    if (lo[v] < lmin) lmin = lo[v];
                    if (lo[v] > lmax) lmax = lo[v];
                }
            const double nd_ = (double)nv;
            double lon0;
            if (lmax - lmin > 180.0) {
                double d = 0.0;
#pragma unroll
                for (int v = 0; v < MAXV; ++v)
                    if (v < nv) d += km_pymod360(lo[v] - first + 180.0) - 180.0;
                lon0 = first + d / nd_;
            } else {
                lon0 = ssum / nd_;
            }
            double lat0 = 0.0;
#pragma unroll
            for (int v = 0; v < MAXV; ++v)
                if (v < nv) lat0 += la[v];
            lat0 /= nd_;
            const double cos_lat0 = km_cos(lat0 * kDeg2Rad);
            double vx[4] = {0, 0, 0, 0}, vy[4] = {0, 0, 0, 0};
#pragma unroll
            for (int v = 0; v < MAXV; ++v)
                if (v < nv) {
                    vx[v] = (km_pymod360(lo[v] - lon0 + 180.0) - 180.0) * kDeg2Rad * R * cos_lat0;
                    vy[v] = (la[v] - lat0) * kDeg2Rad * R;
                }
            const double px = (km_pymod360(lon_p - lon0 + 180.0) - 180.0) * kDeg2Rad * R * cos_lat0;
            const double py = (lat_p - lat0) * kDeg2Rad * R;
            for (int t = 0; t < nv - 2; ++t) {
                const int i1 = t + 1, i2 = t + 2;
                const double ax = vx[0], ay = vy[0], bx = vx[i1], by = vy[i1], cx = vx[i2], cy = vy[i2];
                const double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
                if (km_fabs(denom) < 1e-14) continue;
                const double w1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / denom;
                const double w2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / denom;
                const double w3 = 1.0 - w1 - w2;
                const double eps = -1e-6;
                if (w1 >= eps && w2 >= eps && w3 >= eps) {
                    if (MAXV == 4 && nv == 4) {
                        double wts[4];
                        wachspress4(px, py, vx, vy, wts);
                        double acc = 0.0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const double val = node_value(m, s, z, bz, verts[q]);
                            if (!__builtin_isfinite(val)) return fallback;
                            acc += wts[q] * val;
                        }
                        return acc;
                    }
                    const double v0 = node_value(m, s, z, bz, verts[0]);
                    const double v1 = node_value(m, s, z, bz, verts[i1]);
                    const double v2 = node_value(m, s, z, bz, verts[i2]);
                    if (!(__builtin_isfinite(v0) && __builtin_isfinite(v1) && __builtin_isfinite(v2)))
                        return fallback;
                    return w1 * v0 + w2 * v1 + w3 * v2;
                }
            }
        }
        return fallback;
    }
    // leastsq
    double z0 = km_cos(lat_p * kDeg2Rad) * kSLM;
    if (z0 < 1.0) z0 = 1.0;
    double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int n_fin = 0;
    const int count = for_neighbours<MAXV>(m, nn, [&](int nd) {
        const double val = node_value(m, s, z, bz, nd);
        if (!__builtin_isfinite(val)) return;
        ++n_fin;
        const double2 ll = m.lonlat[nd];
        const double dlon = km_pymod360(ll.x - lon_p + 180.0) - 180.0;
        const double ex = dlon * z0;
        const double ey = (ll.y - lat_p) * kSLM;
        a11 += 1.0;
        a12 += ex;
        a13 += ey;
        a22 += ex * ex;
        a23 += ex * ey;
        a33 += ey * ey;
        f0 += val;
        f1 += val * ex;
        f2 += val * ey;
    });
    if (count < 3 || n_fin < 3) return fallback;
    const double a21 = a12, a31 = a13, a32 = a23;
    const double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                        a12 * a21 * a33 - a11 * a23 * a32);
    if (km_fabs(det) < 1e-12 || km_fabs(det) < 1e-12 * km_fabs(a11 * a22 * a33)) return fallback;
    const double inv = 1.0 / km_fabs(det);
    const double b11 = a22 * a33 - a23 * a32;
    const double b12 = a13 * a32 - a12 * a33;
    const double b13 = a12 * a23 - a13 * a22;
    return (b11 * f0 + b12 * f1 + b13 * f2) * inv;
}

struct SampleState {
    const double* __restrict__ lon;
    const double* __restrict__ lat;
    const double* __restrict__ z;
    const int32_t* __restrict__ status;
    const int32_t* __restrict__ nn;
    int64_t n;
};

template <int MAXV>
__global__ void sample_kernel(DMesh m, SField s, Snap<double, true> fw, int method, SampleState st,
                              double* __restrict__ out)
{
    const int64_t i = xcd_thread_index();
    if (i >= st.n) return;
    const int node = st.nn[i];
    if (st.status[i] != kAlive || node < 0) {
        out[i] = __builtin_nan("");
        return;
    }
    const double lon = st.lon[i], lat = st.lat[i];
    const double z = st.z ? st.z[i] : 0.0;
    double r;
    if (s.n_planes == m.n_levels && s.n_planes > 1) {
        // interface field: the W interpolators on the record copy, time_dir +1
        if (method == 0)
            r = w_linear(m, fw, s.alpha, 0.0, z, node, 1.0);
        else if (method == 1)
            r = w_lsq<MAXV>(m, fw, s.alpha, 0.0, lon, lat, z, node, 1.0);
        else if (method == 3)
            r = w_idw<MAXV>(m, fw, s.alpha, 0.0, lon, lat, z, node, 1.0);
        else
            r = w_bary<MAXV>(m, fw, s.alpha, 0.0, lon, lat, z, node, 1.0);
    } else {
        r = combine<MAXV>(m, s, z, lon, lat, node, method);
    }
    out[i] = r;
}

// Two 1-ring scalars by least squares in one pass (the usual T and S pair):
// the neighbour scan, coordinates and depth brackets are shared, each
// variable keeps its own sums (non-finite values are filtered per
// variable), so each output equals combine(..., method 1) of its variable.
template <int MAXV>
__global__ void sample2_lsq_kernel(DMesh m, SField s0, SField s1, SampleState st, double* __restrict__ out0,
                                   double* __restrict__ out1)
{
    const int64_t i = xcd_thread_index();
    if (i >= st.n) return;
    const int nn = st.nn[i];
    if (st.status[i] != kAlive || nn < 0) {
        out0[i] = __builtin_nan("");
        out1[i] = __builtin_nan("");
        return;
    }
    const double lon_p = st.lon[i], lat_p = st.lat[i];
    const double z = st.z ? st.z[i] : 0.0;
    const bool mid = s0.n_planes > 1 || s1.n_planes > 1;
    const MidBracket bz = mid ? mid_bracket(m, z) : MidBracket{0, -1.0};
    const SField* sv[2] = {&s0, &s1};
    double z0 = km_cos(lat_p * kDeg2Rad) * kSLM;
    if (z0 < 1.0) z0 = 1.0;
    double a[2][6] = {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}};
    double f[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int n_fin[2] = {0, 0};
    double fallback[2];
    const int count = for_neighbours<MAXV>(m, nn, [&](int nd) {
        MidBracket bn = bz;
        if (mid) {
            const double dep = m.depth[nd];
            if (!(z < dep)) bn = mid_bracket(m, dep);
        }
        const double2 ll = m.lonlat[nd];
        const double dlon = km_pymod360(ll.x - lon_p + 180.0) - 180.0;
        const double ex = dlon * z0;
        const double ey = (ll.y - lat_p) * kSLM;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const SField& s = *sv[q];
            const double val = s.n_planes == 1 ? sval(s, nd) : mid_value(m, s, bn, nd);
            if (nd == nn) fallback[q] = val;
            if (!__builtin_isfinite(val)) continue;
            ++n_fin[q];
            a[q][0] += 1.0;
            a[q][1] += ex;
            a[q][2] += ey;
            a[q][3] += ex * ex;
            a[q][4] += ex * ey;
            a[q][5] += ey * ey;
            f[q][0] += val;
            f[q][1] += val * ex;
            f[q][2] += val * ey;
        }
    });
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        double r = fallback[q];
        if (count >= 3 && n_fin[q] >= 3) {
            const double a11 = a[q][0], a12 = a[q][1], a13 = a[q][2], a22 = a[q][3], a23 = a[q][4],
                         a33 = a[q][5];
            const double a21 = a12, a31 = a13, a32 = a23;
            const double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                                a12 * a21 * a33 - a11 * a23 * a32);
            if (!(km_fabs(det) < 1e-12 || km_fabs(det) < 1e-12 * km_fabs(a11 * a22 * a33))) {
                const double inv = 1.0 / km_fabs(det);
                const double b11 = a22 * a33 - a23 * a32;
                const double b12 = a13 * a32 - a12 * a33;
                const double b13 = a12 * a23 - a13 * a22;
                r = (b11 * f[q][0] + b12 * f[q][1] + b13 * f[q][2]) * inv;
            }
        }
        (q == 0 ? out0 : out1)[i] = r;
    }
}

// host planes (n_planes x N, mesh order) into internal node order: plane
// layout (n_planes x N) or, with `column`, one column per node (N x n_planes)
__global__ void permute_planes_kernel(int64_t N, int64_t n_planes, const int32_t* __restrict__ old_of_new,
                                      const double* __restrict__ in, double* __restrict__ out, int column)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * n_planes) return;
    if (column) {
        const int64_t k = i / n_planes, L = i - k * n_planes;
        out[i] = in[L * N + old_of_new[k]];
    } else {
        const int64_t L = i / N, k = i - L * N;
        out[i] = in[L * N + old_of_new[k]];
    }
}

// interface planes (internal order) into (0, 0, top, bottom) records per layer
__global__ void pair_records_kernel(int64_t N, int64_t n_layers, const double* __restrict__ planes,
                                    double4* __restrict__ rec)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * n_layers) return;
    const int64_t L = i / N, k = i - L * N;
    rec[i] = make_double4(0.0, 0.0, planes[L * N + k], planes[(L + 1) * N + k]);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

__global__ void interleave_kernel(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                  double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[2 * i] = a[i];
    out[2 * i + 1] = b[i];
}

// The slot's two mid-layer variables va, vb interleaved per (node, layer),
// so one line holds both at the two layers a depth brackets; built once per
// upload of either variable.  Null when it cannot be allocated (the caller
// then reads the separate columns).
double* slot_pair(kelp_ctx* c, FieldSlot& fs, int va, int vb)
{
    const ScalarField& f0 = fs.scalars[va];
    const ScalarField& f1 = fs.scalars[vb];
    if (fs.pair && fs.pair_vars[0] == va && fs.pair_vars[1] == vb) return fs.pair;
    const int64_t n = c->m.n_nodes * f0.n_planes;
    if (!fs.pair && hipMalloc(&fs.pair, (size_t)n * 2 * sizeof(double)) != hipSuccess) {
        fs.pair = nullptr;
        (void)hipGetLastError();
        return nullptr;
    }
    interleave_kernel<<<grid_of(n), 256, 0, c->stream>>>(n, f0.planes, f1.planes, fs.pair);
    if (hipGetLastError() != hipSuccess) return nullptr;
    fs.pair_vars[0] = va;
    fs.pair_vars[1] = vb;
    return fs.pair;
}

}  // namespace

extern "C" int kelp_scalar_upload(kelp_ctx* c, int slot, int var, int64_t n_planes, const double* host_values)
{
    if (!c || !host_values) return kelp_fail(KELP_EINVAL, "kelp_scalar_upload: NULL argument");
    if (slot < 0 || slot >= kMaxSlots) return kelp_fail(KELP_EINVAL, "kelp_scalar_upload: slot %d out of range", slot);
    if (var < 0 || var >= kMaxScalars)
        return kelp_fail(KELP_EINVAL, "kelp_scalar_upload: variable %d not in [0, %d)", var, kMaxScalars);
    FieldSlot& fs = c->slots[slot];
    if (fs.dtype < 0) return kelp_fail(KELP_ESTATE, "kelp_scalar_upload: slot %d holds no snapshot", slot);
    const int64_t K = c->m.n_levels, N = c->m.n_nodes;
    if (!(n_planes == 1 || (K >= 2 && (n_planes == K - 1 || n_planes == K))))
        return kelp_fail(KELP_EINVAL, "kelp_scalar_upload: %lld planes (1, K-1 or K = %lld expected)",
                         (long long)n_planes, (long long)K);
    HIP_TRY(hipSetDevice(c->device));
    ScalarField& f = fs.scalars[var];
    if (fs.pair_vars[0] == var || fs.pair_vars[1] == var) fs.pair_vars[0] = fs.pair_vars[1] = -1;
    kelp_free_scalar(f);
    const size_t bytes = (size_t)n_planes * N * sizeof(double);
    double* tmp = nullptr;
    hipError_t e = hipMalloc(&tmp, bytes);
    if (e == hipSuccess) e = hipMalloc(&f.planes, bytes);
    const bool ifc = n_planes == K && K >= 2;
    if (e == hipSuccess && ifc) e = hipMalloc(&f.wrec, (size_t)(K - 1) * N * sizeof(double4));
    if (e == hipSuccess) e = hipMemcpyAsync(tmp, host_values, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        permute_planes_kernel<<<grid_of(N * n_planes), 256, 0, c->stream>>>(N, n_planes, c->node_old_of_new, tmp,
                                                                          f.planes, (!ifc && n_planes > 1) ? 1 : 0);
        if (ifc)
            pair_records_kernel<<<grid_of(N * (K - 1)), 256, 0, c->stream>>>(N, K - 1, f.planes,
                                                                           (double4*)f.wrec);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) {
        kelp_free_scalar(f);
        return kelp_fail(e == hipErrorOutOfMemory ? KELP_ENOMEM : KELP_EHIP, "kelp_scalar_upload: %s",
                         hipGetErrorString(e));
    }
    f.n_planes = n_planes;
    return KELP_OK;
}

extern "C" int kelp_field_upload_az(kelp_ctx* c, int slot, int64_t n_planes, const double* host_az)
{
    if (!c || !host_az) return kelp_fail(KELP_EINVAL, "kelp_field_upload_az: NULL argument");
    if (slot < 0 || slot >= kMaxSlots)
        return kelp_fail(KELP_EINVAL, "kelp_field_upload_az: slot %d out of range", slot);
    FieldSlot& fs = c->slots[slot];
    if (fs.dtype < 0 || !fs.is3d)
        return kelp_fail(KELP_ESTATE, "kelp_field_upload_az: slot %d holds no 3-D snapshot", slot);
    const int64_t K = c->m.n_levels, N = c->m.n_nodes;
    // readers/fesom2.py:1607-1632: Az lives on the interfaces (the W layout)
    if (n_planes != K)
        return kelp_fail(KELP_EINVAL, "kelp_field_upload_az: %lld planes, Az needs the K = %lld interfaces",
                         (long long)n_planes, (long long)K);
    HIP_TRY(hipSetDevice(c->device));
    const size_t bytes = (size_t)K * N * sizeof(double);
    hipError_t e = hipSuccess;
    if (!fs.az) e = hipMalloc(&fs.az, bytes);
    double* tmp = nullptr;
    if (e == hipSuccess) e = hipMalloc(&tmp, bytes);
    // on the context stream: after every step enqueued so far, which may
    // read the slot's previous Az
    if (e == hipSuccess) e = hipMemcpyAsync(tmp, host_az, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        permute_planes_kernel<<<grid_of(N * K), 256, 0, c->stream>>>(N, K, c->node_old_of_new, tmp, fs.az, 1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess)
        return kelp_fail(e == hipErrorOutOfMemory ? KELP_ENOMEM : KELP_EHIP, "kelp_field_upload_az: %s",
                         hipGetErrorString(e));
    fs.az_k = K;
    return KELP_OK;
}

extern "C" int kelp_sample(kelp_ctx* c, const kelp_particles* p, int var, int slot_a, int slot_b, double alpha,
                           int method, double* out)
{
    if (!c || !p || !out) return kelp_fail(KELP_EINVAL, "kelp_sample: NULL argument");
    if (!c->has_mesh) return kelp_fail(KELP_ESTATE, "kelp_sample: no mesh uploaded");
    if (var < 0 || var >= kMaxScalars) return kelp_fail(KELP_EINVAL, "kelp_sample: bad variable %d", var);
    if (method < 0 || method > 3) return kelp_fail(KELP_EINVAL, "kelp_sample: method must be 0..3");
    if (slot_a < 0 || slot_a >= kMaxSlots || slot_b < -1 || slot_b >= kMaxSlots)
        return kelp_fail(KELP_EINVAL, "kelp_sample: slot out of range");
    const ScalarField& fa = c->slots[slot_a].scalars[var];
    if (!fa.planes) return kelp_fail(KELP_ESTATE, "kelp_sample: slot %d has no variable %d", slot_a, var);
    const ScalarField* fb = slot_b >= 0 ? &c->slots[slot_b].scalars[var] : nullptr;
    if (alpha != 0.0 && (!fb || !fb->planes || fb->n_planes != fa.n_planes))
        return kelp_fail(KELP_EINVAL, "kelp_sample: a blend needs the variable on slot_b with the same shape");
    const bool ifc = fa.n_planes == c->m.n_levels && c->m.n_levels >= 2;
    if (fa.n_planes > 1 && !p->z) return kelp_fail(KELP_EINVAL, "kelp_sample: a layered field needs z");
    if (!p->lon || !p->lat || !p->status || !p->nn) return kelp_fail(KELP_EINVAL, "kelp_sample: missing particle array");
    if (p->n <= 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    for (int k : {slot_a, slot_b})
        if (k >= 0) {
            int rc = kelp_wait_slot(c, c->slots[k]);
            if (rc) return rc;
        }
    SField s{fa.planes, (fb && fb->planes) ? fb->planes : fa.planes, alpha, fa.n_planes};
    Snap<double, true> fw;
    fw.a = static_cast<const double4*>(fa.wrec);
    fw.b = (fb && fb->wrec) ? static_cast<const double4*>(fb->wrec) : fw.a;
    fw.t = fw.a;
    SampleState st{p->lon, p->lat, p->z, p->status, p->nn, p->n};
    if (c->m.max_v == 3)
        sample_kernel<3><<<grid_of(p->n), 256, 0, c->stream>>>(c->m, s, fw, method, st, out);
    else
        sample_kernel<4><<<grid_of(p->n), 256, 0, c->stream>>>(c->m, s, fw, method, st, out);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}

// Several variables of one record at once: rows of `out` (n_vars x n), in
// the order of `vars`; 1-ring least-squares pairs share one pass.
extern "C" int kelp_sample_vars(kelp_ctx* c, const kelp_particles* p, int n_vars, const int32_t* vars,
                                int slot_a, int slot_b, double alpha, int method, double* out)
{
    if (!c || !p || !vars || !out || n_vars < 0) return kelp_fail(KELP_EINVAL, "kelp_sample_vars: bad argument");
    if (!c->has_mesh) return kelp_fail(KELP_ESTATE, "kelp_sample_vars: no mesh uploaded");
    int k = 0;
    while (k < n_vars) {
        const int va = vars[k];
        if (va < 0 || va >= kMaxScalars) return kelp_fail(KELP_EINVAL, "kelp_sample_vars: bad variable %d", va);
        double* oa = out + (int64_t)k * p->n;
        const bool pair = method == 1 && k + 1 < n_vars && vars[k + 1] >= 0 && vars[k + 1] < kMaxScalars;
        if (pair && slot_a >= 0 && slot_a < kMaxSlots && slot_b >= -1 && slot_b < kMaxSlots) {
            const int vb = vars[k + 1];
            const ScalarField& fa0 = c->slots[slot_a].scalars[va];
            const ScalarField& fa1 = c->slots[slot_a].scalars[vb];
            const ScalarField* fb0 = slot_b >= 0 ? &c->slots[slot_b].scalars[va] : nullptr;
            const ScalarField* fb1 = slot_b >= 0 ? &c->slots[slot_b].scalars[vb] : nullptr;
            const int64_t K = c->m.n_levels;
            const auto ring = [&](const ScalarField& f) { return f.planes && !(f.n_planes == K && K >= 2); };
            const auto blend_ok = [&](const ScalarField& f, const ScalarField* g) {
                return alpha == 0.0 || (g && g->planes && g->n_planes == f.n_planes);
            };
            if (ring(fa0) && ring(fa1) && blend_ok(fa0, fb0) && blend_ok(fa1, fb1) && p->n > 0 &&
                p->lon && p->lat && p->status && p->nn && (p->z || (fa0.n_planes == 1 && fa1.n_planes == 1))) {
                HIP_TRY(hipSetDevice(c->device));
                for (int k : {slot_a, slot_b})
                    if (k >= 0) {
                        int rc = kelp_wait_slot(c, c->slots[k]);
                        if (rc) return rc;
                    }
                SField s0{fa0.planes, (fb0 && fb0->planes) ? fb0->planes : fa0.planes, alpha, fa0.n_planes};
                SField s1{fa1.planes, (fb1 && fb1->planes) ? fb1->planes : fa1.planes, alpha, fa1.n_planes};
                if (fa0.n_planes > 1 && fa1.n_planes == fa0.n_planes) {
                    // two mid-layer columns: read them interleaved (same values)
                    const bool use_b = alpha != 0.0 && slot_b >= 0;
                    if (fa0.planes == fa1.planes) {
                        // one variable asked twice: nothing to interleave
                    } else if (double* pa = slot_pair(c, c->slots[slot_a], va, vb)) {
                        double* pb = use_b ? slot_pair(c, c->slots[slot_b], va, vb) : pa;
                        if (pb) {
                            s0 = SField{pa, pb, alpha, fa0.n_planes, 2};
                            s1 = SField{pa + 1, pb + 1, alpha, fa1.n_planes, 2};
                        }
                    }
                }
                SampleState st{p->lon, p->lat, p->z, p->status, p->nn, p->n};
                double* ob = out + (int64_t)(k + 1) * p->n;
                if (c->m.max_v == 3)
                    sample2_lsq_kernel<3><<<grid_of(p->n), 256, 0, c->stream>>>(c->m, s0, s1, st, oa, ob);
                else
                    sample2_lsq_kernel<4><<<grid_of(p->n), 256, 0, c->stream>>>(c->m, s0, s1, st, oa, ob);
                HIP_TRY(hipGetLastError());
                k += 2;
                continue;
            }
        }
        int rc = kelp_sample(c, p, va, slot_a, slot_b, alpha, method, oa);
        if (rc) return rc;
        k += 1;
    }
    return KELP_OK;
}
