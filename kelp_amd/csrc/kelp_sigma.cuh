// kelp_sigma.cuh -- the sigma-coordinate column (numerics/vertical_sigma.py)
// on the device: column heights, layer lookup, the dia-sigma flux W of one
// stage and the end-of-step depth clamp.
//
// On a sigma mesh the mesh's level array holds the sigma interfaces
// sigma_lev (1 at the surface .. 0 at the bottom); a node's column height is
// H = max(depth + eta, Dmin) and interface k lies at z = (1 - sigma_k) H.
// `eta` is the stage's surface elevation per node (internal node order).
#pragma once

#include "kelp_device.cuh"

namespace kelp {

// vertical_sigma.py:473 _col_H
__device__ __forceinline__ double col_H(const DMesh& m, const double* __restrict__ eta, double Dmin, int nd)
{
    double H = m.depth[nd] + eta[nd];
    if (H < Dmin) H = Dmin;
    return H;
}

// vertical_sigma.py:482 _sigma_frac
__device__ __forceinline__ double sigma_frac(double z, double Hp)
{
    double f = z / Hp;
    if (f < 0.0)
        f = 0.0;
    else if (f > 1.0)
        f = 1.0;
    return f;
}

// vertical_sigma.py:74 find_layer_sigma: mid-layer of depth z in a column of
// height H, clamped to [0, K-2]
__device__ inline int find_layer_sigma(const DMesh& m, double z, double H)
{
    const int K = m.n_levels;
    const double* s = m.zlev;
    if (K < 2) return 0;
    if (z <= (1.0 - s[0]) * H) return 0;
    if (z >= (1.0 - s[K - 1]) * H) return K - 2;
    for (int k = 0; k < K - 1; ++k)
        if (z < (1.0 - s[k + 1]) * H) return k;
    return K - 2;
}

// vertical_sigma.py:237 _loop_clamp_depth_sigma_H for one particle
__device__ __forceinline__ double clamp_sigma(double z, double H, double Dmin, double z0b_min)
{
    if (H < Dmin) H = Dmin;
    if (z < 0.0) z = 0.0;
    double bottom = H - z0b_min;
    if (bottom < 0.0) bottom = 0.0;
    if (z > bottom) z = bottom;
    return z;
}

// The unique 1-ring nodes of `node` in the order of vertical_sigma.py:419
// _collect_neighbour_nodes_sigma (centre first, at most 64).
template <int MAXV, typename F>
__device__ inline int ring_nodes(const DMesh& m, int node, F&& fn)
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

// Containing element among the anchor's elements with its vertex weights
// (vertical_sigma.py:566 _bary_locate: seam-safe origin, R = 6371 km,
// fan-triangle test with eps -1e-6; quads weighted by Wachspress).
struct SigLoc {
    int n;
    int v[4];
    double w[4];
};

template <int MAXV>
__device__ inline bool sigma_bary_locate(const DMesh& m, double xp, double yp, int node, SigLoc& L)
{
    int k0, k1;
    n2e_range(m, node, k0, k1);
    for (int k = k0; k < k1; ++k) {
        const int e = m.n2e[k];
        if (e < 0) continue;
        const int4 q = m.elems[e];
        int nv = 0;
        int vs[4] = {0, 0, 0, 0};
#pragma unroll
        for (int v = 0; v < MAXV; ++v) {
            const int nd = v == 0 ? q.x : (v == 1 ? q.y : (v == 2 ? q.z : q.w));
            if (nd < 0) continue;
            if (nv > 0 && nd == vs[0]) continue;
            put4(vs, nv, nd);
            ++nv;
        }
        if (nv < 3) continue;
        double lo[4] = {0, 0, 0, 0}, la[4] = {0, 0, 0, 0};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (v < nv) {
                const double2 ll = m.lonlat[vs[v]];
                lo[v] = ll.x;
                la[v] = ll.y;
            }
        }
        // core/mesh.py:1974 _seam_safe_lon0
        double lmin = lo[0], lmax = lo[0], ssum = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (v < nv) {
                ssum += lo[v];
                if (lo[v] < lmin) lmin = lo[v];
                if (lo[v] > lmax) lmax = lo[v];
            }
        }
        double lon0;
        if (lmax - lmin > 180.0) {
            double d = 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (v < nv) d += km_pymod360(lo[v] - lo[0] + 180.0) - 180.0;
            lon0 = lo[0] + d / (double)nv;
        } else {
            lon0 = ssum / (double)nv;
        }
        double lat0 = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (v < nv) lat0 += la[v];
        lat0 /= (double)nv;
        const double cosl = km_cos(lat0 * kDeg2Rad);
        const double R = 6371000.0;
        double vx[4] = {0, 0, 0, 0}, vy[4] = {0, 0, 0, 0};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (v < nv) {
                const double dlon = km_pymod360(lo[v] - lon0 + 180.0) - 180.0;
                vx[v] = dlon * kDeg2Rad * R * cosl;
                vy[v] = (la[v] - lat0) * kDeg2Rad * R;
            }
        }
        const double dlon_p = km_pymod360(xp - lon0 + 180.0) - 180.0;
        const double px = dlon_p * kDeg2Rad * R * cosl;
        const double py = (yp - lat0) * kDeg2Rad * R;
        for (int t = 0; t < nv - 2; ++t) {
            const double ax = vx[0], ay = vy[0];
            const double bx = t == 0 ? vx[1] : vx[2], by = t == 0 ? vy[1] : vy[2];
            const double cx = t == 0 ? vx[2] : vx[3], cy = t == 0 ? vy[2] : vy[3];
            const double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
            if (km_fabs(denom) < 1e-14) continue;
            const double w1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / denom;
            const double w2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / denom;
            const double w3 = 1.0 - w1 - w2;
            const double eps = -1e-6;
            if (w1 >= eps && w2 >= eps && w3 >= eps) {
                L.n = nv;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    L.v[v] = vs[v];
                    L.w[v] = 0.0;
                }
                if (nv == 4) {
                    wachspress4(px, py, vx, vy, L.w);
                } else {
                    L.w[0] = w1;
                    L.w[1] = w2;
                    L.w[2] = w3;
                }
                return true;
            }
        }
    }
    L.n = 0;
    return false;
}

// Column height at the particle, interpolated like W (vertical_sigma.py:1035
// _loop_sigma_column_height; method 0 nearest, 1 leastsq, 2 barycentric,
// 3 idw).
template <int MAXV>
__device__ inline double column_height(const DMesh& m, const double* __restrict__ eta, double Dmin, int method,
                                       double xp, double yp, int node)
{
    if (method == 1) {
        // vertical_sigma.py:493 _H_leastsq_single
        double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
        double fw = 0.0, f1w = 0.0, f2w = 0.0;
        double z0 = km_cos(yp * kDeg2Rad) * kSLM;
        if (z0 < 1.0) z0 = 1.0;
        const int count = ring_nodes<MAXV>(m, node, [&](int nd) {
            const double h = col_H(m, eta, Dmin, nd);
            const double2 ll = m.lonlat[nd];
            const double dlon = km_pymod360(ll.x - xp + 180.0) - 180.0;
            const double ex = dlon * z0;
            const double ey = (ll.y - yp) * kSLM;
            a11 += 1.0;
            a12 += ex;
            a13 += ey;
            a22 += ex * ex;
            a23 += ex * ey;
            a33 += ey * ey;
            fw += h;
            f1w += h * ex;
            f2w += h * ey;
        });
        if (count < 3) return col_H(m, eta, Dmin, node);
        const double a21 = a12, a31 = a13, a32 = a23;
        const double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                            a12 * a21 * a33 - a11 * a23 * a32);
        if (km_fabs(det) < 1e-30 || km_fabs(det) < 1e-12 * km_fabs(a11 * a22 * a33))
            return col_H(m, eta, Dmin, node);
        const double inv = 1.0 / km_fabs(det);
        const double b11 = a22 * a33 - a23 * a32;
        const double b12 = a13 * a32 - a12 * a33;
        const double b13 = a12 * a23 - a13 * a22;
        double H = (b11 * fw + b12 * f1w + b13 * f2w) * inv;
        if (H < Dmin) H = Dmin;
        return H;
    }
    if (method == 2) {
        // vertical_sigma.py:636 _H_bary_single
        SigLoc L;
        if (!sigma_bary_locate<MAXV>(m, xp, yp, node, L)) return col_H(m, eta, Dmin, node);
        double H = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (v < L.n) H += L.w[v] * col_H(m, eta, Dmin, L.v[v]);
        return H;
    }
    if (method == 3) {
        // vertical_sigma.py:534 _H_idw_single (a neighbour within 1 m wins)
        double z0 = km_cos(yp * kDeg2Rad) * kSLM;
        if (z0 < 1.0) z0 = 1.0;
        double sw = 0.0, swh = 0.0, hit_h = 0.0;
        bool hit = false;
        ring_nodes<MAXV>(m, node, [&](int nd) {
            if (hit) return;
            const double h = col_H(m, eta, Dmin, nd);
            const double2 ll = m.lonlat[nd];
            const double dlon = km_pymod360(ll.x - xp + 180.0) - 180.0;
            const double dx = dlon * z0;
            const double dy = (ll.y - yp) * kSLM;
            const double d2 = dx * dx + dy * dy;
            if (d2 < 1.0) {
                hit = true;
                hit_h = h;
                return;
            }
            sw += 1.0 / d2;
            swh += h / d2;
        });
        if (hit) return hit_h;
        if (sw < 1e-30) return col_H(m, eta, Dmin, node);
        return swh / sw;
    }
    return col_H(m, eta, Dmin, node);
}

// vertical_sigma.py:297 _interp_w_sigma_linear_single on the stage field
// (the interface pair of record (k, node) blended in time): W linear in
// depth between the bracketing interfaces of a column of height H, times
// time_dir.
template <typename T>
__device__ inline double w_sig_linear(const DMesh& m, const Snap<T, true>& fw, double alpha0, double alpha, double z,
                                      int node, double H, double time_dir)
{
    const int K = m.n_levels;
    const double* s = m.zlev;
    const int64_t N = m.n_nodes;
    double wt, wb;
    if (z <= (1.0 - s[0]) * H) {
        w_pair_blend(fw, node, alpha0, alpha, wt, wb);
        return time_dir * wt;
    }
    if (z >= (1.0 - s[K - 1]) * H) {
        w_pair_blend(fw, (int64_t)(K - 2) * N + node, alpha0, alpha, wt, wb);
        return time_dir * wb;
    }
    int k = K - 2;
    for (int j = 0; j < K - 1; ++j) {
        if (z < (1.0 - s[j + 1]) * H) {
            k = j;
            break;
        }
    }
    const double z_k = (1.0 - s[k]) * H;
    const double z_k1 = (1.0 - s[k + 1]) * H;
    const double dz = z_k1 - z_k;
    w_pair_blend(fw, (int64_t)k * N + node, alpha0, alpha, wt, wb);
    if (dz < 1e-12) return time_dir * wt;
    const double a = (z - z_k) / dz;
    return time_dir * ((1.0 - a) * wt + a * wb);
}

// W of the own column at min(z, H): the fallback of every horizontal method
template <typename T>
__device__ __forceinline__ double w_sig_own(const DMesh& m, const Snap<T, true>& fw, const double* __restrict__ eta,
                                            double Dmin, double alpha0, double alpha, double z, int node,
                                            double time_dir)
{
    const double H = col_H(m, eta, Dmin, node);
    return w_sig_linear(m, fw, alpha0, alpha, z < H ? z : H, node, H, time_dir);
}

// vertical_sigma.py:971 interp_w_sigma_spatial for one particle: the
// dia-sigma flux at (xp, yp, z), every neighbour column sampled at the
// particle's sigma (z / H_p) H_j (method 0 nearest: the anchor's column at z).
template <int MAXV, typename T>
__device__ inline double w_sigma(const DMesh& m, const Snap<T, true>& fw, const double* __restrict__ eta, double Dmin,
                                 int method, double alpha0, double alpha, double xp, double yp, double z, int node,
                                 double time_dir)
{
    if (method == 0) return w_sig_linear(m, fw, alpha0, alpha, z, node, col_H(m, eta, Dmin, node), time_dir);
    if (method == 2) {
        // vertical_sigma.py:767 _interp_w_sigma_barycentric_single
        SigLoc L;
        if (!sigma_bary_locate<MAXV>(m, xp, yp, node, L))
            return w_sig_own(m, fw, eta, Dmin, alpha0, alpha, z, node, time_dir);
        double Hp = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (v < L.n) Hp += L.w[v] * col_H(m, eta, Dmin, L.v[v]);
        const double frac = sigma_frac(z, Hp);
        double acc = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (v < L.n) {
                const double Hv = col_H(m, eta, Dmin, L.v[v]);
                acc += L.w[v] * w_sig_linear(m, fw, alpha0, alpha, frac * Hv, L.v[v], Hv, time_dir);
            }
        }
        return acc;
    }
    double z0 = km_cos(yp * kDeg2Rad) * kSLM;
    if (z0 < 1.0) z0 = 1.0;
    if (method == 3) {
        // vertical_sigma.py:820 _interp_w_sigma_idw_single
        const double frac = sigma_frac(z, column_height<MAXV>(m, eta, Dmin, 3, xp, yp, node));
        double sum_w = 0.0, sum_wv = 0.0, hit_w = 0.0;
        bool hit = false;
        ring_nodes<MAXV>(m, node, [&](int nd) {
            if (hit) return;
            const double H = col_H(m, eta, Dmin, nd);
            const double wv = w_sig_linear(m, fw, alpha0, alpha, frac * H, nd, H, time_dir);
            const double2 ll = m.lonlat[nd];
            const double dlon = km_pymod360(ll.x - xp + 180.0) - 180.0;
            const double dx = dlon * z0;
            const double dy = (ll.y - yp) * kSLM;
            const double d2 = dx * dx + dy * dy;
            if (d2 < 1.0) {
                hit = true;
                hit_w = wv;
                return;
            }
            const double weight = 1.0 / d2;
            sum_w += weight;
            sum_wv += weight * wv;
        });
        if (hit) return hit_w;
        if (sum_w < 1e-30) return w_sig_own(m, fw, eta, Dmin, alpha0, alpha, z, node, time_dir);
        return sum_wv / sum_w;
    }
    // vertical_sigma.py:656 _interp_w_sigma_leastsq_single
    const double frac = sigma_frac(z, column_height<MAXV>(m, eta, Dmin, 1, xp, yp, node));
    double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
    double fw0 = 0.0, f1w = 0.0, f2w = 0.0;
    const int count = ring_nodes<MAXV>(m, node, [&](int nd) {
        const double H = col_H(m, eta, Dmin, nd);
        const double wv = w_sig_linear(m, fw, alpha0, alpha, frac * H, nd, H, time_dir);
        const double2 ll = m.lonlat[nd];
        const double dlon = km_pymod360(ll.x - xp + 180.0) - 180.0;
        const double ex = dlon * z0;
        const double ey = (ll.y - yp) * kSLM;
        a11 += 1.0;
        a12 += ex;
        a13 += ey;
        a22 += ex * ex;
        a23 += ex * ey;
        a33 += ey * ey;
        fw0 += wv;
        f1w += wv * ex;
        f2w += wv * ey;
    });
    if (count < 3) return w_sig_own(m, fw, eta, Dmin, alpha0, alpha, z, node, time_dir);
    const double a21 = a12, a31 = a13, a32 = a23;
    const double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 - a12 * a21 * a33 -
                        a11 * a23 * a32);
    if (km_fabs(det) < 1e-30 || km_fabs(det) < 1e-12 * km_fabs(a11 * a22 * a33))
        return w_sig_own(m, fw, eta, Dmin, alpha0, alpha, z, node, time_dir);
    const double inv = 1.0 / km_fabs(det);
    const double b11 = a22 * a33 - a23 * a32;
    const double b12 = a13 * a32 - a12 * a33;
    const double b13 = a12 * a23 - a13 * a22;
    return (b11 * fw0 + b12 * f1w + b13 * f2w) * inv;
}

}  // namespace kelp
