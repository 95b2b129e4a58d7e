// kelp_device.cuh -- device-side building blocks of the particle step.
//
// One HIP thread owns one particle for a whole step (a wavefront = 64
// particles).  Everything here is __device__ code over the HBM-resident
// mesh layout built by the library at upload time:
//
//   lonlat[N]      double2 (lon, lat) per node: one 16-byte load per vertex
//   gh_xyz[N]      double4 (x, y, z, element-list range) in grid-hash CSR order, so
//                  a row segment of the hash is one contiguous stream
//   gh_offsets     int32 CSR offsets over cells (cj * nx + ci)
//   elems[E]       int4 vertex ids (-1 padded): one 16-byte load per element
//   n2e_off/n2e    node -> element adjacency in CSR form (same order as the
//                  reference's dense -1-padded table)
//
// The floating-point expressions follow the reference's Numba kernels
// operation for operation (same operand order, no contraction, Python `%`
// via km_pymod360, transcendental functions from kmath.h) so the results
// equal the CPU restatement built with the same math bit for bit.  The
// reference file:line each function follows is cited beside it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmath.h"

namespace kelp {

constexpr double kDeg2Rad = 0.017453292519943295;  // np.pi/180, core/geometry.py:31
constexpr double kRad2Deg = 57.29577951308232;     // 180/np.pi, core/geometry.py:32
constexpr double kSLM = 111120.0;                  // METERS_PER_DEG_LAT, core/geometry.py:28
constexpr double kREarthStd = 6371000.0;           // R_EARTH_STANDARD, core/geometry.py:30
constexpr double kBaryEps = -1e-6;                 // _BARY_EPS, core/mesh.py:33
constexpr double kAreaTol = 1e-10;                 // _AREA_TOL, core/mesh.py:32

// MI355X: 8 XCDs, each with its own L2.  The dispatcher hands workgroup b to
// XCD b % 8, so consecutive workgroups -- neighbouring particles once they
// are sorted by grid-hash cell -- would land in eight different L2s.
constexpr int kXcds = 8;

// Global thread index with the workgroups regrouped by XCD: the blocks XCD x
// receives (x, x + 8, x + 16, ...) take the x-th contiguous run of the
// block range, so each L2 serves one stretch of sorted particles.  A
// bijection of [0, gridDim.x) onto itself for any grid size.
__device__ __forceinline__ int64_t xcd_thread_index()
{
    const int64_t nb = gridDim.x, b = blockIdx.x;
    const int64_t per = nb / kXcds, rem = nb % kXcds, x = b % kXcds;
    const int64_t lb = x * per + (x < rem ? x : rem) + b / kXcds;
    return lb * (int64_t)blockDim.x + threadIdx.x;
}
constexpr int kShortRowCells = 32;                 // _GH_EXACT_SHORT_ROW_CELLS, core/mesh.py:1608

constexpr int32_t kAlive = 1;
constexpr int32_t kBeached = 2;
constexpr int32_t kOutside = 3;
constexpr int8_t kFrInterp = 1;
constexpr int8_t kFrOOB = 2;

struct DMesh {
    const double2* __restrict__ lonlat;
    const double4* __restrict__ gh_xyz;
    const int32_t* __restrict__ gh_offsets;
    const int4* __restrict__ elems;
    const int32_t* __restrict__ n2e_off;
    const int32_t* __restrict__ n2e;
    const double* __restrict__ depth;     // node bottom depth (m, positive down)
    const double* __restrict__ zlev;      // K interface depths (3-D)
    const uint8_t* __restrict__ coastal;  // (E,) coastal-element flags (beaching)
    const double2* __restrict__ ecent;    // (E,) element centres (lon, lat); element-layout u/v
    const double* __restrict__ coslat;    // (N,) km_cos(lat * kDeg2Rad) per node
    const double4* __restrict__ eframe;   // (4E,) containment frame per element, one 128-byte line each
    const uint8_t* __restrict__ eflags;   // (E,) bit 0: the interpolation ring is the containment ring
    const int4* __restrict__ enbr;        // (E,) neighbour across edge k of the compacted ring, -1 boundary
                                          //   (core/mesh.py:2824 build_elem_neighbors); null without it
    double cell, lon_min, lat_min, r_earth;
    int32_t nx, ny, wrap, far_rows;
    int32_t n_nodes, n_elems, max_v, n_levels;
    int32_t tile_on;  // 1 only in launches whose kernel stages a GhTile in its dynamic LDS first
};

__device__ __forceinline__ double sqd(double x) { return x * x; }

// Node nn's element list n2e[k0 .. k1), from the w of its grid-hash entry
// (internal node id = CSR slot, so gh_xyz[nn] is the node's entry): the
// line the anchor search has just read, not a separate offsets array.
__device__ __forceinline__ void n2e_range(const DMesh& m, int nn, int& k0, int& k1)
{
    const double w = m.gh_xyz[nn].w;
    k0 = __double2loint(w);
    k1 = __double2hiint(w);
}

// core/mesh.py:1193 _gh_fold_dlon
__device__ inline double gh_fold_dlon(const DMesh& m, double lon_p)
{
    double dlon = km_pymod360(lon_p - m.lon_min);
    if (!m.wrap) {
        double grid_span = m.nx * m.cell;
        if (dlon > grid_span && (dlon - grid_span) > (360.0 - dlon)) dlon -= 360.0;
    }
    return dlon;
}

// core/mesh.py:1211 _gh_cell_of
__device__ inline void gh_cell_of(const DMesh& m, double lon_p, double lat_p, int& ci, int& cj)
{
    double dlon = gh_fold_dlon(m, lon_p);
    ci = (int)(dlon / m.cell);
    cj = (int)((lat_p - m.lat_min) / m.cell);
    if (ci < 0) ci = 0;
    if (ci >= m.nx) ci = m.nx - 1;
    if (cj < 0) cj = 0;
    if (cj >= m.ny) cj = m.ny - 1;
}

struct Query {
    double x, y, z, cos_lat;
};

// core/mesh.py:1281-1286 (query unit vector)
__device__ inline Query make_query(double lon_p, double lat_p)
{
    double lat_r = lat_p * kDeg2Rad;
    double lon_r = lon_p * kDeg2Rad;
    double sl, cl, so, co;
    km_sincos(lat_r, &sl, &cl);
    km_sincos(lon_r, &so, &co);
    Query q;
    q.cos_lat = cl;
    q.x = cl * co;
    q.y = cl * so;
    q.z = sl;
    return q;
}

// ---------------------------------------------------------------------------
// The workgroup's grid-hash tile in LDS.
//
// The particles of one workgroup are neighbours once they are sorted by hash
// cell, so the 3x3 windows that all their anchor searches scan (every RK
// site, the kick, the re-anchoring after a fold) fall inside one small box of
// cells.  A kernel launched with m.tile_on stages that box -- the cells
// around its alive particles' start cells, two cells of margin -- in its
// dynamic LDS once per launch: the CSR offsets of the box's cells and the
// unit vectors of their nodes.  A scan whose row slice lies inside the box
// reads LDS; any other (a particle that left the box, a box too large to
// stage, a window that wraps) reads HBM as before.  Same values, same
// order, same strict `<`: the answer is the same node either way.
// ---------------------------------------------------------------------------

constexpr int kTileRows = 8;       // cell rows a tile can hold
constexpr int kTileCols = 64;      // cells per row
constexpr int kTileNodes = 1536;   // nodes (24 B each in LDS)

struct GhTile {
    int32_t ok, ci0, cj0, nci, ncj;
    int32_t lo_ci, hi_ci, lo_cj, hi_cj;
    int32_t delta[kTileRows];               // LDS slot - CSR slot, per row
    int32_t off[kTileRows][kTileCols + 1];  // gh_offsets of the box's cells (+ the row end)
    double x[kTileNodes], y[kTileNodes], z[kTileNodes];
};

extern __shared__ __align__(16) unsigned char kelp_dyn_lds[];

__device__ __forceinline__ GhTile& gh_tile() { return *reinterpret_cast<GhTile*>(kelp_dyn_lds); }


__device__ __forceinline__ int wave_min_i32(int v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const int o = __shfl_xor(v, s, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const int o = __shfl_xor(v, s, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Stage the tile of this workgroup.  Every thread of the workgroup calls it
// (it synchronises the workgroup); `alive` lanes contribute their position's
// cell to the box.  Launch with sizeof(GhTile) bytes of dynamic LDS.
__device__ inline void stage_gh_tile(const DMesh& m, bool alive, double lon_p, double lat_p)
{
    GhTile& t = gh_tile();
    const int tid = threadIdx.x, nthr = blockDim.x;
    int ci = 0x7fffffff, cj = 0x7fffffff, xi = -1, xj = -1;
    if (alive) {
        gh_cell_of(m, lon_p, lat_p, ci, cj);
        xi = ci;
        xj = cj;
    }
    if (tid == 0) {
        t.lo_ci = 0x7fffffff;
        t.lo_cj = 0x7fffffff;
        t.hi_ci = -1;
        t.hi_cj = -1;
    }
    __syncthreads();
    ci = wave_min_i32(ci);
    cj = wave_min_i32(cj);
    xi = wave_max_i32(xi);
    xj = wave_max_i32(xj);
    if ((tid & 63) == 0) {
        atomicMin(&t.lo_ci, ci);
        atomicMin(&t.lo_cj, cj);
        atomicMax(&t.hi_ci, xi);
        atomicMax(&t.hi_cj, xj);
    }
    __syncthreads();
    if (tid == 0) {
        int ok = t.hi_ci >= 0;
        const int a0 = t.lo_ci - 2 > 0 ? t.lo_ci - 2 : 0;
        const int a1 = t.hi_ci + 2 < m.nx - 1 ? t.hi_ci + 2 : m.nx - 1;
        const int b0 = t.lo_cj - 2 > 0 ? t.lo_cj - 2 : 0;
        const int b1 = t.hi_cj + 2 < m.ny - 1 ? t.hi_cj + 2 : m.ny - 1;
        ok = ok && a1 - a0 + 1 <= kTileCols && b1 - b0 + 1 <= kTileRows;
        t.ci0 = a0;
        t.cj0 = b0;
        t.nci = ok ? a1 - a0 + 1 : 0;
        t.ncj = ok ? b1 - b0 + 1 : 0;
        t.ok = ok;
    }
    __syncthreads();
    const int nci = t.nci, ncj = t.ncj;
    // offsets of cells ci0 .. ci0 + nci (the last one ends the row's slice);
    // ci0 + nci <= nx, so the index stays within the nx * ny + 1 offsets
    for (int e = tid; e < ncj * (nci + 1); e += nthr) {
        const int r = e / (nci + 1), cc = e - r * (nci + 1);
        t.off[r][cc] = m.gh_offsets[(int64_t)(t.cj0 + r) * m.nx + t.ci0 + cc];
    }
    __syncthreads();
    if (tid == 0 && t.ok) {
        int l = 0;
        for (int r = 0; r < ncj; ++r) {
            t.delta[r] = l - t.off[r][0];
            l += t.off[r][nci] - t.off[r][0];
        }
        if (l > kTileNodes) t.ok = 0;
    }
    __syncthreads();
    if (t.ok) {
        for (int r = 0; r < ncj; ++r) {
            const int k0 = t.off[r][0], k1 = t.off[r][nci], d = t.delta[r];
            for (int k = k0 + tid; k < k1; k += nthr) {
                const double4 p = m.gh_xyz[k];
                t.x[k + d] = p.x;
                t.y[k + d] = p.y;
                t.z[k + d] = p.z;
            }
        }
    }
    __syncthreads();
}

// The same scan with eight loads in flight, for the exact search's
// widening: its rectangles can span whole rows, and one load at a time
// makes a lane that widens far a long serial tail of its wavefront.  Same
// order and comparisons as scan_range.
__device__ inline void scan_range_wide(const DMesh& m, int cj2, int c_lo, int c_hi, const Query& q, int& best,
                                       double& best_d2)
{
    const int64_t row = (int64_t)cj2 * m.nx;
    const int32_t start = m.gh_offsets[row + c_lo];
    const int32_t end = m.gh_offsets[row + c_hi + 1];
    int32_t k = start;
    for (; k + 8 <= end; k += 8) {
        double4 p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = m.gh_xyz[k + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            double d2 = sqd(p[u].x - q.x) + sqd(p[u].y - q.y) + sqd(p[u].z - q.z);
            if (d2 < best_d2) {
                best_d2 = d2;
                best = k + u;
            }
        }
    }
    for (; k < end; ++k) {
        double4 p = m.gh_xyz[k];
        double d2 = sqd(p.x - q.x) + sqd(p.y - q.y) + sqd(p.z - q.z);
        if (d2 < best_d2) {
            best_d2 = d2;
            best = k;
        }
    }
}

// core/mesh.py:1570 _gh_scan_row_range: cells c_lo..c_hi of row cj2 are one
// contiguous CSR slice; strict `<` keeps the first node found on a tie.
__device__ inline void scan_range(const DMesh& m, int cj2, int c_lo, int c_hi, const Query& q,
                                  int& best, double& best_d2)
{
    if (m.tile_on) {
        const GhTile& t = gh_tile();
        const int r = cj2 - t.cj0, a = c_lo - t.ci0, b = c_hi - t.ci0;
        if (t.ok && r >= 0 && r < t.ncj && a >= 0 && b < t.nci) {
            const int32_t start = t.off[r][a], end = t.off[r][b + 1], d = t.delta[r];
            for (int32_t k = start; k < end; ++k) {
                double d2 = sqd(t.x[k + d] - q.x) + sqd(t.y[k + d] - q.y) + sqd(t.z[k + d] - q.z);
                if (d2 < best_d2) {
                    best_d2 = d2;
                    best = k;
                }
            }
            return;
        }
    }
    const int64_t row = (int64_t)cj2 * m.nx;
    int32_t start = m.gh_offsets[row + c_lo];
    int32_t end = m.gh_offsets[row + c_hi + 1];
    for (int32_t k = start; k < end; ++k) {
        double4 p = m.gh_xyz[k];
        double d2 = sqd(p.x - q.x) + sqd(p.y - q.y) + sqd(p.z - q.z);
        if (d2 < best_d2) {
            best_d2 = d2;
            best = k;  // the internal node id is the CSR slot
        }
    }
}

// core/mesh.py:1588 _gh_scan_row_window
__device__ inline void scan_row_window(const DMesh& m, int cj2, int ci, int rx, const Query& q,
                                       int& best, double& best_d2)
{
    const int nx = m.nx;
    if (m.wrap && 2 * rx + 1 >= nx) {
        scan_range(m, cj2, 0, nx - 1, q, best, best_d2);
        return;
    }
    int c_lo = ci - rx, c_hi = ci + rx;
    if (m.wrap) {
        if (c_lo < 0) {
            scan_range(m, cj2, c_lo + nx, nx - 1, q, best, best_d2);
            c_lo = 0;
        } else if (c_hi > nx - 1) {
            scan_range(m, cj2, 0, c_hi - nx, q, best, best_d2);
            c_hi = nx - 1;
        }
    } else {
        if (c_lo < 0) c_lo = 0;
        if (c_hi > nx - 1) c_hi = nx - 1;
    }
    scan_range(m, cj2, c_lo, c_hi, q, best, best_d2);
}

// scan_row_window over scan_range_wide
__device__ inline void scan_row_window_wide(const DMesh& m, int cj2, int ci, int rx, const Query& q,
                                       int& best, double& best_d2)
{
    const int nx = m.nx;
    if (m.wrap && 2 * rx + 1 >= nx) {
        scan_range_wide(m, cj2, 0, nx - 1, q, best, best_d2);
        return;
    }
    int c_lo = ci - rx, c_hi = ci + rx;
    if (m.wrap) {
        if (c_lo < 0) {
            scan_range_wide(m, cj2, c_lo + nx, nx - 1, q, best, best_d2);
            c_lo = 0;
        } else if (c_hi > nx - 1) {
            scan_range_wide(m, cj2, 0, c_hi - nx, q, best, best_d2);
            c_hi = nx - 1;
        }
    } else {
        if (c_lo < 0) c_lo = 0;
        if (c_hi > nx - 1) c_hi = nx - 1;
    }
    scan_range_wide(m, cj2, c_lo, c_hi, q, best, best_d2);
}

// core/mesh.py:1236 _grid_hash_nearest_3x3 (best_d2 returned for the exact
// search, which re-derives the same value from the table)
__device__ inline int nearest_3x3(const DMesh& m, int ci, int cj, const Query& q, double& best_d2)
{
    int best = -1;
    best_d2 = 1e30;
    const int nx = m.nx, ny = m.ny;
    for (int dj = -1; dj < 2; ++dj) {
        int cj2 = cj + dj;
        if (cj2 < 0 || cj2 >= ny) continue;
        // the three cells of a row are one CSR slice unless the window is
        // clipped or wraps; scanning cell by cell keeps the reference order
        // either way (cells ascend within the row)
        int lo = ci - 1, hi = ci + 1;
        if (lo >= 0 && hi < nx) {
            scan_range(m, cj2, lo, hi, q, best, best_d2);
            continue;
        }
        for (int di = -1; di < 2; ++di) {
            int ci2 = ci + di;
            if (ci2 < 0)
                ci2 = m.wrap ? ci2 + nx : -1;
            else if (ci2 >= nx)
                ci2 = m.wrap ? ci2 - nx : -1;
            if (ci2 < 0) continue;
            scan_range(m, cj2, ci2, ci2, q, best, best_d2);
        }
    }
    return best;
}

// core/mesh.py:1307 _cap_extents, for the query q and a mesh node.  The
// sines and cosines of both latitudes come from tables instead of being
// recomputed: the query's are those make_query took of the same argument,
// the node's sine is the z of its unit vector and its cosine the node's
// coslat entry, all from km_sincos(lat * kDeg2Rad) -- the same bits.
__device__ inline void cap_extents(const DMesh& m, const Query& q, double lon_p, double lat_p, int node,
                                   double& theta_deg, double& ext_lon)
{
    const double lon_n = m.lonlat[node].x;
    const double sp = q.z, cp = q.cos_lat;
    const double sn = m.gh_xyz[node].z, cn = m.coslat[node];
    double cosd = (sp * sn + cp * cn * km_cos((lon_p - lon_n) * kDeg2Rad));
    if (cosd > 1.0) cosd = 1.0;
    if (cosd < -1.0) cosd = -1.0;
    double theta = km_acos(cosd);
    if (theta <= 0.0) {
        theta_deg = 0.0;
        ext_lon = 0.0;
        return;
    }
    theta_deg = theta * kRad2Deg;
    if (km_fabs(lat_p) + theta_deg >= 90.0) {
        ext_lon = 1.0e30;
        return;
    }
    double r = km_sin(theta) / cp;
    if (r >= 1.0) {
        ext_lon = 1.0e30;
        return;
    }
    ext_lon = km_asin(r) * kRad2Deg;
}

__device__ __forceinline__ int pyimod(int a, int n)
{
    int r = a % n;
    return r < 0 ? r + n : r;
}

// core/mesh.py:1476 _anchor_provable_rect with the node's cap (theta_deg,
// ext_lon) from cap_extents
__device__ inline int anchor_provable_rect(const DMesh& m, double lon_p, double lat_p, double theta_deg,
                                           double ext_lon, int rx, int ry)
{
    if (theta_deg <= 0.0) return 0;
    const double cell = m.cell, lon_min = m.lon_min, lat_min = m.lat_min;
    const int nx = m.nx, ny = m.ny;
    double dlon = gh_fold_dlon(m, lon_p);
    double lon_q = lon_min + dlon;
    int ci_raw = (int)floor(dlon / cell);
    int cj_raw = (int)floor((lat_p - lat_min) / cell);
    const double INF = 1.0e30;
    int cj = cj_raw;
    if (cj < 0) cj = 0;
    if (cj > ny - 1) cj = ny - 1;
    double cov_dn = (cj - ry >= 0) ? lat_p - (lat_min + (cj - ry) * cell) : INF;
    double cov_up = (cj + ry <= ny - 1) ? (lat_min + (cj + ry + 1) * cell) - lat_p : INF;
    int code = 0;
    if (theta_deg > cov_dn || theta_deg > cov_up) code |= 1;
    if (m.wrap) {
        if (2 * rx + 1 >= nx) return code;
        if (ext_lon >= INF) return code | 4;
        double w_last = 360.0 - (nx - 1) * cell;
        int ci = pyimod(ci_raw, nx);
        double left = lon_min + (ci - rx) * cell;
        if (ci - rx < 0) left += cell - w_last;
        double right = lon_min + (ci + rx + 1) * cell;
        if (ci + rx >= nx - 1) right -= cell - w_last;
        if (ext_lon > lon_q - left || ext_lon > right - lon_q) code |= 2;
        return code;
    }
    int ci = ci_raw;
    if (ci < 0) ci = 0;
    if (ci > nx - 1) ci = nx - 1;
    double cov_l = (ci - rx >= 0) ? lon_q - (lon_min + (ci - rx) * cell) : INF;
    double cov_r = (ci + rx <= nx - 1) ? (lon_min + (ci + rx + 1) * cell) - lon_q : INF;
    if (cov_l >= INF && cov_r >= INF) return code;
    if (ext_lon >= INF) return code | 4;
    if (ext_lon > cov_l || ext_lon > cov_r) code |= 2;
    double n_lo = lon_min + cell;
    double n_hi = lon_min + (nx - 1) * cell;
    for (int s = 0; s < 2; ++s) {
        double shift = s == 0 ? -360.0 : 360.0;
        double a = lon_q - ext_lon + shift;
        double b = lon_q + ext_lon + shift;
        if (b >= n_lo && a <= n_hi) return code | 4;
    }
    return code;
}

// core/mesh.py:1365 _grid_hash_nearest_widened (ring widening after an
// empty 3x3 window)
__device__ inline int widen_rings(const DMesh& m, int ci, int cj, const Query& q)
{
    int best = -1;
    double best_d2 = 1e30;
    const int nx = m.nx, ny = m.ny;
    int max_ring = nx > ny ? nx : ny;
    int found_ring = -1;
    for (int r = 2; r <= max_ring; ++r) {
        for (int dj = -r; dj < r + 1; ++dj) {
            int cj2 = cj + dj;
            if (cj2 < 0 || cj2 >= ny) continue;
            int di_step = (dj == -r || dj == r) ? 1 : 2 * r;
            for (int di = -r; di <= r; di += di_step) {
                int ci2 = ci + di;
                if (m.wrap)
                    ci2 = pyimod(ci2, nx);
                else if (ci2 < 0 || ci2 >= nx)
                    continue;
                scan_range(m, cj2, ci2, ci2, q, best, best_d2);
            }
        }
        if (best >= 0) {
            if (found_ring < 0)
                found_ring = r;
            else if (r > found_ring)
                break;
        }
    }
    return best;
}

__device__ inline int nearest_widened(const DMesh& m, double lon_p, double lat_p)
{
    int ci, cj;
    gh_cell_of(m, lon_p, lat_p, ci, cj);
    Query q = make_query(lon_p, lat_p);
    double d2;
    int best = nearest_3x3(m, ci, cj, q, d2);
    if (best >= 0) return best;
    return widen_rings(m, ci, cj, q);
}

// core/mesh.py:1612 _grid_hash_nearest_exact: the 3x3 window, accepted when
// the certificate proves it; else the rectangle grown by the certificate's
// deficits.  The common case (3x3 proven) returns after one window scan.
// The rare paths (this widening, the 2-ring pass, the edge walk) are out of
// line and marked cold: the compiler then keeps the step's live values in
// registers on the common path and saves them around those calls only.
__device__ __noinline__ __attribute__((cold)) int exact_widen(const DMesh& m, double lon_p, double lat_p, int best,
                                        double best_d2, int code, int ci, int cj, const Query& q);

__device__ inline int nearest_exact(const DMesh& m, double lon_p, double lat_p)
{
    const double cell = m.cell;
    const double far = m.far_rows * cell;
    if (lat_p < m.lat_min - far || lat_p > m.lat_min + m.ny * cell + far)
        return nearest_widened(m, lon_p, lat_p);
    int ci, cj;
    gh_cell_of(m, lon_p, lat_p, ci, cj);
    Query q = make_query(lon_p, lat_p);
    double best_d2;
    int best = nearest_3x3(m, ci, cj, q, best_d2);
    int code = 0;
    if (best >= 0) {
        double theta_deg, ext;
        cap_extents(m, q, lon_p, lat_p, best, theta_deg, ext);
        code = anchor_provable_rect(m, lon_p, lat_p, theta_deg, ext, 1, 1);
        if (code == 0) return best;
        if (theta_deg > far) return best;
    }
    return exact_widen(m, lon_p, lat_p, best, best_d2, code, ci, cj, q);
}

__device__ __noinline__ __attribute__((cold)) int exact_widen(const DMesh& m, double lon_p, double lat_p, int best,
                                        double best_d2, int code, int ci, int cj, const Query& q)
{
    const double cell = m.cell;
    const int nx = m.nx, ny = m.ny, wrap = m.wrap;
    const double far = m.far_rows * cell;
    if (best < 0) best_d2 = 1.0e30;
    int rx = 1, ry = 1;
    int short_row = 360.0 * q.cos_lat <= kShortRowCells * cell;
    int rx_max;
    if (wrap) {
        rx_max = nx / 2;
        if (rx_max < 1) rx_max = 1;
    } else {
        rx_max = ci > nx - 1 - ci ? ci : nx - 1 - ci;
    }
    int ry_max = cj > ny - 1 - cj ? cj : ny - 1 - cj;
    double abs_lat = lat_p >= 0.0 ? lat_p : -lat_p;
    double dlon_q = gh_fold_dlon(m, lon_p);
    const double INF = 1.0e30;
    for (;;) {
        if (best >= 0 && code == 0) {
            double theta_deg, ext;
            cap_extents(m, q, lon_p, lat_p, best, theta_deg, ext);
            code = anchor_provable_rect(m, lon_p, lat_p, theta_deg, ext, rx, ry);
            if (code == 0) return best;
        }
        if (best < 0 || 2.0 * km_asin(0.5 * km_sqrt(best_d2)) * kRad2Deg > far) {
            int cols_full = wrap ? (2 * rx + 1 >= nx) : (ci - rx <= 0 && ci + rx >= nx - 1);
            double clear = (ry >= ry_max) ? INF : ry * cell;
            if (!cols_full) {
                double lat_edge = abs_lat + (ry + 1) * cell;
                double clear_lon;
                if (lat_edge >= 90.0) {
                    clear_lon = 0.0;
                } else {
                    double dl = rx * cell;
                    if (!wrap) {
                        double x = dlon_q > 0.0 ? dlon_q : 0.0;
                        double w = nx * cell;
                        if (x > w) x = w;
                        double near_edge = x < w - x ? x : w - x;
                        double dl_cut = 360.0 - w + near_edge;
                        if (dl_cut < dl) dl = dl_cut;
                    }
                    if (dl > 180.0) dl = 180.0;
                    clear_lon = km_asin(km_sin(dl * kDeg2Rad) * km_cos(lat_edge * kDeg2Rad)) * kRad2Deg;
                }
                if (clear_lon < clear) clear = clear_lon;
            }
            if (clear >= far) {
                if (best >= 0) return best;
                return nearest_widened(m, lon_p, lat_p);
            }
        }
        int grow_x, grow_y, all_x;
        if (best >= 0) {
            grow_x = (code & 2) != 0;
            grow_y = (code & 1) != 0;
            all_x = (code & 4) != 0;
            code = 0;
        } else if (short_row) {
            grow_x = 0;
            grow_y = 1;
            all_x = 1;
        } else {
            grow_x = 1;
            grow_y = 1;
            all_x = 0;
        }
        if (rx >= rx_max) {
            grow_x = 0;
            all_x = 0;
        }
        if (ry >= ry_max) grow_y = 0;
        if (!(grow_x || grow_y || all_x)) {
            if (best >= 0) return best;
            return nearest_widened(m, lon_p, lat_p);
        }
        int rx_new = all_x ? rx_max : (grow_x ? rx + 1 : rx);
        int ry_new = grow_y ? ry + 1 : ry;
        int j_lo = cj - ry > 0 ? cj - ry : 0;
        int j_hi = cj + ry < ny - 1 ? cj + ry : ny - 1;
        if (rx_new > rx) {
            if (wrap && 2 * rx_new + 1 >= nx && rx_new - rx > 1) {
                for (int cj2 = j_lo; cj2 < j_hi + 1; ++cj2) {
                    int c_lo = pyimod(ci + rx + 1, nx);
                    int c_hi = pyimod(ci - rx - 1, nx);
                    if (c_lo <= c_hi) {
                        scan_range_wide(m, cj2, c_lo, c_hi, q, best, best_d2);
                    } else {
                        scan_range_wide(m, cj2, c_lo, nx - 1, q, best, best_d2);
                        scan_range_wide(m, cj2, 0, c_hi, q, best, best_d2);
                    }
                }
            } else if (!wrap && rx_new - rx > 1) {
                for (int cj2 = j_lo; cj2 < j_hi + 1; ++cj2) {
                    if (ci - rx - 1 >= 0) {
                        int c_lo = ci - rx_new > 0 ? ci - rx_new : 0;
                        scan_range_wide(m, cj2, c_lo, ci - rx - 1, q, best, best_d2);
                    }
                    if (ci + rx + 1 <= nx - 1) {
                        int c_hi = ci + rx_new < nx - 1 ? ci + rx_new : nx - 1;
                        scan_range_wide(m, cj2, ci + rx + 1, c_hi, q, best, best_d2);
                    }
                }
            } else {
                for (int r = rx + 1; r < rx_new + 1; ++r) {
                    if (wrap) {
                        int c_left = pyimod(ci - r, nx);
                        int two = (2 * r + 1 <= nx);
                        int c_right = pyimod(ci + r, nx);
                        for (int cj2 = j_lo; cj2 < j_hi + 1; ++cj2) {
                            scan_range_wide(m, cj2, c_left, c_left, q, best, best_d2);
                            if (two) scan_range_wide(m, cj2, c_right, c_right, q, best, best_d2);
                        }
                    } else {
                        for (int s = 0; s < 2; ++s) {
                            int c2 = s == 0 ? ci - r : ci + r;
                            if (c2 < 0 || c2 > nx - 1) continue;
                            for (int cj2 = j_lo; cj2 < j_hi + 1; ++cj2)
                                scan_range_wide(m, cj2, c2, c2, q, best, best_d2);
                        }
                    }
                }
            }
        }
        for (int r = ry + 1; r < ry_new + 1; ++r) {
            for (int s = 0; s < 2; ++s) {
                int cj2 = s == 0 ? cj - r : cj + r;
                if (cj2 < 0 || cj2 > ny - 1) continue;
                scan_row_window_wide(m, cj2, ci, rx_new, q, best, best_d2);
            }
        }
        rx = rx_new;
        ry = ry_new;
    }
}

// Anchor of one evaluation site.  mode 1: the provably exact search
// (mesh.exact_nearest); 0: the fused kernels' guard-off widened window
// (integrators.py:1471 _fused_site_anchor); 2: the batch guard-off anchor of
// core/mesh.py:1896 find_nearest_nodes (3x3 window, an empty window healed
// with the true nearest node).
__device__ inline int site_anchor(const DMesh& m, double lon_p, double lat_p, int mode)
{
    if (mode == 1) return nearest_exact(m, lon_p, lat_p);
    if (mode == 2) {
        int ci, cj;
        gh_cell_of(m, lon_p, lat_p, ci, cj);
        Query q = make_query(lon_p, lat_p);
        double d2;
        int b = nearest_3x3(m, ci, cj, q, d2);
        return b >= 0 ? b : nearest_exact(m, lon_p, lat_p);
    }
    return nearest_widened(m, lon_p, lat_p);
}

// ---------------------------------------------------------------------------
// Containment
//
// MAXV is the element width of the mesh (3: triangles, 4: quads and mixed).
// Vertex lists are compacted with predicated, fully unrolled loops so the
// per-lane arrays stay in registers.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int elem_slot(int4 e, int k)
{
    return k == 0 ? e.x : (k == 1 ? e.y : (k == 2 ? e.z : e.w));
}

// set v[q] = x for the q == idx (idx is a runtime value < 4)
__device__ __forceinline__ void put4(int (&v)[4], int idx, int x)
{
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q == idx) v[q] = x;
}

// Valid vertices of element c for the containment test: -1 slots and the
// FESOM-C repeated-first-vertex slot 3 dropped (core/mesh.py:2221-2238).
template <int MAXV>
__device__ __forceinline__ int contain_ring(const DMesh& m, int c, int (&v)[4])
{
    int4 en = m.elems[c];
    int n = 0;
#pragma unroll
    for (int j = 0; j < MAXV; ++j) {
        int nd = elem_slot(en, j);
        bool keep = nd >= 0;
        if (j == 3 && en.w == en.x && en.x >= 0) keep = false;
        if (keep) {
            put4(v, n, nd);
            ++n;
        }
    }
    return n;
}

// Compacted vertex ring for interpolation: -1 slots and any repeat of the
// first vertex dropped (interpolate.py:1376-1386, vertical.py:634-644).
template <int MAXV>
__device__ __forceinline__ int interp_ring(int4 en, int (&v)[4])
{
    int n = 0;
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        int nd = elem_slot(en, k);
        bool keep = nd >= 0 && !(n > 0 && nd == v[0]);
        if (keep) {
            put4(v, n, nd);
            ++n;
        }
    }
    return n;
}

// core/mesh.py:2055 _barycentric_weights_tri
__device__ inline bool bary_tri(double px, double py, double ax, double ay, double bx, double by,
                                double cx, double cy, double& w1, double& w2, double& w3)
{
    double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
    if (km_fabs(denom) < kAreaTol) {
        w1 = w2 = w3 = -1.0;
        return false;
    }
    w1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / denom;
    w2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / denom;
    w3 = 1.0 - w1 - w2;
    return true;
}

__device__ __forceinline__ bool in_tri(double px, double py, double ax, double ay, double bx, double by,
                                       double cx, double cy)
{
    double w1, w2, w3;
    bary_tri(px, py, ax, ay, bx, by, cx, cy, w1, w2, w3);
    return w1 >= kBaryEps && w2 >= kBaryEps && w3 >= kBaryEps;
}

// The particle-independent part of core/mesh.py:2211 _element_contains for
// element c, computed once at mesh upload: the containment ring's first
// longitude lo0, the bbox of the wrapped longitude deltas (dmin, dmax) and
// of the latitudes (lamin, lamax), the projection origin of :2007
// _project_to_local_xy (lon0 seam-gated, lat0, cos(lat0)) and the projected
// vertices.  Stored in one 128-byte line per element: eframe[4c] = (lo0,
// dmin, dmax, lamin), eframe[4c+1] = (lamax, lon0, lat0, cos_lat0), then the
// projected vertices (x_k, y_k) as double2 pairs (frame_xy).  A ring of fewer
// than three vertices gets an empty latitude box (never contains); a
// triangle in a four-wide mesh gets a NaN fourth vertex (its second fan
// triangle never contains).  Every value is the expression the per-query
// test evaluated, so the test's answers are unchanged bit for bit.
// the projected vertices of element c's frame: the second half of its line
__device__ __forceinline__ const double2* frame_xy(const DMesh& m, int64_t c)
{
    return reinterpret_cast<const double2*>(m.eframe + 4 * c + 2);
}

// Element e's vertex ids (elems[e]).  On a triangle mesh the frame's fourth
// vertex slot is free and holds them, so the interpolation after a
// containment test reads the line that test has just brought in instead of
// a second one.
template <int MAXV>
__device__ __forceinline__ int4 elem_nodes(const DMesh& m, int64_t e)
{
    if constexpr (MAXV == 3)
        return *reinterpret_cast<const int4*>(reinterpret_cast<const char*>(m.eframe + 4 * e + 3) + 16);
    else
        return m.elems[e];
}

template <int MAXV>
__device__ inline void elem_geom(const DMesh& m, int c, double4& g0, double4& g1, double2 (&xy)[4],
                                 uint8_t& flags)
{
    int v[4] = {0, 0, 0, 0};
    const int n = contain_ring<MAXV>(m, c, v);
    {
        // the W and scalar interpolators use the interpolation ring; they
        // may use this frame only when it is the same vertex sequence
        int w[4] = {0, 0, 0, 0};
        const int nw = interp_ring<MAXV>(m.elems[c], w);
        bool same = nw == n;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < n && w[k] != v[k]) same = false;
        flags = same ? 1 : 0;
    }
    const double nan = __builtin_nan("");
#pragma unroll
    for (int k = 0; k < 4; ++k) xy[k] = make_double2(nan, nan);
    if (n < 3) {
        g0 = make_double4(0.0, 0.0, 0.0, __builtin_inf());
        g1 = make_double4(-__builtin_inf(), 0.0, 0.0, 0.0);
        return;
    }
    double lo[4] = {0, 0, 0, 0}, la[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < n) {
            double2 ll = m.lonlat[v[k]];
            lo[k] = ll.x;
            la[k] = ll.y;
        }
    }
    const double lon_ref = lo[0];
    double dmax = 0.0, dmin = 0.0;
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < n) {
            double d = km_pymod360(lo[k] - lon_ref + 180.0) - 180.0;
            if (k == 0 || d > dmax) dmax = d;
            if (k == 0 || d < dmin) dmin = d;
        }
    }
    double lamax = la[0], lamin = la[0];
#pragma unroll
    for (int k = 1; k < MAXV; ++k) {
        if (k < n) {
            if (la[k] > lamax) lamax = la[k];
            if (la[k] < lamin) lamin = la[k];
        }
    }
    double lmin = lo[0], lmax = lo[0];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < n) {
            if (lo[k] < lmin) lmin = lo[k];
            if (lo[k] > lmax) lmax = lo[k];
        }
    }
    const double nd = (double)n;
    double lon0;
    if (lmax - lmin > 180.0) {
        double dsum = 0.0;
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k < n) dsum += km_pymod360(lo[k] - lo[0] + 180.0) - 180.0;
        lon0 = lo[0] + dsum / nd;
    } else {
        double ssum = 0.0;
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k < n) ssum += lo[k];
        lon0 = ssum / nd;
    }
    double sl = 0.0;
#pragma unroll
    for (int k = 0; k < MAXV; ++k)
        if (k < n) sl += la[k];
    const double lat0 = sl / nd;
    const double cos_lat0 = km_cos(lat0 * kDeg2Rad);
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < n) {
            double dlon = km_pymod360(lo[k] - lon0 + 180.0) - 180.0;
            double dlat = la[k] - lat0;
            xy[k] = make_double2(dlon * kDeg2Rad * kREarthStd * cos_lat0, dlat * kDeg2Rad * kREarthStd);
        }
    }
    g0 = make_double4(lon_ref, dmin, dmax, lamin);
    g1 = make_double4(lamax, lon0, lat0, cos_lat0);
}

// core/mesh.py:2211 _element_contains (bbox prefilter on wrapped deltas,
// then :2140 _polygon_contains_bool with the :2007 projection) on the
// element's precomputed frame
template <int MAXV>
__device__ inline bool element_contains(const DMesh& m, int c, double lon_p, double lat_p)
{
    const double4 g0 = m.eframe[4 * (int64_t)c];
    const double dlon_p = km_pymod360(lon_p - g0.x + 180.0) - 180.0;
    if (dlon_p > g0.z || dlon_p < g0.y) return false;
    const double4 g1 = m.eframe[4 * (int64_t)c + 1];
    if (lat_p > g1.x || lat_p < g0.w) return false;
    const double2* xy = frame_xy(m, c);
    const double2 a = xy[0], b = xy[1], d = xy[2];
    const double dlp = km_pymod360(lon_p - g1.y + 180.0) - 180.0;
    const double px = dlp * kDeg2Rad * kREarthStd * g1.w;
    const double py = (lat_p - g1.z) * kDeg2Rad * kREarthStd;
    if (in_tri(px, py, a.x, a.y, b.x, b.y, d.x, d.y)) return true;
    if (MAXV == 3) return false;
    // quad: second fan triangle (0, 2, 3); NaN for a triangle never holds
    const double2 e = xy[3];
    return in_tri(px, py, a.x, a.y, d.x, d.y, e.x, e.y);
}

template <int MAXV>
__device__ __forceinline__ bool elem_touches(const DMesh& m, int c, int nn)
{
    int4 e = m.elems[c];
    if (e.x == nn || e.y == nn || e.z == nn) return true;
    return MAXV == 4 && e.w == nn;
}

// The rare 2-ring pass of core/mesh.py:2440-2463, kept out of line so the
// common 1-ring hit stays a short straight path.
template <int MAXV>
__device__ __noinline__ __attribute__((cold)) int containing_2ring(const DMesh& m, double lon_p, double lat_p, int nn)
{
    int k0, k1;
    n2e_range(m, nn, k0, k1);
    for (int k = k0; k < k1; ++k) {
        int e = m.n2e[k];
        int4 en = m.elems[e];
#pragma unroll
        for (int j = 0; j < MAXV; ++j) {
            int v = elem_slot(en, j);
            if (v < 0 || v == nn) continue;
            int q0, q1;
            n2e_range(m, v, q0, q1);
            for (int kk = q0; kk < q1; ++kk) {
                int c = m.n2e[kk];
                if (elem_touches<MAXV>(m, c, nn)) continue;
                if (element_contains<MAXV>(m, c, lon_p, lat_p)) return c;
            }
        }
    }
    return -1;
}

// core/mesh.py:2393 _find_containing_element_local
template <int MAXV>
__device__ inline int find_containing(const DMesh& m, double lon_p, double lat_p, int nn)
{
    int k0, k1;
    n2e_range(m, nn, k0, k1);
    if (k1 == k0) return -1;
    for (int k = k0; k < k1; ++k) {
        int c = m.n2e[k];
        if (element_contains<MAXV>(m, c, lon_p, lat_p)) return c;
    }
    return containing_2ring<MAXV>(m, lon_p, lat_p, nn);
}

// ---------------------------------------------------------------------------
// Edge walk (core/mesh.py:2489 _walk_containing_element, :2582
// _loop_walk_containing): elem_search = walk
// ---------------------------------------------------------------------------

constexpr int kWalkMaxSteps = 256;     // WALK_MAX_STEPS, core/mesh.py:2473
constexpr double kWalkEps = 1e-6;      // _WALK_EPS, core/mesh.py:2478
constexpr double kWalkVerify = 1e-2;   // _WALK_VERIFY_BAND, core/mesh.py:2485

// core/mesh.py:2007 _project_to_local_xy of the compacted ring v[0..n)
template <int MAXV>
__device__ inline void project_ring(const DMesh& m, const int (&v)[4], int n, double lon_p, double lat_p,
                                    double (&x)[4], double (&y)[4], double& px, double& py)
{
    double lo[4] = {0, 0, 0, 0}, la[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < n) {
            const double2 ll = m.lonlat[v[k]];
            lo[k] = ll.x;
            la[k] = ll.y;
        }
    }
    double lmin = lo[0], lmax = lo[0];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < n) {
            if (lo[k] < lmin) lmin = lo[k];
            if (lo[k] > lmax) lmax = lo[k];
        }
    }
    const double nd = (double)n;
    double lon0;
    if (lmax - lmin > 180.0) {
        double dsum = 0.0;
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k < n) dsum += km_pymod360(lo[k] - lo[0] + 180.0) - 180.0;
        lon0 = lo[0] + dsum / nd;
    } else {
        double ssum = 0.0;
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k < n) ssum += lo[k];
        lon0 = ssum / nd;
    }
    double sl = 0.0;
#pragma unroll
    for (int k = 0; k < MAXV; ++k)
        if (k < n) sl += la[k];
    const double lat0 = sl / nd;
    const double cosl = km_cos(lat0 * kDeg2Rad);
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        x[k] = 0.0;
        y[k] = 0.0;
        if (k < n) {
            x[k] = (km_pymod360(lo[k] - lon0 + 180.0) - 180.0) * kDeg2Rad * kREarthStd * cosl;
            y[k] = (la[k] - lat0) * kDeg2Rad * kREarthStd;
        }
    }
    px = (km_pymod360(lon_p - lon0 + 180.0) - 180.0) * kDeg2Rad * kREarthStd * cosl;
    py = (lat_p - lat0) * kDeg2Rad * kREarthStd;
}

template <int MAXV>
__device__ inline int walk_element(const DMesh& m, double lon_p, double lat_p, int start)
{
    int e = start;
    if (e < 0 || e >= m.n_elems) return -1;
    int prev = -1;
    for (int it = 0; it < kWalkMaxSteps; ++it) {
        int v[4] = {0, 0, 0, 0};
        const int4 en = elem_nodes<MAXV>(m, e);
        const int nv = interp_ring<MAXV>(en, v);
        if (nv < 3) return -1;
        double x[4], y[4], px, py;
        if (m.eflags[e] & 1) {
            // the element's precomputed frame is this projection
            const double4 g1 = m.eframe[4 * (int64_t)e + 1];
            const double2* xy = frame_xy(m, e);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                x[k] = k < nv ? xy[k].x : 0.0;
                y[k] = k < nv ? xy[k].y : 0.0;
            }
            px = (km_pymod360(lon_p - g1.y + 180.0) - 180.0) * kDeg2Rad * kREarthStd * g1.w;
            py = (lat_p - g1.z) * kDeg2Rad * kREarthStd;
        } else {
            project_ring<MAXV>(m, v, nv, lon_p, lat_p, x, y, px, py);
        }
        double S = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < nv) {
                const int j = (k + 1 == nv) ? 0 : k + 1;
                S += x[k] * y[j] - x[j] * y[k];
            }
        }
        const double absS = km_fabs(S);
        if (absS < kAreaTol) return -1;
        const double sgn = S > 0.0 ? 1.0 : -1.0;
        double worst = 1.0e100;
        int wk = -1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < nv) {
                const int j = (k + 1 == nv) ? 0 : k + 1;
                const double A = (x[j] - x[k]) * (py - y[k]) - (px - x[k]) * (y[j] - y[k]);
                const double w = A * sgn / absS;
                if (w < worst) {
                    worst = w;
                    wk = k;
                }
            }
        }
        if (worst >= -kWalkEps) {
            if (worst >= kWalkVerify) return e;
            return element_contains<MAXV>(m, e, lon_p, lat_p) ? e : -1;
        }
        const int4 nb4 = m.enbr[e];
        const int nb = wk == 0 ? nb4.x : (wk == 1 ? nb4.y : (wk == 2 ? nb4.z : nb4.w));
        if (nb < 0 || nb == prev) return -1;
        prev = e;
        e = nb;
    }
    return -1;
}

// One particle of _loop_walk_containing: the walk from the carried element,
// else the local search around the widened grid-hash nearest node.
template <int MAXV>
__device__ inline int walk_locate(const DMesh& m, double lon_p, double lat_p, int start)
{
    if (start >= 0) {
        const int e = walk_element<MAXV>(m, lon_p, lat_p, start);
        if (e >= 0) return e;
    }
    const int nn = nearest_widened(m, lon_p, lat_p);
    if (nn < 0) return -1;
    return find_containing<MAXV>(m, lon_p, lat_p, nn);
}

// ---------------------------------------------------------------------------
// Vertex weights (interpolate.py:1376-1404 + core/mesh.py:2077/2120)
// ---------------------------------------------------------------------------

struct Weights {
    int n;
    int v[4];
    double w[4];
};

// product-form Wachspress weights of a convex quad (core/mesh.py:2077)
__device__ __forceinline__ void wachspress4(double px, double py, const double (&vx)[4], const double (&vy)[4],
                                            double (&w)[4])
{
    double A[4], C[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i1 = (k + 1) & 3, im = (k + 3) & 3;
        A[k] = ((vx[k] - px) * (vy[i1] - py) - (vx[i1] - px) * (vy[k] - py));
        C[k] = ((vx[k] - vx[im]) * (vy[i1] - vy[im]) - (vx[i1] - vx[im]) * (vy[k] - vy[im]));
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        w[k] = C[k] * A[(k + 1) & 3] * A[(k + 2) & 3];
        s += w[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = w[k] / s;
}

template <int MAXV>
__device__ inline bool vertex_weights(const DMesh& m, int e, double xp, double yp, Weights& W)
{
    W.v[0] = W.v[1] = W.v[2] = W.v[3] = 0;
    W.w[0] = W.w[1] = W.w[2] = W.w[3] = 0.0;
    const int nv = interp_ring<MAXV>(elem_nodes<MAXV>(m, e), W.v);
    W.n = nv;
    if (nv < 3) return false;
    const double R = m.r_earth;
    double lo[4] = {0, 0, 0, 0}, la[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < nv) {
            double2 ll = m.lonlat[W.v[k]];
            lo[k] = ll.x;
            la[k] = ll.y;
        }
    }
    const double lon_ref = lo[0];
    double dsum = 0.0, lat0 = 0.0;
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < nv) {
            dsum += km_pymod360(lo[k] - lon_ref + 180.0) - 180.0;
            lat0 += la[k];
        }
    }
    const double nd = (double)nv;
    const double lon0 = lon_ref + dsum / nd;
    lat0 /= nd;
    const double cosl = km_cos(lat0 * kDeg2Rad);
    double vx[4] = {0, 0, 0, 0}, vy[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < nv) {
            double dlon = km_pymod360(lo[k] - lon0 + 180.0) - 180.0;
            vx[k] = dlon * kDeg2Rad * R * cosl;
            vy[k] = (la[k] - lat0) * kDeg2Rad * R;
        }
    }
    double dlon_p = km_pymod360(xp - lon0 + 180.0) - 180.0;
    double px = dlon_p * kDeg2Rad * R * cosl;
    double py = (yp - lat0) * kDeg2Rad * R;
    if (nv == 3 || MAXV == 3) {
        bary_tri(px, py, vx[0], vy[0], vx[1], vy[1], vx[2], vy[2], W.w[0], W.w[1], W.w[2]);
        return W.w[0] != -1.0;
    }
    wachspress4(px, py, vx, vy, W.w);
    return W.w[0] != -1.0;
}

// ---------------------------------------------------------------------------
// Snapshot records
//
// A snapshot is stored as one record per (layer, node), node in the mesh's
// internal order: 3-D (u, v, w_top, w_bottom) of layer L -- the layer's u/v and
// the W of its two bounding interfaces L and L+1 -- as float4/double4; 2-D
// (u, v) as float2/double2.  One vertex's stage values are one 16/32-byte load
// instead of four scattered ones.  Values are the uploaded floats unchanged.
// ---------------------------------------------------------------------------

template <typename T, bool IS3D>
struct RecOf;
template <>
struct RecOf<float, true> {
    using type = float4;
};
template <>
struct RecOf<double, true> {
    using type = double4;
};
template <>
struct RecOf<float, false> {
    using type = float2;
};
template <>
struct RecOf<double, false> {
    using type = double2;
};

// bfloat16 storage: the upper half of an IEEE float32, widened exactly
// (through float) to double wherever a record value is read, so the
// arithmetic downstream is that of a float32 field holding the same values
struct Bf16 {
    uint16_t bits;
    __device__ __forceinline__ operator double() const
    {
        return (double)__uint_as_float((uint32_t)bits << 16);
    }
};
struct __align__(8) Bf16x4 {
    Bf16 x, y, z, w;
};
struct __align__(4) Bf16x2 {
    Bf16 x, y;
};
template <>
struct RecOf<Bf16, true> {
    using type = Bf16x4;
};
template <>
struct RecOf<Bf16, false> {
    using type = Bf16x2;
};

// float32 -> bfloat16, round to nearest even (a NaN stays a quiet NaN)
__device__ __forceinline__ Bf16 to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return Bf16{(uint16_t)((u >> 16) | 0x40u)};
    u += 0x7fffu + ((u >> 16) & 1u);
    return Bf16{(uint16_t)(u >> 16)};
}

// float16 storage with a power-of-two scale per level (KELP_DT_F16S): a
// record holds (u, v[, W top, W bottom]) divided by their level's scale as
// IEEE half values; the scales of record row L (one row of n_nodes records
// per layer) are the float4 (s_uv, s_uv, s_w(L), s_w(L + 1)) of the slot's
// scale table.  A value reads back as half -> float -> double times its
// scale, exactly (the scale is a power of two), so the arithmetic downstream
// is that of a float32 field holding the decoded values.
struct F16s {
    uint16_t bits;
};
struct __align__(8) F16sx4 {
    uint16_t x, y, z, w;
};
struct __align__(4) F16sx2 {
    uint16_t x, y;
};
template <>
struct RecOf<F16s, true> {
    using type = F16sx4;
};
template <>
struct RecOf<F16s, false> {
    using type = F16sx2;
};

__device__ __forceinline__ double f16_val(uint16_t h, float scale)
{
    return (double)(float)__builtin_bit_cast(_Float16, h) * (double)scale;
}

// bracket pair (a, b) blended by alpha0 and the stage target t
template <typename T, bool IS3D>
struct Snap {
    using R = typename RecOf<T, IS3D>::type;
    const R* __restrict__ a;
    const R* __restrict__ b;
    const R* __restrict__ t;
};

// float16 records carry their slots' per-row scale tables
template <bool IS3D>
struct Snap<F16s, IS3D> {
    using R = typename RecOf<F16s, IS3D>::type;
    const R* __restrict__ a;
    const R* __restrict__ b;
    const R* __restrict__ t;
    const float4* __restrict__ sa;
    const float4* __restrict__ sb;
    const float4* __restrict__ st;
    uint32_t row_len;  // records per row (the mesh's node count)
};

// numerics/vertical.py:1043 _tblend_scalar == interpolate.py:1419-1437: the
// nested two-level time blend of one value
__device__ __forceinline__ double tblend(double a, double b, double t, double alpha0, double alpha)
{
    if (alpha < 1e-12) {
        if (alpha0 == 0.0) return a;
        return (1.0 - alpha0) * a + alpha0 * b;
    }
    if (alpha > 1.0 - 1e-12) return t;
    if (alpha0 == 0.0) return (1.0 - alpha) * a + alpha * t;
    double a0 = (1.0 - alpha0) * a + alpha0 * b;
    return (1.0 - alpha) * a0 + alpha * t;
}

// blended (u, v) of record idx; only the snapshots the blend uses are read
template <typename T, bool IS3D>
__device__ __forceinline__ void uv_blend(const Snap<T, IS3D>& f, int64_t idx, double alpha0, double alpha,
                                         double& u, double& v)
{
    if (alpha < 1e-12) {
        auto ra = f.a[idx];
        if (alpha0 == 0.0) {
            u = (double)ra.x;
            v = (double)ra.y;
            return;
        }
        auto rb = f.b[idx];
        u = (1.0 - alpha0) * (double)ra.x + alpha0 * (double)rb.x;
        v = (1.0 - alpha0) * (double)ra.y + alpha0 * (double)rb.y;
        return;
    }
    if (alpha > 1.0 - 1e-12) {
        auto rt = f.t[idx];
        u = (double)rt.x;
        v = (double)rt.y;
        return;
    }
    auto ra = f.a[idx];
    auto rt = f.t[idx];
    if (alpha0 == 0.0) {
        u = (1.0 - alpha) * (double)ra.x + alpha * (double)rt.x;
        v = (1.0 - alpha) * (double)ra.y + alpha * (double)rt.y;
        return;
    }
    auto rb = f.b[idx];
    double ua0 = (1.0 - alpha0) * (double)ra.x + alpha0 * (double)rb.x;
    double va0 = (1.0 - alpha0) * (double)ra.y + alpha0 * (double)rb.y;
    u = (1.0 - alpha) * ua0 + alpha * (double)rt.x;
    v = (1.0 - alpha) * va0 + alpha * (double)rt.y;
}

// the same for float16 records: each value times its row's scale
template <bool IS3D>
__device__ __forceinline__ void uv_blend(const Snap<F16s, IS3D>& f, int64_t idx, double alpha0, double alpha,
                                         double& u, double& v)
{
    const uint32_t row = (uint32_t)idx / f.row_len;
    if (alpha < 1e-12) {
        const auto ra = f.a[idx];
        const float4 sa = f.sa[row];
        if (alpha0 == 0.0) {
            u = f16_val(ra.x, sa.x);
            v = f16_val(ra.y, sa.y);
            return;
        }
        const auto rb = f.b[idx];
        const float4 sb = f.sb[row];
        u = (1.0 - alpha0) * f16_val(ra.x, sa.x) + alpha0 * f16_val(rb.x, sb.x);
        v = (1.0 - alpha0) * f16_val(ra.y, sa.y) + alpha0 * f16_val(rb.y, sb.y);
        return;
    }
    if (alpha > 1.0 - 1e-12) {
        const auto rt = f.t[idx];
        const float4 st = f.st[row];
        u = f16_val(rt.x, st.x);
        v = f16_val(rt.y, st.y);
        return;
    }
    const auto ra = f.a[idx];
    const auto rt = f.t[idx];
    const float4 sa = f.sa[row], st = f.st[row];
    if (alpha0 == 0.0) {
        u = (1.0 - alpha) * f16_val(ra.x, sa.x) + alpha * f16_val(rt.x, st.x);
        v = (1.0 - alpha) * f16_val(ra.y, sa.y) + alpha * f16_val(rt.y, st.y);
        return;
    }
    const auto rb = f.b[idx];
    const float4 sb = f.sb[row];
    double ua0 = (1.0 - alpha0) * f16_val(ra.x, sa.x) + alpha0 * f16_val(rb.x, sb.x);
    double va0 = (1.0 - alpha0) * f16_val(ra.y, sa.y) + alpha0 * f16_val(rb.y, sb.y);
    u = (1.0 - alpha) * ua0 + alpha * f16_val(rt.x, st.x);
    v = (1.0 - alpha) * va0 + alpha * f16_val(rt.y, st.y);
}

// blended W of the two interfaces bounding record idx (w_top, w_bottom)
template <typename T>
__device__ __forceinline__ void w_pair_blend(const Snap<T, true>& f, int64_t idx, double alpha0, double alpha,
                                             double& w_top, double& w_bot)
{
    auto ra = f.a[idx];
    auto rb = f.b[idx];
    auto rt = f.t[idx];
    w_top = tblend((double)ra.z, (double)rb.z, (double)rt.z, alpha0, alpha);
    w_bot = tblend((double)ra.w, (double)rb.w, (double)rt.w, alpha0, alpha);
}

__device__ __forceinline__ void w_pair_blend(const Snap<F16s, true>& f, int64_t idx, double alpha0, double alpha,
                                             double& w_top, double& w_bot)
{
    const uint32_t row = (uint32_t)idx / f.row_len;
    const auto ra = f.a[idx];
    const auto rb = f.b[idx];
    const auto rt = f.t[idx];
    const float4 sa = f.sa[row], sb = f.sb[row], st = f.st[row];
    w_top = tblend(f16_val(ra.z, sa.z), f16_val(rb.z, sb.z), f16_val(rt.z, st.z), alpha0, alpha);
    w_bot = tblend(f16_val(ra.w, sa.w), f16_val(rb.w, sb.w), f16_val(rt.w, st.w), alpha0, alpha);
}

// interpolate.py:207 _interp_leastsq_core_tinterp (== :52 for a plain field):
// a plane through the blended element-centre values of the anchor's 1-ring
// in particle-centred metric coordinates (antimeridian gate per centre),
// Sarrus determinant with the scale-relative singularity guard, Cramer's rule
// for the constant term.  `off` is the layer plane's first element record.
template <typename T>
__device__ inline bool interp_lsq(const DMesh& m, const Snap<T, false>& fe, double alpha0, double alpha,
                                  double time_dir, double xp, double yp, int nn, int64_t off, double& U,
                                  double& V)
{
    U = 0.0;
    V = 0.0;
    int k0, k1;
    n2e_range(m, nn, k0, k1);
    if (k1 - k0 < 1) return false;
    const double z0 = km_cos(yp * kDeg2Rad) * kSLM;
    const double xp_m = xp * z0;
    const double yp_m = yp * kSLM;
    double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
    double fUn = 0.0, f1Un = 0.0, f2Un = 0.0, fVn = 0.0, f1Vn = 0.0, f2Vn = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int e = m.n2e[k];
        const double2 cc = m.ecent[e];
        double ex = cc.x * z0 - xp_m;
        const double dlon_raw = cc.x - xp;
        if (dlon_raw > 180.0)
            ex -= 360.0 * z0;
        else if (dlon_raw < -180.0)
            ex += 360.0 * z0;
        const double ey = cc.y * kSLM - yp_m;
        double u, v;
        uv_blend(fe, off + e, alpha0, alpha, u, v);
        const double uu = time_dir * u;
        const double vv = time_dir * v;
        a11 = a11 + 1.0;
        a12 = a12 + ex;
        a13 = a13 + ey;
        a22 = a22 + ex * ex;
        a23 = a23 + ex * ey;
        a33 = a33 + ey * ey;
        fUn = fUn + uu;
        f1Un = f1Un + uu * ex;
        f2Un = f2Un + uu * ey;
        fVn = fVn + vv;
        f1Vn = f1Vn + vv * ex;
        f2Vn = f2Vn + vv * ey;
    }
    const double a21 = a12, a31 = a13, a32 = a23;
    const double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                        a12 * a21 * a33 - a11 * a23 * a32);
    if (km_fabs(det) < 1e-30 || km_fabs(det) < 1e-12 * km_fabs(a11 * a22 * a33)) return false;
    const double inv = 1.0 / km_fabs(det);
    const double b11 = a22 * a33 - a23 * a32;
    const double b12 = a13 * a32 - a12 * a33;
    const double b13 = a12 * a23 - a13 * a22;
    U = (b11 * fUn + b12 * f1Un + b13 * f2Un) * inv;
    V = (b11 * fVn + b12 * f1Vn + b13 * f2Vn) * inv;
    return true;
}

// ---------------------------------------------------------------------------
// Vertical (z-level)
// ---------------------------------------------------------------------------

// First j in [0, K-1) with zl[j+1] > x, K-1 when there is none.  The levels
// are non-decreasing (checked at upload), so the predicate is monotone and a
// binary search returns the index of the reference's linear scans.
__device__ __forceinline__ int first_above(const DMesh& m, double x)
{
    const double* zl = m.zlev;
    int lo = 0, hi = m.n_levels - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (zl[mid + 1] > x)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// numerics/vertical.py:40 find_layer
__device__ inline int find_layer(const DMesh& m, double zp, double bottom)
{
    const int K = m.n_levels;
    const double* zl = m.zlev;
    if (K < 2) return 0;
    if (zp <= zl[0]) return 0;
    int k = first_above(m, bottom);
    int max_layer = k < K - 1 ? (k - 1 > 0 ? k - 1 : 0) : K - 2;
    if (zp >= bottom || zp >= zl[max_layer + 1]) return max_layer;
    const int r = first_above(m, zp);  // > max_layer only for a NaN depth
    return r <= max_layer ? r : max_layer;
}

// The vertical bracket of depth zp among the interfaces, in the branch
// structure of numerics/vertical.py:260 _w_bracket: the record (layer) k and
// either its top (a = -1: zp at or above the surface level, or a degenerate
// layer), its bottom (a = 2: zp at or below the last level) or the blend
// weight a in [0, 1).  Depends on zp only, so one bracket serves every node
// evaluated at the same depth.
struct WBracket {
    int k;
    double a;
};

__device__ __forceinline__ WBracket w_bracket(const DMesh& m, double zp)
{
    const double* zl = m.zlev;
    const int K = m.n_levels;
    if (zp <= zl[0]) return {0, -1.0};
    if (zp >= zl[K - 1]) return {K - 2, 2.0};
    int k = first_above(m, zp);  // < K - 1 here unless zp is NaN
    if (k > K - 2) k = K - 2;
    const double dz = zl[k + 1] - zl[k];
    if (dz < 1e-12) return {k, -1.0};
    return {k, (zp - zl[k]) / dz};
}

// W of one node at a bracket (numerics/vertical.py:282 _w_at_bracket on the
// stage field: the interface pair blended in time, then in depth)
template <typename T>
__device__ __forceinline__ double w_at(const DMesh& m, const Snap<T, true>& fw, double alpha0, double alpha,
                                       const WBracket& b, int node, double time_dir)
{
    double wt, wb;
    w_pair_blend(fw, (int64_t)b.k * m.n_nodes + node, alpha0, alpha, wt, wb);
    if (b.a < 0.0) return time_dir * wt;
    if (b.a == 2.0) return time_dir * wb;
    return time_dir * ((1.0 - b.a) * wt + b.a * wb);
}

// numerics/vertical.py:1065 _interp_w_linear_single_tinterp: W linear in
// depth between the interfaces bracketing zp at one node (records hold the
// interface pair of each layer: level k < K-1 is record k's top, level K-1 the
// last record's bottom)
template <typename T>
__device__ inline double w_linear(const DMesh& m, const Snap<T, true>& fw, double alpha0, double alpha,
                                  double zp, int node, double time_dir)
{
    return w_at(m, fw, alpha0, alpha, w_bracket(m, zp), node, time_dir);
}

// W at node nd, depth clamped to its bottom (the bottom-aware evaluation of
// the barycentric and least-squares W); bz is the bracket of zp itself
template <typename T>
__device__ __forceinline__ double w_clamped(const DMesh& m, const Snap<T, true>& fw, double alpha0, double alpha,
                                            double zp, const WBracket& bz, int nd, double time_dir)
{
    const double dep = m.depth[nd];
    if (zp < dep) return w_at(m, fw, alpha0, alpha, bz, nd, time_dir);
    return w_at(m, fw, alpha0, alpha, w_bracket(m, dep), nd, time_dir);
}

// numerics/vertical.py:745 _interp_w_bary_single_tinterp: first fan
// sub-triangle of a 1-ring element holding the point (projection around the
// seam-safe vertex mean, core/mesh.py:1974), bottom-aware linear W per vertex;
// fallback = linear W at the anchor, depth clamped to its bottom.  The
// projection is the element's precomputed containment frame (the same
// expressions) whenever the two vertex rings coincide.
template <int MAXV, typename T>
__device__ inline double w_bary(const DMesh& m, const Snap<T, true>& fw, double alpha0, double alpha, double xp,
                                double yp, double zp, int node, double time_dir)
{
    const WBracket bz = w_bracket(m, zp);
    int k0, k1;
    n2e_range(m, node, k0, k1);
    for (int k = k0; k < k1; ++k) {
        const int e = m.n2e[k];
        int verts[4] = {0, 0, 0, 0};
        const int nv = interp_ring<MAXV>(elem_nodes<MAXV>(m, e), verts);
        if (nv < 3) continue;
        double vx[4] = {0, 0, 0, 0}, vy[4] = {0, 0, 0, 0};
        double px, py;
        if (m.eflags[e] & 1) {
            const double4 g1 = m.eframe[4 * (int64_t)e + 1];
            const double2* xy = frame_xy(m, e);
#pragma unroll
            for (int v = 0; v < MAXV; ++v) {
                if (v < nv) {
                    const double2 q = xy[v];
                    vx[v] = q.x;
                    vy[v] = q.y;
                }
            }
            const double dlon_p = km_pymod360(xp - g1.y + 180.0) - 180.0;
            px = dlon_p * kDeg2Rad * kREarthStd * g1.w;
            py = (yp - g1.z) * kDeg2Rad * kREarthStd;
        } else {
            double lo[4] = {0, 0, 0, 0}, la[4] = {0, 0, 0, 0};
#pragma unroll
            for (int v = 0; v < MAXV; ++v) {
                if (v < nv) {
                    double2 ll = m.lonlat[verts[v]];
                    lo[v] = ll.x;
                    la[v] = ll.y;
                }
            }
            const double first = lo[0];
            double lmin = first, lmax = first, ssum = 0.0;
#pragma unroll
            for (int v = 0; v < MAXV; ++v) {
                if (v < nv) {
                    ssum += lo[v];
                    if (lo[v] < lmin) lmin = lo[v];
                    if (lo[v] > lmax) lmax = lo[v];
                }
            }
            const double nd = (double)nv;
            double lon0;
            if (lmax - lmin > 180.0) {
                double d = 0.0;
#pragma unroll
                for (int v = 0; v < MAXV; ++v)
                    if (v < nv) d += km_pymod360(lo[v] - first + 180.0) - 180.0;
                lon0 = first + d / nd;
            } else {
                lon0 = ssum / nd;
            }
            double lat0 = 0.0;
#pragma unroll
            for (int v = 0; v < MAXV; ++v)
                if (v < nv) lat0 += la[v];
            lat0 /= nd;
            const double cos_lat0 = km_cos(lat0 * kDeg2Rad);
            const double R = 6371000.0;
#pragma unroll
            for (int v = 0; v < MAXV; ++v) {
                if (v < nv) {
                    double dlon = km_pymod360(lo[v] - lon0 + 180.0) - 180.0;
                    double dlat = la[v] - lat0;
                    vx[v] = dlon * kDeg2Rad * R * cos_lat0;
                    vy[v] = dlat * kDeg2Rad * R;
                }
            }
            const double dlon_p = km_pymod360(xp - lon0 + 180.0) - 180.0;
            px = dlon_p * kDeg2Rad * R * cos_lat0;
            py = (yp - lat0) * kDeg2Rad * R;
        }
#pragma unroll
        for (int t = 0; t < MAXV - 2; ++t) {
            if (t < nv - 2) {
                const int i1 = t + 1, i2 = t + 2;
                const double ax = vx[0], ay = vy[0], bx = vx[i1], by = vy[i1], cx = vx[i2], cy = vy[i2];
                const double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
                if (!(km_fabs(denom) < 1e-14)) {
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
                            for (int q = 0; q < 4; ++q)
                                acc += wts[q] * w_clamped(m, fw, alpha0, alpha, zp, bz, verts[q], time_dir);
                            return acc;
                        }
                        const double wv0 = w_clamped(m, fw, alpha0, alpha, zp, bz, verts[0], time_dir);
                        const double wv1 = w_clamped(m, fw, alpha0, alpha, zp, bz, verts[i1], time_dir);
                        const double wv2 = w_clamped(m, fw, alpha0, alpha, zp, bz, verts[i2], time_dir);
                        return w1 * wv0 + w2 * wv1 + w3 * wv2;
                    }
                }
            }
        }
    }
    return w_clamped(m, fw, alpha0, alpha, zp, bz, node, time_dir);
}

// Did vertex slot (k, v) of the anchor's element list hold node `nd` at an
// earlier position of the scan (an earlier element, or an earlier vertex of
// the same element)?  The 1-ring is re-read from the mesh (cached lines)
// instead of kept in a per-lane list.
template <int MAXV>
__device__ __forceinline__ bool seen_before(const DMesh& m, int k0, int k, int v, int nd)
{
    for (int kk = k0; kk < k; ++kk) {
        const int e2 = m.n2e[kk];
        if (e2 < 0) continue;
        const int4 q = m.elems[e2];
        if (q.x == nd || q.y == nd || q.z == nd || (MAXV == 4 && q.w == nd)) return true;
    }
    const int4 q = m.elems[m.n2e[k]];
    if (v > 0 && q.x == nd) return true;
    if (v > 1 && q.y == nd) return true;
    if (v > 2 && q.z == nd) return true;
    return false;
}

// numerics/vertical.py:493 _interp_w_leastsq_single on the stage field
// (interface values blended in time): the unique 1-ring nodes in the order
// of vertical.py:446 _collect_neighbour_nodes (centre first, then each
// element's vertices, at most 64), W linear in depth at each (clamped to that
// node's bottom), a plane fit evaluated at the particle; fewer than 3 nodes
// or a singular fit fall back to linear W at the anchor.
template <int MAXV, typename T>
__device__ inline double w_lsq(const DMesh& m, const Snap<T, true>& fw, double alpha0, double alpha, double xp,
                               double yp, double zp, int node, double time_dir)
{
    int k0, k1;
    n2e_range(m, node, k0, k1);
    const WBracket bz = w_bracket(m, zp);
    double z0 = km_cos(yp * kDeg2Rad) * kSLM;
    if (z0 < 1.0) z0 = 1.0;
    double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
    double fw_ = 0.0, f1w = 0.0, f2w = 0.0;
    int count = 0;
    // k = k0 - 1 stands for the centre node
    for (int k = k0 - 1; k < k1; ++k) {
        int4 q = make_int4(node, -1, -1, -1);
        int nv = 1;
        if (k >= k0) {
            const int e = m.n2e[k];
            if (e < 0) continue;
            q = m.elems[e];
            nv = MAXV;
        }
#pragma unroll
        for (int v = 0; v < MAXV; ++v) {
            if (v >= nv) break;
            const int nd = v == 0 ? q.x : (v == 1 ? q.y : (v == 2 ? q.z : q.w));
            if (nd < 0) continue;
            if (k >= k0 && (nd == node || seen_before<MAXV>(m, k0, k, v, nd))) continue;
            if (count >= 64) continue;
            ++count;
            const double w_val = w_clamped(m, fw, alpha0, alpha, zp, bz, nd, time_dir);
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
            fw_ += w_val;
            f1w += w_val * ex;
            f2w += w_val * ey;
        }
    }
    const double a21 = a12, a31 = a13, a32 = a23;
    const double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                        a12 * a21 * a33 - a11 * a23 * a32);
    if (count < 3 || km_fabs(det) < 1e-30 || km_fabs(det) < 1e-12 * km_fabs(a11 * a22 * a33))
        return w_clamped(m, fw, alpha0, alpha, zp, bz, node, time_dir);
    const double inv = 1.0 / km_fabs(det);
    const double b11 = a22 * a33 - a23 * a32;
    const double b12 = a13 * a32 - a12 * a33;
    const double b13 = a12 * a23 - a13 * a22;
    return (b11 * fw_ + b12 * f1w + b13 * f2w) * inv;
}

// numerics/vertical.py:1148 _interp_w_idw_single on the stage field: the
// same unique 1-ring nodes and bottom-aware W per node as w_lsq, weighted by
// 1 / d^2 in the local metric; a node within 1 m returns its own W, an empty
// weight sum falls back to linear W at the anchor.
template <int MAXV, typename T>
__device__ inline double w_idw(const DMesh& m, const Snap<T, true>& fw, double alpha0, double alpha, double xp,
                               double yp, double zp, int node, double time_dir)
{
    int k0, k1;
    n2e_range(m, node, k0, k1);
    const WBracket bz = w_bracket(m, zp);
    double z0 = km_cos(yp * kDeg2Rad) * kSLM;
    if (z0 < 1.0) z0 = 1.0;
    double sw = 0.0, swv = 0.0;
    int count = 0;
    // k = k0 - 1 stands for the centre node
    for (int k = k0 - 1; k < k1; ++k) {
        int4 q = make_int4(node, -1, -1, -1);
        int nv = 1;
        if (k >= k0) {
            const int e = m.n2e[k];
            if (e < 0) continue;
            q = m.elems[e];
            nv = MAXV;
        }
#pragma unroll
        for (int v = 0; v < MAXV; ++v) {
            if (v >= nv) break;
            const int nd = v == 0 ? q.x : (v == 1 ? q.y : (v == 2 ? q.z : q.w));
            if (nd < 0) continue;
            if (k >= k0 && (nd == node || seen_before<MAXV>(m, k0, k, v, nd))) continue;
            if (count >= 64) continue;
            ++count;
            const double w_val = w_clamped(m, fw, alpha0, alpha, zp, bz, nd, time_dir);
            const double2 ll = m.lonlat[nd];
            const double dlon = km_pymod360(ll.x - xp + 180.0) - 180.0;
            const double dx = dlon * z0;
            const double dy = (ll.y - yp) * kSLM;
            const double d2 = dx * dx + dy * dy;
            if (d2 < 1.0) return w_val;
            const double w = 1.0 / d2;
            sw += w;
            swv += w * w_val;
        }
    }
    if (sw < 1e-30) return w_clamped(m, fw, alpha0, alpha, zp, bz, node, time_dir);
    return swv / sw;
}

// ---------------------------------------------------------------------------
// Counter RNG (numerics/diffusion.py:181-246, numerics/beaching.py:46)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x = x + 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x = x ^ (x >> 31);
    return x;
}

__device__ __forceinline__ uint64_t hash5(uint64_t k0, uint64_t k1, uint64_t k2, uint64_t k3, uint64_t k4)
{
    uint64_t h = k0;
    h = splitmix64(h ^ k1);
    h = splitmix64(h ^ k2);
    h = splitmix64(h ^ k3);
    h = splitmix64(h ^ k4);
    return h;
}

__device__ inline void normal_pair(uint64_t seed, uint64_t real_id, uint64_t pid, uint64_t step,
                                   uint64_t comp, double& g1, double& g2)
{
    const double INV_2_32 = 1.0 / 4294967296.0;
    const double HALF_INV_2_32 = 0.5 * INV_2_32;
    const double TWO_PI = 6.283185307179586;
    uint64_t h = hash5(seed, real_id, pid, step, comp);
    double high = (double)((h >> 32) & 0xFFFFFFFFULL);
    double low = (double)(h & 0xFFFFFFFFULL);
    double u1 = high * INV_2_32 + HALF_INV_2_32;
    double u2 = low * INV_2_32 + HALF_INV_2_32;
    double r = km_sqrt(-2.0 * km_log(u1));
    double s, c;
    km_sincos(TWO_PI * u2, &s, &c);
    g1 = r * c;
    g2 = r * s;
}

__device__ __forceinline__ double u01(uint64_t seed, uint64_t real_id, uint64_t pid, uint64_t step, uint64_t comp)
{
    uint64_t h = hash5(seed, real_id, pid, step, comp);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

constexpr double kRMetric = 111120.0 / 0.017453292519943295;  // diffusion.py:532

// numerics/diffusion.py:539 sphere_displace
__device__ inline void sphere_displace(double lon, double lat, double dx_m, double dy_m, double& lon_o,
                                       double& lat_o)
{
    double dist = km_sqrt(dx_m * dx_m + dy_m * dy_m);
    if (dist == 0.0) {
        lon_o = lon;
        lat_o = lat;
        return;
    }
    double lam = lon * kDeg2Rad, phi = lat * kDeg2Rad;
    double sl, cl, sp, cp;
    km_sincos(lam, &sl, &cl);
    km_sincos(phi, &sp, &cp);
    double delta = dist / kRMetric;
    double sd, c;
    km_sincos(delta, &sd, &c);
    double s = sd / dist;
    double tx = -sl * dx_m - sp * cl * dy_m;
    double ty = cl * dx_m - sp * sl * dy_m;
    double tz = cp * dy_m;
    double x2 = cp * cl * c + tx * s;
    double y2 = cp * sl * c + ty * s;
    double z2 = sp * c + tz * s;
    if (z2 > 1.0)
        z2 = 1.0;
    else if (z2 < -1.0)
        z2 = -1.0;
    lat_o = km_asin(z2) / kDeg2Rad;
    double lon2 = km_atan2(y2, x2) / kDeg2Rad;
    lon_o = km_pymod360(lon2 + 180.0) - 180.0;
}

}  // namespace kelp
