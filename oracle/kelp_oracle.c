/* kelp_oracle.c -- CPU oracle (test infrastructure; see kelp_oracle.h).
 *
 * Each function restates one reference function; the citation names the
 * file:line of the reference (src/kelp/...) it follows.  Loop order,
 * operand order, comparison strictness and the Python `%` semantics are
 * kept, since they fix tie-breaks and rounding.  Compiled with
 * -O2 -ffp-contract=off (the reference's Numba build has no fastmath, so
 * LLVM never contracts to FMA).
 *
 * Parity status of each function is recorded beside it ("pinned by ...").
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "kelp_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_KMATH
#include "../kelp_amd/csrc/kmath.h"
#define M_SIN km_sin
#define M_COS km_cos
#define M_ASIN km_asin
#define M_ACOS km_acos
#define M_ATAN2 km_atan2
#define M_LOG km_log
#define M_EXP km_exp
#define M_POW km_pow
#define PYMOD360 km_pymod360
#else
#define M_SIN sin
#define M_COS cos
#define M_ASIN asin
#define M_ACOS acos
#define M_ATAN2 atan2
#define M_LOG log
#define M_EXP exp
#define M_POW pow
/* Python float `%` by 360 (CPython float_rem): fmod, then the divisor's
 * sign; a zero remainder is +0.0. */
static double PYMOD360(double x)
{
    double m = fmod(x, 360.0);
    if (m != 0.0) {
        if (m < 0.0) m += 360.0;
    } else {
        m = 0.0;
    }
    return m;
}
#endif

#define DEG2RAD 0.017453292519943295 /* np.pi / 180.0, core/geometry.py:31 */
#define RAD2DEG 57.29577951308232    /* 180.0 / np.pi, core/geometry.py:32 */
#define SLM 111120.0                 /* METERS_PER_DEG_LAT, core/geometry.py:28 */
#define R_EARTH_STANDARD 6371000.0   /* core/geometry.py:30 */
#define BARY_EPS (-1e-6)             /* core/mesh.py:33 */
#define AREA_TOL 1e-10               /* core/mesh.py:32 */
#define EXACT_SHORT_ROW_CELLS 32     /* core/mesh.py:1608 */

#define ST_ALIVE 1
#define ST_OUTSIDE 3
#define FR_INTERP 1
#define FR_OOB 2

static inline double sq(double x) { return x * x; }

static inline double fld(const void *p, int dtype, int64_t i)
{
    if (dtype == OR_F32) return (double)((const float *)p)[i];
    return ((const double *)p)[i];
}

/* ------------------------------------------------------------------------
 * Grid hash: cell mapping and scans
 * ---------------------------------------------------------------------- */

/* core/mesh.py:1193 _gh_fold_dlon */
static double gh_fold_dlon(double lon_p, const or_mesh_t *m)
{
    double dlon = PYMOD360(lon_p - m->gh_lon_min);
    if (!m->gh_wrap) {
        double grid_span = m->gh_nx * m->gh_cell;
        if (dlon > grid_span && (dlon - grid_span) > (360.0 - dlon))
            dlon -= 360.0;
    }
    return dlon;
}

/* core/mesh.py:1211 _gh_cell_of -- int() truncates toward zero */
static void gh_cell_of(double lon_p, double lat_p, const or_mesh_t *m, int *ci_o, int *cj_o)
{
    double dlon = gh_fold_dlon(lon_p, m);
    int ci = (int)(dlon / m->gh_cell);
    int cj = (int)((lat_p - m->gh_lat_min) / m->gh_cell);
    if (ci < 0) ci = 0;
    if (ci >= m->gh_nx) ci = m->gh_nx - 1;
    if (cj < 0) cj = 0;
    if (cj >= m->gh_ny) cj = m->gh_ny - 1;
    *ci_o = ci;
    *cj_o = cj;
}

/* query unit vector, the expressions of core/mesh.py:1281-1286 */
static void query_xyz(double lon_p, double lat_p, double *xp, double *yp, double *zp, double *cos_lat_o)
{
    double lat_r = lat_p * DEG2RAD;
    double lon_r = lon_p * DEG2RAD;
    double cos_lat_p = M_COS(lat_r);
    *xp = cos_lat_p * M_COS(lon_r);
    *yp = cos_lat_p * M_SIN(lon_r);
    *zp = M_SIN(lat_r);
    if (cos_lat_o) *cos_lat_o = cos_lat_p;
}

/* chord^2 to node n, (x-xp)**2 + (y-yp)**2 + (z-zp)**2 (numba: x*x) */
static inline double chord2(const or_mesh_t *m, int n, double xp, double yp, double zp)
{
    const double *q = m->node_xyz + 3 * (int64_t)n;
    return sq(q[0] - xp) + sq(q[1] - yp) + sq(q[2] - zp);
}

/* core/mesh.py:1570 _gh_scan_row_range (also serves _gh_scan_cell :1455
 * with c_lo == c_hi) */
static void gh_scan_row_range(const or_mesh_t *m, int cj2, int c_lo, int c_hi,
                              double xp, double yp, double zp, int *best, double *best_d2)
{
    int32_t start = m->gh_offsets[(int64_t)cj2 * m->gh_nx + c_lo];
    int32_t end = m->gh_offsets[(int64_t)cj2 * m->gh_nx + c_hi + 1];
    for (int32_t k = start; k < end; k++) {
        int n = m->gh_node_ids[k];
        double d2 = chord2(m, n, xp, yp, zp);
        if (d2 < *best_d2) {
            *best_d2 = d2;
            *best = n;
        }
    }
}

/* core/mesh.py:1588 _gh_scan_row_window */
static void gh_scan_row_window(const or_mesh_t *m, int cj2, int ci, int rx,
                               double xp, double yp, double zp, int *best, double *best_d2)
{
    int nx = m->gh_nx;
    if (m->gh_wrap && 2 * rx + 1 >= nx) {
        gh_scan_row_range(m, cj2, 0, nx - 1, xp, yp, zp, best, best_d2);
        return;
    }
    int c_lo = ci - rx, c_hi = ci + rx;
    if (m->gh_wrap) {
        if (c_lo < 0) {
            gh_scan_row_range(m, cj2, c_lo + nx, nx - 1, xp, yp, zp, best, best_d2);
            c_lo = 0;
        } else if (c_hi > nx - 1) {
            gh_scan_row_range(m, cj2, 0, c_hi - nx, xp, yp, zp, best, best_d2);
            c_hi = nx - 1;
        }
    } else {
        if (c_lo < 0) c_lo = 0;
        if (c_hi > nx - 1) c_hi = nx - 1;
    }
    gh_scan_row_range(m, cj2, c_lo, c_hi, xp, yp, zp, best, best_d2);
}

/* core/mesh.py:1236 _grid_hash_nearest_3x3.  pinned by brute-force chord
 * argmin (tests/test_oracle_nearest.py). */
int or_nearest_3x3(const or_mesh_t *m, double lon_p, double lat_p)
{
    int ci, cj;
    gh_cell_of(lon_p, lat_p, m, &ci, &cj);
    double xp, yp, zp;
    query_xyz(lon_p, lat_p, &xp, &yp, &zp, NULL);
    int best = -1;
    double best_d2 = 1e30;
    int nx = m->gh_nx, ny = m->gh_ny;
    for (int dj = -1; dj < 2; dj++) {
        int cj2 = cj + dj;
        if (cj2 < 0 || cj2 >= ny) continue;
        for (int di = -1; di < 2; di++) {
            int ci2 = ci + di;
            if (ci2 < 0)
                ci2 = m->gh_wrap ? ci2 + nx : -1;
            else if (ci2 >= nx)
                ci2 = m->gh_wrap ? ci2 - nx : -1;
            if (ci2 < 0) continue;
            gh_scan_row_range(m, cj2, ci2, ci2, xp, yp, zp, &best, &best_d2);
        }
    }
    return best;
}

/* core/mesh.py:1307 _cap_extents */
static void cap_extents(double lon_p, double lat_p, double lon_n, double lat_n,
                        double *theta_o, double *ext_o)
{
    double lat_pr = lat_p * DEG2RAD;
    double lat_nr = lat_n * DEG2RAD;
    double cosd = (M_SIN(lat_pr) * M_SIN(lat_nr)
                   + M_COS(lat_pr) * M_COS(lat_nr) * M_COS((lon_p - lon_n) * DEG2RAD));
    if (cosd > 1.0) cosd = 1.0;
    if (cosd < -1.0) cosd = -1.0;
    double theta = M_ACOS(cosd);
    if (theta <= 0.0) {
        *theta_o = 0.0;
        *ext_o = 0.0;
        return;
    }
    double theta_deg = theta * RAD2DEG;
    if (fabs(lat_p) + theta_deg >= 90.0) {
        *theta_o = theta_deg;
        *ext_o = 1.0e30;
        return;
    }
    double r = M_SIN(theta) / M_COS(lat_pr);
    if (r >= 1.0) {
        *theta_o = theta_deg;
        *ext_o = 1.0e30;
        return;
    }
    *theta_o = theta_deg;
    *ext_o = M_ASIN(r) * RAD2DEG;
}

/* core/mesh.py:1476 _anchor_provable_rect */
static int anchor_provable_rect(const or_mesh_t *m, double lon_p, double lat_p, int node, int rx, int ry)
{
    double theta_deg, ext_lon;
    cap_extents(lon_p, lat_p, m->nodes_lon[node], m->nodes_lat[node], &theta_deg, &ext_lon);
    if (theta_deg <= 0.0) return 0;
    double cell = m->gh_cell, lon_min = m->gh_lon_min, lat_min = m->gh_lat_min;
    int nx = m->gh_nx, ny = m->gh_ny;
    double dlon = gh_fold_dlon(lon_p, m);
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
    if (m->gh_wrap) {
        if (2 * rx + 1 >= nx) return code;
        if (ext_lon >= INF) return code | 4;
        double w_last = 360.0 - (nx - 1) * cell;
        int ci = ci_raw % nx; /* Python int %: non-negative */
        if (ci < 0) ci += nx;
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
    const double shifts[2] = {-360.0, 360.0};
    for (int s = 0; s < 2; s++) {
        double a = lon_q - ext_lon + shifts[s];
        double b = lon_q + ext_lon + shifts[s];
        if (b >= n_lo && a <= n_hi) return code | 4;
    }
    return code;
}

static inline int pyimod(int a, int n) /* Python int % for n > 0 */
{
    int r = a % n;
    return r < 0 ? r + n : r;
}

/* core/mesh.py:1365 _grid_hash_nearest_widened */
int or_nearest_widened(const or_mesh_t *m, double lon_p, double lat_p)
{
    int best = or_nearest_3x3(m, lon_p, lat_p);
    if (best >= 0) return best;
    int ci, cj;
    gh_cell_of(lon_p, lat_p, m, &ci, &cj);
    double xp, yp, zp;
    query_xyz(lon_p, lat_p, &xp, &yp, &zp, NULL);
    double best_d2 = 1e30;
    int nx = m->gh_nx, ny = m->gh_ny;
    int max_ring = nx > ny ? nx : ny;
    int found_ring = -1;
    for (int r = 2; r <= max_ring; r++) {
        for (int dj = -r; dj < r + 1; dj++) {
            int cj2 = cj + dj;
            if (cj2 < 0 || cj2 >= ny) continue;
            int di_step = (dj == -r || dj == r) ? 1 : 2 * r;
            for (int di = -r; di <= r; di += di_step) {
                int ci2 = ci + di;
                if (m->gh_wrap)
                    ci2 = pyimod(ci2, nx);
                else if (ci2 < 0 || ci2 >= nx)
                    continue;
                gh_scan_row_range(m, cj2, ci2, ci2, xp, yp, zp, &best, &best_d2);
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

/* core/mesh.py:1612 _grid_hash_nearest_exact.  pinned by brute-force chord
 * argmin (tests/test_oracle_nearest.py). */
int or_nearest_exact(const or_mesh_t *m, double lon_p, double lat_p)
{
    double cell = m->gh_cell;
    int nx = m->gh_nx, ny = m->gh_ny, wrap = m->gh_wrap;
    double far = m->gh_far_rows * cell;
    if (lat_p < m->gh_lat_min - far || lat_p > m->gh_lat_min + ny * cell + far)
        return or_nearest_widened(m, lon_p, lat_p);
    int best = or_nearest_3x3(m, lon_p, lat_p);
    int code = 0;
    if (best >= 0) {
        code = anchor_provable_rect(m, lon_p, lat_p, best, 1, 1);
        if (code == 0) return best;
        double theta_deg, ext;
        cap_extents(lon_p, lat_p, m->nodes_lon[best], m->nodes_lat[best], &theta_deg, &ext);
        if (theta_deg > far) return best;
    }
    int ci, cj;
    gh_cell_of(lon_p, lat_p, m, &ci, &cj);
    double xp, yp, zp, cos_lat_p;
    query_xyz(lon_p, lat_p, &xp, &yp, &zp, &cos_lat_p);
    double best_d2 = 1.0e30;
    if (best >= 0) best_d2 = chord2(m, best, xp, yp, zp);
    int rx = 1, ry = 1;
    int short_row = 360.0 * cos_lat_p <= EXACT_SHORT_ROW_CELLS * cell;
    int rx_max;
    if (wrap) {
        rx_max = nx / 2;
        if (rx_max < 1) rx_max = 1;
    } else {
        rx_max = ci > nx - 1 - ci ? ci : nx - 1 - ci;
    }
    int ry_max = cj > ny - 1 - cj ? cj : ny - 1 - cj;
    double abs_lat = lat_p >= 0.0 ? lat_p : -lat_p;
    double dlon_q = gh_fold_dlon(lon_p, m);
    const double INF = 1.0e30;
    for (;;) {
        if (best >= 0 && code == 0) {
            code = anchor_provable_rect(m, lon_p, lat_p, best, rx, ry);
            if (code == 0) return best;
        }
        if (best < 0 || 2.0 * M_ASIN(0.5 * sqrt(best_d2)) * RAD2DEG > far) {
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
                    clear_lon = M_ASIN(M_SIN(dl * DEG2RAD) * M_COS(lat_edge * DEG2RAD)) * RAD2DEG;
                }
                if (clear_lon < clear) clear = clear_lon;
            }
            if (clear >= far) {
                if (best >= 0) return best;
                return or_nearest_widened(m, lon_p, lat_p);
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
            return or_nearest_widened(m, lon_p, lat_p);
        }
        int rx_new = all_x ? rx_max : (grow_x ? rx + 1 : rx);
        int ry_new = grow_y ? ry + 1 : ry;
        int j_lo = cj - ry > 0 ? cj - ry : 0;
        int j_hi = cj + ry < ny - 1 ? cj + ry : ny - 1;
        if (rx_new > rx) {
            if (wrap && 2 * rx_new + 1 >= nx && rx_new - rx > 1) {
                for (int cj2 = j_lo; cj2 < j_hi + 1; cj2++) {
                    int c_lo = pyimod(ci + rx + 1, nx);
                    int c_hi = pyimod(ci - rx - 1, nx);
                    if (c_lo <= c_hi) {
                        gh_scan_row_range(m, cj2, c_lo, c_hi, xp, yp, zp, &best, &best_d2);
                    } else {
                        gh_scan_row_range(m, cj2, c_lo, nx - 1, xp, yp, zp, &best, &best_d2);
                        gh_scan_row_range(m, cj2, 0, c_hi, xp, yp, zp, &best, &best_d2);
                    }
                }
            } else if (!wrap && rx_new - rx > 1) {
                for (int cj2 = j_lo; cj2 < j_hi + 1; cj2++) {
                    if (ci - rx - 1 >= 0) {
                        int c_lo = ci - rx_new > 0 ? ci - rx_new : 0;
                        gh_scan_row_range(m, cj2, c_lo, ci - rx - 1, xp, yp, zp, &best, &best_d2);
                    }
                    if (ci + rx + 1 <= nx - 1) {
                        int c_hi = ci + rx_new < nx - 1 ? ci + rx_new : nx - 1;
                        gh_scan_row_range(m, cj2, ci + rx + 1, c_hi, xp, yp, zp, &best, &best_d2);
                    }
                }
            } else {
                for (int r = rx + 1; r < rx_new + 1; r++) {
                    if (wrap) {
                        int c_left = pyimod(ci - r, nx);
                        int two = (2 * r + 1 <= nx);
                        int c_right = pyimod(ci + r, nx);
                        for (int cj2 = j_lo; cj2 < j_hi + 1; cj2++) {
                            gh_scan_row_range(m, cj2, c_left, c_left, xp, yp, zp, &best, &best_d2);
                            if (two)
                                gh_scan_row_range(m, cj2, c_right, c_right, xp, yp, zp, &best, &best_d2);
                        }
                    } else {
                        int cs[2] = {ci - r, ci + r};
                        for (int s = 0; s < 2; s++) {
                            int c2 = cs[s];
                            if (c2 < 0 || c2 > nx - 1) continue;
                            for (int cj2 = j_lo; cj2 < j_hi + 1; cj2++)
                                gh_scan_row_range(m, cj2, c2, c2, xp, yp, zp, &best, &best_d2);
                        }
                    }
                }
            }
        }
        for (int r = ry + 1; r < ry_new + 1; r++) {
            int cjs[2] = {cj - r, cj + r};
            for (int s = 0; s < 2; s++) {
                int cj2 = cjs[s];
                if (cj2 < 0 || cj2 > ny - 1) continue;
                gh_scan_row_window(m, cj2, ci, rx_new, xp, yp, zp, &best, &best_d2);
            }
        }
        rx = rx_new;
        ry = ry_new;
    }
}

/* Anchor of one evaluation site.  exact = 1: the provably exact search
 * (mesh.exact_nearest, every path); 0: the fused single-particle kernels'
 * guard-off anchor, integrators.py:1471 _fused_site_anchor (widened window);
 * 2: the batch guard-off anchor of core/mesh.py:1896 find_nearest_nodes
 * (3x3 window, an empty window healed with the true nearest node -- the
 * reference asks a KD-tree), used by the unfused paths, the kick and the
 * fused 3-D kernels' rerun. */
int or_site_anchor(const or_mesh_t *m, double lon_p, double lat_p, int exact)
{
    if (exact == 1) return or_nearest_exact(m, lon_p, lat_p);
    if (exact == 2) {
        int b = or_nearest_3x3(m, lon_p, lat_p);
        return b >= 0 ? b : or_nearest_exact(m, lon_p, lat_p);
    }
    return or_nearest_widened(m, lon_p, lat_p);
}

/* ------------------------------------------------------------------------
 * Containment and weights
 * ---------------------------------------------------------------------- */

/* core/mesh.py:2007 _project_to_local_xy (n = 3 or 4 vertices) */
static void project_to_local_xy(const double *lon_n, const double *lat_n, int n,
                                double lon_p, double lat_p,
                                double *x_n, double *y_n, double *x_p, double *y_p)
{
    double lo_min = lon_n[0], lo_max = lon_n[0];
    for (int k = 0; k < n; k++) {
        if (lon_n[k] < lo_min) lo_min = lon_n[k];
        if (lon_n[k] > lo_max) lo_max = lon_n[k];
    }
    double lon0;
    if (lo_max - lo_min > 180.0) {
        double ref = lon_n[0], dsum = 0.0;
        for (int k = 0; k < n; k++) dsum += PYMOD360(lon_n[k] - ref + 180.0) - 180.0;
        lon0 = ref + dsum / n;
    } else {
        double s = 0.0; /* numba np.mean: sequential sum / size */
        for (int k = 0; k < n; k++) s += lon_n[k];
        lon0 = s / n;
    }
    double sl = 0.0;
    for (int k = 0; k < n; k++) sl += lat_n[k];
    double lat0 = sl / n;
    double cos_lat0 = M_COS(lat0 * DEG2RAD);
    for (int k = 0; k < n; k++) {
        double dlon = PYMOD360(lon_n[k] - lon0 + 180.0) - 180.0;
        double dlat = lat_n[k] - lat0;
        x_n[k] = dlon * DEG2RAD * R_EARTH_STANDARD * cos_lat0;
        y_n[k] = dlat * DEG2RAD * R_EARTH_STANDARD;
    }
    double dlon_p = PYMOD360(lon_p - lon0 + 180.0) - 180.0;
    double dlat_p = lat_p - lat0;
    *x_p = dlon_p * DEG2RAD * R_EARTH_STANDARD * cos_lat0;
    *y_p = dlat_p * DEG2RAD * R_EARTH_STANDARD;
}

/* core/mesh.py:2055 _barycentric_weights_tri; returns 0 on the degenerate
 * sentinel [-1, -1, -1] (also written to w) */
static int bary_tri(double px, double py, const double *tx, const double *ty, double *w)
{
    double ax = tx[0], ay = ty[0], bx = tx[1], by = ty[1], cx = tx[2], cy = ty[2];
    double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
    if (fabs(denom) < AREA_TOL) {
        w[0] = w[1] = w[2] = -1.0;
        return 0;
    }
    double w1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / denom;
    double w2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / denom;
    double w3 = 1.0 - w1 - w2;
    w[0] = w1;
    w[1] = w2;
    w[2] = w3;
    return 1;
}

/* core/mesh.py:2077 _wachspress_weights_quad */
static void wachspress_quad(double px, double py, const double *xn, const double *yn, double *out)
{
    double A[4], C[4];
    for (int k = 0; k < 4; k++) {
        int i1 = (k + 1) % 4;
        int im = (k + 3) % 4; /* (k - 1) % 4, Python */
        A[k] = ((xn[k] - px) * (yn[i1] - py) - (xn[i1] - px) * (yn[k] - py));
        C[k] = ((xn[k] - xn[im]) * (yn[i1] - yn[im]) - (xn[i1] - xn[im]) * (yn[k] - yn[im]));
    }
    double s = 0.0;
    for (int k = 0; k < 4; k++) {
        double w = C[k] * A[(k + 1) % 4] * A[(k + 2) % 4];
        out[k] = w;
        s += w;
    }
    for (int k = 0; k < 4; k++) out[k] = out[k] / s;
}

/* core/mesh.py:2120 _polygon_weights: w[0] == -1.0 flags the sentinel */
static void polygon_weights(double px, double py, const double *xn, const double *yn, int nv, double *w)
{
    if (nv == 3) {
        bary_tri(px, py, xn, yn, w);
        return;
    }
    wachspress_quad(px, py, xn, yn, w);
}

static inline int all_ge_eps3(const double *w)
{
    return w[0] >= BARY_EPS && w[1] >= BARY_EPS && w[2] >= BARY_EPS;
}

/* core/mesh.py:2140 _polygon_contains_bool */
static int polygon_contains_bool(const double *lon_n, const double *lat_n, int n, double lon_p, double lat_p)
{
    if (n < 3) return 0;
    double xn[4], yn[4], px, py, w[3];
    project_to_local_xy(lon_n, lat_n, n, lon_p, lat_p, xn, yn, &px, &py);
    if (n == 3) {
        bary_tri(px, py, xn, yn, w);
        return all_ge_eps3(w);
    }
    for (int i = 1; i < n - 1; i++) {
        double tx[3] = {xn[0], xn[i], xn[i + 1]};
        double ty[3] = {yn[0], yn[i], yn[i + 1]};
        bary_tri(px, py, tx, ty, w);
        if (all_ge_eps3(w)) return 1;
    }
    return 0;
}

/* core/mesh.py:2211 _element_contains.  pinned by brute-force containment
 * and closed-form interior points (tests/test_oracle_containment.py). */
int or_element_contains(const or_mesh_t *m, int c, double lon_p, double lat_p)
{
    const int32_t *nodes = m->elem_nodes + (int64_t)c * m->max_v;
    int mv = m->max_v;
    int mask[4];
    for (int j = 0; j < mv; j++) mask[j] = nodes[j] >= 0;
    if (mv == 4 && nodes[3] == nodes[0] && nodes[0] >= 0) mask[3] = 0;
    int n_valid = 0;
    for (int j = 0; j < mv; j++) n_valid += mask[j];
    if (n_valid < 3) return 0;
    double lon_n[4], lat_n[4];
    int idx = 0;
    for (int j = 0; j < mv; j++) {
        if (mask[j]) {
            lon_n[idx] = m->nodes_lon[nodes[j]];
            lat_n[idx] = m->nodes_lat[nodes[j]];
            idx++;
        }
    }
    double lon_ref = lon_n[0];
    double dmax = 0.0, dmin = 0.0;
    for (int k = 0; k < n_valid; k++) {
        double d = PYMOD360(lon_n[k] - lon_ref + 180.0) - 180.0;
        if (k == 0 || d > dmax) dmax = d;
        if (k == 0 || d < dmin) dmin = d;
    }
    double dlon_p = PYMOD360(lon_p - lon_ref + 180.0) - 180.0;
    if (dlon_p > dmax || dlon_p < dmin) return 0;
    double lamax = lat_n[0], lamin = lat_n[0];
    for (int k = 1; k < n_valid; k++) {
        if (lat_n[k] > lamax) lamax = lat_n[k];
        if (lat_n[k] < lamin) lamin = lat_n[k];
    }
    if (lat_p > lamax || lat_p < lamin) return 0;
    return polygon_contains_bool(lon_n, lat_n, n_valid, lon_p, lat_p);
}

/* core/mesh.py:2393 _find_containing_element_local; returns elem or -1 */
int or_find_containing_local(const or_mesh_t *m, double lon_p, double lat_p, int nn)
{
    int count = m->n2e_count[nn];
    if (count == 0) return -1;
    const int32_t *row = m->n2e + (int64_t)nn * m->max_conn;
    for (int k = 0; k < count; k++) {
        int c = row[k];
        if (c < 0) continue;
        if (or_element_contains(m, c, lon_p, lat_p)) return c;
    }
    for (int k = 0; k < count; k++) {
        int e = row[k];
        if (e < 0) continue;
        for (int j = 0; j < m->max_v; j++) {
            int v = m->elem_nodes[(int64_t)e * m->max_v + j];
            if (v < 0 || v == nn) continue;
            int count_v = m->n2e_count[v];
            const int32_t *row_v = m->n2e + (int64_t)v * m->max_conn;
            for (int kk = 0; kk < count_v; kk++) {
                int c = row_v[kk];
                if (c < 0) continue;
                int touches = 0;
                for (int jj = 0; jj < m->max_v; jj++) {
                    if (m->elem_nodes[(int64_t)c * m->max_v + jj] == nn) {
                        touches = 1;
                        break;
                    }
                }
                if (touches) continue;
                if (or_element_contains(m, c, lon_p, lat_p)) return c;
            }
        }
    }
    return -1;
}

/* ------------------------------------------------------------------------
 * Barycentric interpolation with the nested time blend
 * ---------------------------------------------------------------------- */

/* Compacted vertex ring of element e (interpolate.py:1376-1386); returns nv */
static int elem_ring(const or_mesh_t *m, int e, int *verts)
{
    int nv = 0;
    for (int k = 0; k < m->max_v; k++) {
        int nd = m->elem_nodes[(int64_t)e * m->max_v + k];
        if (nd < 0) continue;
        if (nv > 0 && nd == verts[0]) continue;
        verts[nv++] = nd;
    }
    return nv;
}

/* core/mesh.py:2489 _walk_containing_element: (element, 1) or (-1, 0) */
static int walk_element(const or_mesh_t *m, double lon_p, double lat_p, int start)
{
    int e = start;
    if (e < 0 || e >= m->n_elems) return -1;
    int prev = -1;
    for (int it = 0; it < 256; it++) {
        int verts[4];
        int nv = elem_ring(m, e, verts);
        if (nv < 3) return -1;
        double lon_n[4], lat_n[4], x[4], y[4], px, py;
        for (int k = 0; k < nv; k++) {
            lon_n[k] = m->nodes_lon[verts[k]];
            lat_n[k] = m->nodes_lat[verts[k]];
        }
        project_to_local_xy(lon_n, lat_n, nv, lon_p, lat_p, x, y, &px, &py);
        double S = 0.0;
        for (int k = 0; k < nv; k++) {
            int j = (k + 1) % nv;
            S += x[k] * y[j] - x[j] * y[k];
        }
        double absS = fabs(S);
        if (absS < AREA_TOL) return -1;
        double sgn = S > 0.0 ? 1.0 : -1.0;
        double worst = 1.0e100;
        int worst_k = -1;
        for (int k = 0; k < nv; k++) {
            int j = (k + 1) % nv;
            double A = (x[j] - x[k]) * (py - y[k]) - (px - x[k]) * (y[j] - y[k]);
            double w = A * sgn / absS;
            if (w < worst) {
                worst = w;
                worst_k = k;
            }
        }
        if (worst >= -1e-6) {
            if (worst >= 1e-2) return e;
            return or_element_contains(m, e, lon_p, lat_p) ? e : -1;
        }
        int nb = m->elem_neighbors[(int64_t)e * m->max_v + worst_k];
        if (nb < 0 || nb == prev) return -1;
        prev = e;
        e = nb;
    }
    return -1;
}

/* core/mesh.py:2582 _loop_walk_containing for one query: the walk, else
 * the local search around the widened grid-hash nearest node */
int or_walk_locate(const or_mesh_t *m, double lon_p, double lat_p, int start)
{
    if (start >= 0) {
        int e = walk_element(m, lon_p, lat_p, start);
        if (e >= 0) return e;
    }
    int nn = or_nearest_widened(m, lon_p, lat_p);
    if (nn < 0) return -1;
    return or_find_containing_local(m, lon_p, lat_p, nn);
}

/* Vertex weights of (xp, yp) in element e: the projection of
 * interpolate.py:1387-1404; returns 0 when the tri sentinel fires. */
static int elem_vertex_weights(const or_mesh_t *m, int e, double xp, double yp,
                               int *verts, int *nv_o, double *w)
{
    int nv = elem_ring(m, e, verts);
    *nv_o = nv;
    if (nv < 3) return 0;
    const double DEG2R = 0.017453292519943295;
    double R = m->r_earth;
    double lon_ref = m->nodes_lon[verts[0]];
    double dsum = 0.0, lat0 = 0.0;
    for (int k = 0; k < nv; k++) {
        dsum += PYMOD360(m->nodes_lon[verts[k]] - lon_ref + 180.0) - 180.0;
        lat0 += m->nodes_lat[verts[k]];
    }
    double lon0 = lon_ref + dsum / nv;
    lat0 /= nv;
    double cosl = M_COS(lat0 * DEG2R);
    double vx[4], vy[4];
    for (int k = 0; k < nv; k++) {
        double dlon = PYMOD360(m->nodes_lon[verts[k]] - lon0 + 180.0) - 180.0;
        vx[k] = dlon * DEG2R * R * cosl;
        vy[k] = (m->nodes_lat[verts[k]] - lat0) * DEG2R * R;
    }
    double dlon_p = PYMOD360(xp - lon0 + 180.0) - 180.0;
    double px = dlon_p * DEG2R * R * cosl;
    double py = (yp - lat0) * DEG2R * R;
    polygon_weights(px, py, vx, vy, nv, w);
    if (w[0] == -1.0) return 0;
    return 1;
}

/* Blend of one vertex value (interpolate.py:1419-1437); `off` indexes the
 * layer plane */
static inline double vblend(const void *a, const void *b, const void *t, int dtype,
                            int64_t idx, double alpha0, double alpha)
{
    double beta0 = 1.0 - alpha0, beta = 1.0 - alpha;
    if (alpha < 1e-12) {
        if (alpha0 == 0.0) return fld(a, dtype, idx);
        return beta0 * fld(a, dtype, idx) + alpha0 * fld(b, dtype, idx);
    } else if (alpha > 1.0 - 1e-12) {
        return fld(t, dtype, idx);
    } else if (alpha0 == 0.0) {
        return beta * fld(a, dtype, idx) + alpha * fld(t, dtype, idx);
    }
    double a0 = beta0 * fld(a, dtype, idx) + alpha0 * fld(b, dtype, idx);
    return beta * a0 + alpha * fld(t, dtype, idx);
}

/* interpolate.py:1341 _interp_barycentric_core_tinterp (layer < 0: 2-D)
 * and :1444 _interp_barycentric_core_tinterp_3d.  Returns ok. */
/* ehint >= 0: a containing element validated for this position (the edge
 * walk), the search is skipped (interpolate.py:1266) */
static int interp_bary_hint(const or_mesh_t *m, const or_fields_t *f, double xp, double yp,
                            int nn, double time_dir, int layer, double alpha,
                            double *U_o, double *V_o, int *elem_o, int ehint)
{
    *U_o = 0.0;
    *V_o = 0.0;
    int e = ehint >= 0 ? ehint : or_find_containing_local(m, xp, yp, nn);
    if (elem_o) *elem_o = e;
    if (e < 0) return 0;
    int verts[4], nv;
    double w[4];
    if (!elem_vertex_weights(m, e, xp, yp, verts, &nv, w)) return 0;
    int64_t off = layer > 0 ? (int64_t)layer * m->n_nodes : 0;
    double U = 0.0, V = 0.0;
    for (int k = 0; k < nv; k++) {
        int64_t idx = off + verts[k];
        double un = vblend(f->u_a, f->u_b, f->u_t, f->dtype, idx, f->alpha0, alpha);
        double vn = vblend(f->v_a, f->v_b, f->v_t, f->dtype, idx, f->alpha0, alpha);
        U += w[k] * un;
        V += w[k] * vn;
    }
    *U_o = time_dir * U;
    *V_o = time_dir * V;
    return 1;
}

int or_interp_bary_tinterp(const or_mesh_t *m, const or_fields_t *f, double xp, double yp,
                           int nn, double time_dir, int layer, double alpha,
                           double *U_o, double *V_o, int *elem_o)
{
    return interp_bary_hint(m, f, xp, yp, nn, time_dir, layer, alpha, U_o, V_o, elem_o, -1);
}

/* interpolate.py:207 _interp_leastsq_core_tinterp (== :52
 * _interp_leastsq_core for a plain field): a plane through the blended
 * element-centre values of the anchor's 1-ring in particle-centred metric
 * coordinates, the antimeridian gate on each centre, Sarrus determinant with
 * the scale-relative singularity guard, Cramer's rule for the constant term.
 * `layer` selects the (E,) plane of a layered field (< 0: 2-D). */
int or_interp_lsq_tinterp(const or_mesh_t *m, const or_fields_t *f, double xp, double yp,
                          int nn, double time_dir, int layer, double alpha,
                          double *U_o, double *V_o)
{
    const double DEG2R = 0.017453292519943295;
    *U_o = 0.0;
    *V_o = 0.0;
    int count = m->n2e_count[nn];
    if (count < 1) return 0;
    const int32_t *row = m->n2e + (int64_t)nn * m->max_conn;
    double z0 = M_COS(yp * DEG2R) * SLM;
    double xp_m = xp * z0;
    double yp_m = yp * SLM;
    double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
    double fUn = 0.0, f1Un = 0.0, f2Un = 0.0, fVn = 0.0, f1Vn = 0.0, f2Vn = 0.0;
    int64_t off = layer > 0 ? (int64_t)layer * m->n_elems : 0;
    for (int n = 0; n < count; n++) {
        int e = row[n];
        double ex = m->elem_center_lon[e] * z0 - xp_m;
        double dlon_raw = m->elem_center_lon[e] - xp;
        if (dlon_raw > 180.0)
            ex -= 360.0 * z0;
        else if (dlon_raw < -180.0)
            ex += 360.0 * z0;
        double ey = m->elem_center_lat[e] * SLM - yp_m;
        double u = vblend(f->u_a, f->u_b, f->u_t, f->dtype, off + e, f->alpha0, alpha);
        double v = vblend(f->v_a, f->v_b, f->v_t, f->dtype, off + e, f->alpha0, alpha);
        double uu = time_dir * u;
        double vv = time_dir * v;
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
    double a21 = a12, a31 = a13, a32 = a23;
    double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                  a12 * a21 * a33 - a11 * a23 * a32);
    if (fabs(det) < 1e-30 || fabs(det) < 1e-12 * fabs(a11 * a22 * a33)) return 0;
    double inv = 1.0 / fabs(det);
    double b11 = a22 * a33 - a23 * a32;
    double b12 = a13 * a32 - a12 * a33;
    double b13 = a12 * a23 - a13 * a22;
    *U_o = (b11 * fUn + b12 * f1Un + b13 * f2Un) * inv;
    *V_o = (b11 * fVn + b12 * f1Vn + b13 * f2Vn) * inv;
    return 1;
}

/* ------------------------------------------------------------------------
 * Vertical (z-level)
 * ---------------------------------------------------------------------- */

/* numerics/vertical.py:40 find_layer */
static int find_layer(const double *zl, int K, double zp, double bottom)
{
    if (K < 2) return 0;
    if (zp <= zl[0]) return 0;
    int max_layer = K - 2;
    for (int k = 0; k < K - 1; k++) {
        if (zl[k + 1] > bottom) {
            max_layer = k - 1 > 0 ? k - 1 : 0;
            break;
        }
    }
    if (zp >= bottom || zp >= zl[max_layer + 1]) return max_layer;
    for (int k = 0; k < max_layer + 1; k++)
        if (zp < zl[k + 1]) return k;
    return max_layer;
}

/* numerics/vertical.py:1043 _tblend_scalar */
static inline double tblend(double a, double b, double t, double alpha0, double alpha)
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

static inline double wlev(const or_fields_t *f, int64_t idx, double alpha)
{
    return tblend(fld(f->w_a, f->dtype, idx), fld(f->w_b, f->dtype, idx), fld(f->w_t, f->dtype, idx),
                  f->alpha0, alpha);
}

/* numerics/vertical.py:1065 _interp_w_linear_single_tinterp (reduces to
 * :223 _interp_w_linear_single for a plain field) */
static double w_linear(const or_mesh_t *m, const or_fields_t *f, double alpha, double zp, int node,
                       double time_dir)
{
    const double *zl = m->z_levels;
    int K = m->n_levels;
    int64_t N = m->n_nodes;
    if (K < 2) return time_dir * wlev(f, node, alpha);
    if (zp <= zl[0]) return time_dir * wlev(f, node, alpha);
    if (zp >= zl[K - 1]) return time_dir * wlev(f, (int64_t)(K - 1) * N + node, alpha);
    int k = K - 2;
    for (int j = 0; j < K - 1; j++) {
        if (zp < zl[j + 1]) {
            k = j;
            break;
        }
    }
    double dz = zl[k + 1] - zl[k];
    if (dz < 1e-12) return time_dir * wlev(f, (int64_t)k * N + node, alpha);
    double a = (zp - zl[k]) / dz;
    double w_lo = wlev(f, (int64_t)k * N + node, alpha);
    double w_hi = wlev(f, (int64_t)(k + 1) * N + node, alpha);
    return time_dir * ((1.0 - a) * w_lo + a * w_hi);
}

/* core/mesh.py:1974 _seam_safe_lon0 */
static double seam_safe_lon0(const or_mesh_t *m, const int *verts, int n)
{
    double first = m->nodes_lon[verts[0]];
    double lo = first, hi = first, s = 0.0;
    for (int v = 0; v < n; v++) {
        double x = m->nodes_lon[verts[v]];
        s += x;
        if (x < lo) lo = x;
        if (x > hi) hi = x;
    }
    if (hi - lo > 180.0) {
        double d = 0.0;
        for (int v = 0; v < n; v++) d += PYMOD360(m->nodes_lon[verts[v]] - first + 180.0) - 180.0;
        return first + d / n;
    }
    return s / n;
}

/* numerics/vertical.py:745 _interp_w_bary_single_tinterp (== :604
 * _interp_w_barycentric_single on a plain field): the first fan
 * sub-triangle of a 1-ring element that holds the point; bottom-aware
 * linear W per vertex; fallback = linear at the anchor. */
static double w_bary(const or_mesh_t *m, const or_fields_t *f, double alpha, double xp, double yp,
                     double zp, int node, double time_dir)
{
    const double DEG2R = 0.017453292519943295;
    int n_el = m->n2e_count[node];
    if (n_el == 0) {
        double zc = zp < m->depth[node] ? zp : m->depth[node];
        return w_linear(m, f, alpha, zc, node, time_dir);
    }
    const int32_t *row = m->n2e + (int64_t)node * m->max_conn;
    for (int k = 0; k < n_el; k++) {
        int e = row[k];
        if (e < 0) continue;
        int verts[4], nv = 0;
        for (int v = 0; v < m->max_v; v++) {
            int nd = m->elem_nodes[(int64_t)e * m->max_v + v];
            if (nd < 0) continue;
            if (nv > 0 && nd == verts[0]) continue;
            verts[nv++] = nd;
        }
        if (nv < 3) continue;
        double lon0 = seam_safe_lon0(m, verts, nv);
        double lat0 = 0.0;
        for (int v = 0; v < nv; v++) lat0 += m->nodes_lat[verts[v]];
        lat0 /= nv;
        double cos_lat0 = M_COS(lat0 * DEG2R);
        const double R = 6371000.0;
        double vx[4], vy[4];
        for (int v = 0; v < nv; v++) {
            double dlon = PYMOD360(m->nodes_lon[verts[v]] - lon0 + 180.0) - 180.0;
            double dlat = m->nodes_lat[verts[v]] - lat0;
            vx[v] = dlon * DEG2R * R * cos_lat0;
            vy[v] = dlat * DEG2R * R;
        }
        double dlon_p = PYMOD360(xp - lon0 + 180.0) - 180.0;
        double dlat_p = yp - lat0;
        double px = dlon_p * DEG2R * R * cos_lat0;
        double py = dlat_p * DEG2R * R;
        for (int t = 0; t < nv - 2; t++) {
            int i1 = t + 1, i2 = t + 2;
            double ax = vx[0], ay = vy[0], bx = vx[i1], by = vy[i1], cx = vx[i2], cy = vy[i2];
            double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
            if (fabs(denom) < 1e-14) continue;
            double w1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / denom;
            double w2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / denom;
            double w3 = 1.0 - w1 - w2;
            const double eps = -1e-6;
            if (w1 >= eps && w2 >= eps && w3 >= eps) {
                if (nv == 4) {
                    double wts[4];
                    wachspress_quad(px, py, vx, vy, wts);
                    double acc = 0.0;
                    for (int vi = 0; vi < 4; vi++) {
                        int nd = verts[vi];
                        double zn = zp < m->depth[nd] ? zp : m->depth[nd];
                        acc += wts[vi] * w_linear(m, f, alpha, zn, nd, time_dir);
                    }
                    return acc;
                }
                int n0 = verts[0], n1 = verts[i1], n2 = verts[i2];
                double z0n = zp < m->depth[n0] ? zp : m->depth[n0];
                double z1n = zp < m->depth[n1] ? zp : m->depth[n1];
                double z2n = zp < m->depth[n2] ? zp : m->depth[n2];
                double wv0 = w_linear(m, f, alpha, z0n, n0, time_dir);
                double wv1 = w_linear(m, f, alpha, z1n, n1, time_dir);
                double wv2 = w_linear(m, f, alpha, z2n, n2, time_dir);
                return w1 * wv0 + w2 * wv1 + w3 * wv2;
            }
        }
    }
    double zc = zp < m->depth[node] ? zp : m->depth[node];
    return w_linear(m, f, alpha, zc, node, time_dir);
}

/* numerics/vertical.py:493 _interp_w_leastsq_single on the stage field
 * (interface values blended in time, as the unfused path's materialised
 * lerp): the unique 1-ring nodes (centre first, then each element's vertices
 * in order, at most 64 -- vertical.py:446 _collect_neighbour_nodes), W
 * linear in depth at each (clamped to that node's bottom), a plane fit
 * through them evaluated at the particle; fewer than 3 nodes or a singular
 * fit fall back to linear W at the anchor. */
static double w_leastsq(const or_mesh_t *m, const or_fields_t *f, double alpha, double xp, double yp,
                        double zp, int node, double time_dir)
{
    const double DEG2R = 0.017453292519943295;
    int buf[64];
    int count = 0;
    buf[count++] = node;
    int n_el = m->n2e_count[node];
    const int32_t *row = m->n2e + (int64_t)node * m->max_conn;
    for (int k = 0; k < n_el; k++) {
        int e = row[k];
        if (e < 0) continue;
        for (int v = 0; v < m->max_v; v++) {
            int nd = m->elem_nodes[(int64_t)e * m->max_v + v];
            if (nd < 0) continue;
            int found = 0;
            for (int j = 0; j < count; j++)
                if (buf[j] == nd) {
                    found = 1;
                    break;
                }
            if (!found && count < 64) buf[count++] = nd;
        }
    }
    if (count < 3) {
        double zc = zp < m->depth[node] ? zp : m->depth[node];
        return w_linear(m, f, alpha, zc, node, time_dir);
    }
    double z0 = M_COS(yp * DEG2R) * SLM;
    if (z0 < 1.0) z0 = 1.0;
    double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
    double fw = 0.0, f1w = 0.0, f2w = 0.0;
    for (int i = 0; i < count; i++) {
        int nd = buf[i];
        double bottom = m->depth[nd];
        double w_val = w_linear(m, f, alpha, zp > bottom ? bottom : zp, nd, time_dir);
        double dlon = PYMOD360(m->nodes_lon[nd] - xp + 180.0) - 180.0;
        double ex = dlon * z0;
        double ey = (m->nodes_lat[nd] - yp) * SLM;
        a11 += 1.0;
        a12 += ex;
        a13 += ey;
        a22 += ex * ex;
        a23 += ex * ey;
        a33 += ey * ey;
        fw += w_val;
        f1w += w_val * ex;
        f2w += w_val * ey;
    }
    double a21 = a12, a31 = a13, a32 = a23;
    double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                  a12 * a21 * a33 - a11 * a23 * a32);
    if (fabs(det) < 1e-30 || fabs(det) < 1e-12 * fabs(a11 * a22 * a33)) {
        double zc = zp < m->depth[node] ? zp : m->depth[node];
        return w_linear(m, f, alpha, zc, node, time_dir);
    }
    double inv = 1.0 / fabs(det);
    double b11 = a22 * a33 - a23 * a32;
    double b12 = a13 * a32 - a12 * a33;
    double b13 = a12 * a23 - a13 * a22;
    return (b11 * fw + b12 * f1w + b13 * f2w) * inv;
}

/* numerics/vertical.py:1148 _interp_w_idw_single on the stage field: the
 * unique 1-ring nodes as for the least-squares W, W linear in depth at each
 * (clamped to that node's bottom), weights 1 / d^2 in the local metric
 * (z0 = max(cos(lat) * 111120, 1), dateline-safe dlon); a node within 1 m
 * returns its own W; an empty weight sum falls back to linear W at the
 * anchor.  pinned by the reference's on-node test (test_w_spatial_interp.py
 * TestIDW), the 1 m coincidence radius and the inverse-square weights of an
 * equatorial lattice whose ring distances are known in closed form. */
static int collect_neighbours(const or_mesh_t *m, int node, int *buf);

static double w_idw(const or_mesh_t *m, const or_fields_t *f, double alpha, double xp, double yp, double zp,
                    int node, double time_dir)
{
    const double DEG2R = 0.017453292519943295;
    const double EPS_DIST = 1.0;
    int buf[64];
    int count = collect_neighbours(m, node, buf);
    double zc = zp < m->depth[node] ? zp : m->depth[node];
    if (count < 1) return w_linear(m, f, alpha, zc, node, time_dir);
    double z0 = M_COS(yp * DEG2R) * SLM;
    if (z0 < 1.0) z0 = 1.0;
    double sum_w = 0.0, sum_wv = 0.0;
    for (int i = 0; i < count; i++) {
        int nd = buf[i];
        double bottom = m->depth[nd];
        double w_val = w_linear(m, f, alpha, zp > bottom ? bottom : zp, nd, time_dir);
        double dlon = PYMOD360(m->nodes_lon[nd] - xp + 180.0) - 180.0;
        double dx = dlon * z0;
        double dy = (m->nodes_lat[nd] - yp) * SLM;
        double d2 = dx * dx + dy * dy;
        if (d2 < EPS_DIST * EPS_DIST) return w_val;
        double w = 1.0 / d2;
        sum_w += w;
        sum_wv += w * w_val;
    }
    if (sum_w < 1e-30) return w_linear(m, f, alpha, zc, node, time_dir);
    return sum_wv / sum_w;
}

/* ------------------------------------------------------------------------
 * Counter RNG and kicks (numerics/diffusion.py)
 * ---------------------------------------------------------------------- */

/* diffusion.py:181 _splitmix64 */
static inline uint64_t splitmix64(uint64_t x)
{
    x = x + 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x = x ^ (x >> 31);
    return x;
}

/* diffusion.py:194 _philox_uint64 (a keyed SplitMix64 chain) */
uint64_t or_hash5(uint64_t k0, uint64_t k1, uint64_t k2, uint64_t k3, uint64_t k4)
{
    uint64_t h = k0;
    h = splitmix64(h ^ k1);
    h = splitmix64(h ^ k2);
    h = splitmix64(h ^ k3);
    h = splitmix64(h ^ k4);
    return h;
}

/* diffusion.py:219 + :234 _uint64_to_two_uniforms / _normal_pair */
void or_normal_pair(uint64_t seed, uint64_t real_id, uint64_t pid, uint64_t step, uint64_t comp,
                    double *g1, double *g2)
{
    const double INV_2_32 = 1.0 / 4294967296.0;
    const double HALF_INV_2_32 = 0.5 * INV_2_32;
    const double TWO_PI = 6.283185307179586; /* 2.0 * math.pi */
    uint64_t h = or_hash5(seed, real_id, pid, step, comp);
    double high = (double)((h >> 32) & 0xFFFFFFFFULL);
    double low = (double)(h & 0xFFFFFFFFULL);
    double u1 = high * INV_2_32 + HALF_INV_2_32;
    double u2 = low * INV_2_32 + HALF_INV_2_32;
    double r = sqrt(-2.0 * M_LOG(u1));
    *g1 = r * M_COS(TWO_PI * u2);
    *g2 = r * M_SIN(TWO_PI * u2);
}

/* beaching.py:46 _u01 */
double or_u01(uint64_t seed, uint64_t real_id, uint64_t pid, uint64_t step, uint64_t comp)
{
    uint64_t h = or_hash5(seed, real_id, pid, step, comp);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

#define R_METRIC (111120.0 / 0.017453292519943295) /* diffusion.py:532 */

/* diffusion.py:539 sphere_displace */
static void sphere_displace(double lon, double lat, double dx_m, double dy_m, double *lon_o, double *lat_o)
{
    double dist = sqrt(dx_m * dx_m + dy_m * dy_m);
    if (dist == 0.0) {
        *lon_o = lon;
        *lat_o = lat;
        return;
    }
    double lam = lon * DEG2RAD, phi = lat * DEG2RAD;
    double cl = M_COS(lam), sl = M_SIN(lam), cp = M_COS(phi), sp = M_SIN(phi);
    double delta = dist / R_METRIC;
    double c = M_COS(delta);
    double s = M_SIN(delta) / dist;
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
    double lat2 = M_ASIN(z2) / DEG2RAD;
    double lon2 = M_ATAN2(y2, x2) / DEG2RAD;
    lon2 = PYMOD360(lon2 + 180.0) - 180.0;
    *lon_o = lon2;
    *lat_o = lat2;
}

/* ------------------------------------------------------------------------
 * The step
 * ---------------------------------------------------------------------- */

static inline double wrap_lon(double x) { return PYMOD360(x + 180.0) - 180.0; }

static inline double clamp_lat(double y)
{
    if (y > 90.0) return 90.0;
    if (y < -90.0) return -90.0;
    return y;
}

static inline double z0_of(double lat)
{
    double z0 = M_COS(lat * DEG2RAD) * SLM;
    if (z0 < 1.0) z0 = 1.0;
    return z0;
}

/* u, v (and W in 3-D) at one evaluation site: the anchor, the particle's
 * layer at that anchor (vertical.py:40), the layer-sliced barycentric u/v
 * (interpolate.py:1444) and W by the configured method (integrators.py:1260
 * _compute_w).  Returns ok of the u/v interpolation. */
static int eval_site(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a, int stage,
                     double lon, double lat, double zp, int *nn_o, double *U, double *V, double *W,
                     int want_w, int ehint)
{
    int nn = or_site_anchor(m, lon, lat, a->exact);
    *nn_o = nn;
    int layer = -1;
    if (a->is3d) layer = find_layer(m->z_levels, m->n_levels, zp, m->depth[nn]);
    int ok;
    if (f->uv_layout == 1)
        ok = or_interp_lsq_tinterp(m, f, lon, lat, nn, a->time_dir, layer, a->alphas[stage], U, V);
    else
        ok = interp_bary_hint(m, f, lon, lat, nn, a->time_dir, layer, a->alphas[stage], U, V, NULL, ehint);
    *W = 0.0;
    if (a->is3d && want_w) {
        if (a->w_method == 0)
            *W = w_linear(m, f, a->alphas[stage], zp, nn, a->time_dir);
        else if (a->w_method == 1)
            *W = w_leastsq(m, f, a->alphas[stage], lon, lat, zp, nn, a->time_dir);
        else if (a->w_method == 3)
            *W = w_idw(m, f, a->alphas[stage], lon, lat, zp, nn, a->time_dir);
        else
            *W = w_bary(m, f, a->alphas[stage], lon, lat, zp, nn, a->time_dir);
    }
    return ok;
}

static inline double clamp_z(const or_mesh_t *m, double z, int nn)
{
    if (z < m->z_levels[0]) z = m->z_levels[0];
    double bottom = m->depth[nn];
    if (z > bottom) z = bottom;
    return z;
}

/* The deterministic step of one alive particle.  Euler follows
 * integrators.py:4997 _euler_step_unfused, RK2 :2522/:3648 (the fused
 * 2-D/3-D barycentric kernels, bit-identical to :5380), RK4 :2642 (2-D) and
 * :5760 _rk4_step_unfused (3-D).  In place; returns 0 alive, else the
 * failure code. */
static int det_step(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a,
                    double *lon, double *lat, double *z, int32_t *st, int32_t *nn, int8_t *fr,
                    double *Uo, double *Vo, double *Wo, int32_t *elem)
{
    const double dt = a->dt;
    const double lon0 = *lon, lat0 = *lat, zz0 = *z;
    int node_i;
    double U1, V1, W1;
    /* elem_search walk (integrators.py:5020, :5415, :5792): each site's
     * element walked from the previous site's, the carried one first */
    const int e1 = a->walk ? or_walk_locate(m, lon0, lat0, *elem) : -1;
    int e_last = e1;
    int ok = eval_site(m, f, a, 0, lon0, lat0, zz0, &node_i, &U1, &V1, &W1, 1, e1);
    *Uo = U1;
    *Vo = V1;
    *Wo = W1;
    double lon_new = lon0, lat_new = lat0, z_new = zz0;
    if (ok) {
        if (a->order == 1) {
            lon_new = wrap_lon(lon0 + dt * U1 / z0_of(lat0));
            lat_new = clamp_lat(lat0 + dt * V1 / SLM);
            z_new = zz0 + W1 * dt;
        } else if (a->order == 2) {
            double z0 = z0_of(lat0);
            double lon_mid = wrap_lon(lon0 + 0.5 * dt * U1 / z0);
            double lat_mid = clamp_lat(lat0 + 0.5 * dt * V1 / SLM);
            double z_mid = zz0 + W1 * 0.5 * dt;
            int nn_mid;
            double U2, V2, W2;
            if (a->walk) e_last = or_walk_locate(m, lon_mid, lat_mid, e_last);
            ok = eval_site(m, f, a, 1, lon_mid, lat_mid, z_mid, &nn_mid, &U2, &V2, &W2, 1, e_last);
            if (ok) {
                double z0m = z0_of(lat_mid);
                lon_new = wrap_lon(lon0 + dt * U2 / z0m);
                lat_new = clamp_lat(lat0 + dt * V2 / SLM);
                z_new = zz0 + W2 * dt;
            }
        } else {
            double dlon1 = dt * U1 / z0_of(lat0);
            double dlat1 = dt * V1 / SLM;
            double dz1 = dt * W1;
            double lon2 = wrap_lon(lon0 + 0.5 * dlon1);
            double lat2 = clamp_lat(lat0 + 0.5 * dlat1);
            double z2 = zz0 + 0.5 * dz1;
            int nn2, nn3, nn4;
            double U2, V2, W2, U3, V3, W3, U4, V4, W4;
            if (a->walk) e_last = or_walk_locate(m, lon2, lat2, e_last);
            ok = eval_site(m, f, a, 1, lon2, lat2, z2, &nn2, &U2, &V2, &W2, 1, e_last);
            if (ok) {
                double dlon2 = dt * U2 / z0_of(lat2);
                double dlat2 = dt * V2 / SLM;
                double dz2 = dt * W2;
                double lon3 = wrap_lon(lon0 + 0.5 * dlon2);
                double lat3 = clamp_lat(lat0 + 0.5 * dlat2);
                double z3 = zz0 + 0.5 * dz2;
                if (a->walk) e_last = or_walk_locate(m, lon3, lat3, e_last);
                ok = eval_site(m, f, a, 2, lon3, lat3, z3, &nn3, &U3, &V3, &W3, 1, e_last);
                if (ok) {
                    double dlon3 = dt * U3 / z0_of(lat3);
                    double dlat3 = dt * V3 / SLM;
                    double dz3 = dt * W3;
                    double lon4 = wrap_lon(lon0 + dlon3);
                    double lat4 = clamp_lat(lat0 + dlat3);
                    double z4 = zz0 + dz3;
                    if (a->walk) e_last = or_walk_locate(m, lon4, lat4, e_last);
                    ok = eval_site(m, f, a, 3, lon4, lat4, z4, &nn4, &U4, &V4, &W4, 1, e_last);
                    if (ok) {
                        double dlon4 = dt * U4 / z0_of(lat4);
                        double dlat4 = dt * V4 / SLM;
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
        *st = ST_OUTSIDE;
        *fr = FR_INTERP;
        *nn = node_i;
        if (a->walk) *elem = e1;
        return FR_INTERP;
    }
    int nn_new = or_site_anchor(m, lon_new, lat_new, a->exact);
    int e = a->walk ? or_walk_locate(m, lon_new, lat_new, e_last)
                    : or_find_containing_local(m, lon_new, lat_new, nn_new);
    if (e < 0) {
        *st = ST_OUTSIDE;
        *fr = FR_OOB;
        *nn = node_i;
        if (a->walk) *elem = e1;
        return FR_OOB;
    }
    *lon = lon_new;
    *lat = lat_new;
    if (a->is3d) *z = clamp_z(m, z_new, nn_new);
    *nn = nn_new;
    if (a->walk) *elem = e;
    return 0;
}

/* ---------------------------------------------------------------------
 * rk45: integrators.py:6233 _rk45_kernel_2d -- per-particle adaptive
 * Dormand-Prince 5(4) over exactly dt (2-D, u/v at element centres by least
 * squares, steady or time-blended), then the re-validation of rk45_step
 * (:6586-6611).  The tableau is Dormand & Prince (1980); the error weights
 * are b5 - b4 as exact fractions.  No reference golden holds an rk45 run:
 * pinned by the uniform-flow closed form (one exact substep) and by agreeing
 * with RK4 on 64 substeps at a tight tolerance, steady and blended, forward
 * and backward (tests/test_oracle_rk45.py).
 * --------------------------------------------------------------------- */
#define DP_C2 (1.0 / 5.0)
#define DP_C3 (3.0 / 10.0)
#define DP_C4 (4.0 / 5.0)
#define DP_C5 (8.0 / 9.0)

/* a trial position: lon wrapped only when it left [-180, 180), lat clamped
 * (integrators.py:6182 _norm_lonlat) */
static void norm_lonlat(double *lon, double *lat)
{
    if (*lon >= 180.0 || *lon < -180.0) *lon = PYMOD360(*lon + 180.0) - 180.0;
    if (*lat > 90.0) *lat = 90.0;
    else if (*lat < -90.0) *lat = -90.0;
}

/* integrators.py:6197 _rk45_eval: velocity and degree rates at (lon, lat)
 * with the stage fraction alpha */
static int rk45_eval(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a, double lon, double lat,
                     double alpha, double *U, double *V, double *kx, double *ky)
{
    int nn = or_site_anchor(m, lon, lat, a->exact == 1 ? 1 : 0);
    *U = *V = *kx = *ky = 0.0;
    if (nn < 0) return 0;
    double u, v;
    if (!or_interp_lsq_tinterp(m, f, lon, lat, nn, a->time_dir, -1, alpha, &u, &v)) return 0;
    double z0 = z0_of(lat);
    *U = u;
    *V = v;
    *kx = u / z0;
    *ky = v / SLM;
    return 1;
}

static inline double stage_alpha(double t, double inv_dtn)
{
    double al = t * inv_dtn;
    return al > 1.0 ? 1.0 : al;
}

static void rk45_det(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a, double *lon, double *lat,
                     int32_t *st, int32_t *nn, int8_t *fr, double *Uo, double *Vo)
{
    const double dt = a->dt, tol = a->rk45_tol_m, h_min = a->rk45_h_min, inv = a->rk45_inv_dtn;
    double y0 = *lon, y1 = *lat, t = 0.0, h = dt;
    double Uv, Vv, k1x, k1y;
    int ok_out = 1;
    *Uo = *Vo = 0.0;
    if (!rk45_eval(m, f, a, y0, y1, stage_alpha(t, inv), &Uv, &Vv, &k1x, &k1y)) {
        ok_out = 0;
    } else {
        int64_t nsub = 0;
        double k2x = 0, k2y = 0, k3x = 0, k3y = 0, k4x = 0, k4y = 0, k5x = 0, k5y = 0, k6x = 0, k6y = 0, k7x = 0,
               k7y = 0, U7 = 0, V7 = 0, y5x = y0, y5y = y1, du, dv;
        while (t < dt * (1.0 - 1e-12)) {
            if (h > dt - t) h = dt - t;
            int clamped = 0;
            if (h < h_min) {
                h = h_min;
                if (h > dt - t) h = dt - t;
                clamped = 1;
            }
            int okst = 1;
            double px, py;
            px = y0 + h * (1.0 / 5.0) * k1x;
            py = y1 + h * (1.0 / 5.0) * k1y;
            norm_lonlat(&px, &py);
            okst = rk45_eval(m, f, a, px, py, stage_alpha(t + DP_C2 * h, inv), &du, &dv, &k2x, &k2y);
            if (okst) {
                px = y0 + h * ((3.0 / 40.0) * k1x + (9.0 / 40.0) * k2x);
                py = y1 + h * ((3.0 / 40.0) * k1y + (9.0 / 40.0) * k2y);
                norm_lonlat(&px, &py);
                okst = rk45_eval(m, f, a, px, py, stage_alpha(t + DP_C3 * h, inv), &du, &dv, &k3x, &k3y);
            }
            if (okst) {
                px = y0 + h * ((44.0 / 45.0) * k1x + (-56.0 / 15.0) * k2x + (32.0 / 9.0) * k3x);
                py = y1 + h * ((44.0 / 45.0) * k1y + (-56.0 / 15.0) * k2y + (32.0 / 9.0) * k3y);
                norm_lonlat(&px, &py);
                okst = rk45_eval(m, f, a, px, py, stage_alpha(t + DP_C4 * h, inv), &du, &dv, &k4x, &k4y);
            }
            if (okst) {
                px = y0 + h * ((19372.0 / 6561.0) * k1x + (-25360.0 / 2187.0) * k2x + (64448.0 / 6561.0) * k3x +
                               (-212.0 / 729.0) * k4x);
                py = y1 + h * ((19372.0 / 6561.0) * k1y + (-25360.0 / 2187.0) * k2y + (64448.0 / 6561.0) * k3y +
                               (-212.0 / 729.0) * k4y);
                norm_lonlat(&px, &py);
                okst = rk45_eval(m, f, a, px, py, stage_alpha(t + DP_C5 * h, inv), &du, &dv, &k5x, &k5y);
            }
            const double a6 = stage_alpha(t + h, inv);
            if (okst) {
                px = y0 + h * ((9017.0 / 3168.0) * k1x + (-355.0 / 33.0) * k2x + (46732.0 / 5247.0) * k3x +
                               (49.0 / 176.0) * k4x + (-5103.0 / 18656.0) * k5x);
                py = y1 + h * ((9017.0 / 3168.0) * k1y + (-355.0 / 33.0) * k2y + (46732.0 / 5247.0) * k3y +
                               (49.0 / 176.0) * k4y + (-5103.0 / 18656.0) * k5y);
                norm_lonlat(&px, &py);
                okst = rk45_eval(m, f, a, px, py, a6, &du, &dv, &k6x, &k6y);
            }
            if (okst) {
                y5x = y0 + h * ((35.0 / 384.0) * k1x + (500.0 / 1113.0) * k3x + (125.0 / 192.0) * k4x +
                                (-2187.0 / 6784.0) * k5x + (11.0 / 84.0) * k6x);
                y5y = y1 + h * ((35.0 / 384.0) * k1y + (500.0 / 1113.0) * k3y + (125.0 / 192.0) * k4y +
                                (-2187.0 / 6784.0) * k5y + (11.0 / 84.0) * k6y);
                norm_lonlat(&y5x, &y5y);
                okst = rk45_eval(m, f, a, y5x, y5y, a6, &U7, &V7, &k7x, &k7y);
            }
            nsub += 1;
            if (!okst) {
                /* a failed stage is a rejected attempt, terminal only at the
                 * h_min floor or the cap */
                if (nsub > a->rk45_max_sub || h <= h_min * (1.0 + 1e-12)) {
                    ok_out = 0;
                    break;
                }
                h = h * 0.2;
                if (h < h_min) h = h_min;
                continue;
            }
            double errx = h * ((71.0 / 57600.0) * k1x + (-71.0 / 16695.0) * k3x + (71.0 / 1920.0) * k4x +
                               (-17253.0 / 339200.0) * k5x + (22.0 / 525.0) * k6x + (-1.0 / 40.0) * k7x);
            double erry = h * ((71.0 / 57600.0) * k1y + (-71.0 / 16695.0) * k3y + (71.0 / 1920.0) * k4y +
                               (-17253.0 / 339200.0) * k5y + (22.0 / 525.0) * k6y + (-1.0 / 40.0) * k7y);
            double z0m = z0_of(y1);
            double err = sqrt(sq(errx * z0m) + sq(erry * SLM));
            if (nsub > a->rk45_max_sub) {
                ok_out = 0;
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
                fac = 0.9 * M_POW(tol / err, 0.2);
                if (fac < 0.2) fac = 0.2;
                else if (fac > 5.0) fac = 5.0;
            } else {
                fac = 5.0;
            }
            h = h * fac;
        }
    }
    if (!ok_out) {
        *st = ST_OUTSIDE;
        *fr = FR_INTERP;
        return;
    }
    int nn_new = or_site_anchor(m, y0, y1, a->exact == 1 ? 1 : 2);
    if (or_find_containing_local(m, y0, y1, nn_new) < 0) {
        *st = ST_OUTSIDE;
        *fr = FR_OOB;
        return;
    }
    *lon = y0;
    *lat = y1;
    *nn = nn_new;
    *Uo = Uv;
    *Vo = Vv;
}

/* ---------------------------------------------------------------------
 * Sigma meshes (numerics/vertical_sigma.py and the sigma branches of the
 * multi-pass steppers, integrators.py:1310-1441, :4997, :5380, :5760).
 * The mesh's levels are the sigma interfaces; interface k of a column of
 * height H lies at z = (1 - sigma_k) H.
 * --------------------------------------------------------------------- */

static int collect_neighbours(const or_mesh_t *m, int node, int *buf);

/* vertical_sigma.py:473 _col_H */
static inline double col_H(const or_mesh_t *m, const double *eta, double Dmin, int nd)
{
    double H = m->depth[nd] + eta[nd];
    if (H < Dmin) H = Dmin;
    return H;
}

/* vertical_sigma.py:482 _sigma_frac */
static inline double sigma_frac(double z, double Hp)
{
    double f = z / Hp;
    if (f < 0.0)
        f = 0.0;
    else if (f > 1.0)
        f = 1.0;
    return f;
}

/* vertical_sigma.py:74 find_layer_sigma */
static int find_layer_sigma(const double *s, int K, double z, double H)
{
    if (K < 2) return 0;
    double z_top = (1.0 - s[0]) * H;
    if (z <= z_top) return 0;
    double z_bot = (1.0 - s[K - 1]) * H;
    if (z >= z_bot) return K - 2;
    for (int k = 0; k < K - 1; k++) {
        double z_next = (1.0 - s[k + 1]) * H;
        if (z < z_next) return k;
    }
    return K - 2;
}

/* vertical_sigma.py:297 _interp_w_sigma_linear_single on the stage field */
static double w_sig_linear(const or_mesh_t *m, const or_fields_t *f, double alpha, double z, int node,
                           double H, double time_dir)
{
    const double *s = m->z_levels;
    int K = m->n_levels;
    int64_t N = m->n_nodes;
    if (K < 2) return time_dir * wlev(f, node, alpha);
    double z_top = (1.0 - s[0]) * H;
    double z_bot = (1.0 - s[K - 1]) * H;
    if (z <= z_top) return time_dir * wlev(f, node, alpha);
    if (z >= z_bot) return time_dir * wlev(f, (int64_t)(K - 1) * N + node, alpha);
    int k = K - 2;
    for (int j = 0; j < K - 1; j++) {
        double z_next = (1.0 - s[j + 1]) * H;
        if (z < z_next) {
            k = j;
            break;
        }
    }
    double z_k = (1.0 - s[k]) * H;
    double z_k1 = (1.0 - s[k + 1]) * H;
    double dz = z_k1 - z_k;
    if (dz < 1e-12) return time_dir * wlev(f, (int64_t)k * N + node, alpha);
    double al = (z - z_k) / dz;
    double w_val = (1.0 - al) * wlev(f, (int64_t)k * N + node, alpha) + al * wlev(f, (int64_t)(k + 1) * N + node, alpha);
    return time_dir * w_val;
}

/* the own column at min(z, H): every horizontal method's fallback */
static double w_sig_own(const or_mesh_t *m, const or_fields_t *f, const double *eta, double Dmin, double alpha,
                        double z, int node, double time_dir)
{
    double H = col_H(m, eta, Dmin, node);
    double zc = z < H ? z : H;
    return w_sig_linear(m, f, alpha, zc, node, H, time_dir);
}

/* vertical_sigma.py:566 _bary_locate: the containing element among the
 * node's elements and its vertex weights; returns n_valid (0: none) */
static int bary_locate(const or_mesh_t *m, double xp, double yp, int node, int *verts, double *wts)
{
    const double DEG2R = 0.017453292519943295;
    int n_el = m->n2e_count[node];
    const int32_t *row = m->n2e + (int64_t)node * m->max_conn;
    for (int k = 0; k < n_el; k++) {
        int e = row[k];
        if (e < 0) continue;
        int nv = 0;
        for (int v = 0; v < m->max_v; v++) {
            int nd = m->elem_nodes[(int64_t)e * m->max_v + v];
            if (nd < 0) continue;
            if (nv > 0 && nd == verts[0]) continue;
            verts[nv++] = nd;
        }
        if (nv < 3) continue;
        double lon0 = seam_safe_lon0(m, verts, nv);
        double lat0 = 0.0;
        for (int v = 0; v < nv; v++) lat0 += m->nodes_lat[verts[v]];
        lat0 /= nv;
        double cos_lat0 = M_COS(lat0 * DEG2R);
        const double R = 6371000.0;
        double vx[4], vy[4];
        for (int v = 0; v < nv; v++) {
            double dlon = PYMOD360(m->nodes_lon[verts[v]] - lon0 + 180.0) - 180.0;
            vx[v] = dlon * DEG2R * R * cos_lat0;
            vy[v] = (m->nodes_lat[verts[v]] - lat0) * DEG2R * R;
        }
        double dlon_p = PYMOD360(xp - lon0 + 180.0) - 180.0;
        double px = dlon_p * DEG2R * R * cos_lat0;
        double py = (yp - lat0) * DEG2R * R;
        for (int t = 0; t < nv - 2; t++) {
            double ax = vx[0], ay = vy[0], bx = vx[t + 1], by = vy[t + 1], cx = vx[t + 2], cy = vy[t + 2];
            double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
            if (fabs(denom) < 1e-14) continue;
            double w1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / denom;
            double w2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / denom;
            double w3 = 1.0 - w1 - w2;
            const double eps = -1e-6;
            if (w1 >= eps && w2 >= eps && w3 >= eps) {
                for (int v = 0; v < 4; v++) wts[v] = 0.0;
                if (nv == 4) {
                    wachspress_quad(px, py, vx, vy, wts);
                } else {
                    wts[0] = w1;
                    wts[t + 1] = w2;
                    wts[t + 2] = w3;
                }
                return nv;
            }
        }
    }
    return 0;
}

/* vertical_sigma.py:1035 _loop_sigma_column_height for one particle
 * (:493 leastsq, :636 barycentric, :534 idw, else the anchor's column) */
static double column_height(const or_mesh_t *m, const double *eta, double Dmin, int method, double xp,
                            double yp, int node)
{
    const double DEG2R = 0.017453292519943295;
    if (method == 1) {
        int buf[64];
        int count = collect_neighbours(m, node, buf);
        if (count < 3) return col_H(m, eta, Dmin, node);
        double z0 = M_COS(yp * DEG2R) * SLM;
        if (z0 < 1.0) z0 = 1.0;
        double a11 = 0, a12 = 0, a13 = 0, a22 = 0, a23 = 0, a33 = 0, fw = 0, f1w = 0, f2w = 0;
        for (int i = 0; i < count; i++) {
            int nd = buf[i];
            double h = col_H(m, eta, Dmin, nd);
            double dlon = PYMOD360(m->nodes_lon[nd] - xp + 180.0) - 180.0;
            double ex = dlon * z0;
            double ey = (m->nodes_lat[nd] - yp) * SLM;
            a11 += 1.0; a12 += ex; a13 += ey;
            a22 += ex * ex; a23 += ex * ey; a33 += ey * ey;
            fw += h; f1w += h * ex; f2w += h * ey;
        }
        double a21 = a12, a31 = a13, a32 = a23;
        double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 - a12 * a21 * a33 -
                      a11 * a23 * a32);
        if (fabs(det) < 1e-30 || fabs(det) < 1e-12 * fabs(a11 * a22 * a33)) return col_H(m, eta, Dmin, node);
        double inv = 1.0 / fabs(det);
        double b11 = a22 * a33 - a23 * a32;
        double b12 = a13 * a32 - a12 * a33;
        double b13 = a12 * a23 - a13 * a22;
        double H = (b11 * fw + b12 * f1w + b13 * f2w) * inv;
        if (H < Dmin) H = Dmin;
        return H;
    }
    if (method == 2) {
        int verts[4];
        double wts[4];
        int nv = bary_locate(m, xp, yp, node, verts, wts);
        if (nv == 0) return col_H(m, eta, Dmin, node);
        double H = 0.0;
        for (int v = 0; v < nv; v++) H += wts[v] * col_H(m, eta, Dmin, verts[v]);
        return H;
    }
    if (method == 3) {
        int buf[64];
        int count = collect_neighbours(m, node, buf);
        if (count < 1) return col_H(m, eta, Dmin, node);
        double z0 = M_COS(yp * DEG2R) * SLM;
        if (z0 < 1.0) z0 = 1.0;
        double sw = 0.0, swh = 0.0;
        for (int i = 0; i < count; i++) {
            int nd = buf[i];
            double h = col_H(m, eta, Dmin, nd);
            double dlon = PYMOD360(m->nodes_lon[nd] - xp + 180.0) - 180.0;
            double dx = dlon * z0;
            double dy = (m->nodes_lat[nd] - yp) * SLM;
            double d2 = dx * dx + dy * dy;
            if (d2 < 1.0) return h;
            sw += 1.0 / d2;
            swh += h / d2;
        }
        if (sw < 1e-30) return col_H(m, eta, Dmin, node);
        return swh / sw;
    }
    return col_H(m, eta, Dmin, node);
}

/* vertical_sigma.py:1063 sigma_column_height over P particles (method 0
 * nearest, 1 leastsq, 2 barycentric, 3 idw); entries of particles that are
 * not alive are 1.0, as there.  Test entry point for the pins of the
 * column-height interpolators. */
void or_column_height_batch(const or_mesh_t *m, const double *eta, double Dmin, int method, int64_t P,
                            const double *lon, const double *lat, const int32_t *nn, const int32_t *alive,
                            double *out)
{
    for (int64_t i = 0; i < P; i++)
        out[i] = alive[i] ? column_height(m, eta, Dmin, method, lon[i], lat[i], nn[i]) : 1.0;
}

/* vertical_sigma.py:971 interp_w_sigma_spatial for one particle */
static double w_sigma(const or_mesh_t *m, const or_fields_t *f, const double *eta, double Dmin, int method,
                      double alpha, double xp, double yp, double z, int node, double time_dir)
{
    const double DEG2R = 0.017453292519943295;
    if (method == 0) return w_sig_linear(m, f, alpha, z, node, col_H(m, eta, Dmin, node), time_dir);
    if (method == 2) {
        int verts[4];
        double wts[4];
        int nv = bary_locate(m, xp, yp, node, verts, wts);
        if (nv == 0) return w_sig_own(m, f, eta, Dmin, alpha, z, node, time_dir);
        double Hp = 0.0;
        for (int v = 0; v < nv; v++) Hp += wts[v] * col_H(m, eta, Dmin, verts[v]);
        double frac = sigma_frac(z, Hp);
        double acc = 0.0;
        for (int v = 0; v < nv; v++) {
            double Hv = col_H(m, eta, Dmin, verts[v]);
            acc += wts[v] * w_sig_linear(m, f, alpha, frac * Hv, verts[v], Hv, time_dir);
        }
        return acc;
    }
    int buf[64];
    int count = collect_neighbours(m, node, buf);
    if (method == 3) {
        if (count < 1) return w_sig_own(m, f, eta, Dmin, alpha, z, node, time_dir);
        double z0 = M_COS(yp * DEG2R) * SLM;
        if (z0 < 1.0) z0 = 1.0;
        double frac = sigma_frac(z, column_height(m, eta, Dmin, 3, xp, yp, node));
        double sum_w = 0.0, sum_wv = 0.0;
        for (int i = 0; i < count; i++) {
            int nd = buf[i];
            double H = col_H(m, eta, Dmin, nd);
            double wv = w_sig_linear(m, f, alpha, frac * H, nd, H, time_dir);
            double dlon = PYMOD360(m->nodes_lon[nd] - xp + 180.0) - 180.0;
            double dx = dlon * z0;
            double dy = (m->nodes_lat[nd] - yp) * SLM;
            double d2 = dx * dx + dy * dy;
            if (d2 < 1.0) return wv;
            double weight = 1.0 / d2;
            sum_w += weight;
            sum_wv += weight * wv;
        }
        if (sum_w < 1e-30) return w_sig_own(m, f, eta, Dmin, alpha, z, node, time_dir);
        return sum_wv / sum_w;
    }
    /* vertical_sigma.py:656 leastsq */
    if (count < 3) return w_sig_own(m, f, eta, Dmin, alpha, z, node, time_dir);
    double z0 = M_COS(yp * DEG2R) * SLM;
    if (z0 < 1.0) z0 = 1.0;
    double frac = sigma_frac(z, column_height(m, eta, Dmin, 1, xp, yp, node));
    double a11 = 0, a12 = 0, a13 = 0, a22 = 0, a23 = 0, a33 = 0, fw = 0, f1w = 0, f2w = 0;
    for (int i = 0; i < count; i++) {
        int nd = buf[i];
        double H = col_H(m, eta, Dmin, nd);
        double wv = w_sig_linear(m, f, alpha, frac * H, nd, H, time_dir);
        double dlon = PYMOD360(m->nodes_lon[nd] - xp + 180.0) - 180.0;
        double ex = dlon * z0;
        double ey = (m->nodes_lat[nd] - yp) * SLM;
        a11 += 1.0; a12 += ex; a13 += ey;
        a22 += ex * ex; a23 += ex * ey; a33 += ey * ey;
        fw += wv; f1w += wv * ex; f2w += wv * ey;
    }
    double a21 = a12, a31 = a13, a32 = a23;
    double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 - a12 * a21 * a33 -
                  a11 * a23 * a32);
    if (fabs(det) < 1e-30 || fabs(det) < 1e-12 * fabs(a11 * a22 * a33))
        return w_sig_own(m, f, eta, Dmin, alpha, z, node, time_dir);
    double inv = 1.0 / fabs(det);
    double b11 = a22 * a33 - a23 * a32;
    double b12 = a13 * a32 - a12 * a33;
    double b13 = a12 * a23 - a13 * a22;
    return (b11 * fw + b12 * f1w + b13 * f2w) * inv;
}

/* One sigma stage: anchor, column height with the stage surface, the
 * stage depth (z, or s H_p), the layer in that column, u/v, W. */
static int sig_site(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a, int stage, double lon,
                    double lat, int from_s, double s_or_z, int *nn_o, double *H_o, double *U, double *V,
                    double *W)
{
    int nn = or_site_anchor(m, lon, lat, a->exact);
    *nn_o = nn;
    const double *eta = a->sig_eta[stage];
    double H = column_height(m, eta, a->Dmin, a->w_method, lon, lat, nn);
    *H_o = H;
    double zs = from_s ? s_or_z * H : s_or_z;
    double Hl = H < a->Dmin ? a->Dmin : H;
    int layer = find_layer_sigma(m->z_levels, m->n_levels, zs, Hl);
    int ok;
    if (f->uv_layout == 1)
        ok = or_interp_lsq_tinterp(m, f, lon, lat, nn, a->time_dir, layer, a->alphas[stage], U, V);
    else
        ok = or_interp_bary_tinterp(m, f, lon, lat, nn, a->time_dir, layer, a->alphas[stage], U, V, NULL);
    *W = w_sigma(m, f, eta, a->Dmin, a->w_method, a->alphas[stage], lon, lat, zs, nn, a->time_dir);
    return ok;
}

/* vertical_sigma.py:237 _loop_clamp_depth_sigma_H */
static double clamp_sigma(double z, double H, double Dmin, double z0b_min)
{
    if (H < Dmin) H = Dmin;
    if (z < 0.0) z = 0.0;
    double bottom = H - z0b_min;
    if (bottom < 0.0) bottom = 0.0;
    if (z > bottom) z = bottom;
    return z;
}

/* The sigma step of one alive particle: Euler :4997, RK2 :5380, RK4 :5760
 * (sigma branches), in place; returns 0 alive, else the failure code. */
static int sigma_det_step(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a, double *lon,
                          double *lat, double *z, int32_t *st, int32_t *nn, int8_t *fr, double *Uo, double *Vo,
                          double *Wo)
{
    const double dt = a->dt;
    const double lon0 = *lon, lat0 = *lat, zz0 = *z;
    int n1;
    double H1, U1, V1, W1;
    int ok = sig_site(m, f, a, 0, lon0, lat0, 0, zz0, &n1, &H1, &U1, &V1, &W1);
    *Uo = U1;
    *Vo = V1;
    *Wo = W1;
    double lon_new = lon0, lat_new = lat0, s_new = 0.0;
    if (ok) {
        if (a->order == 1) {
            lon_new = wrap_lon(lon0 + dt * U1 / z0_of(lat0));
            lat_new = clamp_lat(lat0 + dt * V1 / SLM);
            s_new = (zz0 + W1 * dt) / H1;
        } else if (a->order == 2) {
            double z0 = z0_of(lat0);
            double lon_mid = wrap_lon(lon0 + 0.5 * dt * U1 / z0);
            double lat_mid = clamp_lat(lat0 + 0.5 * dt * V1 / SLM);
            double s_mid = (zz0 + W1 * 0.5 * dt) / H1;
            int n2;
            double H2, U2, V2, W2;
            ok = sig_site(m, f, a, 1, lon_mid, lat_mid, 1, s_mid, &n2, &H2, &U2, &V2, &W2);
            if (ok) {
                lon_new = wrap_lon(lon0 + dt * U2 / z0_of(lat_mid));
                lat_new = clamp_lat(lat0 + dt * V2 / SLM);
                s_new = zz0 / H1 + dt * W2 / H2;
            }
        } else {
            double dlon1 = dt * U1 / z0_of(lat0);
            double dlat1 = dt * V1 / SLM;
            double s0 = zz0 / H1;
            double f1 = W1 / H1;
            double lon2 = wrap_lon(lon0 + 0.5 * dlon1);
            double lat2 = clamp_lat(lat0 + 0.5 * dlat1);
            int n2, n3, n4;
            double H2, U2, V2, W2, H3, U3, V3, W3, H4, U4, V4, W4;
            ok = sig_site(m, f, a, 1, lon2, lat2, 1, s0 + 0.5 * dt * f1, &n2, &H2, &U2, &V2, &W2);
            if (ok) {
                double f2 = W2 / H2;
                double dlon2 = dt * U2 / z0_of(lat2);
                double dlat2 = dt * V2 / SLM;
                double lon3 = wrap_lon(lon0 + 0.5 * dlon2);
                double lat3 = clamp_lat(lat0 + 0.5 * dlat2);
                ok = sig_site(m, f, a, 2, lon3, lat3, 1, s0 + 0.5 * dt * f2, &n3, &H3, &U3, &V3, &W3);
                if (ok) {
                    double f3 = W3 / H3;
                    double dlon3 = dt * U3 / z0_of(lat3);
                    double dlat3 = dt * V3 / SLM;
                    double lon4 = wrap_lon(lon0 + dlon3);
                    double lat4 = clamp_lat(lat0 + dlat3);
                    ok = sig_site(m, f, a, 3, lon4, lat4, 1, s0 + dt * f3, &n4, &H4, &U4, &V4, &W4);
                    if (ok) {
                        double f4 = W4 / H4;
                        double dlon4 = dt * U4 / z0_of(lat4);
                        double dlat4 = dt * V4 / SLM;
                        double dlon = (dlon1 + 2.0 * dlon2 + 2.0 * dlon3 + dlon4) / 6.0;
                        double dlat = (dlat1 + 2.0 * dlat2 + 2.0 * dlat3 + dlat4) / 6.0;
                        lon_new = wrap_lon(lon0 + dlon);
                        lat_new = clamp_lat(lat0 + dlat);
                        s_new = s0 + dt * (f1 + 2.0 * f2 + 2.0 * f3 + f4) / 6.0;
                    }
                }
            }
        }
    }
    if (!ok) {
        *st = ST_OUTSIDE;
        *fr = FR_INTERP;
        *nn = n1;
        return FR_INTERP;
    }
    int nn_new = or_site_anchor(m, lon_new, lat_new, a->exact);
    if (or_find_containing_local(m, lon_new, lat_new, nn_new) < 0) {
        *st = ST_OUTSIDE;
        *fr = FR_OOB;
        *nn = n1;
        return FR_OOB;
    }
    *lon = lon_new;
    *lat = lat_new;
    double H_end = column_height(m, a->sig_eta[4], a->Dmin, a->w_method, lon_new, lat_new, nn_new);
    *z = clamp_sigma(s_new * H_end, H_end, a->Dmin, a->z0b_min);
    *nn = nn_new;
    return 0;
}

/* integrators.py:673 _apply_split_diffusion_kick for one particle alive
 * after the deterministic step; pre_* is the start-of-step state. */
static void kick_one(const or_mesh_t *m, const or_step_args_t *a, uint64_t pid,
                     double pre_lon, double pre_lat, double pre_z, int32_t pre_nn,
                     double *lon, double *lat, double *z, int32_t *st, int32_t *nn, int8_t *fr)
{
    double g1, g2;
    or_normal_pair(a->seed, a->real_id, pid, a->step_i, 0, &g1, &g2);
    double abs_dt = a->dt >= 0.0 ? a->dt : -a->dt;
    double sqrt_abs_dt = sqrt(abs_dt);
    double sigma_x = sqrt(2.0 * a->kh_x);
    double sigma_y = sqrt(2.0 * a->kh_y);
    double dx = sigma_x * sqrt_abs_dt * g1;
    double dy = sigma_y * sqrt_abs_dt * g2;
    if (a->kh_x != a->kh_y) {
        double la = *lat;
        if (la < -89.9) la = -89.9;
        if (la > 89.9) la = 89.9;
        double corr = (a->kh_x - a->kh_y) * (M_SIN(la * DEG2RAD) / M_COS(la * DEG2RAD)) / R_METRIC * fabs(a->dt);
        dy = dy + corr;
    }
    double lon_k, lat_k;
    sphere_displace(*lon, *lat, dx, dy, &lon_k, &lat_k);
    int nn_k = or_site_anchor(m, lon_k, lat_k, a->exact == 1 ? 1 : 2); /* find_nearest_nodes */
    int e = or_find_containing_local(m, lon_k, lat_k, nn_k);
    if (e < 0) {
        *lon = pre_lon;
        *lat = pre_lat;
        *z = pre_z;
        *nn = pre_nn;
        *st = ST_OUTSIDE;
        *fr = FR_OOB;
        return;
    }
    if (a->is3d && m->vertical_coord == 1) {
        /* integrators.py:784-801: keep s = z / H_p in the end-of-step
         * surface, clamp in the kicked column */
        double Hp = column_height(m, a->sig_eta[4], a->Dmin, a->w_method, *lon, *lat, *nn);
        double Hk = column_height(m, a->sig_eta[4], a->Dmin, a->w_method, lon_k, lat_k, nn_k);
        *z = clamp_sigma(*z / Hp * Hk, Hk, a->Dmin, a->z0b_min);
    } else if (a->is3d) {
        *z = clamp_z(m, *z, nn_k);
    }
    *lon = lon_k;
    *lat = lat_k;
    *nn = nn_k;
}

#ifdef ORACLE_KMATH
#define REFLECT_DEPTH km_reflect_depth
#define PYMOD_POS km_pymod_pos
#else
/* Python float `%` for a positive divisor (CPython float_rem): fmod is
 * exact; the remainder takes the divisor's sign and a zero is +0.0. */
static double PYMOD_POS(double x, double d)
{
    double m = fmod(x, d);
    if (m != 0.0) {
        if (m < 0.0) m += d;
    } else {
        m = 0.0;
    }
    return m;
}

/* numerics/diffusion.py:334 reflect_depth: specular fold into [top, bot]. */
static double REFLECT_DEPTH(double z, double top, double bot)
{
    double L = bot - top;
    if (L <= 0.0) return top;
    double y = PYMOD_POS(z - top, 2.0 * L);
    if (y < 0.0) y += 2.0 * L;
    if (y > L) y = 2.0 * L - y;
    return top + y;
}
#endif

/* integrators.py:842 _apply_split_az_kick with a constant source, z-level
 * mesh: numerics/diffusion.py:427 _loop_az_kick_zlevel (gradient 0) and
 * :302 diffusion_az_step_meters, one particle alive after the horizontal
 * kick.  The single normal is the first Box-Muller half of component 1. */
/* numerics/diffusion.py:459 _loop_az_kick_sigma (integrators.py:909-928):
 * folded into [0, max(H_p, Dmin) - z0b_min] of the end-of-step column */
static void az_sigma_one(const or_mesh_t *m, const or_step_args_t *a, uint64_t pid, double lon, double lat,
                         int32_t node, double *z)
{
    double az = a->az_const < 0.0 ? 0.0 : a->az_const;
    double H = column_height(m, a->sig_eta[4], a->Dmin, a->w_method, lon, lat, node);
    if (H < a->Dmin) H = a->Dmin;
    double bot = H - a->z0b_min;
    if (bot < 0.0) bot = 0.0;
    double g = 0.0;
    double xi, unused;
    or_normal_pair(a->seed, a->real_id, pid, a->step_i, 1, &xi, &unused);
    double sdt = a->time_dir * a->dt;
    double abs_dt = sdt >= 0.0 ? sdt : -sdt;
    double dz = g * abs_dt + sqrt(2.0 * az * abs_dt) * xi;
    *z = REFLECT_DEPTH(*z + dz, 0.0, bot);
}

/* numerics/diffusion.py:378 az_column_value_gradient: Az and dAz/dz on the
 * column of `node` at depth z clamped to [z_levels[0], min(depth, last
 * level)]; the value linear between interfaces (:357 _az_interp_column),
 * the gradient the slope of the segment the layer search picks (the one
 * above an interface the depth sits on).  No reference golden holds an
 * Az-field run: pinned by the linear-profile closed form, the constant-field
 * identity and the well-mixed condition (tests/test_oracle_az.py). */
static void az_column(const or_mesh_t *m, const double *azf, int32_t node, double z, double *val, double *grad)
{
    const int K = m->n_levels;
    const double *zl = m->z_levels;
    const int64_t N = m->n_nodes;
#define AZ(k) azf[(int64_t)(k) * N + node]
    double top = zl[0];
    double bot = m->depth[node] < zl[K - 1] ? m->depth[node] : zl[K - 1];
    if (bot < top) bot = top;
    double ze = z;
    if (ze < top) ze = top;
    if (ze > bot) ze = bot;
    if (K < 2) {
        *val = AZ(0);
        *grad = 0.0;
        return;
    }
    if (ze <= zl[0]) {
        *val = AZ(0);
    } else if (ze >= zl[K - 1]) {
        *val = AZ(K - 1);
    } else {
        int k = 0;
        while (!(ze <= zl[k + 1])) ++k;
        double dzl = zl[k + 1] - zl[k];
        if (dzl <= 0.0) {
            *val = AZ(k);
        } else {
            double frac = (ze - zl[k]) / dzl;
            *val = (1.0 - frac) * AZ(k) + frac * AZ(k + 1);
        }
    }
    int ks = K - 2;
    for (int kk = 0; kk < K - 1; ++kk) {
        if (ze <= zl[kk + 1]) {
            ks = kk;
            break;
        }
    }
    double dz = zl[ks + 1] - zl[ks];
    *grad = dz <= 0.0 ? 0.0 : (AZ(ks + 1) - AZ(ks)) / dz;
#undef AZ
}

static void az_one(const or_mesh_t *m, const or_step_args_t *a, uint64_t pid, int32_t node, double *z)
{
    double az = a->az_const < 0.0 ? 0.0 : a->az_const;
    double g = 0.0;
    if (a->az_field) {
        /* numerics/diffusion.py:448-454: a negative value kills the kick */
        az_column(m, a->az_field, node, *z, &az, &g);
        if (az < 0.0) {
            az = 0.0;
            g = 0.0;
        }
    }
    double xi, unused;
    or_normal_pair(a->seed, a->real_id, pid, a->step_i, 1, &xi, &unused);
    double sdt = a->time_dir * a->dt;
    double abs_dt = sdt >= 0.0 ? sdt : -sdt;
    double dz = g * abs_dt + sqrt(2.0 * az * abs_dt) * xi;
    *z = REFLECT_DEPTH(*z + dz, m->z_levels[0], m->depth[node]);
}

/* tracking.py:1089 _apply_periodic_wrap for one ALIVE particle: fold a
 * coordinate outside [x0, x1) back with numpy's mod (Python % semantics),
 * keep it strictly inside, and re-anchor a folded particle with the batch
 * search (find_nearest_nodes). */
static void periodic_one(const or_mesh_t *m, const or_step_args_t *a, double *lon, double *lat, int32_t *nn,
                         int32_t *elem)
{
    int moved = 0;
    if (a->periodic_lon && (*lon < a->px0 || *lon >= a->px1)) {
        double fo = a->px0 + PYMOD_POS(*lon - a->px0, a->px1 - a->px0);
        if (fo >= a->px1) fo = a->px0;
        *lon = fo;
        moved = 1;
    }
    if (a->periodic_lat && (*lat < a->py0 || *lat >= a->py1)) {
        double fo = a->py0 + PYMOD_POS(*lat - a->py0, a->py1 - a->py0);
        if (fo >= a->py1) fo = a->py0;
        *lat = fo;
        moved = 1;
    }
    if (moved) {
        *nn = or_site_anchor(m, *lon, *lat, a->exact == 1 ? 1 : 2);
        *elem = -1; /* tracking.py:1143: the carried element is dropped */
    }
}

/* tracking.py:199 _apply_beaching_hook + beaching.py:95 evaluate_beaching
 * for one particle (prev = status at the step start). */
static void beach_one(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a, uint64_t pid,
                      int32_t prev, double lon, double lat, int32_t *st, int32_t *nn, int8_t *fr, int32_t *ws)
{
    const int ALIVE = 1, BEACHED = 2, OUTSIDE = 3;
    int need_topo = a->use_coastal || a->use_stochastic;
    int newly_failed = (*st == OUTSIDE) && (prev == ALIVE);
    if (newly_failed && *nn < 0) *nn = or_site_anchor(m, lon, lat, a->exact == 1 ? 1 : 2);
    int n = *nn;
    int containing = -1;
    if (need_topo && (*st == ALIVE || *st == BEACHED || newly_failed) && n >= 0)
        containing = or_find_containing_local(m, lon, lat, n);
    int coastal = need_topo && containing >= 0 && m->elem_coastal && m->elem_coastal[containing];
    double h_raw = n >= 0 ? m->depth[n] : INFINITY;
    /* + vel.eta[nn] of the step's field: the bracket pair blended by alpha0
     * (core/io.py:386 LazyLerpField), or the plain snapshot */
    if (a->use_eta && n >= 0 && f->eta_a) {
        double eta = f->eta_a[n];
        if (f->alpha0 != 0.0 && f->eta_b) eta = (1.0 - f->alpha0) * f->eta_a[n] + f->alpha0 * f->eta_b[n];
        h_raw = h_raw + eta;
    }
    if (a->on_failure && newly_failed) {
        int context = h_raw < a->zone_depth || coastal;
        if (context) {
            *st = BEACHED;
            *fr = 0;
        }
    }
    int alive = *st == ALIVE, beached = *st == BEACHED;
    int to_beach = 0;
    if (a->use_depth) to_beach |= h_raw < a->depth_min;
    if (a->use_coastal) to_beach |= coastal;
    if (a->use_stochastic && a->t_beach > 0) {
        int in_zone = h_raw < a->zone_depth || coastal;
        double p = 1.0 - M_EXP(-fabs(a->beach_dt) / a->t_beach);
        if (alive && in_zone && or_u01(a->beach_seed, a->real_id, pid, a->step_i, 10) < p) to_beach = 1;
    }
    to_beach &= alive;
    int to_release = 0;
    int w = *ws;
    if (a->resuspension != 0 && beached) {
        int rewet = h_raw > (a->depth_min + a->rewet_hysteresis) && n >= 0;
        if (rewet)
            w += 1;
        else
            w = 0;
        if (a->resuspension == 1 || a->resuspension == 3) to_release |= w >= a->min_wet_steps;
        if ((a->resuspension == 2 || a->resuspension == 3) && a->t_resuspend > 0) {
            double p_r = 1.0 - M_EXP(-fabs(a->beach_dt) / a->t_resuspend);
            if (rewet && or_u01(a->beach_seed, a->real_id, pid, a->step_i, 11) < p_r) to_release = 1;
        }
    }
    if (to_release) w = 0;
    if (!beached) w = 0;
    *ws = w;
    if (to_beach) *st = BEACHED;
    if (to_release) {
        *st = ALIVE;
        *fr = 0;
    }
}

/* One full step over all particles, in place: deterministic step (dead
 * particles keep everything and get u = v = w = 0), split Kh kick, beaching
 * hook.  fr is the sticky failure reason; pid the RNG key; ws the
 * beaching wet streak. */
void or_step(const or_mesh_t *m, const or_fields_t *f, const or_step_args_t *a, int64_t P,
             double *lon, double *lat, double *z, int32_t *status, int32_t *nn, int8_t *fr,
             double *U, double *V, double *W, const int64_t *pid, int32_t *ws, int32_t *elem)
{
    /* particles are independent within a step (the reference's prange,
     * integrators.py fused kernels): one thread per block of particles in
     * the threaded build, identical results in any thread count */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < P; i++) {
        int32_t prev = status[i];
        int32_t el_dummy = -1;
        int32_t *el = elem ? &elem[i] : &el_dummy;
        double pre_lon = lon[i], pre_lat = lat[i], pre_z = z[i];
        int32_t pre_nn = nn[i];
        U[i] = 0.0;
        V[i] = 0.0;
        W[i] = 0.0;
        if (prev == ST_ALIVE && a->is3d && m->vertical_coord == 1) {
            sigma_det_step(m, f, a, &lon[i], &lat[i], &z[i], &status[i], &nn[i], &fr[i], &U[i], &V[i], &W[i]);
            if (a->kick && status[i] == ST_ALIVE && !(a->kh_x == 0.0 && a->kh_y == 0.0))
                kick_one(m, a, (uint64_t)pid[i], pre_lon, pre_lat, pre_z, pre_nn,
                         &lon[i], &lat[i], &z[i], &status[i], &nn[i], &fr[i]);
            if (a->az && a->az_const > 0.0 && status[i] == ST_ALIVE)
                az_sigma_one(m, a, (uint64_t)pid[i], lon[i], lat[i], nn[i], &z[i]);
            if ((a->periodic_lon || a->periodic_lat) && status[i] == ST_ALIVE)
                periodic_one(m, a, &lon[i], &lat[i], &nn[i], el);
        } else if (prev == ST_ALIVE) {
            if (a->order == 5)
                rk45_det(m, f, a, &lon[i], &lat[i], &status[i], &nn[i], &fr[i], &U[i], &V[i]);
            else
                det_step(m, f, a, &lon[i], &lat[i], &z[i], &status[i], &nn[i], &fr[i], &U[i], &V[i], &W[i], el);
            if (a->kick && status[i] == ST_ALIVE && !(a->kh_x == 0.0 && a->kh_y == 0.0))
                kick_one(m, a, (uint64_t)pid[i], pre_lon, pre_lat, pre_z, pre_nn,
                         &lon[i], &lat[i], &z[i], &status[i], &nn[i], &fr[i]);
            if (a->az && a->is3d && (a->az_field || a->az_const > 0.0) && status[i] == ST_ALIVE)
                az_one(m, a, (uint64_t)pid[i], nn[i], &z[i]);
            if ((a->periodic_lon || a->periodic_lat) && status[i] == ST_ALIVE)
                periodic_one(m, a, &lon[i], &lat[i], &nn[i], el);
        }
        if (a->beach) beach_one(m, f, a, (uint64_t)pid[i], prev, lon[i], lat[i], &status[i], &nn[i], &fr[i], &ws[i]);
    }
}

/* 2-D convenience form used by the first tests (no kick, no beaching). */
void or_step2d(const or_mesh_t *m, const or_fields_t *f, int order, const double *alphas,
               double dt, double time_dir, int exact, int64_t P,
               const double *lon, const double *lat, const int32_t *status, const int32_t *nn,
               double *lon_out, double *lat_out, int32_t *status_out, int32_t *nn_out,
               double *U_out, double *V_out, int8_t *fr_out)
{
    or_step_args_t a;
    memset(&a, 0, sizeof a);
    a.order = order;
    a.exact = exact;
    a.dt = dt;
    a.time_dir = time_dir;
    for (int k = 0; k < 4; k++) a.alphas[k] = alphas[k];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < P; i++) {
        lon_out[i] = lon[i];
        lat_out[i] = lat[i];
        status_out[i] = status[i];
        nn_out[i] = nn[i];
        U_out[i] = 0.0;
        V_out[i] = 0.0;
        fr_out[i] = 0;
        if (status[i] != ST_ALIVE) continue;
        double zz = 0.0, ww;
        int32_t el = -1;
        det_step(m, f, &a, &lon_out[i], &lat_out[i], &zz, &status_out[i], &nn_out[i], &fr_out[i],
                 &U_out[i], &V_out[i], &ww, &el);
    }
}

/* Batch nearest (core/mesh.py:1863 _grid_hash_nearest_batch) */
void or_nearest_batch(const or_mesh_t *m, int64_t P, const double *lon, const double *lat,
                      int exact, int32_t *out)
{
    for (int64_t i = 0; i < P; i++)
        out[i] = exact ? or_nearest_exact(m, lon[i], lat[i]) : or_nearest_3x3(m, lon[i], lat[i]);
}

/* Batch containment (core/mesh.py:2267 _loop_find_containing_elements) */
void or_contain_batch(const or_mesh_t *m, int64_t P, const double *lon, const double *lat,
                      const int32_t *nn, int32_t *elem_out)
{
    for (int64_t i = 0; i < P; i++) elem_out[i] = or_find_containing_local(m, lon[i], lat[i], nn[i]);
}

/* Node unit-vector table (core/mesh.py:713 _node_xyz_fill) */
void or_node_xyz(int64_t N, const double *lon, const double *lat, double *out)
{
    for (int64_t n = 0; n < N; n++) {
        double lat_r = lat[n] * DEG2RAD, lon_r = lon[n] * DEG2RAD;
        double c = M_COS(lat_r);
        out[3 * n + 0] = c * M_COS(lon_r);
        out[3 * n + 1] = c * M_SIN(lon_r);
        out[3 * n + 2] = M_SIN(lat_r);
    }
}

/* Elementwise probe of the transcendental functions this build uses
 * (fn: 0 sin, 1 cos, 2 asin, 3 acos, 4 atan2, 5 log, 6 exp, 7 pymod360). */
void or_math_probe(int fn, int64_t n, const double *x, const double *y, double *out)
{
    for (int64_t i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = M_SIN(x[i]); break;
        case 1: out[i] = M_COS(x[i]); break;
        case 2: out[i] = M_ASIN(x[i]); break;
        case 3: out[i] = M_ACOS(x[i]); break;
        case 4: out[i] = M_ATAN2(x[i], y[i]); break;
        case 5: out[i] = M_LOG(x[i]); break;
        case 6: out[i] = M_EXP(x[i]); break;
        case 7: out[i] = PYMOD360(x[i]); break;
        case 9: out[i] = PYMOD_POS(x[i], y[i]); break;
        case 10: out[i] = REFLECT_DEPTH(x[i], 0.0, y[i]); break;
        case 13: out[i] = M_POW(x[i], y[i]); break;
        case 11:
        case 12: {
            double g1, g2;
            or_normal_pair(20240501ull, 0ull, (uint64_t)x[i], (uint64_t)y[i], 0ull, &g1, &g2);
            out[i] = fn == 11 ? g1 : g2;
            break;
        }
        default: out[i] = sqrt(x[i]); break;
        }
    }
}

/* ------------------------------------------------------------------------
 * Record-point sampling of node scalars (numerics/sampling.py)
 * ---------------------------------------------------------------------- */

/* A node scalar of one record: planes of N values (1: 2-D, K-1: mid-layer,
 * K: interface), snapshot a blended towards b by alpha (core/io.py:68
 * _blend on the LazyLerpField scalars; alpha 0 = the plain snapshot). */
typedef struct {
    const double *a, *b;
    double alpha;
    int n_planes;
} or_scalar_t;

static inline double sval(const or_scalar_t *s, int64_t idx)
{
    if (s->alpha == 0.0) return s->a[idx];
    return (1.0 - s->alpha) * s->a[idx] + s->alpha * s->b[idx];
}

/* sampling.py:433 _interp_mid_linear_single at one node */
static double mid_linear(const or_mesh_t *m, const or_scalar_t *s, double z, int node)
{
    const double *zl = m->z_levels;
    int K1 = m->n_levels - 1;
    int64_t N = m->n_nodes;
    if (K1 <= 0) return 0.0;
    if (K1 == 1) return sval(s, node);
    double zm0 = 0.5 * (zl[0] + zl[1]);
    double zml = 0.5 * (zl[K1 - 1] + zl[K1]);
    if (z <= zm0) return sval(s, node);
    if (z >= zml) return sval(s, (int64_t)(K1 - 1) * N + node);
    for (int k = 0; k < K1 - 1; k++) {
        double zk = 0.5 * (zl[k] + zl[k + 1]);
        double zk1 = 0.5 * (zl[k + 1] + zl[k + 2]);
        if (z < zk1) {
            double dz = zk1 - zk;
            if (dz < 1e-12) return sval(s, (int64_t)k * N + node);
            double al = (z - zk) / dz;
            return (1.0 - al) * sval(s, (int64_t)k * N + node) + al * sval(s, (int64_t)(k + 1) * N + node);
        }
    }
    return sval(s, (int64_t)(K1 - 1) * N + node);
}

/* value of the field at node nd for the horizontal combination: 2-D the
 * node value, mid-layer linear in depth at min(z, depth[nd])
 * (sampling.py:517-520) */
static inline double node_value(const or_mesh_t *m, const or_scalar_t *s, double z, int nd)
{
    if (s->n_planes == 1) return sval(s, nd);
    double ze = z < m->depth[nd] ? z : m->depth[nd];
    return mid_linear(m, s, ze, nd);
}

static int collect_neighbours(const or_mesh_t *m, int node, int *buf)
{
    int count = 0;
    buf[count++] = node;
    int n_el = m->n2e_count[node];
    const int32_t *row = m->n2e + (int64_t)node * m->max_conn;
    for (int k = 0; k < n_el; k++) {
        int e = row[k];
        if (e < 0) continue;
        for (int v = 0; v < m->max_v; v++) {
            int nd = m->elem_nodes[(int64_t)e * m->max_v + v];
            if (nd < 0) continue;
            int found = 0;
            for (int j = 0; j < count; j++)
                if (buf[j] == nd) {
                    found = 1;
                    break;
                }
            if (!found && count < 64) buf[count++] = nd;
        }
    }
    return count;
}

/* sampling.py:83 _combine_horizontal (0 nearest, 1 leastsq, 2 barycentric,
 * 3 idw) over the 1-ring values */
static double combine_horizontal(const or_mesh_t *m, const or_scalar_t *s, double z, const int *neigh,
                                 int n_neigh, const double *vals, double lon_p, double lat_p, int nn,
                                 int method)
{
    const double DEG2R = 0.017453292519943295;
    double fallback = vals[0];
    if (method == 0) return fallback;
    if (method == 3) {
        double z0 = M_COS(lat_p * DEG2R) * SLM;
        if (z0 < 1.0) z0 = 1.0;
        double sw = 0.0, swv = 0.0;
        for (int i = 0; i < n_neigh; i++) {
            int nd = neigh[i];
            double dlon = PYMOD360(m->nodes_lon[nd] - lon_p + 180.0) - 180.0;
            double dx = dlon * z0;
            double dy = (m->nodes_lat[nd] - lat_p) * SLM;
            double d2 = dx * dx + dy * dy;
            if (d2 < 1e-12) return vals[i];
            if (!isfinite(vals[i])) continue;
            double w = 1.0 / d2;
            sw += w;
            swv += w * vals[i];
        }
        if (sw < 1e-30) return fallback;
        return swv / sw;
    }
    if (method == 2) {
        const double R = 6371000.0;
        int n_el = m->n2e_count[nn];
        const int32_t *row = m->n2e + (int64_t)nn * m->max_conn;
        for (int k = 0; k < n_el; k++) {
            int e = row[k];
            if (e < 0) continue;
            int verts[4], nv = 0;
            for (int v = 0; v < m->max_v; v++) {
                int nd = m->elem_nodes[(int64_t)e * m->max_v + v];
                if (nd < 0) continue;
                if (nv > 0 && nd == verts[0]) continue;
                verts[nv++] = nd;
            }
            if (nv < 3) continue;
            double lon0 = seam_safe_lon0(m, verts, nv);
            double lat0 = 0.0;
            for (int v = 0; v < nv; v++) lat0 += m->nodes_lat[verts[v]];
            lat0 /= nv;
            double cos_lat0 = M_COS(lat0 * DEG2R);
            double vx[4], vy[4];
            for (int v = 0; v < nv; v++) {
                double dlon = PYMOD360(m->nodes_lon[verts[v]] - lon0 + 180.0) - 180.0;
                double dlat = m->nodes_lat[verts[v]] - lat0;
                vx[v] = dlon * DEG2R * R * cos_lat0;
                vy[v] = dlat * DEG2R * R;
            }
            double dlon_p = PYMOD360(lon_p - lon0 + 180.0) - 180.0;
            double px = dlon_p * DEG2R * R * cos_lat0;
            double py = (lat_p - lat0) * DEG2R * R;
            for (int t = 0; t < nv - 2; t++) {
                int i1 = t + 1, i2 = t + 2;
                double ax = vx[0], ay = vy[0], bx = vx[i1], by = vy[i1], cx = vx[i2], cy = vy[i2];
                double denom = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
                if (fabs(denom) < 1e-14) continue;
                double w1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / denom;
                double w2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / denom;
                double w3 = 1.0 - w1 - w2;
                const double eps = -1e-6;
                if (w1 >= eps && w2 >= eps && w3 >= eps) {
                    if (nv == 4) {
                        double wts[4];
                        wachspress_quad(px, py, vx, vy, wts);
                        double acc = 0.0;
                        for (int vi = 0; vi < 4; vi++) {
                            int j;
                            for (j = 0; j < n_neigh; j++)
                                if (neigh[j] == verts[vi]) break;
                            if (j == n_neigh || !isfinite(vals[j])) return fallback;
                            acc += wts[vi] * vals[j];
                        }
                        return acc;
                    }
                    int idx[3] = {verts[0], verts[i1], verts[i2]};
                    double vv[3];
                    for (int q = 0; q < 3; q++) {
                        int j;
                        for (j = 0; j < n_neigh; j++)
                            if (neigh[j] == idx[q]) break;
                        if (j == n_neigh) return fallback;
                        vv[q] = vals[j];
                    }
                    if (!(isfinite(vv[0]) && isfinite(vv[1]) && isfinite(vv[2]))) return fallback;
                    return w1 * vv[0] + w2 * vv[1] + w3 * vv[2];
                }
            }
        }
        return fallback;
    }
    /* leastsq */
    if (n_neigh < 3) return fallback;
    double z0 = M_COS(lat_p * DEG2R) * SLM;
    if (z0 < 1.0) z0 = 1.0;
    double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int n_fin = 0;
    for (int i = 0; i < n_neigh; i++) {
        int nd = neigh[i];
        double val = vals[i];
        if (!isfinite(val)) continue;
        n_fin++;
        double dlon = PYMOD360(m->nodes_lon[nd] - lon_p + 180.0) - 180.0;
        double ex = dlon * z0;
        double ey = (m->nodes_lat[nd] - lat_p) * SLM;
        a11 += 1.0;
        a12 += ex;
        a13 += ey;
        a22 += ex * ex;
        a23 += ex * ey;
        a33 += ey * ey;
        f0 += val;
        f1 += val * ex;
        f2 += val * ey;
    }
    if (n_fin < 3) return fallback;
    double a21 = a12, a31 = a13, a32 = a23;
    double det = (a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31 -
                  a12 * a21 * a33 - a11 * a23 * a32);
    if (fabs(det) < 1e-12 || fabs(det) < 1e-12 * fabs(a11 * a22 * a33)) return fallback;
    double inv = 1.0 / fabs(det);
    double b11 = a22 * a33 - a23 * a32;
    double b12 = a13 * a32 - a12 * a33;
    double b13 = a12 * a23 - a13 * a22;
    (void)s;
    (void)z;
    return (b11 * f0 + b12 * f1 + b13 * f2) * inv;
}

/* sampling.py:727 sample_field on a z-level mesh for each particle:
 * NaN where not alive or without an anchor; 2-D and mid-layer fields by the
 * 1-ring combination, interface fields by the W interpolators
 * (vertical.py:1330 interp_w_spatial, time_dir +1).  method: 0 nearest,
 * 1 leastsq, 2 barycentric, 3 idw. */
void or_sample(const or_mesh_t *m, const double *fa, const double *fb, double alpha, int n_planes,
               int method, int64_t P, const double *lon, const double *lat, const double *z,
               const int32_t *status, const int32_t *nn, double *out)
{
    or_scalar_t s = {fa, fb ? fb : fa, alpha, n_planes};
    or_fields_t fw;
    memset(&fw, 0, sizeof fw);
    fw.w_a = fa;
    fw.w_b = fb ? fb : fa;
    fw.w_t = fa;
    fw.dtype = OR_F64;
    fw.alpha0 = alpha;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < P; i++) {
        if (status[i] != ST_ALIVE || nn[i] < 0) {
            out[i] = NAN;
            continue;
        }
        int node = nn[i];
        if (n_planes == m->n_levels && n_planes > 1) {
            if (method == 0)
                out[i] = w_linear(m, &fw, 0.0, z[i], node, 1.0);
            else if (method == 1)
                out[i] = w_leastsq(m, &fw, 0.0, lon[i], lat[i], z[i], node, 1.0);
            else if (method == 3)
                out[i] = w_idw(m, &fw, 0.0, lon[i], lat[i], z[i], node, 1.0);
            else
                out[i] = w_bary(m, &fw, 0.0, lon[i], lat[i], z[i], node, 1.0);
            continue;
        }
        int buf[64];
        double vals[64];
        int count = collect_neighbours(m, node, buf);
        for (int k = 0; k < count; k++) vals[k] = node_value(m, &s, z ? z[i] : 0.0, buf[k]);
        out[i] = combine_horizontal(m, &s, z ? z[i] : 0.0, buf, count, vals, lon[i], lat[i], node, method);
    }
}

/* numerics/temporal_reconstruct.py:168 _fc_slope_scalar */
static double fc_slope(double d_prev, double d_next, double h_prev, double h_next)
{
    if (d_prev * d_next > 0.0) {
        double w1 = 2.0 * h_next + h_prev;
        double w2 = h_next + 2.0 * h_prev;
        return (w1 + w2) / (w1 / d_prev + w2 / d_next);
    }
    return 0.0;
}

/* temporal_reconstruct.py:184 _pchip_value_scalar over n values of four
 * snapshots (float64 inputs) at t0 < t1 < t2 < t3, evaluated at tq.
 * pinned by the reference golden fesom2_node_rk2_fused_pchip (fwd/bwd) and
 * by scipy's PchipInterpolator on the central interval. */
void or_pchip4(int64_t n, const double *f0, const double *f1, const double *f2, const double *f3, double t0,
               double t1, double t2, double t3, double tq, double *out)
{
    for (int64_t i = 0; i < n; i++) {
        double h0 = t1 - t0, h1 = t2 - t1, h2 = t3 - t2;
        double d0 = (f1[i] - f0[i]) / h0;
        double d1 = (f2[i] - f1[i]) / h1;
        double d2 = (f3[i] - f2[i]) / h2;
        double m1 = fc_slope(d0, d1, h0, h1);
        double m2 = fc_slope(d1, d2, h1, h2);
        double s = (tq - t1) / h1;
        double s2 = s * s;
        double s3 = s2 * s;
        double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
        double h10 = s3 - 2.0 * s2 + s;
        double h01 = -2.0 * s3 + 3.0 * s2;
        double h11 = s3 - s2;
        out[i] = h00 * f1[i] + h10 * h1 * m1 + h01 * f2[i] + h11 * h1 * m2;
    }
}

/* Thread count of the threaded build's particle loops (1 in the others);
 * returns the count in effect. */
int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * The occupancy cell rule's per-sample statistics (core/mesh.py:386
 * _gh_sample_stats; the weights _GH_COST_CELL = 1, _GH_COST_CERT = 20 of
 * :351-352), against a candidate grid given as the fields of a mesh copy
 * (gh_cell, gh_lon_min, gh_lat_min, gh_nx, gh_ny, gh_wrap) and its node
 * counts as the 2-D prefix sum cum ((ny+1) x (nx+1)).
 * pinned by tests/test_cell_rule.py: occupancy and the approximate-anchor
 * flag against a brute-force count and a brute-force nearest node, the
 * widening flag against the 3x3 certificate, the cost against its formula
 * on hand-built cases.
 * ---------------------------------------------------------------------- */

/* core/mesh.py:357 _gh_rect_sum */
static int64_t gh_rect_sum(const int32_t *cum, int nx, int ny, int wrap, int ci, int cj, int rx, int ry)
{
#define CUM(j, i) ((int64_t)cum[(int64_t)(j) * (nx + 1) + (i)])
    int j0 = cj - ry;
    if (j0 < 0) j0 = 0;
    int j1 = cj + ry + 1;
    if (j1 > ny) j1 = ny;
    if (wrap && 2 * rx + 1 >= nx) return CUM(j1, nx) - CUM(j0, nx);
    int i0 = ci - rx, i1 = ci + rx + 1;
    if (wrap) {
        if (i0 < 0)
            return (CUM(j1, i1) - CUM(j0, i1)) + (CUM(j1, nx) - CUM(j0, nx)) - (CUM(j1, i0 + nx) - CUM(j0, i0 + nx));
        if (i1 > nx)
            return (CUM(j1, nx) - CUM(j0, nx)) - (CUM(j1, i0) - CUM(j0, i0)) + (CUM(j1, i1 - nx) - CUM(j0, i1 - nx));
    } else {
        if (i0 < 0) i0 = 0;
        if (i1 > nx) i1 = nx;
    }
    return (CUM(j1, i1) - CUM(j0, i1)) - (CUM(j1, i0) - CUM(j0, i0));
#undef CUM
}

void or_gh_sample_stats(const or_mesh_t *m, int64_t n, const double *lon, const double *lat, const int32_t *nn,
                        const int32_t *cum, int64_t *occ_out, double *cost_out, int8_t *unprov_out,
                        int8_t *approx_out)
{
    const int nx = m->gh_nx, ny = m->gh_ny, wrap = m->gh_wrap;
    const double cell = m->gh_cell;
    for (int64_t i = 0; i < n; i++) {
        int ci, cj, cin, cjn;
        gh_cell_of(lon[i], lat[i], m, &ci, &cj);
        int64_t occ = gh_rect_sum(cum, nx, ny, wrap, ci, cj, 1, 1);
        occ_out[i] = occ;
        int node = nn[i];
        gh_cell_of(m->nodes_lon[node], m->nodes_lat[node], m, &cin, &cjn);
        int di = cin - ci;
        if (di < 0) di = -di;
        if (wrap && nx - di < di) di = nx - di;
        int dj = cjn - cj;
        if (dj < 0) dj = -dj;
        approx_out[i] = (di > 1 || dj > 1) ? 1 : 0;
        double theta, ext;
        cap_extents(lon[i], lat[i], m->nodes_lon[node], m->nodes_lat[node], &theta, &ext);
        int provable = occ > 0 && approx_out[i] == 0;
        if (provable) provable = anchor_provable_rect(m, lon[i], lat[i], node, 1, 1) == 0;
        if (provable) {
            cost_out[i] = (double)occ + 9.0 * 1.0 + 20.0;
            unprov_out[i] = 0;
        } else {
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
    }
}
