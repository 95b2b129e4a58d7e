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

#ifdef ORACLE_KMATH
#include "../kelp_amd/csrc/kmath.h"
#define M_SIN km_sin
#define M_COS km_cos
#define M_ASIN km_asin
#define M_ACOS km_acos
#define M_ATAN2 km_atan2
#define M_LOG km_log
#define M_EXP km_exp
#define PYMOD360 km_pymod360
#else
#define M_SIN sin
#define M_COS cos
#define M_ASIN asin
#define M_ACOS acos
#define M_ATAN2 atan2
#define M_LOG log
#define M_EXP exp
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

/* integrators.py:1471 _fused_site_anchor */
int or_site_anchor(const or_mesh_t *m, double lon_p, double lat_p, int exact)
{
    if (!exact) return or_nearest_widened(m, lon_p, lat_p);
    return or_nearest_exact(m, lon_p, lat_p);
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
int or_interp_bary_tinterp(const or_mesh_t *m, const or_fields_t *f, double xp, double yp,
                           int nn, double time_dir, int layer, double alpha,
                           double *U_o, double *V_o, int *elem_o)
{
    *U_o = 0.0;
    *V_o = 0.0;
    int e = or_find_containing_local(m, xp, yp, nn);
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

/* ------------------------------------------------------------------------
 * Fused 2-D steps (node-layout barycentric)
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

/* One particle of integrators.py:2522 _fused_rk2_step_2d_bary_tinterp,
 * :2642 _fused_rk4_step_2d_bary_tinterp and the Euler restatement of
 * _euler_step_unfused (:4997) for barycentric node fields (stage alpha 0).
 * order: 1 Euler, 2 RK2, 4 RK4.  alphas: per-stage blend fractions. */
static void step2d_one(const or_mesh_t *m, const or_fields_t *f, int order, const double *alphas,
                       double dt, double time_dir, int exact,
                       double lon0, double lat0, int32_t nn_prev,
                       double *lon_o, double *lat_o, int32_t *st_o, int32_t *nn_o,
                       double *U_o, double *V_o, int8_t *fr_o)
{
    int node_i = or_site_anchor(m, lon0, lat0, exact);
    (void)nn_prev;
    double U1, V1;
    int ok1 = or_interp_bary_tinterp(m, f, lon0, lat0, node_i, time_dir, -1, alphas[0], &U1, &V1, NULL);
    *U_o = U1;
    *V_o = V1;
    if (!ok1) {
        *lon_o = lon0;
        *lat_o = lat0;
        *st_o = ST_OUTSIDE;
        *fr_o = FR_INTERP;
        *nn_o = node_i;
        return;
    }
    double lon_new, lat_new;
    if (order == 1) {
        lon_new = lon0 + dt * U1 / z0_of(lat0);
        lat_new = lat0 + dt * V1 / SLM;
        lon_new = wrap_lon(lon_new);
        lat_new = clamp_lat(lat_new);
    } else if (order == 2) {
        double z0 = z0_of(lat0);
        double lon_mid = lon0 + 0.5 * dt * U1 / z0;
        double lat_mid = lat0 + 0.5 * dt * V1 / SLM;
        lon_mid = wrap_lon(lon_mid);
        lat_mid = clamp_lat(lat_mid);
        int nn_mid = or_site_anchor(m, lon_mid, lat_mid, exact);
        double U2, V2;
        int ok2 = or_interp_bary_tinterp(m, f, lon_mid, lat_mid, nn_mid, time_dir, -1, alphas[1], &U2, &V2, NULL);
        if (!ok2) {
            *lon_o = lon0;
            *lat_o = lat0;
            *st_o = ST_OUTSIDE;
            *fr_o = FR_INTERP;
            *nn_o = node_i;
            return;
        }
        double z0m = z0_of(lat_mid);
        lon_new = lon0 + dt * U2 / z0m;
        lat_new = lat0 + dt * V2 / SLM;
        lon_new = wrap_lon(lon_new);
        lat_new = clamp_lat(lat_new);
    } else {
        double dlon1 = dt * U1 / z0_of(lat0);
        double dlat1 = dt * V1 / SLM;
        double lon2 = wrap_lon(lon0 + 0.5 * dlon1);
        double lat2 = clamp_lat(lat0 + 0.5 * dlat1);
        int nn2 = or_site_anchor(m, lon2, lat2, exact);
        double U2, V2;
        if (!or_interp_bary_tinterp(m, f, lon2, lat2, nn2, time_dir, -1, alphas[1], &U2, &V2, NULL))
            goto interp_fail;
        double dlon2 = dt * U2 / z0_of(lat2);
        double dlat2 = dt * V2 / SLM;
        double lon3 = wrap_lon(lon0 + 0.5 * dlon2);
        double lat3 = clamp_lat(lat0 + 0.5 * dlat2);
        int nn3 = or_site_anchor(m, lon3, lat3, exact);
        double U3, V3;
        if (!or_interp_bary_tinterp(m, f, lon3, lat3, nn3, time_dir, -1, alphas[2], &U3, &V3, NULL))
            goto interp_fail;
        double dlon3 = dt * U3 / z0_of(lat3);
        double dlat3 = dt * V3 / SLM;
        double lon4 = wrap_lon(lon0 + dlon3);
        double lat4 = clamp_lat(lat0 + dlat3);
        int nn4 = or_site_anchor(m, lon4, lat4, exact);
        double U4, V4;
        if (!or_interp_bary_tinterp(m, f, lon4, lat4, nn4, time_dir, -1, alphas[3], &U4, &V4, NULL))
            goto interp_fail;
        double dlon4 = dt * U4 / z0_of(lat4);
        double dlat4 = dt * V4 / SLM;
        double dlon = (dlon1 + 2.0 * dlon2 + 2.0 * dlon3 + dlon4) / 6.0;
        double dlat = (dlat1 + 2.0 * dlat2 + 2.0 * dlat3 + dlat4) / 6.0;
        lon_new = wrap_lon(lon0 + dlon);
        lat_new = clamp_lat(lat0 + dlat);
    }
    {
        int nn_new = or_site_anchor(m, lon_new, lat_new, exact);
        int e = or_find_containing_local(m, lon_new, lat_new, nn_new);
        if (e < 0) {
            *lon_o = lon0;
            *lat_o = lat0;
            *st_o = ST_OUTSIDE;
            *fr_o = FR_OOB;
            *nn_o = node_i;
        } else {
            *lon_o = lon_new;
            *lat_o = lat_new;
            *st_o = ST_ALIVE;
            *nn_o = nn_new;
        }
    }
    return;
interp_fail:
    *lon_o = lon0;
    *lat_o = lat0;
    *st_o = ST_OUTSIDE;
    *fr_o = FR_INTERP;
    *nn_o = node_i;
}

/* One fused 2-D step over all particles.  Dead particles (status != ALIVE)
 * keep position/status/nn; their U, V are 0 and fr 0, as the runner's
 * output allocation (integrators.py:3393-3400).  fr_out holds this step's
 * kernel failure codes only (the caller merges them stickily). */
void or_step2d(const or_mesh_t *m, const or_fields_t *f, int order, const double *alphas,
               double dt, double time_dir, int exact, int64_t P,
               const double *lon, const double *lat, const int32_t *status, const int32_t *nn,
               double *lon_out, double *lat_out, int32_t *status_out, int32_t *nn_out,
               double *U_out, double *V_out, int8_t *fr_out)
{
    for (int64_t i = 0; i < P; i++) {
        lon_out[i] = lon[i];
        lat_out[i] = lat[i];
        status_out[i] = status[i];
        nn_out[i] = nn[i];
        U_out[i] = 0.0;
        V_out[i] = 0.0;
        fr_out[i] = 0;
        if (status[i] != ST_ALIVE) continue;
        step2d_one(m, f, order, alphas, dt, time_dir, exact, lon[i], lat[i], nn[i],
                   &lon_out[i], &lat_out[i], &status_out[i], &nn_out[i], &U_out[i], &V_out[i], &fr_out[i]);
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
        default: out[i] = sqrt(x[i]); break;
        }
    }
}
