// kelp_sigma_step.cuh -- one particle's step on a sigma mesh (FESOM-C).
//
// The reference runs sigma meshes through its multi-pass steppers
// (integrators.py:4997 _euler_step_unfused, :5380 _rk2_step_unfused, :5760
// _rk4_step_unfused, sigma branches): every stage anchors, interpolates the
// column height H_p like W (vertical_sigma.py:1063 sigma_column_height),
// looks the layer up in that column, reads u/v on the layer and the
// dia-sigma flux W, and the step integrates the sigma fraction s = z / H_p;
// the particle ends at z = s_new H_p(end) clamped into its end column.  Here
// one thread carries one particle through all stages.  Used by the sigma
// step kernel and by the batched kernel.
#pragma once

#include "kelp_sigma.cuh"
#include "kelp_step_impl.cuh"

namespace kelp {

// One stage: anchor, column height with the stage surface, the stage depth
// (z itself, or s H_p), layer, u/v and W.
template <int MAXV, typename T, bool ELEM>
__device__ inline bool sig_site(const DMesh& m, const Snap<T, true>& f, const Snap<T, false>& fe,
                                const StepCfg& c, int k, double lon, double lat, bool from_s, double s_or_z,
                                int& nn, double& H, double& U, double& V, double& W)
{
    if (nn < 0) nn = site_anchor(m, lon, lat, c.exact);
    const double* eta = c.sig_eta[k];
    H = column_height<MAXV>(m, eta, c.Dmin, c.w_method, lon, lat, nn);
    const double zs = from_s ? s_or_z * H : s_or_z;
    const int layer = find_layer_sigma(m, zs, H < c.Dmin ? c.Dmin : H);
    bool ok;
    if constexpr (ELEM)
        ok = interp_lsq(m, fe, c.alpha0, c.alphas[k], c.time_dir, lon, lat, nn, (int64_t)layer * m.n_elems, U, V);
    else
        ok = interp_uv<MAXV>(m, f, c.alpha0, c.alphas[k], c.time_dir, lon, lat, nn, (int64_t)layer * m.n_nodes, U,
                             V);
    W = w_sigma<MAXV>(m, f, eta, c.Dmin, c.w_method, c.alpha0, c.alphas[k], lon, lat, zs, nn, c.time_dir);
    return ok;
}

// Returns the failure code (0 = alive); on failure the position is kept and
// nn is the start position's anchor.
template <int ORDER, int MAXV, typename T, bool ELEM>
__device__ inline int sigma_det_step(const DMesh& m, const Snap<T, true>& f, const Snap<T, false>& fe,
                                     const StepCfg& c, double& lon, double& lat, double& z, int& nn, double& Uo,
                                     double& Vo, double& Wo)
{
    const double dt = c.dt;
    const double lon0 = lon, lat0 = lat, zz0 = z;
    int n1 = (c.reuse_anchor && nn >= 0) ? nn : -1;
    double H1, U1, V1, W1;
    bool ok = sig_site<MAXV, T, ELEM>(m, f, fe, c, 0, lon0, lat0, false, zz0, n1, H1, U1, V1, W1);
    Uo = U1;
    Vo = V1;
    Wo = W1;
    double lon_new = lon0, lat_new = lat0, s_new = 0.0;
    if (ok) {
        if (ORDER == 1) {
            lon_new = wrap_lon(lon0 + dt * U1 / z0_of(lat0));
            lat_new = clamp_lat(lat0 + dt * V1 / kSLM);
            s_new = (zz0 + W1 * dt) / H1;
        } else if (ORDER == 2) {
            const double z0 = z0_of(lat0);
            const double lon_mid = wrap_lon(lon0 + 0.5 * dt * U1 / z0);
            const double lat_mid = clamp_lat(lat0 + 0.5 * dt * V1 / kSLM);
            const double s_mid = (zz0 + W1 * 0.5 * dt) / H1;
            int n2 = -1;
            double H2, U2, V2, W2;
            ok = sig_site<MAXV, T, ELEM>(m, f, fe, c, 1, lon_mid, lat_mid, true, s_mid, n2, H2, U2, V2, W2);
            if (ok) {
                lon_new = wrap_lon(lon0 + dt * U2 / z0_of(lat_mid));
                lat_new = clamp_lat(lat0 + dt * V2 / kSLM);
                s_new = zz0 / H1 + dt * W2 / H2;
            }
        } else {
            const double dlon1 = dt * U1 / z0_of(lat0);
            const double dlat1 = dt * V1 / kSLM;
            const double s0 = zz0 / H1;
            const double f1 = W1 / H1;
            const double lon2 = wrap_lon(lon0 + 0.5 * dlon1);
            const double lat2 = clamp_lat(lat0 + 0.5 * dlat1);
            int n2 = -1, n3 = -1, n4 = -1;
            double H2, U2, V2, W2, H3, U3, V3, W3, H4, U4, V4, W4;
            ok = sig_site<MAXV, T, ELEM>(m, f, fe, c, 1, lon2, lat2, true, s0 + 0.5 * dt * f1, n2, H2, U2, V2, W2);
            if (ok) {
                const double f2 = W2 / H2;
                const double dlon2 = dt * U2 / z0_of(lat2);
                const double dlat2 = dt * V2 / kSLM;
                const double lon3 = wrap_lon(lon0 + 0.5 * dlon2);
                const double lat3 = clamp_lat(lat0 + 0.5 * dlat2);
                ok = sig_site<MAXV, T, ELEM>(m, f, fe, c, 2, lon3, lat3, true, s0 + 0.5 * dt * f2, n3, H3, U3, V3,
                                             W3);
                if (ok) {
                    const double f3 = W3 / H3;
                    const double dlon3 = dt * U3 / z0_of(lat3);
                    const double dlat3 = dt * V3 / kSLM;
                    const double lon4 = wrap_lon(lon0 + dlon3);
                    const double lat4 = clamp_lat(lat0 + dlat3);
                    ok = sig_site<MAXV, T, ELEM>(m, f, fe, c, 3, lon4, lat4, true, s0 + dt * f3, n4, H4, U4, V4,
                                                 W4);
                    if (ok) {
                        const double f4 = W4 / H4;
                        const double dlon4 = dt * U4 / z0_of(lat4);
                        const double dlat4 = dt * V4 / kSLM;
                        const double dlon = (dlon1 + 2.0 * dlon2 + 2.0 * dlon3 + dlon4) / 6.0;
                        const double dlat = (dlat1 + 2.0 * dlat2 + 2.0 * dlat3 + dlat4) / 6.0;
                        lon_new = wrap_lon(lon0 + dlon);
                        lat_new = clamp_lat(lat0 + dlat);
                        s_new = s0 + dt * (f1 + 2.0 * f2 + 2.0 * f3 + f4) / 6.0;
                    }
                }
            }
        }
    }
    if (!ok) {
        nn = n1;
        return kFrInterp;
    }
    const int nn_new = site_anchor(m, lon_new, lat_new, c.exact);
    if (find_containing<MAXV>(m, lon_new, lat_new, nn_new) < 0) {
        nn = n1;
        return kFrOOB;
    }
    lon = lon_new;
    lat = lat_new;
    const double H_end = column_height<MAXV>(m, c.sig_eta[4], c.Dmin, c.w_method, lon_new, lat_new, nn_new);
    z = clamp_sigma(s_new * H_end, H_end, c.Dmin, c.z0b_min);
    nn = nn_new;
    return 0;
}

// One particle's full sigma step at slot i (status `prev` at the step
// start; draws keyed by pid[i] & key_mask).
template <int ORDER, int MAXV, typename T, bool ELEM>
__device__ __forceinline__ void sigma_particle(const DMesh& m, const Snap<T, true>& f, const Snap<T, false>& fe,
                                               const StepCfg& c, const PState& s, int64_t i, int32_t prev,
                                               uint64_t key_mask)
{
    double U = 0.0, V = 0.0, W = 0.0;
    if (prev != kAlive) {
        // the diagnostic velocity of a particle that does not move is 0;
        // beaching still looks at beached particles (resuspension)
        int32_t st = prev;
        int nn = s.nn[i];
        int8_t fr = s.fr[i];
        if (c.beach) {
            beach_one<MAXV>(m, c, (uint64_t)s.pid[i] & key_mask, prev, s.lon[i], s.lat[i], st, nn, fr, s.ws[i]);
            s.status[i] = st;
            s.nn[i] = nn;
            s.fr[i] = fr;
        }
        s.u[i] = U;
        s.v[i] = V;
        if (s.w) s.w[i] = W;
        return;
    }
    int32_t st = prev;
    double lon = s.lon[i], lat = s.lat[i], z = s.z[i];
    int nn = s.nn[i];
    int8_t fr = s.fr[i];
    const double pre_lon = lon, pre_lat = lat, pre_z = z;
    const int pre_nn = nn;
    const int code = sigma_det_step<ORDER, MAXV, T, ELEM>(m, f, fe, c, lon, lat, z, nn, U, V, W);
    if (code) {
        st = kOutside;
        fr = (int8_t)code;
    } else if (c.kick && !(c.kh_x == 0.0 && c.kh_y == 0.0)) {
        kick_one<MAXV, true, true>(m, c, (uint64_t)s.pid[i] & key_mask, pre_lon, pre_lat, pre_z, pre_nn, lon, lat, z, st, nn,
                                   fr);
    }
    if (c.az && st == kAlive) az_sigma_one<MAXV>(m, c, (uint64_t)s.pid[i] & key_mask, lon, lat, nn, z);
    int elem = -1;
    if ((c.periodic_lon || c.periodic_lat) && st == kAlive) periodic_one(m, c, lon, lat, nn, elem);
    if (c.beach) beach_one<MAXV>(m, c, (uint64_t)s.pid[i] & key_mask, prev, lon, lat, st, nn, fr, s.ws[i]);
    s.lon[i] = lon;
    s.lat[i] = lat;
    s.z[i] = z;
    s.status[i] = st;
    s.nn[i] = nn;
    s.fr[i] = fr;
    s.u[i] = U;
    s.v[i] = V;
    if (s.w) s.w[i] = W;
}


}  // namespace kelp
