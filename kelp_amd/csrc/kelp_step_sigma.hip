// kelp_step_sigma.hip -- the 3-D step on sigma meshes (FESOM-C).
//
// The reference runs sigma meshes through its multi-pass steppers
// (integrators.py:4997 _euler_step_unfused, :5380 _rk2_step_unfused, :5760
// _rk4_step_unfused, sigma branches): every stage anchors, interpolates the
// column height H_p like W (vertical_sigma.py:1063 sigma_column_height),
// looks the layer up in that column, reads u/v on the layer and the
// dia-sigma flux W, and the step integrates the sigma fraction s = z / H_p;
// the particle ends at z = s_new H_p(end) clamped into its end column.  Here
// one thread carries one particle through all stages.
//
// Compiled once per (field dtype, u/v layout): -DKS_F32=0|1 -DKS_ELEM=0|1;
// the (double, node) unit also holds the stage-surface kernel.
#include "kelp_sigma_step.cuh"

#if !defined(KS_F32) || !defined(KS_ELEM)
#error "compile with -DKS_F32=0|1 -DKS_ELEM=0|1"
#endif

namespace kelp {

namespace {

template <int ORDER, int MAXV, typename T, bool ELEM>
__global__ __launch_bounds__(256) void sigma_step_kernel(DMesh m, Snap<T, true> f, Snap<T, false> fe, StepCfg c,
                                                         PState s)
{
    const int64_t i = xcd_thread_index();
    if (i >= s.n) return;
    const int32_t prev = s.status[i];
    count_alive(s.alive_count, prev == kAlive);
    sigma_particle<ORDER, MAXV, T, ELEM>(m, f, fe, c, s, i, prev, ~0ull);
}

}  // namespace

template <int MAXV, typename T, bool ELEM>
void launch_sigma_step(hipStream_t stream, int order, const DMesh& m, const Snap<T, true>& f,
                       const Snap<T, false>& fe, const StepCfg& cfg, const PState& ps)
{
    const unsigned g = (unsigned)((ps.n + 255) / 256);
    if (order == 1)
        sigma_step_kernel<1, MAXV, T, ELEM><<<g, 256, 0, stream>>>(m, f, fe, cfg, ps);
    else if (order == 2)
        sigma_step_kernel<2, MAXV, T, ELEM><<<g, 256, 0, stream>>>(m, f, fe, cfg, ps);
    else
        sigma_step_kernel<4, MAXV, T, ELEM><<<g, 256, 0, stream>>>(m, f, fe, cfg, ps);
}

#if KS_F32
using KsT = float;
#else
using KsT = double;
#endif
template void launch_sigma_step<3, KsT, (bool)KS_ELEM>(hipStream_t, int, const DMesh&, const Snap<KsT, true>&,
                                                       const Snap<KsT, false>&, const StepCfg&, const PState&);
template void launch_sigma_step<4, KsT, (bool)KS_ELEM>(hipStream_t, int, const DMesh&, const Snap<KsT, true>&,
                                                       const Snap<KsT, false>&, const StepCfg&, const PState&);

#if !KS_F32 && !KS_ELEM
namespace {
// core/io.py:68 _blend chains of one node: the value of the stage field
// (nested time blend, the same short-circuits as tblend), sources read only
// where the blend uses them.
__device__ __forceinline__ double eta_of(const EtaSrc& e, int64_t n)
{
    if (e.alpha < 1e-12) {
        if (e.alpha0 == 0.0) return e.a[n];
        return (1.0 - e.alpha0) * e.a[n] + e.alpha0 * e.b[n];
    }
    if (e.alpha > 1.0 - 1e-12) return e.t[n];
    if (e.alpha0 == 0.0) return (1.0 - e.alpha) * e.a[n] + e.alpha * e.t[n];
    const double a0 = (1.0 - e.alpha0) * e.a[n] + e.alpha0 * e.b[n];
    return (1.0 - e.alpha) * a0 + e.alpha * e.t[n];
}

// integrators.py:1369 _eta_blend
__device__ __forceinline__ double eta_blend(double e0, double e1, double c)
{
    if (c <= 0.0) return e0;
    if (c >= 1.0) return e1;
    return (1.0 - c) * e0 + c * e1;
}

// the stage surfaces of integrators.py:5426-5498 (RK2) and :5806-5816 (RK4)
__global__ void sigma_eta_kernel(int64_t N, int order, int steady, EtaSrc start, EtaSrc end, int end_is_start,
                                 EtaSrc stages, double a1, double a2, double a3, double* __restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double e0 = eta_of(start, n);
    const double ee = end_is_start ? e0 : eta_of(end, n);
    out[n] = e0;
    out[4 * N + n] = ee;
    const double al[3] = {a1, a2, a3};
    const double cc[3] = {0.5, 0.5, 1.0};
    for (int k = 1; k < order; ++k) {
        double v;
        if (steady) {
            v = eta_blend(e0, ee, order == 2 ? 0.5 : cc[k - 1]);
        } else {
            EtaSrc s = stages;
            s.alpha = al[k - 1];
            v = eta_of(s, n);
        }
        out[(int64_t)k * N + n] = v;
    }
}

}  // namespace

void launch_sigma_eta(hipStream_t stream, int64_t n_nodes, int order, int steady, EtaSrc start, EtaSrc end,
                      int end_is_start, EtaSrc stages, const double* stage_alphas, double* out)
{
    const unsigned g = (unsigned)((n_nodes + 255) / 256);
    sigma_eta_kernel<<<g, 256, 0, stream>>>(n_nodes, order, steady, start, end, end_is_start, stages,
                                            stage_alphas[1], stage_alphas[2], stage_alphas[3], out);
}

#endif

}  // namespace kelp
