// kelp_batch.hip -- many independent particle clouds advanced in one kernel
// per step, and the HIP graph that replays a run of such steps.
//
// The reference runs each release as its own loop (run_single_realization,
// src/kelp/tracking.py:1805, once per realization; utils/split_seeds.py:89
// shards a release into such runs).  A batch holds many such runs at once:
// every particle carries its cloud in the high bits of its pid, every cloud
// its own direction, start step, duration, time on the snapshot axis,
// random-walk diffusivity and RNG seed.  Each step runs two kernels:
//
//  * batch_bracket_kernel (one workgroup): for every cloud, whether it runs
//    this step, its cloud-local step and its snapshot bracket at its own
//    time (slots, alpha0, stage fractions), from the device step counter,
//    which it then advances -- so a captured graph replays any number of
//    steps without host work;
//  * batch_step_kernel: one particle per thread, the same per-particle step
//    as the plain step kernel (step_particle / sigma_particle) with the
//    cloud's values in place of the step-wide ones; 2-D and 3-D clouds in
//    the same launch (a wavefront holds one cloud's particles except at the
//    cloud boundaries).
//
// A particle's draws are keyed by its cloud-local id, the cloud's seed and
// the cloud-local step, so its result is the one of its cloud run alone.
//
// Compiled once per (field dtype, vertical coordinate): -DKB_F32=1 (with
// -DKB_SIGMA=0|1), -DKB_BF16=1 or -DKB_F16S=1 (z-level).
#include <type_traits>

#include "kelp_batch.cuh"
#include "kelp_sigma_step.cuh"

#if !defined(KB_F32) || !defined(KB_SIGMA)
#error "compile with -DKB_F32=0|1 -DKB_SIGMA=0|1 [-DKB_BF16=1]"
#endif

namespace kelp {

constexpr uint64_t kLocalMask = (uint64_t(1) << KELP_CLOUD_SHIFT) - 1;

// the records of a cloud's bracket slots (and their row scales for float16)
template <typename T, bool IS3D>
__device__ __forceinline__ Snap<T, IS3D> batch_snap(const SlotRecs& sr, int a, int b, int t)
{
    using R = typename RecOf<T, IS3D>::type;
    Snap<T, IS3D> f;
    f.a = static_cast<const R*>(sr.rec[a]);
    f.b = static_cast<const R*>(sr.rec[b]);
    f.t = static_cast<const R*>(sr.rec[t]);
    if constexpr (std::is_same<T, F16s>::value) {
        f.sa = sr.scale[a];
        f.sb = sr.scale[b];
        f.st = sr.scale[t];
        f.row_len = sr.row_len;
    }
    return f;
}

template <int ORDER, int MAXV, typename T, bool SIGMA>
__global__ __launch_bounds__(256) KELP_STEP_ATTR void batch_step_kernel(DMesh m, SlotRecs sr, StepCfg c, PState s,
                                                                        const kelp_cloud* __restrict__ clouds,
                                                                        const BatchStep* __restrict__ steps,
                                                                        int64_t n_clouds)
{
    const int64_t i = xcd_thread_index();
    if (i >= s.n) return;
    const uint64_t pid = (uint64_t)s.pid[i];
    const int64_t cl = (int64_t)(pid >> KELP_CLOUD_SHIFT);
    if (cl >= n_clouds) return;
    const BatchStep bs = steps[cl];
    if (!bs.active) return;
    const kelp_cloud cd = clouds[cl];
    const int32_t prev = s.status[i];
    count_alive(s.alive_count, prev == kAlive);
    StepCfg cc = c;
    cc.time_dir = cd.time_dir;
    cc.alpha0 = bs.alpha0;
    for (int k = 0; k < 4; ++k) cc.alphas[k] = bs.alphas[k];
    cc.kick = (cd.flags & KELP_CLOUD_KICK) ? 1 : 0;
    cc.kh_x = cd.kh_x;
    cc.kh_y = cd.kh_y;
    cc.seed = cd.seed;
    cc.beach_seed = cd.seed;
    cc.real_id = cd.real_id;
    cc.step_i = bs.step_i;
    const int sb = bs.slot_b >= 0 ? bs.slot_b : bs.slot_a;
    const int st = bs.slot_t >= 0 ? bs.slot_t : bs.slot_a;
    if (cd.flags & KELP_CLOUD_3D) {
        const Snap<T, true> f = batch_snap<T, true>(sr, bs.slot_a, sb, st);
        const Snap<T, false> fe{};
        if constexpr (SIGMA)
            sigma_particle<ORDER, MAXV, T, false>(m, f, fe, cc, s, i, prev, kLocalMask);
        else
            step_particle<ORDER, MAXV, true, T, false>(m, f, fe, cc, s, i, prev, kLocalMask);
    } else {
        const Snap<T, false> f = batch_snap<T, false>(sr, bs.slot_a, sb, st);
        const Snap<T, false> fe{};
        step_particle<ORDER, MAXV, false, T, false>(m, f, fe, cc, s, i, prev, kLocalMask);
    }
}

template <typename T, bool SIGMA>
void launch_batch(hipStream_t stream, int order, int block, const DMesh& m, const SlotRecs& sr, const StepCfg& c,
                  const PState& ps, const kelp_cloud* clouds, const BatchStep* steps, int64_t n_clouds)
{
    const unsigned g = (unsigned)((ps.n + block - 1) / block);
#define KB_LAUNCH(O, V) \
    batch_step_kernel<O, V, T, SIGMA><<<g, block, 0, stream>>>(m, sr, c, ps, clouds, steps, n_clouds)
    if (m.max_v == 3) {
        if (order == 1) KB_LAUNCH(1, 3);
        else if (order == 2) KB_LAUNCH(2, 3);
        else KB_LAUNCH(4, 3);
    } else {
        if (order == 1) KB_LAUNCH(1, 4);
        else if (order == 2) KB_LAUNCH(2, 4);
        else KB_LAUNCH(4, 4);
    }
#undef KB_LAUNCH
}

#if defined(KB_F16S) && KB_F16S
using KbT = F16s;
#elif defined(KB_BF16) && KB_BF16
using KbT = Bf16;
#elif KB_F32
using KbT = float;
#else
using KbT = double;
#endif

template void launch_batch<KbT, (bool)KB_SIGMA>(hipStream_t, int, int, const DMesh&, const SlotRecs&,
                                                const StepCfg&, const PState&, const kelp_cloud*,
                                                const BatchStep*, int64_t);

}  // namespace kelp
