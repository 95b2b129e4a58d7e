// kelp_step_variant.hip -- the step kernels of one (3-D, field dtype, u/v
// layout) combination, chosen by -DKV_IS3D=0|1 -DKV_F32=0|1 -DKV_ELEM=0|1
// (-DKV_BF16=1: bfloat16 records, -DKV_F16S=1: float16 records with row
// scales, both node layout); the build compiles this file
// once per combination, in parallel.
#include "kelp_step_impl.cuh"

#if !defined(KV_IS3D) || !defined(KV_F32) || !defined(KV_ELEM)
#error "compile with -DKV_IS3D=0|1 -DKV_F32=0|1 -DKV_ELEM=0|1"
#endif

namespace kelp {

#if defined(KV_F16S) && KV_F16S
using KvT = F16s;
#elif defined(KV_BF16) && KV_BF16
using KvT = Bf16;
#elif KV_F32
using KvT = float;
#else
using KvT = double;
#endif

template void launch_step_v<3, (bool)KV_IS3D, KvT, (bool)KV_ELEM>(hipStream_t, int, const DMesh&,
                                                                  const Snap<KvT, (bool)KV_IS3D>&,
                                                                  const Snap<KvT, false>&, const StepCfg&,
                                                                  const PState&);
template void launch_step_v<4, (bool)KV_IS3D, KvT, (bool)KV_ELEM>(hipStream_t, int, const DMesh&,
                                                                  const Snap<KvT, (bool)KV_IS3D>&,
                                                                  const Snap<KvT, false>&, const StepCfg&,
                                                                  const PState&);

#if !KV_IS3D && KV_ELEM && !(defined(KV_BF16) && KV_BF16) && !(defined(KV_F16S) && KV_F16S)
// rk45 (2-D, element-centre u/v) lives with the 2-D element-layout steps
template void launch_rk45_v<3, KvT>(hipStream_t, const DMesh&, const Snap<KvT, false>&, const StepCfg&,
                                    const PState&);
template void launch_rk45_v<4, KvT>(hipStream_t, const DMesh&, const Snap<KvT, false>&, const StepCfg&,
                                    const PState&);
#endif

}  // namespace kelp
