// kelp_batch_host.hip -- the C ABI of batched clouds: the cloud table, the
// per-step bracket kernel, the step sequence (eager or as a captured HIP
// graph) and the collection of a finished cloud's particles.  The batched
// step kernel itself lives in kelp_batch.hip (one unit per field dtype and
// vertical coordinate).
#include <cstring>
#include <vector>

#include "kelp_batch.cuh"

using namespace kelp;

#define fail kelp_fail

namespace {

// The payload of time t on a cycled axis of n snapshots snap_dt apart
// (tracking.py:3136 _build_vel_payload with time.on_exhausted: cycle):
// bracket (ia, ib) with alpha0 snapped to 0 / 1 within 1e-12, stage target
// it (b forward, a backward) and the time to it.
__device__ inline void cloud_bracket(double t, double snap_dt, int n, double dir, int& ia, int& ib, double& a0,
                                     int& it, double& dtn)
{
    int64_t k = (int64_t)floor(t / snap_dt);
    double ta = (double)k * snap_dt;
    double tb = ta + snap_dt;
    if (dir < 0.0 && t == ta) {
        k -= 1;
        tb = ta;
        ta = ta - snap_dt;
    }
    const int ka = (int)(((k % n) + n) % n);
    const int kb = (int)((((k + 1) % n) + n) % n);
    const double alpha = (t - ta) / (tb - ta);
    if (dir > 0.0) {
        dtn = tb - t;
        it = kb;
    } else {
        dtn = t - ta;
        it = ka;
    }
    if (fabs(alpha) < 1e-12) {
        ia = ka;
        ib = -1;
        a0 = 0.0;
    } else if (fabs(alpha - 1.0) < 1e-12) {
        ia = kb;
        ib = -1;
        a0 = 0.0;
    } else {
        ia = ka;
        ib = kb;
        a0 = alpha;
    }
}

// Every cloud's step at the device step counter, then the counter advances
// (one workgroup: the counter is read by every thread before thread 0
// writes it).
__global__ __launch_bounds__(256) void batch_bracket_kernel(const kelp_cloud* __restrict__ clouds, int64_t n_clouds,
                                                            kelp_batch_axis ax, double dt, int order,
                                                            int64_t* __restrict__ counter,
                                                            BatchStep* __restrict__ out)
{
    const int64_t g = *counter;
    for (int64_t c = threadIdx.x; c < n_clouds; c += blockDim.x) {
        const kelp_cloud cd = clouds[c];
        BatchStep bs{};
        const int64_t k = g - cd.start_step;
        bs.active = (cd.flags & KELP_CLOUD_VALID) && k >= 0 && k < cd.n_steps ? 1 : 0;
        bs.slot_a = bs.slot_b = bs.slot_t = -1;
        if (bs.active) {
            const bool d3 = (cd.flags & KELP_CLOUD_3D) != 0;
            const int slot0 = d3 ? ax.slot0_3d : ax.slot0_2d;
            const int n = d3 ? ax.n_snap_3d : ax.n_snap_2d;
            bs.step_i = (uint64_t)k;
            if (n <= 1) {
                bs.slot_a = slot0;
            } else {
                const double t = cd.t0 + cd.time_dir * ((double)k * dt);
                int ia, ib, it;
                double a0, dtn;
                cloud_bracket(t, ax.snap_dt, n, cd.time_dir, ia, ib, a0, it, dtn);
                bs.slot_a = slot0 + ia;
                bs.slot_b = ib >= 0 ? slot0 + ib : -1;
                bs.alpha0 = a0;
                // tracking.py:2560: a step longer than the time to the
                // target runs on the bracket alone (steady RK)
                if (!(dt > dtn)) {
                    bs.slot_t = slot0 + it;
                    const double cs[4] = {0.0, 0.5, 0.5, 1.0};
                    for (int j = 0; j < order && j < 4; ++j) bs.alphas[j] = cs[j] * dt / dtn;
                }
            }
        }
        out[c] = bs;
    }
    __syncthreads();
    if (threadIdx.x == 0) *counter = g + 1;
}

constexpr uint64_t kLocalMask = (uint64_t(1) << KELP_CLOUD_SHIFT) - 1;

__global__ void batch_collect_kernel(int64_t n, int64_t* __restrict__ pid, int32_t* __restrict__ status,
                                     const double* __restrict__ lon, const double* __restrict__ lat,
                                     const double* __restrict__ z, const int8_t* __restrict__ fr, int64_t cloud,
                                     int64_t n_out, double* __restrict__ olon, double* __restrict__ olat,
                                     double* __restrict__ oz, int32_t* __restrict__ ostatus,
                                     int8_t* __restrict__ ofr, int release)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t p = (uint64_t)pid[i];
    if ((int64_t)(p >> KELP_CLOUD_SHIFT) != cloud) return;
    const int64_t local = (int64_t)(p & kLocalMask);
    if (local < n_out) {
        if (olon) olon[local] = lon[i];
        if (olat) olat[local] = lat[i];
        if (oz && z) oz[local] = z[i];
        if (ostatus) ostatus[local] = status[i];
        if (ofr) ofr[local] = fr[i];
    }
    if (release) {
        status[i] = 0;
        pid[i] = (int64_t)((uint64_t)KELP_MAX_CLOUDS << KELP_CLOUD_SHIFT);
    }
}

bool slot_ok(kelp_ctx* c, int k, int dtype, bool is3d)
{
    if (k < 0 || k >= kMaxSlots) return false;
    const FieldSlot& s = c->slots[k];
    return s.dtype == dtype && s.layout == 0 && s.is3d == is3d;
}

// Validate a batched run against the base arguments and the axes; fill the
// kernel configuration.  Waits (eagerly) for pending refills of the axis
// slots and builds a sigma mesh's surfaces.
int batch_prepare(kelp_ctx* c, const kelp_particles* p, const kelp_step_args* base, StepCfg& cfg, PState& ps,
                  SlotRecs& sr, int& block, int& dtype, bool& sigma)
{
    if (!c || !p || !base) return fail(KELP_EINVAL, "kelp_batch: NULL argument");
    if (base->order != 1 && base->order != 2 && base->order != 4)
        return fail(KELP_EINVAL, "kelp_batch: order must be 1, 2 or 4 (got %d)", base->order);
    if (base->walk) return fail(KELP_EINVAL, "kelp_batch: the edge walk is not batched");
    if (!p->pid) return fail(KELP_EINVAL, "kelp_batch: batched particles need pid (cloud << 40 | local id)");
    if (c->batch_n_clouds <= 0 || !c->batch_counter.p)
        return fail(KELP_ESTATE, "kelp_batch: no cloud table (kelp_batch_set_clouds)");
    block = base->block_size ? base->block_size : 256;
    if (block != 64 && block != 128 && block != 256)
        return fail(KELP_EINVAL, "kelp_batch: block_size must be 64, 128 or 256 (got %d)", base->block_size);
    int rc = kelp_step_prepare(c, p, base, cfg, ps);
    if (rc) return rc;
    const kelp_batch_axis& ax = c->batch_axis;
    dtype = c->slots[base->slot_a].dtype;
    sigma = c->vertical_coord == 1;
    if (c->slots[base->slot_a].layout != 0) return fail(KELP_EINVAL, "kelp_batch: node-layout u/v only");
    if (sigma && ax.n_snap_3d > 1) return fail(KELP_EINVAL, "kelp_batch: on a sigma mesh the 3-D axis must be steady");
    if (ax.n_snap_3d <= 0 && ax.n_snap_2d <= 0) return fail(KELP_EINVAL, "kelp_batch: no snapshot axis");
    if (ax.n_snap_3d > 1 || ax.n_snap_2d > 1) {
        if (!(ax.snap_dt > 0.0)) return fail(KELP_EINVAL, "kelp_batch: snap_dt must be positive");
    }
    for (int d3 = 0; d3 < 2; ++d3) {
        const int s0 = d3 ? ax.slot0_3d : ax.slot0_2d, n = d3 ? ax.n_snap_3d : ax.n_snap_2d;
        for (int k = 0; k < n; ++k) {
            if (!slot_ok(c, s0 + k, dtype, d3 != 0))
                return fail(KELP_ESTATE, "kelp_batch: %s axis slot %d is empty or differs in dtype, layout or "
                            "dimension", d3 ? "3-D" : "2-D", s0 + k);
            if (d3 && c->slots[s0 + k].n_planes != c->m.n_levels - 1)
                return fail(KELP_EINVAL, "kelp_batch: 3-D axis slot %d does not match the mesh's levels", s0 + k);
            rc = kelp_wait_slot(c, c->slots[s0 + k]);
            if (rc) return rc;
        }
    }
    if (ax.n_snap_3d > 0 && !p->z) return fail(KELP_EINVAL, "kelp_batch: 3-D clouds need z");
    if (dtype == KELP_DT_F64) return fail(KELP_EINVAL, "kelp_batch: float64 snapshots are not batched (float32, "
                                          "bfloat16 or float16 with level scales)");
    if ((dtype == KELP_DT_BF16 || dtype == KELP_DT_F16S) && sigma)
        return fail(KELP_EINVAL, "kelp_batch: bfloat16 / float16 snapshots on a sigma mesh");
    for (int k = 0; k < kMaxSlots; ++k) {
        sr.rec[k] = c->slots[k].rec;
        sr.scale[k] = c->slots[k].scale;
    }
    sr.row_len = (uint32_t)c->m.n_nodes;
    return KELP_OK;
}

int batch_enqueue(kelp_ctx* c, const kelp_step_args* base, const StepCfg& cfg, const PState& ps,
                  const SlotRecs& sr, int block, int dtype, bool sigma, int n_steps)
{
    const kelp_cloud* clouds = static_cast<const kelp_cloud*>(c->batch_clouds.p);
    BatchStep* steps = static_cast<BatchStep*>(c->batch_steps.p);
    int64_t* counter = static_cast<int64_t*>(c->batch_counter.p);
    for (int s = 0; s < n_steps; ++s) {
        batch_bracket_kernel<<<1, 256, 0, c->stream>>>(clouds, c->batch_n_clouds, c->batch_axis, base->dt,
                                                       base->order, counter, steps);
        if (ps.n > 0) {
            if (dtype == KELP_DT_BF16)
                launch_batch<Bf16, false>(c->stream, base->order, block, c->m, sr, cfg, ps, clouds, steps,
                                          c->batch_n_clouds);
            else if (dtype == KELP_DT_F16S)
                launch_batch<F16s, false>(c->stream, base->order, block, c->m, sr, cfg, ps, clouds, steps,
                                          c->batch_n_clouds);
            else if (dtype == KELP_DT_F32 && sigma)
                launch_batch<float, true>(c->stream, base->order, block, c->m, sr, cfg, ps, clouds, steps,
                                          c->batch_n_clouds);
            else
                launch_batch<float, false>(c->stream, base->order, block, c->m, sr, cfg, ps, clouds, steps,
                                           c->batch_n_clouds);
        }
        HIP_TRY(hipGetLastError());
    }
    return KELP_OK;
}

void drop_graph(kelp_ctx* c)
{
    if (c->batch_graph) (void)hipGraphExecDestroy(c->batch_graph);
    c->batch_graph = nullptr;
    c->batch_graph_key.clear();
}

template <typename V>
void put(std::vector<unsigned char>& k, const V& v)
{
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&v);
    k.insert(k.end(), b, b + sizeof(V));
}

}  // namespace

int kelp_batch_set_clouds(kelp_ctx* c, int64_t first, int64_t n, const kelp_cloud* host_clouds,
                          const kelp_batch_axis* axis)
{
    if (!c || (n > 0 && !host_clouds)) return fail(KELP_EINVAL, "kelp_batch_set_clouds: NULL argument");
    if (first < 0 || n < 0 || first + n > KELP_MAX_CLOUDS)
        return fail(KELP_EINVAL, "kelp_batch_set_clouds: entries [%lld, %lld) outside [0, %d)", (long long)first,
                    (long long)(first + n), KELP_MAX_CLOUDS);
    for (int64_t k = 0; k < n; ++k) {
        const kelp_cloud& cd = host_clouds[k];
        if ((cd.flags & KELP_CLOUD_VALID) &&
            (cd.n_steps < 0 || (cd.time_dir != 1.0 && cd.time_dir != -1.0) || !(cd.kh_x >= 0.0) || !(cd.kh_y >= 0.0)))
            return fail(KELP_EINVAL, "kelp_batch_set_clouds: cloud %lld: n_steps >= 0, time_dir +-1, Kh >= 0",
                        (long long)(first + k));
    }
    HIP_TRY(hipSetDevice(c->device));
    const int64_t need = first + n > c->batch_n_clouds ? first + n : c->batch_n_clouds;
    const int64_t cap = (int64_t)(c->batch_clouds.bytes / sizeof(kelp_cloud));
    if (need > cap || !c->batch_clouds.p) {
        int64_t ncap = cap * 2 > need ? cap * 2 : need;
        if (ncap < 256) ncap = 256;
        if (ncap > KELP_MAX_CLOUDS) ncap = KELP_MAX_CLOUDS;
        DevBuf nb;
        HIP_TRY(hipMalloc(&nb.p, (size_t)ncap * sizeof(kelp_cloud)));
        nb.bytes = (size_t)ncap * sizeof(kelp_cloud);
        hipError_t e = hipMemsetAsync(nb.p, 0, nb.bytes, c->stream);
        if (e == hipSuccess && c->batch_clouds.p && c->batch_n_clouds > 0)
            e = hipMemcpyAsync(nb.p, c->batch_clouds.p, (size_t)c->batch_n_clouds * sizeof(kelp_cloud),
                               hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipFree(nb.p);
            return fail(KELP_EHIP, "kelp_batch_set_clouds: %s", hipGetErrorString(e));
        }
        kelp_release(c->batch_clouds);
        c->batch_clouds = nb;
        int rc = kelp_reserve(c->batch_steps, (size_t)ncap * sizeof(BatchStep));
        if (rc) return rc;
        drop_graph(c);
    }
    if (!c->batch_counter.p) {
        int rc = kelp_reserve(c->batch_counter, sizeof(int64_t));
        if (rc) return rc;
        HIP_TRY(hipMemsetAsync(c->batch_counter.p, 0, sizeof(int64_t), c->stream));
    }
    kelp_cloud* d = static_cast<kelp_cloud*>(c->batch_clouds.p);
    if (first > c->batch_n_clouds)
        HIP_TRY(hipMemsetAsync(d + c->batch_n_clouds, 0, (size_t)(first - c->batch_n_clouds) * sizeof(kelp_cloud),
                               c->stream));
    if (n > 0)
        HIP_TRY(hipMemcpyAsync(d + first, host_clouds, (size_t)n * sizeof(kelp_cloud), hipMemcpyHostToDevice,
                               c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (need != c->batch_n_clouds) drop_graph(c);
    c->batch_n_clouds = need;
    if (axis) {
        if (axis->n_snap_3d < 0 || axis->n_snap_2d < 0 ||
            (axis->n_snap_3d > 0 && (axis->slot0_3d < 0 || axis->slot0_3d + axis->n_snap_3d > kMaxSlots)) ||
            (axis->n_snap_2d > 0 && (axis->slot0_2d < 0 || axis->slot0_2d + axis->n_snap_2d > kMaxSlots)))
            return fail(KELP_EINVAL, "kelp_batch_set_clouds: axis slots outside [0, %d)", kMaxSlots);
        if (std::memcmp(axis, &c->batch_axis, sizeof(kelp_batch_axis)) != 0) drop_graph(c);
        c->batch_axis = *axis;
    }
    return KELP_OK;
}

int kelp_batch_reset(kelp_ctx* c, int64_t n_clouds)
{
    if (!c) return fail(KELP_EINVAL, "kelp_batch_reset: NULL context");
    if (n_clouds < 0 || n_clouds > KELP_MAX_CLOUDS)
        return fail(KELP_EINVAL, "kelp_batch_reset: n_clouds outside [0, %d]", KELP_MAX_CLOUDS);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graph(c);
    c->batch_n_clouds = 0;
    c->batch_axis = kelp_batch_axis{};
    if (n_clouds > 0) {
        std::vector<kelp_cloud> empty((size_t)n_clouds);
        std::memset(empty.data(), 0, empty.size() * sizeof(kelp_cloud));
        int rc = kelp_batch_set_clouds(c, 0, n_clouds, empty.data(), nullptr);
        if (rc) return rc;
    }
    return kelp_batch_set_step(c, 0);
}

int kelp_batch_set_step(kelp_ctx* c, int64_t step)
{
    if (!c) return fail(KELP_EINVAL, "kelp_batch_set_step: NULL context");
    HIP_TRY(hipSetDevice(c->device));
    int rc = kelp_reserve(c->batch_counter, sizeof(int64_t));
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->batch_counter.p, &step, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return KELP_OK;
}

int kelp_batch_get_step(kelp_ctx* c, int64_t* step)
{
    if (!c || !step) return fail(KELP_EINVAL, "kelp_batch_get_step: NULL argument");
    if (!c->batch_counter.p) return fail(KELP_ESTATE, "kelp_batch_get_step: no cloud table");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(step, c->batch_counter.p, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return KELP_OK;
}

int kelp_batch_step(kelp_ctx* c, const kelp_particles* p, const kelp_step_args* base)
{
    StepCfg cfg{};
    PState ps{};
    SlotRecs sr{};
    int block = 256, dtype = 0;
    bool sigma = false;
    int rc = batch_prepare(c, p, base, cfg, ps, sr, block, dtype, sigma);
    if (rc) return rc;
    ps.n = p->n;
    return batch_enqueue(c, base, cfg, ps, sr, block, dtype, sigma, 1);
}

int kelp_batch_run_graph(kelp_ctx* c, const kelp_particles* p, const kelp_step_args* base, int n_steps)
{
    if (n_steps < 1 || n_steps > 4096) return fail(KELP_EINVAL, "kelp_batch_run_graph: n_steps in [1, 4096]");
    StepCfg cfg{};
    PState ps{};
    SlotRecs sr{};
    int block = 256, dtype = 0;
    bool sigma = false;
    int rc = batch_prepare(c, p, base, cfg, ps, sr, block, dtype, sigma);
    if (rc) return rc;
    ps.n = p->n;
    std::vector<unsigned char> key;
    put(key, *p);
    put(key, *base);
    put(key, n_steps);
    put(key, c->batch_clouds.p);
    put(key, c->batch_n_clouds);
    put(key, c->batch_axis);
    put(key, sr);
    if (!c->batch_graph || key != c->batch_graph_key) {
        drop_graph(c);
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        rc = batch_enqueue(c, base, cfg, ps, sr, block, dtype, sigma, n_steps);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(KELP_EHIP, "kelp_batch_run_graph: capture: %s", hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(KELP_EHIP, "kelp_batch_run_graph: instantiate: %s", hipGetErrorString(e));
        c->batch_graph = x;
        c->batch_graph_key = key;
    }
    HIP_TRY(hipGraphLaunch(c->batch_graph, c->stream));
    return KELP_OK;
}

int kelp_batch_collect(kelp_ctx* c, const kelp_particles* p, int64_t cloud, int64_t n_out, double* lon,
                       double* lat, double* z, int32_t* status, int8_t* fr, int release)
{
    if (!c || !p) return fail(KELP_EINVAL, "kelp_batch_collect: NULL argument");
    if (!p->pid || !p->status || !p->lon || !p->lat || !p->fr)
        return fail(KELP_EINVAL, "kelp_batch_collect: missing particle array");
    if (cloud < 0 || cloud >= KELP_MAX_CLOUDS || n_out < 0)
        return fail(KELP_EINVAL, "kelp_batch_collect: cloud %lld outside [0, %d)", (long long)cloud, KELP_MAX_CLOUDS);
    if (p->n <= 0) return KELP_OK;
    HIP_TRY(hipSetDevice(c->device));
    const unsigned g = (unsigned)((p->n + 255) / 256);
    batch_collect_kernel<<<g, 256, 0, c->stream>>>(p->n, p->pid, p->status, p->lon, p->lat, p->z, p->fr, cloud,
                                                   n_out, lon, lat, z, status, fr, release ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return KELP_OK;
}
