// kelp_internal.h -- context and error plumbing shared by the library's
// translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <vector>

#include "../../include/kelp_hip.h"
#include "kelp_device.cuh"
#include "kelp_step.cuh"

int kelp_fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return kelp_fail(e_ == hipErrorOutOfMemory ? KELP_ENOMEM : KELP_EHIP, "%s: %s (%s:%d)", \
                             #expr, hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

constexpr int kMaxSlots = 16;
constexpr int kMaxEvents = 1024;
constexpr int kMaxScalars = 8;

// A node scalar of one snapshot for record-point sampling: n_planes planes
// of N doubles in internal node order (1: 2-D, K-1: mid-layer, K:
// interface); an interface field also as (0, 0, top, bottom) records per
// (layer, node), the layout the W interpolators read.
struct ScalarField {
    int64_t n_planes = 0;
    double* planes = nullptr;
    void* wrec = nullptr;
};

// One snapshot in HBM as packed records (see kelp_device.cuh, Snap): 3-D
// (u, v, w_top, w_bottom) per (layer, node), 2-D (u, v) per node.  With u/v
// at element centres (layout 1) the node records carry only the W pair (u, v
// zero) and `erec` holds (u, v) per (layer, element).
struct FieldSlot {
    int dtype = -1;  // -1: empty
    int layout = 0;  // 0: u/v at nodes, 1: u/v at element centres
    int64_t n_planes = 0, n_w_planes = 0;
    bool is3d = false;
    void* rec = nullptr;
    void* erec = nullptr;
    double* eta = nullptr;  // (N,) surface elevation, internal node order, or null
    double* az = nullptr;   // Az at the interfaces, one column of az_k per node, or null
    // two mid-layer scalars interleaved per (node, layer) for paired
    // sampling, built on first use; pair_vars = {-1, -1} when stale
    double* pair = nullptr;
    int pair_vars[2] = {-1, -1};
    // KELP_DT_F16S: per-row scales (float4 per layer), then the row maxima
    float4* scale = nullptr;
    int64_t az_k = 0;       // 0: no Az for the slot's current snapshot
    ScalarField scalars[kMaxScalars];
    // asynchronous refill (kelp_field_upload_async): device staging of the
    // host planes, and the copy stream's event the readers of the slot wait
    // on while `pending`
    void* stage[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t stage_bytes[4] = {0, 0, 0, 0};
    hipEvent_t ready = nullptr;
    bool pending = false;
};

// growable device buffer owned by a context
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct kelp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // snapshot refills (default priority)
    hipEvent_t compute_mark = nullptr;  // "work enqueued so far" on `stream`
    hipEvent_t events[kMaxEvents] = {};
    bool has_mesh = false;
    kelp::DMesh m{};
    std::vector<void*> mesh_bufs;  // owned device buffers of the mesh
    // internal renumbering (node id = grid-hash CSR slot); device arrays
    int32_t* node_old_of_new = nullptr;
    int32_t* node_new_of_old = nullptr;
    int32_t* elem_old_of_new = nullptr;
    int32_t* elem_new_of_old = nullptr;
    int vertical_coord = 0;        // 0 z-level, 1 sigma
    double* sig_eta = nullptr;     // sigma meshes: 7 x N (mesh-owned): the stage surfaces 0..3 and the
                                   // end surface of one step, then the given start and end surfaces
    bool sig_given[2] = {false, false};
    FieldSlot slots[kMaxSlots];
    DevBuf sort_tmp, sort_keys, sort_keys2, sort_idx, sort_perm, sort_stage;
    bool gh_tile = true;           // plain steps stage grid-hash tiles in LDS (KELP_GH_TILE=0: off)
    // batched clouds (kelp_batch_*): the cloud table, the per-cloud step
    // records the bracket kernel writes, the device step counter, and the
    // captured graph of a run of batched steps with what it was built for
    DevBuf batch_clouds, batch_steps, batch_counter;
    int64_t batch_n_clouds = 0;
    kelp_batch_axis batch_axis{};
    hipGraphExec_t batch_graph = nullptr;
    std::vector<unsigned char> batch_graph_key;
};

void kelp_free_scalar(ScalarField& f);

// Validation and kernel configuration of one step (kelp_step without its
// launch); the sigma stage surfaces are built on the context stream.
extern "C" int kelp_step_prepare(kelp_ctx* c, const kelp_particles* p, const kelp_step_args* a,
                                 kelp::StepCfg& cfg, kelp::PState& ps);

// Order `c->stream` after a pending refill of slot s (no-op otherwise).
int kelp_wait_slot(kelp_ctx* c, FieldSlot& s);

// (re)allocate `b` to at least `bytes` (contents not preserved)
int kelp_reserve(DevBuf& b, size_t bytes);
void kelp_release(DevBuf& b);
