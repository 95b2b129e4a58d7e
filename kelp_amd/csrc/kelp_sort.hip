// kelp_sort.hip -- periodic re-sort of the particle state by grid-hash cell.
//
// The step kernel is a gather over the mesh: each particle reads the nodes
// of its 3x3 hash window, the elements around its anchor and the snapshot
// values of their vertices.  With particles in release order, the 64 lanes of
// a wavefront touch 64 unrelated places and every gather pulls its own
// cache lines from HBM.  Sorting the state by the hash cell of each particle
// (row-major cell order, so a wavefront covers a few adjacent cells) makes
// neighbouring lanes share windows, elements and field lines in L2/MALL.
// Alive particles sort first, then beached ones, then the rest, then the
// free entries of a particle pool, so the tail of the grid holds the cheap
// non-alive lanes.
//
// Results do not depend on the order: every particle is computed
// independently and its random draws are keyed by its global id (pid), not
// by its slot.  Callers restore the release order from pid.
#include <rocprim/device/device_radix_sort.hpp>

#include <mutex>

#include "kelp_internal.h"

using namespace kelp;

namespace {

__global__ void sort_key_kernel(DMesh m, int64_t n, const double* __restrict__ lon,
                                const double* __restrict__ lat, const int32_t* __restrict__ status,
                                uint32_t* __restrict__ keys, int32_t* __restrict__ idx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ci, cj;
    gh_cell_of(m, lon[i], lat[i], ci, cj);
    uint32_t n_cells = (uint32_t)m.nx * (uint32_t)m.ny;
    uint32_t cell = (uint32_t)cj * (uint32_t)m.nx + (uint32_t)ci;
    int32_t st = status[i];
    // free entries of a particle pool (status 0, see kelp_batch_collect) last
    uint32_t band = st == kAlive ? 0u : (st == kBeached ? 1u : (st == 0 ? 3u : 2u));
    keys[i] = band * n_cells + cell;
    idx[i] = (int32_t)i;
}

template <typename T>
__global__ void gather_kernel(int64_t n, const int32_t* __restrict__ perm, const T* __restrict__ src,
                              T* __restrict__ dst)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    dst[j] = src[perm[j]];
}

unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

// dst <- src[perm] through the staging buffer, then copied back in place
template <typename T>
int permute(kelp_ctx* c, int64_t n, const int32_t* perm, T* arr)
{
    if (!arr) return KELP_OK;
    T* stage = static_cast<T*>(c->sort_stage.p);
    gather_kernel<T><<<grid_of(n), 256, 0, c->stream>>>(n, perm, arr, stage);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(arr, stage, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    return KELP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// The same re-sort without any device allocation: keys, permutation,
// staging and the radix sort's temporary storage live in one file-scope
// device array.  Each device holds its own copy of it for the life of the
// process; its address on a device comes from hipGetSymbolAddress (the
// array's name in host code is a stand-in, not a device address).
// ---------------------------------------------------------------------------

namespace {

constexpr int64_t kNoAllocMaxParticles = int64_t(1) << 20;
// keys, keys2, idx, perm (4 B each) + staging (8 B) per particle, the rest
// for the radix sort's temporary storage
constexpr size_t kScratchBytes = size_t(48) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// this device's copy of `scratch` (looked up once per device, with the
// device current)
int scratch_on(int device, unsigned char** out)
{
    if (device < 0 || device >= kMaxDevices)
        return kelp_fail(KELP_EINVAL, "kelp_sort_particles_noalloc: device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return kelp_fail(KELP_EHIP, "kelp_sort_particles_noalloc: no address for the scratch array");
        g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return KELP_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename T>
int permute_via(kelp_ctx* c, int64_t n, const int32_t* perm, T* arr, void* stage_p)
{
    if (!arr) return KELP_OK;
    T* stage = static_cast<T*>(stage_p);
    gather_kernel<T><<<grid_of(n), 256, 0, c->stream>>>(n, perm, arr, stage);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(arr, stage, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    return KELP_OK;
}

}  // namespace

extern "C" int kelp_sort_particles_noalloc(kelp_ctx* c, const kelp_particles* p)
{
    if (!c || !p) return kelp_fail(KELP_EINVAL, "kelp_sort_particles_noalloc: NULL argument");
    if (!c->has_mesh) return kelp_fail(KELP_ESTATE, "kelp_sort_particles_noalloc: no mesh uploaded");
    const int64_t n = p->n;
    if (n < 0) return kelp_fail(KELP_EINVAL, "kelp_sort_particles_noalloc: bad particle count");
    if (n > kNoAllocMaxParticles)
        return kelp_fail(KELP_EINVAL, "kelp_sort_particles_noalloc: %lld particles exceed the maximum of %lld",
                         (long long)n, (long long)kNoAllocMaxParticles);
    if (n <= 1) return KELP_OK;
    if (!p->lon || !p->lat || !p->status)
        return kelp_fail(KELP_EINVAL, "kelp_sort_particles_noalloc: lon, lat and status are required");
    const uint64_t n_cells = (uint64_t)c->m.nx * (uint64_t)c->m.ny;
    if (3 * n_cells >= (1ull << 32))
        return kelp_fail(KELP_EINVAL, "kelp_sort_particles_noalloc: too many hash cells");
    unsigned end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) < 3 * n_cells) ++end_bit;
    HIP_TRY(hipSetDevice(c->device));
    unsigned char* base = nullptr;
    int rc = scratch_on(c->device, &base);
    if (rc) return rc;
    const size_t b4 = align256((size_t)n * 4), b8 = align256((size_t)n * 8);
    uint32_t* keys = reinterpret_cast<uint32_t*>(base);
    uint32_t* keys2 = reinterpret_cast<uint32_t*>(base + b4);
    int32_t* idx = reinterpret_cast<int32_t*>(base + 2 * b4);
    int32_t* perm = reinterpret_cast<int32_t*>(base + 3 * b4);
    void* stage = base + 4 * b4;
    unsigned char* tmp = base + 4 * b4 + b8;
    const size_t used = 4 * b4 + b8;
    size_t tmp_bytes = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, idx, perm, (unsigned int)n, 0u, end_bit,
                                      c->stream));
    if (used + tmp_bytes > kScratchBytes)
        return kelp_fail(KELP_EINVAL, "kelp_sort_particles_noalloc: %lld particles need %zu scratch bytes (%zu held)",
                         (long long)n, used + tmp_bytes, kScratchBytes);
    sort_key_kernel<<<grid_of(n), 256, 0, c->stream>>>(c->m, n, p->lon, p->lat, p->status, keys, idx);
    HIP_TRY(hipGetLastError());
    HIP_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, idx, perm, (unsigned int)n, 0u, end_bit,
                                      c->stream));
    if ((rc = permute_via(c, n, perm, p->lon, stage)) || (rc = permute_via(c, n, perm, p->lat, stage)) ||
        (rc = permute_via(c, n, perm, p->z, stage)) || (rc = permute_via(c, n, perm, p->status, stage)) ||
        (rc = permute_via(c, n, perm, p->nn, stage)) || (rc = permute_via(c, n, perm, p->fr, stage)) ||
        (rc = permute_via(c, n, perm, p->u, stage)) || (rc = permute_via(c, n, perm, p->v, stage)) ||
        (rc = permute_via(c, n, perm, p->w, stage)) || (rc = permute_via(c, n, perm, p->pid, stage)) ||
        (rc = permute_via(c, n, perm, p->wet_streak, stage)) || (rc = permute_via(c, n, perm, p->elem, stage)))
        return rc;
    return KELP_OK;
}

extern "C" int kelp_sort_particles(kelp_ctx* c, const kelp_particles* p)
{
    if (!c || !p) return kelp_fail(KELP_EINVAL, "kelp_sort_particles: NULL argument");
    if (!c->has_mesh) return kelp_fail(KELP_ESTATE, "kelp_sort_particles: no mesh uploaded");
    const int64_t n = p->n;
    if (n < 0 || n >= INT32_MAX) return kelp_fail(KELP_EINVAL, "kelp_sort_particles: bad particle count");
    if (n <= 1) return KELP_OK;
    if (!p->lon || !p->lat || !p->status)
        return kelp_fail(KELP_EINVAL, "kelp_sort_particles: lon, lat and status are required");
    const uint64_t n_cells = (uint64_t)c->m.nx * (uint64_t)c->m.ny;
    if (3 * n_cells >= (1ull << 32)) return kelp_fail(KELP_EINVAL, "kelp_sort_particles: too many hash cells");
    unsigned end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) < 3 * n_cells) ++end_bit;
    HIP_TRY(hipSetDevice(c->device));
    int rc;
    if ((rc = kelp_reserve(c->sort_keys, n * sizeof(uint32_t))) ||
        (rc = kelp_reserve(c->sort_keys2, n * sizeof(uint32_t))) ||
        (rc = kelp_reserve(c->sort_idx, n * sizeof(int32_t))) ||
        (rc = kelp_reserve(c->sort_perm, n * sizeof(int32_t))) ||
        (rc = kelp_reserve(c->sort_stage, n * sizeof(double))))
        return rc;
    uint32_t* keys = static_cast<uint32_t*>(c->sort_keys.p);
    uint32_t* keys2 = static_cast<uint32_t*>(c->sort_keys2.p);
    int32_t* idx = static_cast<int32_t*>(c->sort_idx.p);
    int32_t* perm = static_cast<int32_t*>(c->sort_perm.p);
    sort_key_kernel<<<grid_of(n), 256, 0, c->stream>>>(c->m, n, p->lon, p->lat, p->status, keys, idx);
    HIP_TRY(hipGetLastError());
    size_t tmp_bytes = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, idx, perm, (unsigned int)n, 0u,
                                      end_bit, c->stream));
    if ((rc = kelp_reserve(c->sort_tmp, tmp_bytes))) return rc;
    HIP_TRY(rocprim::radix_sort_pairs(c->sort_tmp.p, tmp_bytes, keys, keys2, idx, perm, (unsigned int)n, 0u,
                                      end_bit, c->stream));
    if ((rc = permute(c, n, perm, p->lon)) || (rc = permute(c, n, perm, p->lat)) ||
        (rc = permute(c, n, perm, p->z)) || (rc = permute(c, n, perm, p->status)) ||
        (rc = permute(c, n, perm, p->nn)) || (rc = permute(c, n, perm, p->fr)) ||
        (rc = permute(c, n, perm, p->u)) || (rc = permute(c, n, perm, p->v)) ||
        (rc = permute(c, n, perm, p->w)) || (rc = permute(c, n, perm, p->pid)) ||
        (rc = permute(c, n, perm, p->wet_streak)) || (rc = permute(c, n, perm, p->elem)))
        return rc;
    return KELP_OK;
}
