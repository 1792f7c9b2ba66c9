// Allocation-free network query: drw_solver_sample_device_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the network query
// of drw_solver_sample_device with its per-query intermediates (selector
// features, face probabilities, predicted W, drawn face) in a fixed device
// array the library holds per GPU until the process ends, instead of the
// solver's grown scratch buffer.  The kernels are the same; only where the
// intermediates live differs.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/dnnsolver_c.h"
#include "common.h"
#include "query.h"
#include "runtime.h"
#include "solver.h"

namespace {

// Largest batch one call takes, and the intermediates of one query at most
// (gradient: 512 selector features f32, 6 probabilities f32, W f32, face i32;
// the layout query_nn_launch uses).
constexpr int kNoAllocMaxB = 8192;
constexpr size_t kNoAllocPerQuery = 512 * 4 + 6 * 4 + 4 + 4;
constexpr size_t kNoAllocBytes = (size_t)kNoAllocMaxB * kNoAllocPerQuery;  // 17,039,360 B (16.25 MiB)

// Each GPU that loads the library holds its own copy of this array.  In host
// code its name is not a device address: host code uses the address that
// hipGetSymbolAddress returned for the current device (g_scratch_ptr).
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

uint8_t* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

namespace drw {

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(DRW_EINVAL, "noalloc scratch: device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    int cur = -1;
    DRW_HIP(hipGetDevice(&cur));
    if (cur != device) return set_error(DRW_EINVAL, "noalloc scratch: device %d is not current (%d)", device, cur);
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(DRW_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                         hipGetErrorString(e));
    g_scratch_ptr[device] = (uint8_t*)p;
    return 0;
}

}  // namespace drw

extern "C" int drw_solver_sample_device_noalloc(drw_solver* s, int B, const float* layers, const float* cuboids,
                                                const int32_t* axis, float* out, void* stream)
{
    using namespace drw;
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "sample_noalloc: null solver");
    if (B < 0) return set_error(DRW_EINVAL, "sample_noalloc: B < 0");
    if (B > kNoAllocMaxB)
        return set_error(DRW_EINVAL, "sample_noalloc: B = %d exceeds the allocation-free maximum of %d queries", B,
                         kNoAllocMaxB);
    if (B == 0) return 0;
    if (!layers || !cuboids || !out) return set_error(DRW_EINVAL, "sample_noalloc: null pointer");
    if (s->kind == DRW_GRADIENT && !axis) return set_error(DRW_EINVAL, "Gradient solver requires axis parameter");
    if (s->backend != DRW_BACKEND_NN || !s->nets)
        return set_error(DRW_ESTATE, "sample_noalloc: network backend not loaded (the FDM backend allocates)");
    DRW_HIP(hipSetDevice(s->device));
    uint8_t* buf = scratch_ptr(s->device);
    if (!buf) {
        // solvers look the array up when they are created; this covers none
        const int rc = noalloc_scratch_init(s->device);
        if (rc) return rc;
        buf = scratch_ptr(s->device);
    }
    // B <= kNoAllocMaxB, so the intermediates query_nn_launch lays out fit in
    // kNoAllocBytes and it never grows (frees, allocates) the buffer
    size_t cap = kNoAllocBytes;
    QueryArgs q{};
    q.B = B;
    q.layers = layers;
    q.cuboids = cuboids;
    q.axis = s->kind == DRW_GRADIENT ? axis : nullptr;
    q.out = out;
    q.seed = s->seed;
    q.counter = s->counter;
    q.timer = nullptr;  // stage events are created on demand: not here
    if (!s->fb_ctr.p) return set_error(DRW_ESTATE, "network fallback buffers missing: load the models first");
    nn_fallback_attach(s, q);
    uint8_t* p = buf;
    const hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    int rc = query_nn_launch(s->nets, q, &p, &cap, st);
    if (!rc) rc = nn_fallback_enqueue(s, q, st, nullptr);  // its buffers come with the models
    if (rc) return rc;
    s->counter += (uint64_t)B;
    return 0;
}
