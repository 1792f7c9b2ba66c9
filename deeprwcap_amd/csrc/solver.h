// The solver object behind the C ABI's drw_solver handle (private to the
// library: capi.cpp and noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/dnnsolver_c.h"
#include "common.h"
#include "fdm.h"
#include "query.h"
#include "runtime.h"

namespace drw {

// Device buffer grown on demand (hipMalloc), kept until released.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int ensure(size_t want)
    {
        if (want <= n) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        const size_t cap = want + want / 4 + 16;
        DRW_HIP(hipMalloc(&p, sizeof(T) * cap));
        n = cap;
        return 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Looks up this device's allocation-free scratch array (noalloc.hip); call
// with the device current.  Idempotent per device.
int noalloc_scratch_init(int device);

}  // namespace drw

struct drw_solver {
    int kind = DRW_POISSON;
    int device = 0;
    hipStream_t stream = nullptr;
    int backend = -1;  // none until load_models / use_fdm
    int fdm_N = 23;
    double fdm_tol = 1e-6;
    uint64_t seed = 0x5eed5eedULL;
    uint64_t counter = 0;
    drw::NetWeights* nets = nullptr;
    drw::DevBuf<float> layers, cuboids, out, eps, cen;
    drw::DevBuf<int32_t> axis, Nq, code;
    drw::DevBuf<int64_t> eps_off, G_off;
    drw::DevBuf<double> G;
    drw::DevBuf<uint8_t> scratch;
    drw::StageTimer timer;
    // FDM path: first-solve limits, per-cube status, acceptance list, bad flags
    int fdm_max_rounds = 12, fdm_max_inner = 2000;
    drw::DevBuf<int32_t> status, list, count;
    drw::DevBuf<uint8_t> bad;
    // network path: FDM fallback slots (allocated with the models)
    drw::DevBuf<float> fb_eps;
    drw::DevBuf<double> fb_G, fb_y;
    drw::DevBuf<int32_t> fb_list, fb_count;
    drw::DevBuf<uint32_t> fb_ctr;
    // [0] FDM re-solves, [1] failed queries, [2] network fallbacks
    uint64_t* stats = nullptr;
};

namespace drw {
constexpr int kFallbackSlots = 32;     // workgroups (and buffer slots) of the network-path fallback
constexpr int kFallbackCap = 16384;    // failed network queries per batch that the fallback takes

// Lets the face kernels of q append their failed queries to the fallback's
// list; nn_fallback_enqueue must follow query_nn_launch(q).
inline void nn_fallback_attach(drw_solver* s, QueryArgs& q)
{
    q.fb_list = s->fb_list.p;
    q.fb_count = s->fb_count.p;
    q.fb_cap = kFallbackCap;
    q.fb_stats = s->stats;
}

// Enqueues the network-path FDM fallback for the batch q just enqueued
// (queries whose result holds a NaN are solved by the FDM and drawn with their
// own counters); allocates nothing.
inline int nn_fallback_enqueue(drw_solver* s, const QueryArgs& q, hipStream_t st, StageTimer* tm)
{
    if (!s->fb_ctr.p) return set_error(DRW_ESTATE, "network fallback buffers missing: load the models first");
    NnFallbackArgs f{};
    f.B = q.B;
    f.width = s->kind == DRW_GRADIENT ? 5 : 3;
    f.cap = kFallbackCap;
    f.out = q.out;
    f.layers = q.layers;
    f.cuboids = q.cuboids;
    f.axis = q.axis;
    f.list = s->fb_list.p;
    f.count = s->fb_count.p;
    f.stats = s->stats;
    f.seed = q.seed;
    f.counter = q.counter;
    f.qid = q.qid;
    f.ctr_dev = q.ctr_dev;
    f.tol = s->fdm_tol;
    f.grid = kFallbackSlots;
    f.eps = s->fb_eps.p;
    f.G = s->fb_G.p;
    f.y_ws = s->fb_y.p;
    f.ctr = s->fb_ctr.p;
    int rc;
    if (tm && (rc = tm->begin(6, st))) return rc;
    if ((rc = nn_fallback_launch(f, st))) return rc;
    return tm ? tm->end(st) : 0;
}
}  // namespace drw
