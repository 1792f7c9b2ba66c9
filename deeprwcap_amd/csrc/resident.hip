// Resident serving kernel (protocol in resident.h).  One workgroup serves
// one N = 23 network query at a time, end to end, with the same device code
// as the batched path (nn_sel_body.h, nn_face_body.h), so a query's result
// is bit-identical to the batched one for the same Philox counter:
//   description -> VoxDesc; selector backbone (FaceSelector or, with the
//   axis pre-rotation, FaceSelectorWeight); Linear head for this query;
//   softmax + face draw (word 0); the face's predictor (PoissonPredictor,
//   Gradient1Predictor for faces 0 and 1, Gradient2Predictor for 2..5);
//   output head and gradient post-process; tile draw (word 1); locate.
// Query flow: dnn_solver_poisson.cpp:86-111, dnn_solver_gradient.cpp:75-164.
#include "resident.h"

#include <cstddef>

#include "common.h"
#include "nn_common.h"
#include "nn_face_body.h"
#include "nn_sel_body.h"
#include "query.h"
#include "sampling.h"

namespace drw {

namespace {


// LDS of one resident workgroup.  The selector phase (cube, im2col offsets,
// then the head's activations) and the face phase (FaceSmem) share one
// region; the description, the selector features and the request stay
// outside it.
struct ResSmem {
    union Phase {
        FaceSmem<32> f32;  // Gradient2Predictor (FaceSolverPlus)
        FaceSmem<16> f16;  // PoissonPredictor, Gradient1Predictor
        struct {
            alignas(16) int koff[27 * 16];
            half_t cube[SelSmem<8, 16>::kRegion];
        } sel;
        struct {
            alignas(16) half_t x[512];
            alignas(16) half_t y1[256];
            alignas(16) half_t y2[128];
            alignas(16) float logits[16];
        } head;
    } u;
    VoxDesc desc;
    alignas(16) float feat[512];
    alignas(16) uint32_t raw[sizeof(ResReq) / 4];  // the request as read from the ring
    alignas(16) unsigned long long resp[kResWords];
    uint64_t seq, qid;
    int kind, axis, face, quit;
    float weight;
};

__device__ __forceinline__ unsigned long long ctl_load(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t bcast64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)hi << 32 | lo;
}

// Scout (wave 0, when one is due): reads the seq words of the ring slots from
// ctl.published on and advances it over the contiguous published prefix;
// copies the host's stop flag.  At most one scout per scout_gap ticks over
// all workgroups.
__device__ bool scout(const ResidentArgs& a, uint64_t now)
{
    const int lane = threadIdx.x & 63;
    uint32_t due = 0;
    if (lane == 0) {
        unsigned long long sc = ctl_load(&a.ctl->scout_next);
        if (now >= sc)
            due = __hip_atomic_compare_exchange_strong(&a.ctl->scout_next, &sc, now + a.scout_gap, __ATOMIC_RELAXED,
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!__builtin_amdgcn_readfirstlane(due)) return false;
    // lane 0 reads the next slot's seq word; only when it is published do all
    // lanes read the next 64 (one round trip over the bus while idle)
    uint64_t p = 0, w0 = 0;
    if (lane == 0) {
        p = ctl_load(&a.ctl->published);
        w0 = __hip_atomic_load(&a.req[p & (kResRing - 1)].seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    p = bcast64(p);
    int k = 0;
    if (bcast64(w0) == p + 1) {
        const uint64_t s = p + (uint64_t)lane;
        const uint64_t w = __hip_atomic_load(&a.req[s & (kResRing - 1)].seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        const uint64_t m = __ballot(w == s + 1);
        k = m == ~0ull ? 64 : __builtin_ctzll(~m);
    }
    if (lane == 0) {
        if (k) __hip_atomic_fetch_max(&a.ctl->published, (unsigned long long)(p + (uint64_t)k), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_load(a.host_stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
            __hip_atomic_store(&a.ctl->stop, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// Wave 0 only: the next sequence this workgroup serves, or ~0 when it
// leaves.  The workgroup takes a ticket h = head++ first (a fetch-and-add
// always succeeds, so workgroups never contend for a query) and then waits
// until h is published: the waiting workgroups form a queue of servers, the
// one holding h serves the next query as soon as it is published.  It leaves
// on the stop flag, or as a safety net after idle_ticks without its ticket
// being published; a ticket held by a leaving workgroup is served after the
// host restarts the kernel with head rewound to the oldest unanswered
// sequence (queries between that and the old head may be served twice; the
// second response repeats the first).
__device__ uint64_t claim_next(const ResidentArgs& a, uint64_t idle0, unsigned long long& scouts)
{
    const int lane = threadIdx.x & 63;
    uint64_t h = 0;
    if (lane == 0) h = __hip_atomic_fetch_add(&a.ctl->head, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    h = bcast64(h);
    for (int spin = 0;; spin++) {
        const uint64_t now = wall_clock64();
        scouts += scout(a, now);
        uint64_t pub = 0, stop = 0;
        if (lane == 0) {
            pub = ctl_load(&a.ctl->published);
            stop = ctl_load(&a.ctl->stop);
        }
        if (bcast64(pub) > h) return h;
        if (bcast64(stop)) {
            if (lane == 0) atomicAdd(&a.ctl->exits_stop, 1ull);
            return ~0ull;
        }
        if (now - idle0 > a.idle_ticks) {
            if (lane == 0) atomicAdd(&a.ctl->exits_idle, 1ull);
            return ~0ull;
        }
        // short waits first (the ticket is usually published within
        // microseconds under load), then a poll every few microseconds
        if (spin < 64) __builtin_amdgcn_s_sleep(8);
        else __builtin_amdgcn_s_sleep(127);
    }
}

static_assert(sizeof(VoxDesc) % 4 == 0, "the description is copied as 32-bit words");

// Face phase of a query whose face is drawn: predictor MODE on FaceSmem<BCS>,
// output head, tile draw (word 1) and the result (thread 0 writes r).

template <int NTH, int MODE, int BCS>
__device__ __forceinline__ void serve_face(ResSmem& sm, FaceSmem<BCS>& f, const FaceNetDev& w, uint64_t seed,
                                           float* r)
{
    const int face = sm.face, axis = sm.axis;
    for (int i = threadIdx.x; i < (int)(sizeof(VoxDesc) / 4); i += NTH)
        ((uint32_t*)&f.desc)[i] = ((const uint32_t*)&sm.desc)[i];
    if (threadIdx.x < 4) f.zero[threadIdx.x] = 0.f;
    for (int i = threadIdx.x; i < kGeluTabEntries; i += NTH) f.gtab[i] = w.gtab[i];
    __syncthreads();
    if constexpr (MODE == 2) face_solver_plus<NTH>(f, w, face, axis); else face_solver<NTH, MODE>(f, w, face, axis);
    __syncthreads();
    const float* kern = face_output_kernel<NTH, MODE>(f, face, axis);
    if (threadIdx.x < 64) {
        uint32_t u[4];
        query_uniforms(seed, sm.qid, 0u, u);
        const int t = wave_cdf_draw<float>(kern, kNN, u[1]);
        if (threadIdx.x == 0) face_result<MODE>(kern, f.desc, t, face, axis, sm.weight, r);
    }
}

// The five networks' device views, in device memory: read where each phase
// uses them (as kernel arguments, about 1.3 KB of pointers, they would be
// held in scalar registers and spilled).
struct ResNets {
    SelNetDev ps, gs;
    FaceNetDev pp, g1, g2;
};

// The resident workgroup's LDS, at module scope so that serve_query, a
// called function, addresses it as LDS.
__shared__ ResSmem g_res;

// One query, end to end (all threads).  Not inlined into the kernel's loop:
// the index arithmetic and weight loads of the layers stay inside the query
// instead of being hoisted out of the loop and held (spilled) across it.
template <int NTH>
__device__ __attribute__((noinline)) bool serve_query(const ResidentArgs& a, const ResNets* nets, uint64_t seq)
{
    ResSmem& sm = g_res;
    const ResReq* rq = a.req + (seq & (kResRing - 1));
    // the request, read with system-coherent 16-byte loads (a handful of bus
    // requests per query) after its seq word was seen by the scout that
    // published it
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    constexpr int kChunks = (int)(sizeof(ResReq) / 16);
    if (threadIdx.x < kChunks)
        ((u4*)sm.raw)[threadIdx.x] = *((const volatile u4*)rq + threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        sm.kind = (int)sm.raw[offsetof(ResReq, kind) / 4];
        sm.axis = (int)sm.raw[offsetof(ResReq, axis) / 4];
        sm.qid = (uint64_t)sm.raw[offsetof(ResReq, qid) / 4] | (uint64_t)sm.raw[offsetof(ResReq, qid) / 4 + 1] << 32;
    }
    __syncthreads();
    const bool grad = sm.kind == 1;  // DRW_GRADIENT
    const int axis = grad ? sm.axis : 2;
    const SelNetDev& ps = nets->ps;
    const SelNetDev& gs = nets->gs;
    const SelNetDev& sw = grad ? gs : ps;
    if (threadIdx.x < 64)
        vox_desc_build(sm.desc, (const float*)sm.raw + offsetof(ResReq, layers) / 4,
                       (const float*)sm.raw + offsetof(ResReq, cuboids) / 4, 0);
    else
        for (int i = threadIdx.x - 64; i < kSelGeluTab; i += NTH - 64) ((float2*)sm.u.sel.koff)[i] = sw.gtab[i];
    __syncthreads();
    if (grad) selector_backbone_body<true, NTH>(sm.desc, sm.u.sel.cube, sm.u.sel.koff, true, axis, gs, sm.feat);
    else selector_backbone_body<false, NTH>(sm.desc, sm.u.sel.cube, sm.u.sel.koff, false, 2, ps, sm.feat);
    // Linear head of this query (the backbone ended with a barrier)
    const int F = grad ? 512 : 128;
    for (int i = threadIdx.x; i < F; i += NTH) sm.u.head.x[i] = (half_t)sm.feat[i];
    __syncthreads();
    if (grad) {
        linear_row<512, 256, true, NTH>(sm.u.head.x, sm.u.head.y1, nullptr, gs.frag[3], gs.bias[3]);
        __syncthreads();
        linear_row<256, 128, true, NTH>(sm.u.head.y1, sm.u.head.y2, nullptr, gs.frag[4], gs.bias[4]);
        __syncthreads();
        linear_row<128, 7, false, NTH>(sm.u.head.y2, nullptr, sm.u.head.logits, gs.frag[5], gs.bias[5]);
    } else {
        linear_row<128, 64, true, NTH>(sm.u.head.x, sm.u.head.y1, nullptr, ps.frag[3], ps.bias[3]);
        __syncthreads();
        linear_row<64, 6, false, NTH>(sm.u.head.y1, nullptr, sm.u.head.logits, ps.frag[4], ps.bias[4]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float p[6];
        sm.face = selector_probs_draw(sm.u.head.logits, a.seed, sm.qid, p);
        sm.weight = grad ? sm.u.head.logits[6] : 0.f;
        sm.axis = axis;
    }
    __syncthreads();
    float r[5];
    const int face = sm.face;
    if (face < 0) {
        for (int j = 0; j < 5; j++) r[j] = __builtin_nanf("");
    } else if (!grad) {
        serve_face<NTH, 0>(sm, sm.u.f16, nets->pp, a.seed, r);
    } else if (face <= 1) {
        serve_face<NTH, 1>(sm, sm.u.f16, nets->g1, a.seed, r);
    } else {
        serve_face<NTH, 2>(sm, sm.u.f32, nets->g2, a.seed, r);
    }
    if (threadIdx.x == 0) {
        const int width = grad ? 5 : 3;
        bool nan = false;
        for (int j = 0; j < width; j++) nan |= isnan(r[j]);
        const uint64_t tag = (uint64_t)(uint32_t)(seq + 1) << 32;
        for (int j = 0; j < kResWords; j++) {
            const uint32_t v = j < 5 ? (j < width ? __builtin_bit_cast(uint32_t, r[j]) : 0u) : (nan ? 1u : 0u);
            sm.resp[j] = tag | v;
        }
    }
    __syncthreads();
    // six self-identifying words in three system-coherent 16-byte stores
    if (threadIdx.x < kResWords / 2) {
        typedef unsigned long long u2 __attribute__((ext_vector_type(2)));
        volatile u2* rp = (volatile u2*)(a.resp + (seq & (kResRing - 1)));
        rp[threadIdx.x] = ((const u2*)sm.resp)[threadIdx.x];
    }
    __syncthreads();  // every wave is done with this query's LDS
    return grad;
}

template <int NTH>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(NTH == 512 ? 4 : 2))) void resident_query_kernel(
    ResidentArgs a, const ResNets* nets)
{
    uint64_t idle0 = wall_clock64();
    unsigned long long served[2] = {0, 0}, scouts = 0;  // this workgroup's counts, added as it leaves
    for (;;) {
        if (threadIdx.x < 64) {
            const uint64_t s = claim_next(a, idle0, scouts);
            if (threadIdx.x == 0) {
                g_res.quit = s == ~0ull;
                g_res.seq = s;
            }
        }
        __syncthreads();
        if (g_res.quit) break;
        served[serve_query<NTH>(a, nets, g_res.seq) ? 1 : 0]++;
        idle0 = wall_clock64();
    }
    if (threadIdx.x == 0) {
        if (served[0]) atomicAdd(&a.ctl->served[0], served[0]);
        if (served[1]) atomicAdd(&a.ctl->served[1], served[1]);
        if (scouts) atomicAdd(&a.ctl->scouts, scouts);
    }
}

}  // namespace

int resident_grid(int device, int threads, int* grid)
{
    if (threads != 256 && threads != 512) return set_error(-2, "resident kernel: %d threads (256 or 512)", threads);
    int per_cu = 0, cus = 0;
    if (threads == 256)
        DRW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, resident_query_kernel<256>, 256, 0));
    else
        DRW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, resident_query_kernel<512>, 512, 0));
    DRW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (per_cu < 1) return set_error(-2, "resident kernel does not fit a CU");
    *grid = per_cu * cus;
    return 0;
}

int resident_nets_create(const NetWeights* poisson, const NetWeights* gradient, void** out)
{
    *out = nullptr;
    if (!poisson || !gradient) return set_error(-4, "resident kernel: models not loaded");
    ResNets h{};
    h.ps = net_selector(poisson);
    h.gs = net_selector(gradient);
    h.pp = net_face(poisson, 0);
    h.g1 = net_face(gradient, 1);
    h.g2 = net_face(gradient, 2);
    void* d = nullptr;
    DRW_HIP(hipMalloc(&d, sizeof h));
    const hipError_t e = hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return set_error(-3, "resident kernel: %s", hipGetErrorString(e));
    }
    *out = d;
    return 0;
}

int resident_launch(const ResidentArgs& a, const void* nets, int threads, int grid, hipStream_t stream)
{
    if (!nets) return set_error(-4, "resident kernel: models not loaded");
    if (grid <= 0) return set_error(-2, "resident kernel: empty grid");
    const ResNets* n = (const ResNets*)nets;
    if (threads == 512)
        hipLaunchKernelGGL(resident_query_kernel<512>, dim3(grid), dim3(512), 0, stream, a, n);
    else
        hipLaunchKernelGGL(resident_query_kernel<256>, dim3(grid), dim3(256), 0, stream, a, n);
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw
