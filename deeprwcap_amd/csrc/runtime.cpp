// Host runtime: thread-local error state and per-(device, stream) workspaces.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "common.h"
#include "runtime.h"

namespace drw {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

const char* last_error() { return g_last_error.c_str(); }
void clear_error() { g_last_error.clear(); }

// Launch-shape rules: kRules slots per kernel, each min_items << 32 |
// threads (0: empty), read lock-free at every launch and written under a
// mutex.  Face kernels: 256 or 512 threads per (query, face) workgroup.
// The built-in rules are the autotuner's choice on MI355X (python -m
// deeprwcap_amd.autotune): 512 threads where the kernel is latency-bound
// (small batches; the FaceSolverPlus kernel, whose 74 KB of LDS allow two
// workgroups per CU, at every size), 256 where the Poisson kernel fills the
// GPU; the fused Poisson query kernel below 4096 queries (a third less time
// than the separate kernels at small batches: 0.035 against 0.054 ms for 64
// queries), the separate kernels from 4096 on.  A kernel without rules takes
// 256 threads at every size.
static constexpr int kRules = 8;
static std::atomic<uint64_t> g_rules[kShapeCount][kRules];
static std::mutex g_rules_mu;
struct ShapeRule {
    int which, min_items, threads;
};
static constexpr ShapeRule kDefaultRules[] = {
    {kShapeFacePoisson, 0, 512},    {kShapeFacePoisson, 1024, 256}, {kShapeFacePoisson, 1536, 512},
    {kShapeFacePoisson, 2048, 256}, {kShapeFaceGrad1, 0, 512},      {kShapeFaceGrad2, 0, 512},
    {kShapePoissonFused, 0, 512},   {kShapePoissonFused, 1024, 256}, {kShapePoissonFused, 1536, 512},
    {kShapePoissonFused, 4096, 1}};

static uint64_t pack_rule(int min_items, int threads) { return (uint64_t)(uint32_t)min_items << 32 | (uint32_t)threads; }

static void default_rules(int which)
{
    int k = 0;
    for (auto& r : g_rules[which]) r.store(0);
    for (const auto& d : kDefaultRules)
        if (d.which == which && k < kRules) g_rules[which][k++].store(pack_rule(d.min_items, d.threads));
}

static bool rules_init = [] {
    for (int w = 0; w < kShapeCount; w++) default_rules(w);
    return true;
}();

int launch_shape(int which, int items)
{
    if (which < 0 || which >= kShapeCount) return 0;
    int best_min = -1, threads = 256;
    for (int i = 0; i < kRules; i++) {
        const uint64_t r = g_rules[which][i].load(std::memory_order_relaxed);
        if (!r) continue;
        const int mn = (int)(r >> 32), th = (int)(uint32_t)r;
        if (mn <= items && mn > best_min) best_min = mn, threads = th;
    }
    return threads;
}

int set_launch_shape(int which, int min_items, int threads)
{
    if (which < 0 || which >= kShapeCount) return set_error(-2, "launch shape %d: no such kernel", which);
    if (threads != 0 && threads != 256 && threads != 512 && !(which == kShapePoissonFused && threads == 1))
        return set_error(-2, "launch shape %d: %d threads per workgroup (256 or 512%s)", which, threads,
                         which == kShapePoissonFused ? ", or 1: separate kernels" : "");
    std::lock_guard<std::mutex> lk(g_rules_mu);
    if (min_items < 0) {  // one shape for every size, or (threads 0) the built-in rules
        if (!threads) {
            default_rules(which);
        } else {
            for (auto& r : g_rules[which]) r.store(0);
            g_rules[which][0].store(pack_rule(0, threads));
        }
        return 0;
    }
    int slot = -1, empty = -1;
    for (int i = 0; i < kRules; i++) {
        const uint64_t r = g_rules[which][i].load();
        if (r && (int)(r >> 32) == min_items) slot = i;
        if (!r && empty < 0) empty = i;
    }
    if (threads == 0) {
        if (slot >= 0) g_rules[which][slot].store(0);
        return 0;
    }
    if (slot < 0) slot = empty;
    if (slot < 0) return set_error(-2, "launch shape %d: more than %d rules", which, kRules);
    g_rules[which][slot].store(pack_rule(min_items, threads));
    return 0;
}

static std::mutex g_ws_mu;
static std::map<std::pair<int, hipStream_t>, Workspace*> g_ws;

int get_workspace(hipStream_t stream, Workspace** out)
{
    int dev = 0;
    DRW_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto key = std::make_pair(dev, stream);
    auto it = g_ws.find(key);
    if (it != g_ws.end()) { *out = it->second; return 0; }
    Workspace* w = new Workspace();
    w->device = dev;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    w->fdm_grid = cus;
    DRW_HIP(hipMalloc(&w->y_ws, sizeof(double) * (size_t)cus * kMaxCells));
    DRW_HIP(hipMalloc(&w->s_ws, sizeof(float) * (size_t)cus * kMaxCells));
    DRW_HIP(hipMalloc(&w->counters, sizeof(uint32_t) * 8));
    DRW_HIP(hipMemset(w->counters, 0, sizeof(uint32_t) * 8));
    DRW_HIP(hipDeviceSynchronize());
    g_ws[key] = w;
    *out = w;
    return 0;
}

static int take_event(std::vector<hipEvent_t>& pool, hipEvent_t* e)
{
    if (!pool.empty()) {
        *e = pool.back();
        pool.pop_back();
        return 0;
    }
    DRW_HIP(hipEventCreate(e));
    return 0;
}

int StageTimer::begin(int stage, hipStream_t st)
{
    if (!enabled) return 0;
    if (stage < 0 || stage >= kStages) return set_error(-2, "StageTimer: bad stage %d", stage);
    int rc;
    if ((rc = take_event(pool, &open.a)) || (rc = take_event(pool, &open.b))) return rc;
    open.stage = stage;
    is_open = true;
    DRW_HIP(hipEventRecord(open.a, st));
    return 0;
}

int StageTimer::end(hipStream_t st)
{
    if (!enabled || !is_open) return 0;
    DRW_HIP(hipEventRecord(open.b, st));
    pending.push_back(open);
    is_open = false;
    return 0;
}

int StageTimer::collect()
{
    for (auto& p : pending) {
        DRW_HIP(hipEventSynchronize(p.b));
        float ms = 0.f;
        DRW_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        total_ms[p.stage] += ms;
        count[p.stage] += 1;
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
    return 0;
}

void StageTimer::release()
{
    for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
    pending.clear();
    pool.clear();
}

}  // namespace drw
