// Host runtime: thread-local error state and per-(device, stream) workspaces.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
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
