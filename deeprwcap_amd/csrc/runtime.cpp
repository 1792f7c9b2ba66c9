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

}  // namespace drw
