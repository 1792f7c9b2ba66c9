// The solver object behind the C ABI's drw_solver handle (private to the
// library: capi.cpp and noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/dnnsolver_c.h"
#include "common.h"
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
};
