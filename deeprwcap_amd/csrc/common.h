// Shared definitions for the MI355X (gfx950) transition-kernel library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drw {

constexpr int kMaxLayers = 15;    // inference_cpp/src/dnn_solver.h:29
constexpr int kMaxCuboids = 15;   // inference_cpp/src/dnn_solver.h:28
constexpr float kPad = -9999.0f;  // inference_cpp/src/cuda_kernels.cu:9
constexpr int kNetN = 23;         // networks exist at N = 23 only (dnn_solver.h:45)
constexpr int kMaxN = 23;
constexpr int kMaxCells = kMaxN * kMaxN * kMaxN;  // 12167

// FDM right-hand-side codes (ggft/src/ggft_solver.h:119)
enum SolveCode : int { kGFT = 0, kWVTZ = 1, kWVTX = 2, kWVTY = 3 };

// Records the first error of the calling thread; returns the code.
int set_error(int code, const char* fmt, ...);

#define DRW_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::drw::set_error(-3, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

// ---------------------------------------------------------------- reductions
// Wave64 butterfly sum; every lane receives the total.
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace drw
