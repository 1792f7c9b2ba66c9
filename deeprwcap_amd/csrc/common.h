// Shared definitions for the MI355X (gfx950) transition-kernel library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "query_limits.h"

namespace drw {

constexpr int kNetN = 23;         // networks exist at N = 23 only (dnn_solver.h:45)
constexpr int kMaxN = 23;
constexpr int kMaxCells = kMaxN * kMaxN * kMaxN;  // 12167

// FDM right-hand-side codes (ggft/src/ggft_solver.h:119)
enum SolveCode : int { kGFT = 0, kWVTZ = 1, kWVTX = 2, kWVTY = 3 };

// Records the first error of the calling thread; returns the code.
int set_error(int code, const char* fmt, ...);

// Launch shapes chosen at run time, process-wide (drw_set_launch_shape):
// threads per workgroup of the tunable kernels as rules (min_items,
// threads); a launch of `items` workgroups takes the rule with the largest
// min_items <= items, else 256 threads.
// kShapePoissonFused: the fused Poisson query kernel's threads (256, 512),
// or 1 for the separate selector, head and predictor kernels.
enum LaunchShape : int {
    kShapeFacePoisson = 0,
    kShapeFaceGrad1 = 1,
    kShapeFaceGrad2 = 2,
    kShapePoissonFused = 3,
    kShapeCount = 4
};
int launch_shape(int which, int items);
int set_launch_shape(int which, int min_items, int threads);

#define DRW_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::drw::set_error(-3, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

// ---------------------------------------------------------------- reductions
// Wave64 reductions on DPP (data-parallel primitives: a VALU operand read
// from another lane of the same 16-lane row) instead of LDS permutes: the
// xor-1 and xor-2 quad swaps and the row rotations by 4 and 8 leave every lane
// of a row with the row's total, and four readlanes add the rows.  Every lane
// of the wave must be active.  DPP controls (GFX9 encoding): quad_perm
// [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E, row_shr:n = 0x110 + n, row_ror:n =
// 0x120 + n, row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    // lanes whose source is outside the row (row_shr) or whose row is masked
    // off read 0
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, true);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __builtin_bit_cast(float, dpp_u32<CTRL>(__builtin_bit_cast(uint32_t, v)));
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t r = (uint64_t)dpp_u32<CTRL>((uint32_t)u) | ((uint64_t)dpp_u32<CTRL>((uint32_t)(u >> 32)) << 32);
    return __builtin_bit_cast(double, r);
}

__device__ __forceinline__ float readlane_f32(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t r = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l) |
                       ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l) << 32);
    return __builtin_bit_cast(double, r);
}

// A value every active lane holds alike, moved to scalar registers (so the
// compiler neither keeps a VGPR copy per lane nor spills one).
__device__ __forceinline__ float uniform_f32(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p)
{
    const uint64_t u = (uint64_t)p;
    return (T*)((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u) |
                ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32)) << 32));
}

__device__ __forceinline__ double uniform_f64(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t r = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u) |
                       ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32)) << 32);
    return __builtin_bit_cast(double, r);
}

// Sum over the wave; every lane receives the total (the same bits in all).
__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_f32<0xB1>(v);
    v += dpp_f32<0x4E>(v);
    v += dpp_f32<0x124>(v);
    v += dpp_f32<0x128>(v);
    return (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
}

__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x124>(v);
    v += dpp_f64<0x128>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Maximum over the wave (float or double, no NaN handling beyond fmax's).
__device__ __forceinline__ float wave_max(float v)
{
    v = fmaxf(v, dpp_f32<0xB1>(v));
    v = fmaxf(v, dpp_f32<0x4E>(v));
    v = fmaxf(v, dpp_f32<0x124>(v));
    v = fmaxf(v, dpp_f32<0x128>(v));
    return fmaxf(fmaxf(readlane_f32(v, 0), readlane_f32(v, 16)), fmaxf(readlane_f32(v, 32), readlane_f32(v, 48)));
}

__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_f64<0xB1>(v));
    v = fmax(v, dpp_f64<0x4E>(v));
    v = fmax(v, dpp_f64<0x124>(v));
    v = fmax(v, dpp_f64<0x128>(v));
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

__device__ __forceinline__ double wave_min(double v)
{
    v = fmin(v, dpp_f64<0xB1>(v));
    v = fmin(v, dpp_f64<0x4E>(v));
    v = fmin(v, dpp_f64<0x124>(v));
    v = fmin(v, dpp_f64<0x128>(v));
    return fmin(fmin(readlane_f64(v, 0), readlane_f64(v, 16)), fmin(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Inclusive prefix sum over the wave's lanes (uint32, wrapping): shifts by 1,
// 2, 4, 8 within each row, then row 0's last lane is added to row 1 and row
// 2's to row 3 (row_bcast:15), then lane 31 to rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    v += dpp_u32<0x111>(v);
    v += dpp_u32<0x112>(v);
    v += dpp_u32<0x114>(v);
    v += dpp_u32<0x118>(v);
    v += dpp_u32<0x142, 0xa>(v);
    v += dpp_u32<0x143, 0xc>(v);
    return v;
}

}  // namespace drw
