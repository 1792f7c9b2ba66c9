// Device helpers shared by the sampler and the fused query kernels:
// Philox4x32-10, the wave64 fixed-point inverse-CDF draw, index remaps and
// the panel-centre table.
#pragma once

#include "common.h"

namespace drw {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), published multipliers
// and Weyl increments.  counter = (q_lo, q_hi, stream, 0), key = seed.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__host__ __device__ inline void query_uniforms(uint64_t seed, uint64_t q, uint32_t stream, uint32_t out[4])
{
    out[0] = (uint32_t)q;
    out[1] = (uint32_t)(q >> 32);
    out[2] = stream;
    out[3] = 0u;
    philox4x32_10(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// ---------------------------------------------------------------- wave ops
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    const uint32_t lo = __shfl((int)(uint32_t)v, src, 64), hi = __shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int d)
{
    const uint32_t lo = __shfl_up((int)(uint32_t)v, d, 64), hi = __shfl_up((int)(uint32_t)(v >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t quant(float a, int s) { return (uint32_t)ldexpf(fabsf(a), s); }
__device__ __forceinline__ uint32_t quant(double a, int s) { return (uint32_t)ldexp(fabs(a), s); }
__device__ __forceinline__ int frexp_exp(float m) { int e; frexpf(m, &e); return e; }
__device__ __forceinline__ int frexp_exp(double m) { int e; frexp(m, &e); return e; }

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((int)(uint32_t)v, o, 64), hi = __shfl_xor((int)(uint32_t)(v >> 32), o, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// Fixed-point inverse-CDF draw over |w[0..M)| by one wave64 (all lanes call,
// all lanes get the result).  q_i = trunc(|w_i| 2^s) with the power of two s
// that puts max|w| in [2^23, 2^24); the prefix sums are exact integers, so the
// drawn bin (smallest i with cumsum_i > floor(u total / 2^32)) is independent
// of the summation order and bit-identical to the CPU sampler.  Returns -1 if
// every weight is zero or the largest is not finite.  w may live in any
// address space (global or LDS).
//
// Every pass reads w in rows of 64 consecutive bins, one bin per lane
// (coalesced): max, total, then an inclusive scan row by row (row sums
// < 64 * 2^24 fit in 32 bits) that stops at the row holding the drawn bin.
template <typename T, typename Ptr>
__device__ int wave_cdf_draw(Ptr w, int M, uint32_t u)
{
    const int lane = threadIdx.x & 63;
    T m = 0;
    for (int i = lane; i < M; i += 64) {
        const T a = w[i] < 0 ? -w[i] : w[i];
        m = a > m ? a : m;
    }
    m = wave_max(m);
    if (!(m > 0) || isinf(m)) return -1;
    const int s = 24 - frexp_exp(m);
    uint64_t total = 0;
    for (int i = lane; i < M; i += 64) total += quant(w[i], s);
    total = wave_sum_u64(total);
    const uint64_t lo = (uint64_t)u * (total & 0xFFFFFFFFull);
    const uint64_t hi = (uint64_t)u * (total >> 32);
    const uint64_t r = hi + (lo >> 32);  // floor(u * total / 2^32), total < 2^48
    uint64_t base = 0;  // cumulative sum of the rows before j0
    for (int j0 = 0; j0 < M; j0 += 64) {
        const int i = j0 + lane;
        const uint32_t incl = wave_inclusive_scan(i < M ? quant(w[i], s) : 0u);
        const uint64_t hit = __ballot(base + incl > r);
        if (hit) return j0 + __ffsll((unsigned long long)hit) - 1;
        base += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
    return M - 1;
}

// wave_cdf_draw over a global array of at most 64 * MPL doubles, read from
// HBM once: each lane keeps its bins i = lane + 64 k in registers for the
// max, total and scan passes (the passes and their summation orders are
// wave_cdf_draw's, so the drawn bin is the same).  `absum` receives the
// lane-strided sum of |w_i| reduced over the wave (gradient queries' W), in
// the order of a loop over i = lane, lane + 64, ...
template <int MPL>
__device__ int wave_cdf_draw_regs(const double* __restrict__ w, int M, uint32_t u, double& absum)
{
    const int lane = threadIdx.x & 63;
    double g[MPL];
    double m = 0.0, sa = 0.0;
#pragma unroll
    for (int k = 0; k < MPL; k++) {
        const int i = lane + 64 * k;
        g[k] = i < M ? w[i] : 0.0;
        const double a = g[k] < 0 ? -g[k] : g[k];
        m = a > m ? a : m;
        if (i < M) sa += fabs(g[k]);
    }
    absum = wave_sum(sa);
    m = wave_max(m);
    if (!(m > 0) || isinf(m)) return -1;
    const int s = 24 - frexp_exp(m);
    uint32_t qv[MPL];
    uint64_t total = 0;
#pragma unroll
    for (int k = 0; k < MPL; k++) {
        qv[k] = lane + 64 * k < M ? quant(g[k], s) : 0u;
        total += qv[k];
    }
    total = wave_sum_u64(total);
    const uint64_t lo = (uint64_t)u * (total & 0xFFFFFFFFull);
    const uint64_t hi = (uint64_t)u * (total >> 32);
    const uint64_t r = hi + (lo >> 32);  // floor(u * total / 2^32), total < 2^48
    uint64_t base = 0;
#pragma unroll
    for (int k = 0; k < MPL; k++) {
        if (64 * k >= M) break;  // uniform
        const uint32_t incl = wave_inclusive_scan(qv[k]);
        const uint64_t hit = __ballot(base + incl > r);
        if (hit) return 64 * k + __ffsll((unsigned long long)hit) - 1;
        base += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
    return M - 1;
}

// Panel centre of tile t on face f (locate_index_simple_kernel,
// inference_cpp/src/cuda_kernels.cu:219-264), for any N.
__device__ __forceinline__ void panel_point(int N, int t, int f, float* o)
{
    const int xid = t % N, yid = t / N;
    const float t0 = (2.0f * xid + 1.0f) / N - 1.0f, t1 = (2.0f * yid + 1.0f) / N - 1.0f;
    switch (f) {
    case 0: o[0] = t0; o[1] = t1; o[2] = -1.0f; break;
    case 1: o[0] = t0; o[1] = t1; o[2] = 1.0f; break;
    case 2: o[0] = t0; o[1] = -1.0f; o[2] = t1; break;
    case 3: o[0] = t0; o[1] = 1.0f; o[2] = t1; break;
    case 4: o[0] = -1.0f; o[1] = t0; o[2] = t1; break;
    default: o[0] = 1.0f; o[1] = t0; o[2] = t1; break;
    }
}

// Face index in the original frame of face f of the axis-pre-rotated cube
// (locate_index_kernel, cuda_kernels.cu:161-173).
__host__ __device__ __forceinline__ int axis_face_remap(int f, int axis)
{
    if (axis == 0) return f == 0 ? 4 : f == 1 ? 5 : f == 4 ? 1 : f == 5 ? 0 : f;
    if (axis == 1) return f == 0 ? 2 : f == 1 ? 3 : f == 2 ? 1 : f == 3 ? 0 : f;
    return f;
}

// Flat inverse-CDF next hop over one query's 6 N^2 FDM kernel by one wave
// (the sampling semantics of ggft_solver::postprocess, ggft_solver.cpp:
// 638-685: normalise |G| and invert its CDF in GGFT panel order), with the
// query's first Philox word.  All 64 lanes call.
//   Poisson  (!grad): o = (x, y, z)
//   gradient (grad):  o = (x, y, z, W sign(G_i), eps_centre), W = sum |G|
// A failed query (bad, or no drawable bin) gets NaN in every output.
// REGS: G is read once into registers (wave_cdf_draw_regs, for kernels with
// registers to spare); otherwise every pass reads it from memory.
template <bool REGS = false>
__device__ inline void flat_draw_wave(const double* G, int N, bool grad, float center_eps, uint64_t seed,
                                      uint64_t counter, bool bad, float* o)
{
    const int lane = threadIdx.x & 63;
    const int NN = N * N, M = 6 * NN;
    const int width = grad ? 5 : 3;
    uint32_t u[4];
    query_uniforms(seed, counter, 0u, u);
    double wsum_regs = 0.0;
    int i;
    if constexpr (REGS) i = bad ? -1 : wave_cdf_draw_regs<(6 * kMaxN * kMaxN + 63) / 64>(G, M, u[0], wsum_regs);
    else i = bad ? -1 : wave_cdf_draw<double>(G, M, u[0]);
    if (i < 0) {
        if (lane < width) o[lane] = __builtin_nanf("");
        return;
    }
    const int f = i / NN, t = i - f * NN;
    if (!grad) {
        if (lane == 0) {
            float p[3];
            panel_point(N, t, f, p);
            o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
        }
        return;
    }
    double wsum = wsum_regs;
    if constexpr (!REGS) {
        for (int k = lane; k < M; k += 64) wsum += fabs(G[k]);
        wsum = wave_sum(wsum);
    }
    if (lane == 0) {
        float p[3];
        panel_point(N, t, f, p);
        const double g = G[i];
        const float sg = g > 0 ? 1.0f : (g < 0 ? -1.0f : 0.0f);
        o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
        o[3] = (float)wsum * sg;
        o[4] = center_eps;
    }
}

}  // namespace drw
