// Device building blocks of the network queries: on-the-fly voxelization of
// the structure description, index remaps, MFMA fragments, GELU.
#pragma once

#include "common.h"

namespace drw {

typedef _Float16 half_t;
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kN = kNetN;            // 23
constexpr int kNN = kN * kN;         // 529
constexpr int kNNN = kNN * kN;       // 12167
constexpr int kPix = 544;            // 529 pixels padded to 34 MFMA row tiles

// ------------------------------------------------------------------------
// Structure description staged for voxelization.  fill_structure_kernel
// (inference_cpp/src/cuda_kernels.cu:21-84) picks the last z-layer with
// min_z <= z_c and lets the first cuboid whose closed box holds the voxel
// centre override it.  Both tests are separable per axis, so for every grid
// index we keep the 15-bit set of cuboids whose x (y, z) interval holds that
// coordinate; the voxel value is cv[ctz(bx & by & bz)] or the layer value.
// ------------------------------------------------------------------------
struct VoxDesc {
    uint16_t bx[kN], by[kN], bz[kN];
    float lz[kN];          // layer value at each z (normalised)
    float cv[kMaxCuboids]; // cuboid values (normalised)
    float maxv;            // per-query max eps (dnn_solver.cpp:274-282)
    uint32_t lzs[kN];               // lz and cv split into fp16 pairs hi | lo << 16,
    uint32_t cvs[kMaxCuboids + 1];  // hi + lo = the fp32 value to ~2^-22 relative; cvs[15] = 0 is
                                    // read by lookups with no cuboid hit (mask | 1 << 15 scans to 15)
};

// fp16 pair (hi | lo << 16) with hi = fp16(v), lo = fp16(v - hi): an MFMA
// operand in two fp16 parts carries v to about fp32 precision.
__device__ __forceinline__ uint32_t half_split(float v)
{
    const half_t hi = (half_t)v;
    const half_t lo = (half_t)(v - (float)hi);
    return (uint32_t)__builtin_bit_cast(uint16_t, hi) | ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

// Builds the description of query b with one wave (lanes 0..63 of the
// calling wave); other waves must not call it.  Normalises by the max.
__device__ inline void vox_desc_build(VoxDesc& d, const float* layers, const float* cuboids, int b)
{
    const int lane = threadIdx.x & 63;
    const float* Ly = layers + (size_t)b * kMaxLayers * 2;
    const float* Cb = cuboids + (size_t)b * kMaxCuboids * 7;
    // max over the 15 layer eps and 15 cuboid eps (padding values included)
    float m = lane < kMaxLayers ? Ly[2 * lane + 1] : (lane < kMaxLayers + kMaxCuboids ? Cb[7 * (lane - kMaxLayers) + 6]
                                                                                         : -3.0e38f);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    // number of live cuboids (list ends at the first padded minx)
    int ncub = kMaxCuboids;
    for (int c = 0; c < kMaxCuboids; c++)
        if (Cb[7 * c] == kPad) { ncub = c; break; }
    if (lane < kN) {
        const float cc = (2.0f * (float)lane + 1.0f) / (float)kN - 1.0f;
        uint32_t mx = 0, my = 0, mz = 0;
        for (int c = 0; c < ncub; c++) {
            const float* q = Cb + 7 * c;
            if (cc >= q[0] && cc <= q[3]) mx |= 1u << c;
            if (cc >= q[1] && cc <= q[4]) my |= 1u << c;
            if (cc >= q[2] && cc <= q[5]) mz |= 1u << c;
        }
        d.bx[lane] = (uint16_t)mx;
        d.by[lane] = (uint16_t)my;
        d.bz[lane] = (uint16_t)mz;
        int li = 0;
        for (int l = 0; l < kMaxLayers; l++) {
            const float mzv = Ly[2 * l];
            if (mzv == kPad) break;
            if (cc >= mzv) li = l; else break;
        }
        const float lv = Ly[2 * li + 1] / m;
        d.lz[lane] = lv;
        d.lzs[lane] = half_split(lv);
    }
    if (lane < kMaxCuboids) {
        const float cv = Cb[7 * lane + 6] / m;
        d.cv[lane] = cv;
        d.cvs[lane] = half_split(cv);
    }
    if (lane == kMaxCuboids) d.cvs[lane] = 0u;
    if (lane == 0) d.maxv = m;
}

// Normalised permittivity of voxel (z, y, x) of the query cube.
__device__ __forceinline__ float vox_value(const VoxDesc& d, int z, int y, int x)
{
    const uint32_t hit = (uint32_t)d.bx[x] & (uint32_t)d.by[y] & (uint32_t)d.bz[z];
    return hit ? d.cv[__builtin_ctz(hit)] : d.lz[z];
}

// ------------------------------------------------------------------------
// Index remaps, all as out[z][y][x] = in[src] (cuda_kernels.cu:299-486 and
// the rot90 of dnn_solver_gradient.cpp:228-238).
// ------------------------------------------------------------------------
// Written as selects (no switch) so that the coordinates stay in registers
// whatever the face / axis.
__device__ __forceinline__ void remap_poisson(int f, int& z, int& y, int& x)
{
    const int Z = z, Y = y, X = x;
    // f: 0 identity; 1 (22-Z, Y, X); 2 (Y, Z, X); 3 (Y, 22-Z, X); 4 (Y, X, Z); 5 (Y, X, 22-Z)
    z = f == 0 ? Z : f == 1 ? kN - 1 - Z : Y;
    y = f <= 1 ? Y : f == 2 ? Z : f == 3 ? kN - 1 - Z : X;
    x = f <= 3 ? X : f == 4 ? Z : kN - 1 - Z;
}

__device__ __forceinline__ void remap_gradient(int f, int& z, int& y, int& x)
{
    const int Z = z, Y = y, X = x;
    // f: 0 (22-Z, Y, X); 1, 2 identity; 3 (Z, 22-Y, X); 4 (Z, X, Y); 5 (Z, X, 22-Y)
    z = f == 0 ? kN - 1 - Z : Z;
    y = f == 3 ? kN - 1 - Y : f >= 4 ? X : Y;
    x = f == 4 ? Y : f == 5 ? kN - 1 - Y : X;
}

__device__ __forceinline__ void remap_axis(int axis, int& z, int& y, int& x)
{
    const int Z = z, Y = y, X = x;
    // axis 0: (22-X, Y, Z); axis 1: (22-Y, Z, X); axis 2: identity
    z = axis == 0 ? kN - 1 - X : axis == 1 ? kN - 1 - Y : Z;
    y = axis == 1 ? Z : Y;
    x = axis == 0 ? Z : X;
}

// K-slot order of the face predictors' input 1x1 conv (25 channels: cube
// slices 0..22, then the w and h position planes 23, 24).  In an MFMA
// operand, lane group g = lane/16 holds slots q = 4 kt + j (kt < 2, j < 4).
// Group g takes slices 6g..6g+5, so every lane builds six slices; group 3's
// sixth slot is plane 23, group 0's seventh is plane 24, the rest are zero
// (-1).  The host packs the weights in the same order.
__host__ __device__ constexpr int input_slot_channel(int g, int q)
{
    return q < 6 ? 6 * g + q : (q == 6 && g == 0 ? kN + 1 : -1);
}

// ------------------------------------------------------------------------
// GELU(x) = x Phi(x) (torch.nn.GELU, exact erf form).  Phi(-|x|) =
// erfc(|x|/sqrt 2)/2 with erfc from Abramowitz & Stegun, Handbook of
// Mathematical Functions, eq. 7.1.26 (Hastings' rational approximation,
// |error| <= 1.5e-7): erfc(u) ~ t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) e^-u^2,
// t = 1 / (1 + p u).  The constants are folded so that the hardware exp2
// takes the argument directly: z = |x| k with k^2 = log2(e) / 2 gives
// e^(-x^2/2) = 2^(-z^2) and t = 1 / (1 + p' z), p' = p / (k sqrt 2); the
// factor 1/2 is folded into the coefficients.  GELU's absolute error stays
// below 5e-7 on all of R (checked against the exact form); 15 VALU
// instructions, two of them transcendental.
// ------------------------------------------------------------------------
__device__ __forceinline__ float phi_neg_abs(float x)  // Phi(-|x|)
{
    const float z = fabsf(x) * 0.8493218f;
    const float t = __builtin_amdgcn_rcpf(fmaf(0.27273747f, z, 1.0f));
    float p = fmaf(0.5307027f, t, -0.72657603f);
    p = fmaf(p, t, 0.7107069f);
    p = fmaf(p, t, -0.14224836f);
    p = fmaf(p, t, 0.1274148f);
    return p * t * __builtin_amdgcn_exp2f(-(z * z));
}

__device__ __forceinline__ float gelu(float x)
{
    const float xq = x * phi_neg_abs(x);
    return x >= 0.f ? x - xq : xq;
}

// GELU by table, two forms.
//  * Face predictors, every layer (gelu_tab): Phi is linear in x on each of
//    kGeluTab intervals of [kGeluLo, -kGeluLo] = [-5, 5]; entry i + 1 holds
//    the line (a, b) of interval i, Phi(x) ~ a + b x, placed midway between
//    the largest deviations of Phi above and below its chord on that interval
//    (the host builds it in double precision, nn_host.cpp, phi_line_table),
//    and GELU(x) = x (a + b x).  Entry 0 is (0, 0) for x < -5 and entry
//    kGeluTab + 1 is (1, 0) for x >= 5, where |GELU(x) - max(x, 0)| =
//    |x| Phi(-|x|) <= 1.5e-6.  The error is |x| times that of the line,
//    <= 7.1e-6 on all of R (largest near |x| = 1.4, about 2e-6 on average
//    over N(0, 1) inputs), in 6 VALU instructions and one LDS read.
//  * Selectors, first conv (gelu_tab_sel): Phi(x) is linear between NTAB + 1
//    nodes on [lo, -lo], read from an NTAB-entry (Phi(x_i), Phi(x_i+1) -
//    Phi(x_i)) table; outside the range the end entries give x Phi(-lo) and
//    x Phi(lo).  216 entries on [-4.5, 4.5] (h = 9/216, error <= h^2/8 max
//    x^2 phi(x) = 0.294 h^2/8 = 6.4e-5, tails x Phi(-4.5) = 3.4e-6 x), the
//    size of the LDS the im2col offsets of the later layers use, which is
//    free during the first; those outputs are stored as fp16 (relative
//    rounding 4.9e-4).
constexpr int kGeluTab = 512;
constexpr int kGeluTabEntries = kGeluTab + 2;
constexpr float kGeluLo = -5.0f, kGeluInvH = kGeluTab / 10.0f;
constexpr int kSelGeluTab = 216;
constexpr float kSelGeluLo = -4.5f, kSelGeluInvH = kSelGeluTab / 9.0f;

template <int NTAB>
__device__ __forceinline__ float gelu_lut(float x, const float2* tab, float inv_h, float lo)
{
    float u = fmaf(x, inv_h, -lo * inv_h);
    u = __builtin_amdgcn_fmed3f(u, 0.0f, (float)NTAB - 6.1e-5f);
    const int i = (int)u;
    const float2 e = tab[i];
    return x * fmaf(e.y, __builtin_amdgcn_fractf(u), e.x);
}

__device__ __forceinline__ float gelu_tab(float x, const float2* tab)
{
    // u in [1, kGeluTab + 1) on the table's range; below it entry 0, above it
    // entry kGeluTab + 1
    float u = fmaf(x, kGeluInvH, 1.0f - kGeluLo * kGeluInvH);
    u = __builtin_amdgcn_fmed3f(u, 0.0f, (float)kGeluTab + 1.5f);
    const float2 e = tab[(int)u];
    return x * fmaf(e.y, x, e.x);
}

__device__ __forceinline__ float gelu_tab_sel(float x, const float2* tab)
{
    return gelu_lut<kSelGeluTab>(x, tab, kSelGeluInvH, kSelGeluLo);
}

// MFMA 16x16x16 fp16 -> fp32.  A: lane l holds row l%16, k = 4(l/16)..+3;
// B: lane l holds column l%16, k = 4(l/16)..+3; C: lane l holds rows
// 4(l/16)..+3 of column l%16.
__device__ __forceinline__ float4_t mfma16(half4_t a, half4_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// MFMA 16x16x32 fp16 -> fp32, gfx950's double-rate shape: one instruction for
// two K = 16 steps.  A and B spread K over the four lane groups in the same
// way (8 values per lane), so a product only needs both operands to put the
// same (k-slot) -> channel assignment in lane group l/16, element j; D is laid
// out as for mfma16.
__device__ __forceinline__ float4_t mfma32(half8_t a, half8_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Two half4 K-steps as one half8 operand (slots 0..3, then 4..7).
__device__ __forceinline__ half8_t cat8(half4_t lo, half4_t hi)
{
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

}  // namespace drw
