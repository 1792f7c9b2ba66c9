// Device bodies of the face predictors (FaceSolver / FaceSolverPlus,
// training_pytorch/src/architecture.py:11-173): layers, input build and the
// two architectures, shared by the batched face kernel (nn_face.hip) and the
// resident serving kernel (resident.hip).
//
// Activations live in LDS as fp16 [pixel][channel] (pixel = 23 h + w, 529
// pixels padded to 544 = 34 MFMA row tiles) up to the last MFMA layer and as
// fp32 in the tail layers (<= 8 channels).  1x1 convolutions with >= 8
// outputs run on MFMA (v_mfma_f32_16x16x32_f16 where a lane holds 8 K slots,
// else 16x16x16; fp32 accumulate); depthwise 3x3, the small 1x1 layers,
// batch-norm (folded into weights/bias on the host) and GELU (LDS table) run
// on the VALU in fp32.
#pragma once

#include <type_traits>

#include "common.h"
#include "nn_common.h"
#include "nn_face.h"
#include "sampling.h"

namespace drw {

// -------------------------------------------------------------- layer bodies
// 1x1 conv on MFMA: in[544][CSI] (CIN_PAD channels read) -> out[544][CSO].
// The product is formed transposed, D[channel][pixel] = W x X^T (weights as
// the A operand, activations as B; the packed fragments are the same
// registers in either role), so a lane ends with 4 consecutive channels of
// one pixel and writes them with one 8-byte store.  The weights enter in two
// fp16 parts (frag = fp16(W), frag_lo = W - fp16(W)): with fp16 weights
// alone the FaceSolverPlus kernels come within a few percent of the 1e-3 L1
// bound against the fp32 forward.  Both parts accumulate in one chain that
// starts from the bias, so the epilogue is the GELU alone.
template <int NTH, int CIN_PAD, int COUT, int COUT_PAD, int CSI, int CSO>
__device__ void pw_mfma(const half_t* in, half_t* out, const half4_t* frag, const half4_t* frag_lo, const float* bias,
                        const float2* gtab)
{
    constexpr int KT = CIN_PAD / 16, NT = COUT_PAD / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = 4 * (lane >> 4);
    half4_t bw[NT][KT], bl[NT][KT];
    float4_t bb[NT];  // bias of the accumulator rows (channels kq..kq+3): the chains start from it
#pragma unroll
    for (int n = 0; n < NT; n++) {
#pragma unroll
        for (int k = 0; k < KT; k++) {
            bw[n][k] = frag[(n * KT + k) * 64 + lane];
            bl[n][k] = frag_lo[(n * KT + k) * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) bb[n][i] = n * 16 + kq + i < COUT ? bias[n * 16 + kq + i] : 0.f;
    }
    for (int m = wave; m < kPix / 16; m += NTH / 64) {
        half4_t a[KT];
#pragma unroll
        for (int k = 0; k < KT; k++) a[k] = *(const half4_t*)(in + (m * 16 + r) * CSI + k * 16 + kq);
        half_t* orow = out + (m * 16 + r) * CSO;
#pragma unroll
        for (int n = 0; n < NT; n++) {
            // one chain: bias, then the small W - fp16(W) part, then fp16(W)
            float4_t acc = bb[n];
            if constexpr (KT == 2) {
                const half8_t a8 = cat8(a[0], a[1]);
                acc = mfma32(cat8(bl[n][0], bl[n][1]), a8, acc);
                acc = mfma32(cat8(bw[n][0], bw[n][1]), a8, acc);
            } else {
#pragma unroll
                for (int k = 0; k < KT; k++) acc = mfma16(bl[n][k], a[k], acc);
#pragma unroll
                for (int k = 0; k < KT; k++) acc = mfma16(bw[n][k], a[k], acc);
            }
            constexpr int VFULL = 16;
            const int V = COUT - n * 16 < VFULL ? COUT - n * 16 : VFULL;  // valid channels of this tile
            if (V == VFULL) {
                half4_t o;
#pragma unroll
                for (int i = 0; i < 4; i++) o[i] = (half_t)gelu_tab(acc[i], gtab);
                *(half4_t*)(orow + n * 16 + kq) = o;
            } else {
                static_assert(COUT % 16 == 0 || COUT % 16 == 8, "tiles hold 16 or 8 valid channels");
                // channels 0..7 sit in lanes 0..31 (elements 0..3); lanes 32..63
                // take elements 2, 3 of lane - 32 so every lane evaluates two GELUs
                const float t2 = __shfl(acc[2], lane & 31, 64), t3 = __shfl(acc[3], lane & 31, 64);
                const bool hi = lane >= 32;
                const int c0 = (hi ? 2 : 0) + 4 * ((lane >> 4) & 1);
                typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
                h2_t o;
                o[0] = (half_t)gelu_tab(hi ? t2 : acc[0], gtab);
                o[1] = (half_t)gelu_tab(hi ? t3 : acc[1], gtab);
                *(h2_t*)(orow + n * 16 + c0) = o;
            }
        }
    }
}

// Activation pairs as fp16 or fp32.  The tail layers (8 channels and fewer)
// keep their activations in fp32: with so few channels per output, fp16
// rounding of their inputs would dominate the kernel's error budget, and
// their storage is small.
__device__ __forceinline__ void load_pair(const half_t* s, float& a, float& b)
{
    const half2_t c = *(const half2_t*)s;
    a = (float)c[0];
    b = (float)c[1];
}
__device__ __forceinline__ void load_pair(const float* s, float& a, float& b)
{
    const float2 c = *(const float2*)s;
    a = c.x;
    b = c.y;
}
__device__ __forceinline__ void store_pair(half_t* d, float a, float b)
{
    half2_t o;
    o[0] = (half_t)a;
    o[1] = (half_t)b;
    *(half2_t*)d = o;
}
__device__ __forceinline__ void store_pair(float* d, float a, float b) { *(float2*)d = make_float2(a, b); }


// Depthwise 3x3 conv (zero padding = dilation) + folded BN + GELU, one
// output (pixel, channel pair) per item with all 9 taps read from LDS: every
// item is independent, so all four waves stay busy.  Channel pairs run
// fastest within a wave, so each tap is one contiguous 64-lane read.  The
// pixels are taken in two passes:
//  * interior pixels (all taps inside the image, (23 - 2 DIL)^2 of them): the
//    nine taps sit at fixed offsets from the pixel, so an item costs its
//    index, nine loads, the products and the GELUs; at dilation 1 an item is
//    three adjacent pixels of a row (the 21 interior columns make 7 groups);
//  * border pixels: at dilation 1 the edges are taken three pixels per item
//    as well (rows along x, columns along y, each reading only its two
//    in-image window lines), the corners one by one; at dilation 2 and 3,
//    and at the corners, a tap outside the image reads `zero` (a zeroed LDS
//    slot) instead, chosen per tap by a select, so the nine loads of an item
//    still issue back to back.
// All paths sum the taps in the same order.  A thread's channel pair, and
// so its 18 weights, stays fixed: items advance by the largest multiple of
// C/2 up to 256 (252 threads at 24 channels, the other 4 idle here).
template <int NTH, int C, int DIL, int CSI, int CSO, typename TI, typename TO>
__device__ void dw_conv_direct(const TI* __restrict__ in, TO* __restrict__ out, const float* w, const float* bias,
                               const TI* zero, const float2* gtab)
{
    constexpr int CP = C / 2;
    constexpr int STRIDE = (NTH / CP) * CP;
    if (threadIdx.x >= STRIDE) return;
    constexpr int NI = kN - 2 * DIL;                 // interior side
    constexpr int NB = kNN - NI * NI;                // border pixels
    constexpr int NTOP = DIL * kN;                   // pixels in the top (and in the bottom) DIL rows
    const int cp = threadIdx.x % CP;
    float w0[9], w1[9];
#pragma unroll
    for (int k = 0; k < 9; k++) { w0[k] = w[(2 * cp) * 9 + k]; w1[k] = w[(2 * cp + 1) * 9 + k]; }
    const float b0 = bias[2 * cp], b1 = bias[2 * cp + 1];
    auto finish = [&](int p, const float* v0, const float* v1) {
        float a0 = b0, a1 = b1;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            a0 += w0[k] * v0[k];
            a1 += w1[k] * v1[k];
        }
        store_pair(out + p * CSO + 2 * cp, gelu_tab(a0, gtab), gelu_tab(a1, gtab));
    };
    if constexpr (DIL == 1 && NI % 3 == 0) {
        // three adjacent pixels per item: their taps are the 3 x 5 window
        // around them (15 loads instead of 27), the index costs once per three
        constexpr int G = 3, NG = NI / G;
        for (int item = threadIdx.x; item < CP * NI * NG; item += STRIDE) {
            const int q = item / CP;
            const int iy = q / NG;
            const int p = (iy + 1) * kN + 1 + G * (q - iy * NG);  // first pixel of the group
            const TI* c = in + (p - kN - 1) * CSI + 2 * cp;      // top-left tap of the window
            float u0[3][G + 2], u1[3][G + 2];
#pragma unroll
            for (int ky = 0; ky < 3; ky++) {
#pragma unroll
                for (int kx = 0; kx < G + 2; kx++) load_pair(c + (ky * kN + kx) * CSI, u0[ky][kx], u1[ky][kx]);
            }
#pragma unroll
            for (int j = 0; j < G; j++) {
                float v0[9], v1[9];
#pragma unroll
                for (int ky = 0; ky < 3; ky++) {
#pragma unroll
                    for (int kx = 0; kx < 3; kx++) {
                        v0[3 * ky + kx] = u0[ky][j + kx];
                        v1[3 * ky + kx] = u1[ky][j + kx];
                    }
                }
                finish(p + j, v0, v1);
            }
        }
    } else {
#pragma unroll 2
        for (int item = threadIdx.x; item < CP * NI * NI; item += STRIDE) {
            const int q = item / CP;
            const int iy = q / NI;
            const int p = q + iy * (kN - NI) + DIL * (kN + 1);  // (iy + DIL) kN + (q - iy NI) + DIL
            const TI* c = in + p * CSI + 2 * cp;
            float v0[9], v1[9];
#pragma unroll
            for (int ky = 0; ky < 3; ky++) {
#pragma unroll
                for (int kx = 0; kx < 3; kx++)
                    load_pair(c + ((ky - 1) * DIL * kN + (kx - 1) * DIL) * CSI, v0[3 * ky + kx], v1[3 * ky + kx]);
            }
            finish(p, v0, v1);
        }
    }
    if constexpr (DIL == 1 && NI % 3 == 0) {
        // Border at dilation 1, three pixels per item like the interior: the
        // top and bottom rows (x = 1..21) in groups along x with their two
        // in-image window rows (2 x 5 loads), the left and right columns
        // (y = 1..21) in groups along y with their two in-image window
        // columns (5 x 2 loads), then the four corners one by one.  The taps
        // outside the image are left out of the sums (the general border
        // path adds them as 0 x w); the others are summed in the same
        // (ky, kx) order.
        constexpr int G = 3, NG = NI / G;
        for (int item = threadIdx.x; item < CP * (4 * NG + 4); item += STRIDE) {
            const int q = item / CP;
            if (q < 2 * NG) {
                const bool bot = q >= NG;
                const int x0 = 1 + G * (q - (bot ? NG : 0));
                const int rA = bot ? kN - 2 : 0;  // upper in-image window row
                const TI* c = in + (rA * kN + x0 - 1) * CSI + 2 * cp;
                float a0[G + 2], a1[G + 2], b0v[G + 2], b1v[G + 2];
#pragma unroll
                for (int kx = 0; kx < G + 2; kx++) {
                    load_pair(c + kx * CSI, a0[kx], a1[kx]);
                    load_pair(c + (kN + kx) * CSI, b0v[kx], b1v[kx]);
                }
                // window row A is tap row 1 (top) or 0 (bottom), row B the next
                float wa0[3], wa1[3], wb0[3], wb1[3];
#pragma unroll
                for (int kx = 0; kx < 3; kx++) {
                    wa0[kx] = bot ? w0[kx] : w0[3 + kx];
                    wa1[kx] = bot ? w1[kx] : w1[3 + kx];
                    wb0[kx] = bot ? w0[3 + kx] : w0[6 + kx];
                    wb1[kx] = bot ? w1[3 + kx] : w1[6 + kx];
                }
                const int py = bot ? kN - 1 : 0;
#pragma unroll
                for (int j = 0; j < G; j++) {
                    float s0 = b0, s1 = b1;
#pragma unroll
                    for (int kx = 0; kx < 3; kx++) {
                        s0 += wa0[kx] * a0[j + kx];
                        s1 += wa1[kx] * a1[j + kx];
                    }
#pragma unroll
                    for (int kx = 0; kx < 3; kx++) {
                        s0 += wb0[kx] * b0v[j + kx];
                        s1 += wb1[kx] * b1v[j + kx];
                    }
                    store_pair(out + (py * kN + x0 + j) * CSO + 2 * cp, gelu_tab(s0, gtab), gelu_tab(s1, gtab));
                }
            } else if (q < 4 * NG) {
                const bool right = q >= 3 * NG;
                const int y0 = 1 + G * (q - (right ? 3 * NG : 2 * NG));
                const int cA = right ? kN - 2 : 0;  // left in-image window column
                const TI* c = in + ((y0 - 1) * kN + cA) * CSI + 2 * cp;
                float a0[G + 2], a1[G + 2], b0v[G + 2], b1v[G + 2];
#pragma unroll
                for (int ky = 0; ky < G + 2; ky++) {
                    load_pair(c + ky * kN * CSI, a0[ky], a1[ky]);
                    load_pair(c + (ky * kN + 1) * CSI, b0v[ky], b1v[ky]);
                }
                // window column A is tap column 1 (left) or 0 (right), B the next
                float wa0[3], wa1[3], wb0[3], wb1[3];
#pragma unroll
                for (int ky = 0; ky < 3; ky++) {
                    wa0[ky] = right ? w0[3 * ky] : w0[3 * ky + 1];
                    wa1[ky] = right ? w1[3 * ky] : w1[3 * ky + 1];
                    wb0[ky] = right ? w0[3 * ky + 1] : w0[3 * ky + 2];
                    wb1[ky] = right ? w1[3 * ky + 1] : w1[3 * ky + 2];
                }
                const int px = right ? kN - 1 : 0;
#pragma unroll
                for (int j = 0; j < G; j++) {
                    float s0 = b0, s1 = b1;
#pragma unroll
                    for (int ky = 0; ky < 3; ky++) {
                        s0 += wa0[ky] * a0[j + ky];
                        s1 += wa1[ky] * a1[j + ky];
                        s0 += wb0[ky] * b0v[j + ky];
                        s1 += wb1[ky] * b1v[j + ky];
                    }
                    store_pair(out + ((y0 + j) * kN + px) * CSO + 2 * cp, gelu_tab(s0, gtab), gelu_tab(s1, gtab));
                }
            } else {
                const int k = q - 4 * NG;
                const int y = (k >> 1) ? kN - 1 : 0, x = (k & 1) ? kN - 1 : 0;
                const int p = y * kN + x;
                const TI* c = in + p * CSI + 2 * cp;
                const bool rv[3] = {y >= 1, true, y + 1 < kN};
                const bool cv[3] = {x >= 1, true, x + 1 < kN};
                float v0[9], v1[9];
#pragma unroll
                for (int ky = 0; ky < 3; ky++) {
#pragma unroll
                    for (int kx = 0; kx < 3; kx++) {
                        const TI* src = (rv[ky] && cv[kx]) ? c + ((ky - 1) * kN + (kx - 1)) * CSI : zero;
                        load_pair(src, v0[3 * ky + kx], v1[3 * ky + kx]);
                    }
                }
                finish(p, v0, v1);
            }
        }
        return;
    }
    for (int item = threadIdx.x; item < CP * NB; item += STRIDE) {
        const int q = item / CP;
        // top DIL rows, bottom DIL rows, then the DIL left and DIL right
        // pixels of each middle row
        int p;
        if (q < NTOP) {
            p = q;
        } else if (q < 2 * NTOP) {
            p = (kN - DIL) * kN + (q - NTOP);
        } else {
            const int r = q - 2 * NTOP, row = DIL + r / (2 * DIL), k = r - (row - DIL) * (2 * DIL);
            p = row * kN + (k < DIL ? k : kN - 2 * DIL + k);
        }
        const int y = p / kN, x = p - y * kN;
        const TI* c = in + p * CSI + 2 * cp;
        const bool rv[3] = {y >= DIL, true, y + DIL < kN};
        const bool cv[3] = {x >= DIL, true, x + DIL < kN};
        float v0[9], v1[9];
#pragma unroll
        for (int ky = 0; ky < 3; ky++) {
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const TI* s = (rv[ky] && cv[kx]) ? c + ((ky - 1) * DIL * kN + (kx - 1) * DIL) * CSI : zero;
                load_pair(s, v0[3 * ky + kx], v1[3 * ky + kx]);
            }
        }
        finish(p, v0, v1);
    }
}

// Depthwise layer (the direct form for every channel count).
template <int NTH, int C, int DIL, int CSI, int CSO, typename TI, typename TO>
__device__ __forceinline__ void dw_conv(const TI* in, TO* out, const float* w, const float* bias, const float* zero,
                                        const float2* gtab)
{
    dw_conv_direct<NTH, C, DIL, CSI, CSO>(in, out, w, bias, (const TI*)zero, gtab);
}

// Small 1x1 conv + folded BN + GELU on the VALU (CIN <= 8, COUT <= 4), fp32
// activations.
template <int NTH, int CIN, int COUT, int CSI, int CSO>
__device__ void pw_valu(const float* in, float* out, const float* w, const float* bias, const float2* gtab)
{
    for (int p = threadIdx.x; p < kNN; p += NTH) {
        float v[CIN];
#pragma unroll
        for (int c = 0; c < CIN; c++) v[c] = (float)in[p * CSI + c];
#pragma unroll
        for (int o = 0; o < COUT; o++) {
            float a = bias[o];
#pragma unroll
            for (int c = 0; c < CIN; c++) a += w[o * CIN + c] * v[c];
            out[p * CSO + o] = gelu_tab(a, gtab);
        }
    }
}

// Head 1x1 conv CIN -> 1 with bias, no activation: fp32 per pixel.
template <int NTH, int CIN, int CSI>
__device__ void head_conv(const float* in, float* hout, const float* w, float b)
{
    for (int p = threadIdx.x; p < kNN; p += NTH) {
        float a = b;
#pragma unroll
        for (int c = 0; c < CIN; c++) a += w[c] * in[p * CSI + c];
        hout[p] = a;
    }
}

// ------------------------------------------------------------ input layer
// Source voxel (z, y, x) of input element T[c][h][w] of the face predictors'
// first layer: mode 0: rotate_poisson_f(cube); mode 1:
// rotate_gradient_f(axis_rotate(cube)); mode 2: the same, permuted so that
// channels are y (SignedFacePredictor channel_dim = 1).
template <int MODE>
__device__ __forceinline__ void input_src(int face, int axis, int c, int h, int w, int& z, int& y, int& x)
{
    if (MODE == 2) { z = h; y = c; x = w; } else { z = c; y = h; x = w; }
    if (MODE == 0) {
        remap_poisson(face, z, y, x);
    } else {
        remap_gradient(face, z, y, x);
        remap_axis(axis, z, y, x);
    }
}

// First 1x1 conv (25 input channels -> COUT) with its operand built in
// registers: the MFMA B fragment of lane l is pixel m*16 + l%16, K slots
// 4 kt + j of lane group l/16 in input_slot_channel order.  Every remap is a
// permutation with flips, so the source voxel is an affine function of
// (c, h, w) in which exactly one source axis moves with c, and its offset and
// slopes depend on the query's face and axis only: the moving axis' cuboid
// masks (and layer values, when it is z) of a lane's six slices are loaded
// once, and a pixel only ANDs them with the masks of its two fixed axes.
// The product runs in split precision, W a ~= Wh ah + Wh al + Wl ah with
// fp16 parts of the fp32 operands: the input layer's rounding is otherwise
// the largest share of the network's error against the fp32 forward.
template <int NTH, int MODE, int COUT, int CSO>
__device__ void pw_input_mfma(const VoxDesc& d, int face, int axis, half_t* out, const half4_t* frag,
                              const half4_t* frag_lo, const float* bias, const float2* gtab)
{
    static_assert(COUT % 16 == 0, "full MFMA column tiles only");
    constexpr int KT = 2, NT = COUT / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    half4_t bw[NT][KT], bl[NT][KT];
    float4_t bb[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) {
#pragma unroll
        for (int k = 0; k < KT; k++) {
            bw[n][k] = frag[(n * KT + k) * 64 + lane];
            bl[n][k] = frag_lo[(n * KT + k) * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) bb[n][i] = bias[n * 16 + 4 * g + i];
    }
    int oz, oy, ox, cz, cy, cx, hz, hy, hx, wz, wy, wx;
    input_src<MODE>(face, axis, 0, 0, 0, oz, oy, ox);
    input_src<MODE>(face, axis, 1, 0, 0, cz, cy, cx);
    input_src<MODE>(face, axis, 0, 1, 0, hz, hy, hx);
    input_src<MODE>(face, axis, 0, 0, 1, wz, wy, wx);
    cz -= oz; cy -= oy; cx -= ox;
    hz -= oz; hy -= oy; hx -= ox;
    wz -= oz; wy -= oy; wx -= ox;
    const bool zmov = cz != 0, ymov = cy != 0, xmov = cx != 0;
    const int vbase = zmov ? oz : ymov ? oy : ox, vdir = cz + cy + cx;
    const uint16_t* vm = zmov ? d.bz : ymov ? d.by : d.bx;
    uint32_t vmask[6], vlz[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int c = min(6 * g + q, kN - 1);  // group 3's sixth slot is a position plane
        const int k = vbase + vdir * c;
        vmask[q] = vm[k];
        vlz[q] = d.lzs[k];
    }
    // Rows past the last pixel (p >= 529 in the last tile) repeat pixel 528's
    // inputs: their outputs are finite and no later layer reads them (the
    // depthwise taps stay inside the image, the tail and the head loop over
    // p < 529).  The loop is instantiated for a moving z axis (layer value per
    // slice) and a fixed one (layer value per pixel).
    auto tiles = [&](auto zmov_c) {
        constexpr bool ZM = decltype(zmov_c)::value;
        for (int m = wave; m < kPix / 16; m += NTH / 64) {
            const int p = m * 16 + r;
            const int pc = min(p, kNN - 1);
            const int h = pc / kN, w = pc - h * kN;
            // fixed-axis coordinates (the moving axis' value is in range and masked out)
            const int zs = oz + hz * h + wz * w, ys = oy + hy * h + wy * w, xs = ox + hx * h + wx * w;
            const uint32_t fixedm = (ZM ? 0xffffu : (uint32_t)d.bz[zs]) & (ymov ? 0xffffu : (uint32_t)d.by[ys]) &
                                    (xmov ? 0xffffu : (uint32_t)d.bx[xs]);
            const uint32_t lzp = ZM ? 0u : d.lzs[zs];
            uint32_t v[8];  // fp16 pairs hi | lo << 16
            // Branch-free: the six cuboid reads always issue (slot 15 when no
            // cuboid holds the voxel) and the empty asm, which takes all six
            // values, keeps the compiler from moving each read into a branch.
            uint32_t mm[6], cvv[6];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                mm[q] = fixedm & vmask[q];
                cvv[q] = d.cvs[__builtin_ctz(mm[q] | (1u << kMaxCuboids))];
            }
            asm volatile("" : "+v"(cvv[0]), "+v"(cvv[1]), "+v"(cvv[2]), "+v"(cvv[3]), "+v"(cvv[4]), "+v"(cvv[5]));
#pragma unroll
            for (int q = 0; q < 6; q++) v[q] = mm[q] ? cvv[q] : (ZM ? vlz[q] : lzp);
            // position planes (PosEmbed2D: linspace(0, 1, 23) over w, then h)
            v[5] = g == 3 ? half_split((float)w * (1.0f / (kN - 1))) : v[5];
            v[6] = g == 0 ? half_split((float)h * (1.0f / (kN - 1))) : 0u;
            v[7] = 0u;
            half4_t ah[KT], al[KT];
#pragma unroll
            for (int kt = 0; kt < KT; kt++) {
                const uint32_t* e = v + 4 * kt;
                const uint2 hi2 = make_uint2((e[0] & 0xffffu) | (e[1] << 16), (e[2] & 0xffffu) | (e[3] << 16));
                const uint2 lo2 = make_uint2((e[0] >> 16) | (e[1] & 0xffff0000u), (e[2] >> 16) | (e[3] & 0xffff0000u));
                ah[kt] = __builtin_bit_cast(half4_t, hi2);
                al[kt] = __builtin_bit_cast(half4_t, lo2);
            }
            // transposed product (see pw_mfma): 4 consecutive channels of pixel p
#pragma unroll
            for (int n = 0; n < NT; n++) {
                // the lane's eight K slots form one K = 32 operand; one chain
                // from the bias: the two small cross terms first, then Wh ah
                static_assert(KT == 2, "eight K slots per lane");
                const half8_t a8 = cat8(ah[0], ah[1]), l8 = cat8(al[0], al[1]), w8 = cat8(bw[n][0], bw[n][1]);
                float4_t acc = mfma32(cat8(bl[n][0], bl[n][1]), a8, bb[n]);
                acc = mfma32(w8, l8, acc);
                acc = mfma32(w8, a8, acc);
                half4_t o;
#pragma unroll
                for (int i = 0; i < 4; i++) o[i] = (half_t)gelu_tab(acc[i], gtab);
                *(half4_t*)(out + p * CSO + n * 16 + 4 * g) = o;
            }
        }
    };
    if (zmov) tiles(std::true_type{}); else tiles(std::false_type{});
}

// ------------------------------------------------------------ architectures
// Two activation buffers X, Y [544][CS] (fp16).  The first layer reads its
// input from registers; every later 1x1 layer runs in place (a wave loads
// the 16 rows of a tile into registers before it writes them, and a row's
// outputs depend on that row only), every depthwise layer moves X <-> Y.
// The fp32 tail (<= 8 channels, 32 B per pixel) fits the same buffers.
template <int CS>
struct FaceSmem {
    alignas(16) half_t X[kPix * CS];  // 16-byte aligned: read as half4 / float2
    alignas(16) half_t Y[kPix * CS];
    alignas(16) float zero[4];  // read by out-of-image depthwise taps; zeroed at kernel start
    alignas(16) float2 gtab[kGeluTabEntries];  // GELU table, copied from the network's device copy at kernel start
    VoxDesc desc;
    float red[8];
    int ibuf[4];
};

// Head output: float[529] in Y.
template <int NTH, int MODE>
__device__ __forceinline__ void face_solver(FaceSmem<16>& s, const FaceNetDev& w, int face, int axis)
{
    pw_input_mfma<NTH, MODE, 16, 16>(s.desc, face, axis, s.X, w.frag[0], w.frag_lo[0], w.pw_bias[0], s.gtab); __syncthreads();
    dw_conv<NTH, 16, 1, 16, 16>(s.X, s.Y, w.dw_w[0], w.dw_b[0], s.zero, s.gtab); __syncthreads();
    pw_mfma<NTH, 16, 16, 16, 16, 16>(s.Y, s.Y, w.frag[1], w.frag_lo[1], w.pw_bias[1], s.gtab); __syncthreads();
    dw_conv<NTH, 16, 1, 16, 16>(s.Y, s.X, w.dw_w[1], w.dw_b[1], s.zero, s.gtab); __syncthreads();
    pw_mfma<NTH, 16, 8, 16, 16, 16>(s.X, s.X, w.frag[2], w.frag_lo[2], w.pw_bias[2], s.gtab); __syncthreads();
    // fp32 tail
    float* fX = (float*)s.X;
    float* fY = (float*)s.Y;
    dw_conv<NTH, 8, 2, 16, 8>(s.X, fY, w.dw_w[2], w.dw_b[2], s.zero, s.gtab); __syncthreads();
    pw_valu<NTH, 8, 4, 8, 4>(fY, fX, w.pw_w[3], w.pw_bias[3], s.gtab); __syncthreads();
    dw_conv<NTH, 4, 3, 4, 4>(fX, fY, w.dw_w[3], w.dw_b[3], s.zero, s.gtab); __syncthreads();
    pw_valu<NTH, 4, 2, 4, 2>(fY, fX, w.pw_w[4], w.pw_bias[4], s.gtab); __syncthreads();
    head_conv<NTH, 2, 2>(fX, fY, w.head_w, w.head_b);
}

template <int NTH>
__device__ __forceinline__ void face_solver_plus(FaceSmem<32>& s, const FaceNetDev& w, int face, int axis)
{
    pw_input_mfma<NTH, 2, 32, 32>(s.desc, face, axis, s.X, w.frag[0], w.frag_lo[0], w.pw_bias[0], s.gtab); __syncthreads();
    dw_conv<NTH, 32, 1, 32, 32>(s.X, s.Y, w.dw_w[0], w.dw_b[0], s.zero, s.gtab); __syncthreads();
    pw_mfma<NTH, 32, 32, 32, 32, 32>(s.Y, s.Y, w.frag[1], w.frag_lo[1], w.pw_bias[1], s.gtab); __syncthreads();
    dw_conv<NTH, 32, 1, 32, 32>(s.Y, s.X, w.dw_w[1], w.dw_b[1], s.zero, s.gtab); __syncthreads();
    // from 24 channels on the activations use a pixel stride of their channel
    // count (24, then 16 halves; the front of each buffer): the 16 rows of an
    // MFMA tile and the depthwise reads of consecutive pixels then spread
    // over the banks, where the 32-channel stride aliases them.  A 24-channel
    // row read with the layer's 32-wide K reads 8 halves of the next row for
    // the K padding, whose weights are zero; every half the layers read
    // there was written by an earlier layer, so the products are exact zeros.
    pw_mfma<NTH, 32, 24, 32, 32, 24>(s.X, s.Y, w.frag[2], w.frag_lo[2], w.pw_bias[2], s.gtab); __syncthreads();
    dw_conv<NTH, 24, 1, 24, 24>(s.Y, s.X, w.dw_w[2], w.dw_b[2], s.zero, s.gtab); __syncthreads();
    pw_mfma<NTH, 32, 24, 32, 24, 24>(s.X, s.X, w.frag[3], w.frag_lo[3], w.pw_bias[3], s.gtab); __syncthreads();
    dw_conv<NTH, 24, 1, 24, 24>(s.X, s.Y, w.dw_w[3], w.dw_b[3], s.zero, s.gtab); __syncthreads();
    pw_mfma<NTH, 32, 16, 16, 24, 16>(s.Y, s.X, w.frag[4], w.frag_lo[4], w.pw_bias[4], s.gtab); __syncthreads();
    dw_conv<NTH, 16, 2, 16, 16>(s.X, s.Y, w.dw_w[4], w.dw_b[4], s.zero, s.gtab); __syncthreads();
    pw_mfma<NTH, 16, 16, 16, 16, 16>(s.Y, s.Y, w.frag[5], w.frag_lo[5], w.pw_bias[5], s.gtab); __syncthreads();
    dw_conv<NTH, 16, 2, 16, 16>(s.Y, s.X, w.dw_w[5], w.dw_b[5], s.zero, s.gtab); __syncthreads();
    pw_mfma<NTH, 16, 8, 16, 16, 16>(s.X, s.X, w.frag[6], w.frag_lo[6], w.pw_bias[6], s.gtab); __syncthreads();
    float* fX = (float*)s.X;
    float* fY = (float*)s.Y;
    dw_conv<NTH, 8, 3, 16, 8>(s.X, fY, w.dw_w[6], w.dw_b[6], s.zero, s.gtab); __syncthreads();
    pw_valu<NTH, 8, 4, 8, 4>(fY, fX, w.pw_w[7], w.pw_bias[7], s.gtab); __syncthreads();
    head_conv<NTH, 4, 4>(fX, fY, w.head_w, w.head_b);
}

// Sum of v over the first four waves (the other waves' v is ignored): the
// summation order is the same for 256- and 512-thread workgroups, so a
// kernel's normalisation, and with it the drawn tile, does not depend on the
// launch shape.
__device__ __forceinline__ float sum_first4_waves(float v, float* red)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0 && threadIdx.x < 256) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const float s = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();
    return s;
}

// Output head of the face predictors (architecture.py:122-132: relu + 1e-10
// then / sum for the Poisson kernel, x / sum |x| for the signed gradient
// kernels) and, for the gradient predictors, post_process_gradient
// (cuda_kernels.cu:516-578).  Reads the head output (fp32 [529] in Y) and
// returns the normalised kernel of the face, fp32 [529] in X.  Ends with a
// barrier.
template <int NTH, int MODE, int BCS>
__device__ __forceinline__ float* face_output_kernel(FaceSmem<BCS>& s, int face, int axis)
{
    const float* hout = (const float*)s.Y;
    // the sum runs on 256 threads whatever NTH (launch-shape independent)
    float loc = 0.f;
    if (threadIdx.x < 256)
        for (int p = threadIdx.x; p < kNN; p += 256) {
            const float v = hout[p];
            loc += MODE == 0 ? (fmaxf(v, 0.f) + 1e-10f) : fabsf(v);
        }
    const float tot = sum_first4_waves(loc, s.red);
    float* kern = (float*)s.X;
    for (int p = threadIdx.x; p < kNN; p += NTH) {
        float v;
        if (MODE == 0) {
            v = (fmaxf(hout[p], 0.f) + 1e-10f) / tot;
        } else {
            const int y = p / kN, x = p - y * kN;
            const bool axX = axis == 0, axY = axis == 1;
            const bool f01 = face == 0 || face == 1, f45 = face == 4 || face == 5;
            const bool rot = axX || (axY && f45), flip = (axX && f45) || (axY && f01);
            int sy = y, sx = x;
            if (rot && flip) { sy = x; sx = y; }
            else if (rot) { sy = x; sx = kN - 1 - y; }
            else if (flip) { sy = kN - 1 - y; }
            v = hout[sy * kN + sx] / tot;
            if (face == 0) v = -v;
        }
        kern[p] = v;
    }
    __syncthreads();
    return kern;
}

// Result of a query whose tile t was drawn from the normalised kernel `kern`
// of `face` (t < 0: no drawable tile, a NaN result): the panel centre on the
// face (locate_index_simple, cuda_kernels.cu:219-264; gradient faces in the
// cube's own frame), and for gradient queries the signed weight W sign(g_t)
// and the centre permittivity of the axis-rotated cube times the max
// (dnn_solver_gradient.cpp:136-162, 260-266).  r: 3 or 5 values.
template <int MODE>
__device__ __forceinline__ void face_result(const float* kern, const VoxDesc& d, int t, int face, int axis, float weight,
                                            float* r)
{
    const int width = MODE == 0 ? 3 : 5;
    if (t < 0) {
        for (int j = 0; j < width; j++) r[j] = __builtin_nanf("");
        return;
    }
    float pt[3];
    panel_point(kN, t, MODE == 0 ? face : axis_face_remap(face, axis), pt);
    r[0] = pt[0]; r[1] = pt[1]; r[2] = pt[2];
    if (MODE != 0) {
        const float g = kern[t];
        const float sg = g > 0.f ? 1.f : (g < 0.f ? -1.f : 0.f);
        r[3] = weight * sg;
        int z = kN / 2, y = kN / 2, x = kN / 2;
        remap_axis(axis, z, y, x);
        r[4] = vox_value(d, z, y, x) * d.maxv;
    }
}

}  // namespace drw
