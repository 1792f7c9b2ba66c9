// Device bodies of the face selectors (FaceSelector / FaceSelectorWeight,
// training_pytorch/src/architecture.py:176-240): the Conv3d backbone, the
// Linear head and the face draw, shared by the batched selector kernels
// (nn_sel.hip) and the resident serving kernel (resident.hip).  NTH: threads
// of the calling workgroup (256 or 512); every function is called by all of
// them.
#pragma once

#include "common.h"
#include "nn_common.h"
#include "nn_sel.h"
#include "sampling.h"

namespace drw {

// Conv3d(CIN -> COUT, k3, s2, no padding) over in[z][y][x][ci] (fp16,
// channel innermost, DIN^3 positions) as an implicit GEMM (rows = output
// positions, k = tap*CIN + ci, tap = 9 tz + 3 ty + tx) with bias + GELU,
// into out[oz][oy][ox][co] (fp16, channel innermost, for the next layer) or
// out[co][oz*DOUT^2+oy*DOUT+ox] (fp32, FP32_OUT: the flatten order of the
// Linear head); `out` must not overlap `in`.  A lane's four K slots of an
// MFMA step are four consecutive channels of one tap, so its operand is one
// 8-byte LDS read at the tap's offset (koff, LDS, 28 ints: -1 past the last
// tap).  With COUT < 16 only COUT of the 16 MFMA columns hold outputs; the
// valid ones are first moved across lanes (ds_bpermute) so that every lane
// evaluates 16*COUT/64 GELUs instead of 4.  Ends with a barrier: outputs
// visible, koff free for the next layer.
template <int CIN, int COUT, int DIN, int DOUT, bool FP32_OUT, int NTH = 256>
__device__ void conv3d_mfma(const half_t* in, void* outp, const half4_t* frag, const float* bias, int* koff)
{
    constexpr int DO2 = DOUT * DOUT, M = DOUT * DO2;
    constexpr int K = CIN * 27, KT = (K + 15) / 16;
    constexpr int NT = (COUT + 15) / 16;
    constexpr int VPT = COUT < 16 ? 16 * COUT / 64 : 4 * NT;  // outputs per lane per tile
    constexpr int NW = NTH / 64;
    constexpr int TP = (M + 15) / 16, TPW = (TP + NW - 1) / NW;  // tiles / tiles per wave
    static_assert(COUT >= 16 || (16 * COUT) % 64 == 0, "compacted tiles need COUT a multiple of 4");
    static_assert(CIN % 4 == 0, "a lane's four K slots are channels of one tap");
    static_assert(KT * 16 <= 28 * CIN, "koff covers every tap a step can reach");
    for (int t = threadIdx.x; t < 28; t += NTH)
        koff[t] = t < 27 ? ((t / 9) * DIN * DIN + ((t / 3) % 3) * DIN + t % 3) * CIN : -1;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = 4 * (lane >> 4);
    half4_t bw[NT][KT > 2 ? 1 : KT];  // fragments of short K kept in registers
    if constexpr (KT <= 2) {
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int kt = 0; kt < KT; kt++) bw[n][kt] = frag[(n * KT + kt) * 64 + lane];
    }
    // operand of output row m (base: its window's first element, -1 past M)
    // for step kt: channels c0..c0+3 of tap (kt*16 + kq) / CIN
    auto gather = [&](int base, int kt) __attribute__((always_inline)) {
        const int k0 = kt * 16 + kq, tap = k0 / CIN, c0 = k0 - tap * CIN;
        const int o = koff[tap];
        half4_t a = {(half_t)0.f, (half_t)0.f, (half_t)0.f, (half_t)0.f};
        if (base >= 0 && o >= 0) a = *(const half4_t*)(in + base + o + c0);
        return a;
    };
    auto row_base = [](int m) __attribute__((always_inline)) {
        if (m >= M) return -1;
        const int oz = m / DO2, oy = (m / DOUT) % DOUT, ox = m % DOUT;
        return (2 * oz * DIN * DIN + 2 * oy * DIN + 2 * ox) * CIN;
    };
    auto store = [&](int ch, int mm, float g) __attribute__((always_inline)) {
        if (mm < M) {
            if (FP32_OUT) ((float*)outp)[ch * M + mm] = g;
            else ((half_t*)outp)[mm * COUT + ch] = (half_t)g;
        }
    };
    if constexpr (NT > 1 && TP < NW) {
        // fewer row tiles than waves (FaceSelectorWeight's last layer: one
        // tile of 8 positions, 64 channels): one (row tile, channel tile)
        // per wave, so every wave works; each product is the same MFMA chain
        // over k as below
        constexpr int ITEMS = TP * NT, IPW = (ITEMS + NW - 1) / NW;
#pragma unroll
        for (int u = 0; u < IPW; u++) {
            const int item = wave + NW * u;
            if (item >= ITEMS) continue;
            const int t = item / NT, n = item - t * NT;
            const int base = row_base(t * 16 + r);
            float4_t acc = {0.f, 0.f, 0.f, 0.f};
            for (int kt = 0; kt < KT; kt++) acc = mfma16(gather(base, kt), frag[(n * KT + kt) * 64 + lane], acc);
            const int ch = n * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; i++) store(ch, t * 16 + kq + i, gelu(acc[i] + bias[ch]));
        }
        __syncthreads();
        return;
    }
#pragma unroll
    for (int u = 0; u < TPW; u++) {
        const int t = wave + NW * u;
        if (t >= TP) continue;
        const int base = row_base(t * 16 + r);
        float4_t acc[NT];
#pragma unroll
        for (int n = 0; n < NT; n++) acc[n] = (float4_t){0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < KT; kt++) {
            const half4_t a = gather(base, kt);
#pragma unroll
            for (int n = 0; n < NT; n++) {
                half4_t b;
                if constexpr (KT <= 2) b = bw[n][kt];
                else b = frag[(n * KT + kt) * 64 + lane];
                acc[n] = mfma16(a, b, acc[n]);
            }
        }
#pragma unroll
        for (int j = 0; j < VPT; j++) {
            float v;
            int ch, mm;
            if constexpr (COUT < 16) {
                // output o = lane*VPT + j -> (row, col) = (o / COUT, o % COUT),
                // held by lane col + 16*(row/4) as element row % 4
                const int o = lane * VPT + j, row = o / COUT;
                ch = o % COUT;
                mm = t * 16 + row;
                const int src = ch + 16 * (row >> 2), el = row & 3;
                v = 0.f;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float tv = __shfl(acc[0][i], src, 64);
                    v = el == i ? tv : v;
                }
            } else {
                ch = (j >> 2) * 16 + r;
                mm = t * 16 + kq + (j & 3);
                v = acc[j >> 2][j & 3];
            }
            store(ch, mm, gelu(v + bias[ch]));
        }
    }
    __syncthreads();
}

// First Conv3d (1 -> C1, k3, s2, no padding) + folded BN + GELU (gelu_tab_sel) on the VALU,
// one output z-slab per phase, in place: slab oz reads input planes
// 2oz..2oz+2, so after the phase's barrier the planes below 2oz+2 are dead
// and the slab's outputs (held in registers until then) are written over
// them as out[oz][oy*11+ox][c] (channel innermost, C1 * 121 <= 2 * 529).  An item is one output
// pixel of the slab and one channel pair; its 27 taps lie inside the cube
// for every output, so they are LDS reads at fixed offsets from one base
// address.  The slab base offset is even and the plane and row strides
// (529, 23) are odd, so in tap row (tz, ty) the x-taps 0, 1 (tz + ty even)
// or 1, 2 (odd) form a 4-byte-aligned pair: a row is one 32-bit and one
// 16-bit read, and the products are v_dot2_f32_f16 (fp16 inputs and
// weights, fp32 accumulate) over the nine pairs and the nine remaining taps
// paired up in row order (conv1_pairs on the host lays the weights out the
// same way).  With one input channel (K = 27) and 4 or 8 outputs, the
// implicit-GEMM form spends most of its instructions on the gather and on
// moving the valid columns out of the 16-wide MFMA tile.
template <int C1, int NTH = 256>
__device__ void conv1_direct(const half_t* cube, half_t* out, const half_t* w, const float* bias, const float2* gtab)
{
    constexpr int DO = 11, DO2 = DO * DO, CP = C1 / 2, NP = 14;
    constexpr int IPT = (DO2 * CP + NTH - 1) / NTH;  // items per thread and phase
    static_assert(C1 % 2 == 0 && NTH % CP == 0, "a thread's channel pair stays fixed");
    static_assert(C1 * DO2 <= 2 * kNN, "slab outputs must fit the freed planes");
    const int cp = threadIdx.x % CP;
    half2_t w0[NP], w1[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
        w0[k] = ((const half2_t*)w)[(2 * cp) * NP + k];
        w1[k] = ((const half2_t*)w)[(2 * cp + 1) * NP + k];
    }
    const float b0 = bias[2 * cp], b1 = bias[2 * cp + 1];
    for (int oz = 0; oz < DO; oz++) {
        float held[IPT][2];
        int pix[IPT];
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            const int mm = (threadIdx.x + NTH * u) / CP;
            pix[u] = mm;
            float a0 = b0, a1 = b1;
            if (mm < DO2) {
                const int oy = mm / DO, ox = mm - oy * DO;
                const half_t* src = cube + 2 * oz * kNN + 2 * oy * kN + 2 * ox;
                half_t single[9];
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    const int tz = r / 3, ty = r % 3, par = (tz + ty) & 1;
                    const half_t* row = src + tz * kNN + ty * kN;
                    const half2_t pr = *(const half2_t*)(row + par);
                    single[r] = row[par ? 0 : 2];
                    a0 = __builtin_amdgcn_fdot2(pr, w0[r], a0, false);
                    a1 = __builtin_amdgcn_fdot2(pr, w1[r], a1, false);
                }
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const half2_t sp = {single[2 * k], k < 4 ? single[2 * k + 1] : (half_t)0.f};
                    a0 = __builtin_amdgcn_fdot2(sp, w0[9 + k], a0, false);
                    a1 = __builtin_amdgcn_fdot2(sp, w1[9 + k], a1, false);
                }
            }
            held[u][0] = gelu_tab_sel(a0, gtab);
            held[u][1] = gelu_tab_sel(a1, gtab);
        }
        __syncthreads();  // this slab's input planes fully read
        half_t* o = out + oz * (C1 * DO2);
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            if (pix[u] < DO2) {
                half2_t h;
                h[0] = (half_t)held[u][0];
                h[1] = (half_t)held[u][1];
                *(half2_t*)(o + pix[u] * C1 + 2 * cp) = h;
            }
        }
    }
    __syncthreads();  // outputs visible
}

// LDS of one query: the cube, the first conv's output (written slab by slab
// over the cube planes it no longer needs, [oz][11*11][c]) and the second
// conv's output ([125][c], after the first's) share one region:
// 26.4 KB in all for FaceSelector (6 workgroups per CU), 27.4 KB for
// FaceSelectorWeight (5 per CU).
template <int C1, int C2>
struct SelSmem {
    static constexpr int kRegion = kNNN > C1 * 1331 + C2 * 125 ? kNNN : C1 * 1331 + C2 * 125;
    alignas(16) int koff[27 * 16];  //  1,728 B  tap offsets of the current layer (28 ints); the
                                    //           first conv's GELU table before that
    half_t cube[kRegion];           // >= 24,334 B
    VoxDesc desc;
};

// Backbone of one query: the (axis pre-rotated) cube built into `cube` from
// the description d, then the three Conv3d layers; writes the flattened
// features (fp32 [C3 * 8]) to `feat`.  On entry d is built and the first
// conv's GELU table (kSelGeluTab entries) sits in `koff`, visible to all
// threads.  Ends with a barrier.
template <bool WEIGHTED, int NTH>
__device__ __forceinline__ void selector_backbone_body(const VoxDesc& d, half_t* cube, int* koff, bool has_axis,
                                                       int axis, const SelNetDev& w, float* feat)
{
    constexpr int C1 = WEIGHTED ? 8 : 4, C2 = WEIGHTED ? 16 : 8, C3 = WEIGHTED ? 64 : 16;
    // Cube in the (axis pre-rotated) frame, built row by row: along an
    // x-row of the output exactly one source axis varies (+-1 per voxel), so
    // the cuboid masks of the two fixed source axes are AND-ed once per row
    // chunk; items are quarter rows (6, 6, 6, 5 voxels) to balance the waves.
    for (int it = threadIdx.x; it < kNN * 4; it += NTH) {
        const int r = it >> 2, q = it & 3;  // row r = 23 z + y of the output frame
        const int zo = r / kN, yo = r - zo * kN;
        // source of output voxel (zo, yo, x) under remap_axis: axis 0 ->
        // (22 - x, yo, zo), axis 1 -> (22 - yo, zo, x), none / 2 -> (zo, yo, x)
        const uint16_t* vm;
        int base, dir, zfix;
        uint32_t fixedm;
        const bool zvar = has_axis && axis == 0;
        if (zvar) {
            vm = d.bz; base = kN - 1; dir = -1; zfix = 0;
            fixedm = (uint32_t)d.by[yo] & d.bx[zo];
        } else {
            const bool ax1 = has_axis && axis == 1;
            const int zs = ax1 ? kN - 1 - yo : zo, ys = ax1 ? zo : yo;
            vm = d.bx; base = 0; dir = 1; zfix = zs;
            fixedm = (uint32_t)d.bz[zs] & d.by[ys];
        }
        const float lfix = d.lz[zfix];
        const int xa = 6 * q, xb = q == 3 ? kN : xa + 6;
        half_t* dst = cube + r * kN;
        for (int x = xa; x < xb; x++) {
            const int k = base + dir * x;
            const uint32_t m = fixedm & vm[k];
            dst[x] = (half_t)(m ? d.cv[__builtin_ctz(m)] : (zvar ? d.lz[k] : lfix));
        }
    }
    __syncthreads();
    half_t* l1 = cube;                 // [11][121][C1], over the cube
    half_t* l2 = cube + C1 * 1331;     // [125][C2], after l1
    conv1_direct<C1, NTH>(cube, l1, w.conv1_w, w.bias[0], (const float2*)koff);
    conv3d_mfma<C1, C2, 11, 5, false, NTH>(l1, l2, w.frag[1], w.bias[1], koff);
    conv3d_mfma<C2, C3, 5, 2, true, NTH>(l2, feat, w.frag[2], w.bias[2], koff);
}

// Linear layer on MFMA for 16 rows: X[16][K] (fp16, LDS) -> Y[16][N].
template <int K, int N, bool ACT>
__device__ void linear16(const half_t* X, half_t* Y, float* Yf, const half4_t* frag, const float* bias)
{
    constexpr int KT = K / 16, NT = (N + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = 4 * (lane >> 4);
    for (int n = wave; n < NT; n += 4) {
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < KT; kt++) {
            const half4_t av = *(const half4_t*)(X + r * K + kt * 16 + kq);
            acc = mfma16(av, frag[(n * KT + kt) * 64 + lane], acc);
        }
        const int col = n * 16 + r;
        if (col >= N) continue;
        const float bb = bias[col];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = kq + i;
            float v = acc[i] + bb;
            if (ACT) v = gelu(v);
            if (Yf) Yf[row * 16 + col] = v;
            else Y[row * N + col] = (half_t)v;
        }
    }
}

// Face draw over six probabilities, sequentially (same arithmetic as the
// wave-parallel draw: fixed-point quantisation + exact cumulative sums).
__device__ inline int small_cdf_draw(const float* w, int M, uint32_t u)
{
    float m = 0.f;
    for (int i = 0; i < M; i++) m = fmaxf(m, fabsf(w[i]));
    if (!(m > 0.f) || isinf(m)) return -1;
    int E;
    frexpf(m, &E);
    const int sh = 24 - E;
    uint64_t total = 0;
    for (int i = 0; i < M; i++) total += (uint32_t)ldexpf(fabsf(w[i]), sh);
    const uint64_t r = (uint64_t)u * (total >> 32) + (((uint64_t)u * (total & 0xFFFFFFFFull)) >> 32);
    uint64_t acc = 0;
    for (int i = 0; i < M; i++) {
        acc += (uint32_t)ldexpf(fabsf(w[i]), sh);
        if (acc > r) return i;
    }
    return M - 1;
}

// Linear layer on MFMA for one query: x[K] (fp16, LDS) -> y[N] (fp16) or
// yf[N] (fp32).  The row is broadcast to the 16 A rows, so every row of the
// product is the query's, computed as linear16 computes a row of a batch
// (same operands, same MFMA chain, same epilogue).
template <int K, int N, bool ACT, int NTH>
__device__ void linear_row(const half_t* x, half_t* y, float* yf, const half4_t* frag, const float* bias)
{
    constexpr int KT = K / 16, NT = (N + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = 4 * (lane >> 4);
    for (int n = wave; n < NT; n += NTH / 64) {
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < KT; kt++) {
            const half4_t av = *(const half4_t*)(x + kt * 16 + kq);
            acc = mfma16(av, frag[(n * KT + kt) * 64 + lane], acc);
        }
        const int col = n * 16 + r;
        if (kq != 0 || col >= N) continue;
        float v = acc[0] + bias[col];
        if (ACT) v = gelu(v);
        if (yf) yf[col] = v;
        else y[col] = (half_t)v;
    }
}

// Softmax over the six face logits and the face draw with the query's first
// uniform word (dnn_solver_poisson.cpp:91-95, dnn_solver_gradient.cpp:
// 240-244).  Returns the face, -1 when no face is drawable (non-finite or
// zero probabilities; the query then gets a NaN result and goes to the FDM
// fallback).  One thread.
__device__ __forceinline__ int selector_probs_draw(const float* lg, uint64_t seed, uint64_t qid, float* p)
{
    float mx = lg[0];
    for (int i = 1; i < 6; i++) mx = fmaxf(mx, lg[i]);
    float sum = 0.f;
    for (int i = 0; i < 6; i++) { p[i] = expf(lg[i] - mx); sum += p[i]; }
    for (int i = 0; i < 6; i++) p[i] /= sum;
    uint32_t u[4];
    query_uniforms(seed, qid, 0u, u);
    return small_cdf_draw(p, 6, u[0]);
}

}  // namespace drw
