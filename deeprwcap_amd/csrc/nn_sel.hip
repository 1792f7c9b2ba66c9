// Face selectors: FaceSelector (PoissonSelector) and FaceSelectorWeight
// (GradientSelectorWeight), training_pytorch/src/architecture.py:176-240.
//
// selector_backbone_kernel: one 256-thread workgroup per query voxelizes the
//   (axis pre-rotated) cube into LDS as fp16 and runs the three
//   Conv3d(k3, s2) + folded BN + GELU layers as implicit GEMMs on
//   v_mfma_f32_16x16x16f16 (rows = output positions, k = ci*27 + tap).
//   Writes the flattened features [B][C3 * 8].
// selector_head_kernel: 16 queries per workgroup, the Linear layers as MFMA
//   GEMMs, softmax over the six face logits, the face draw with the query's
//   first uniform word (dnn_solver_poisson.cpp:91-95, dnn_solver_gradient.cpp:
//   240-244) and, for FaceSelectorWeight, the raw weight logit.
#include "common.h"
#include "nn_common.h"
#include "nn_sel.h"
#include "sampling.h"

namespace drw {

// Conv3d(CIN -> COUT, k3, s2, no padding) over in[CIN][DIN^3] (fp16) as an
// implicit GEMM; out[COUT][DOUT^3] fp16 after bias + GELU, or fp32 features.
// koff (LDS, >= KT*16 ints) receives the im2col offset of every k = ci*27 +
// tap (-1 past K), so a gathered element costs one add and one LDS read.
// With COUT < 16 only COUT of the 16 MFMA columns hold outputs; the valid
// ones are first moved across lanes (ds_bpermute) so that every lane
// evaluates 16*COUT/64 GELUs instead of 4.
template <int CIN, int COUT, int DIN, int DOUT, bool FP32_OUT>
__device__ void conv3d_mfma(const half_t* in, void* outp, const half4_t* frag, const float* bias, int* koff)
{
    constexpr int M = DOUT * DOUT * DOUT, MT = (M + 15) / 16;
    constexpr int K = CIN * 27, KT = (K + 15) / 16;
    constexpr int NT = (COUT + 15) / 16;
    constexpr int D2 = DIN * DIN, DO2 = DOUT * DOUT;
    for (int k = threadIdx.x; k < KT * 16; k += 256) {
        int o = -1;
        if (k < K) {
            const int ci = k / 27, tap = k - 27 * ci;
            o = ci * DIN * D2 + (tap / 9) * D2 + ((tap / 3) % 3) * DIN + tap % 3;
        }
        koff[k] = o;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = 4 * (lane >> 4);
    for (int mt = wave; mt < MT; mt += 4) {
        const int m = mt * 16 + r;
        int base = -1;
        if (m < M) {
            const int oz = m / DO2, oy = (m / DOUT) % DOUT, ox = m % DOUT;
            base = 2 * oz * D2 + 2 * oy * DIN + 2 * ox;
        }
        float4_t acc[NT];
#pragma unroll
        for (int n = 0; n < NT; n++) acc[n] = (float4_t){0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < KT; kt++) {
            const int4 o4 = *(const int4*)(koff + kt * 16 + kq);
            const int oo[4] = {o4.x, o4.y, o4.z, o4.w};
            half4_t a;
#pragma unroll
            for (int j = 0; j < 4; j++) a[j] = (base >= 0 && oo[j] >= 0) ? in[base + oo[j]] : (half_t)0.f;
#pragma unroll
            for (int n = 0; n < NT; n++) acc[n] = mfma16(a, frag[(n * KT + kt) * 64 + lane], acc[n]);
        }
        if constexpr (COUT < 16) {
            // output o = lane*E + e  ->  (row, col) = (o / COUT, o % COUT), held by
            // lane col + 16*(row/4) as element row % 4
            constexpr int E = 16 * COUT / 64;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int o = lane * E + e, row = o / COUT, col = o % COUT;
                const int src = col + 16 * (row >> 2), el = row & 3;
                float v = 0.f;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float t = __shfl(acc[0][i], src, 64);
                    v = el == i ? t : v;
                }
                const int mm = mt * 16 + row;
                if (mm < M) {
                    const float g = gelu(v + bias[col]);
                    if (FP32_OUT) ((float*)outp)[col * M + mm] = g;
                    else ((half_t*)outp)[col * M + mm] = (half_t)g;
                }
            }
        } else {
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int ch = n * 16 + r;
                const float bb = bias[ch];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int mm = mt * 16 + kq + i;
                    if (mm >= M) continue;
                    const float v = gelu(acc[n][i] + bb);
                    if (FP32_OUT) ((float*)outp)[ch * M + mm] = v;
                    else ((half_t*)outp)[ch * M + mm] = (half_t)v;
                }
            }
        }
    }
    __syncthreads();  // koff is rewritten by the next layer
}

// LDS of one query, sized per architecture (FaceSelector: C1 = 4, C2 = 8,
// 36.9 KB; FaceSelectorWeight: C1 = 8, C2 = 16, 47.6 KB).  The second conv
// output reuses the cube's storage: the cube is dead once the first conv has
// finished (conv3d_mfma ends with a barrier).
template <int C1, int C2>
struct SelSmem {
    alignas(16) int koff[27 * 16];  //  1,728 B  im2col offsets of the current layer (read as int4)
    union {
        half_t cube[kNNN];  // 24,334 B
        half_t l2[C2 * 125];
    };
    half_t l1[C1 * 1331];
    VoxDesc desc;
};

template <bool WEIGHTED>
__global__ __launch_bounds__(256) void selector_backbone_kernel(SelBackboneArgs a, SelNetDev w)
{
    constexpr int C1 = WEIGHTED ? 8 : 4, C2 = WEIGHTED ? 16 : 8, C3 = WEIGHTED ? 64 : 16;
    static_assert(C2 * 125 <= kNNN, "second conv output must fit the cube's storage");
    __shared__ SelSmem<C1, C2> s;
    const int b = blockIdx.x;
    if (threadIdx.x < 64) vox_desc_build(s.desc, a.layers, a.cuboids, b);
    __syncthreads();
    const int axis = a.axis ? a.axis[b] : 2;
    // Cube in the (axis pre-rotated) frame, built row by row: along an
    // x-row of the output exactly one source axis varies (+-1 per voxel), so
    // the cuboid masks of the two fixed source axes are AND-ed once per row
    // chunk; items are quarter rows (6, 6, 6, 5 voxels) to balance the waves.
    for (int it = threadIdx.x; it < kNN * 4; it += 256) {
        const int r = it >> 2, q = it & 3;  // row r = 23 z + y of the output frame
        const int zo = r / kN, yo = r - zo * kN;
        int z0 = zo, y0 = yo, x0 = 0, z1 = zo, y1 = yo, x1 = 1;
        if (a.axis) {
            remap_axis(axis, z0, y0, x0);
            remap_axis(axis, z1, y1, x1);
        }
        const VoxDesc& d = s.desc;
        const uint16_t* vm;
        int base, dir;
        uint32_t fixedm;
        const bool zvar = z1 != z0;
        if (zvar) { vm = d.bz; base = z0; dir = z1 - z0; fixedm = (uint32_t)d.by[y0] & d.bx[x0]; }
        else if (y1 != y0) { vm = d.by; base = y0; dir = y1 - y0; fixedm = (uint32_t)d.bz[z0] & d.bx[x0]; }
        else { vm = d.bx; base = x0; dir = x1 - x0; fixedm = (uint32_t)d.bz[z0] & d.by[y0]; }
        const float lfix = d.lz[z0];
        const int xa = 6 * q, xb = q == 3 ? kN : xa + 6;
        half_t* dst = s.cube + r * kN;
        for (int x = xa; x < xb; x++) {
            const int k = base + dir * x;
            const uint32_t m = fixedm & vm[k];
            dst[x] = (half_t)(m ? d.cv[__builtin_ctz(m)] : (zvar ? d.lz[k] : lfix));
        }
    }
    __syncthreads();
    conv3d_mfma<1, C1, 23, 11, false>(s.cube, s.l1, w.frag[0], w.bias[0], s.koff);
    conv3d_mfma<C1, C2, 11, 5, false>(s.l1, s.l2, w.frag[1], w.bias[1], s.koff);
    conv3d_mfma<C2, C3, 5, 2, true>(s.l2, a.feat + (size_t)b * C3 * 8, w.frag[2], w.bias[2], s.koff);
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
__device__ int small_cdf_draw(const float* w, int M, uint32_t u)
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

template <bool WEIGHTED>
__global__ __launch_bounds__(256) void selector_head_kernel(SelHeadArgs a, SelNetDev w)
{
    constexpr int F = WEIGHTED ? 512 : 128;
    constexpr int H1 = WEIGHTED ? 256 : 64;
    constexpr int H2 = 128;
    __shared__ half_t X[16 * F];
    __shared__ half_t Y1[16 * H1];
    __shared__ half_t Y2[16 * H2];
    __shared__ float logits[16 * 16];
    const int b0 = blockIdx.x * 16;
    for (int i = threadIdx.x; i < 16 * F; i += 256) {
        const int row = i / F, b = b0 + row;
        X[i] = (half_t)(b < a.B ? a.feat[(size_t)b * F + (i - row * F)] : 0.f);
    }
    __syncthreads();
    linear16<F, H1, true>(X, Y1, nullptr, w.frag[3], w.bias[3]);
    __syncthreads();
    if (WEIGHTED) {
        linear16<H1, H2, true>(Y1, Y2, nullptr, w.frag[4], w.bias[4]);
        __syncthreads();
        linear16<H2, 7, false>(Y2, nullptr, logits, w.frag[5], w.bias[5]);
    } else {
        linear16<H1, 6, false>(Y1, nullptr, logits, w.frag[4], w.bias[4]);
    }
    __syncthreads();
    const int row = threadIdx.x;
    if (row < 16 && b0 + row < a.B) {
        const int b = b0 + row;
        const float* lg = logits + row * 16;
        float mx = lg[0];
        for (int i = 1; i < 6; i++) mx = fmaxf(mx, lg[i]);
        float p[6], sum = 0.f;
        for (int i = 0; i < 6; i++) { p[i] = expf(lg[i] - mx); sum += p[i]; }
        for (int i = 0; i < 6; i++) p[i] /= sum;
        if (a.probs)
            for (int i = 0; i < 6; i++) a.probs[(size_t)b * 6 + i] = p[i];
        if (WEIGHTED && a.weight) a.weight[b] = lg[6];
        if (a.face) {
            uint32_t u[4];
            const uint64_t base = a.ctr_dev ? *a.ctr_dev : a.counter;
            query_uniforms(a.seed, a.qid ? a.qid[b] : base + (uint64_t)b, 0u, u);
            const int f = small_cdf_draw(p, 6, u[0]);
            a.face[b] = f < 0 ? 0 : f;
        }
    }
}

int selector_backbone_launch(bool weighted, const SelBackboneArgs& ba, const SelNetDev& w, hipStream_t stream)
{
    if (ba.B <= 0) return 0;
    if (weighted) hipLaunchKernelGGL(selector_backbone_kernel<true>, dim3(ba.B), dim3(256), 0, stream, ba, w);
    else hipLaunchKernelGGL(selector_backbone_kernel<false>, dim3(ba.B), dim3(256), 0, stream, ba, w);
    DRW_HIP(hipGetLastError());
    return 0;
}

int selector_head_launch(bool weighted, const SelHeadArgs& ha, const SelNetDev& w, hipStream_t stream)
{
    if (ha.B <= 0) return 0;
    const dim3 grid((ha.B + 15) / 16);
    if (weighted) hipLaunchKernelGGL(selector_head_kernel<true>, grid, dim3(256), 0, stream, ha, w);
    else hipLaunchKernelGGL(selector_head_kernel<false>, grid, dim3(256), 0, stream, ha, w);
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw
