// Face selectors: FaceSelector (PoissonSelector) and FaceSelectorWeight
// (GradientSelectorWeight), training_pytorch/src/architecture.py:176-240.
//
// selector_backbone_kernel: one 256-thread workgroup per query voxelizes the
//   (axis pre-rotated) cube into LDS as fp16 and runs the three
//   Conv3d(k3, s2) + folded BN + GELU layers as implicit GEMMs on
//   v_mfma_f32_16x16x16f16 (rows = output positions, k = ci*27 + tap),
//   except the first conv (one input channel), which runs on the VALU
//   (conv1_direct).
//   Writes the flattened features [B][C3 * 8].
// selector_head_kernel: 16 queries per workgroup, the Linear layers as MFMA
//   GEMMs, softmax over the six face logits, the face draw with the query's
//   first uniform word (dnn_solver_poisson.cpp:91-95, dnn_solver_gradient.cpp:
//   240-244) and, for FaceSelectorWeight, the raw weight logit.
#include "common.h"
#include "nn_common.h"
#include "nn_sel.h"
#include "nn_sel_body.h"
#include "sampling.h"

namespace drw {

template <bool WEIGHTED>
__global__ __launch_bounds__(256) void selector_backbone_kernel(SelBackboneArgs a, SelNetDev w)
{
    static_assert(sizeof(SelSmem<4, 8>::koff) == kSelGeluTab * sizeof(float2), "GELU table fills the offset slot");
    constexpr int C1 = WEIGHTED ? 8 : 4, C2 = WEIGHTED ? 16 : 8, C3 = WEIGHTED ? 64 : 16;
    static_assert(C1 * 1331 + C2 * 125 <= SelSmem<C1, C2>::kRegion, "conv outputs must fit the shared region");
    __shared__ SelSmem<C1, C2> s;
    const int b = blockIdx.x;
    // wave 0 builds the description; the others stage the first conv's GELU
    // table in the im2col-offset slot, which only the later convs use
    if (threadIdx.x < 64) vox_desc_build(s.desc, a.layers, a.cuboids, b);
    else
        for (int i = threadIdx.x - 64; i < kSelGeluTab; i += 192) ((float2*)s.koff)[i] = w.gtab[i];
    __syncthreads();
    selector_backbone_body<WEIGHTED, 256>(s.desc, s.cube, s.koff, a.axis != nullptr, a.axis ? a.axis[b] : 2, w,
                                          a.feat + (size_t)b * C3 * 8);
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
        const uint64_t base = a.ctr_dev ? *a.ctr_dev : a.counter;
        float p[6];
        const int face = selector_probs_draw(lg, a.seed, a.qid ? a.qid[b] : base + (uint64_t)b, p);
        if (a.probs)
            for (int i = 0; i < 6; i++) a.probs[(size_t)b * 6 + i] = p[i];
        if (WEIGHTED && a.weight) a.weight[b] = lg[6];
        if (a.face) a.face[b] = face;
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
