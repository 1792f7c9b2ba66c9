// Fused face-predictor kernel: on-the-fly voxelization in the rotated frame
// + FaceSolver / FaceSolverPlus + output head + post-process + tile draw +
// locate, one 256-thread workgroup per (query, face).
//
// Network definitions: training_pytorch/src/architecture.py:11-173
//   FaceSolver      (PoissonPredictor, Gradient1Predictor)  :44-72
//   FaceSolverPlus  (Gradient2Predictor, input permuted)     :75-105, 135-173
// Query flow: dnn_solver_poisson.cpp:86-111, dnn_solver_gradient.cpp:75-164.
//
// Activations live in LDS as fp16 [pixel][channel] (pixel = 23 h + w, 529
// pixels padded to 544 = 34 MFMA row tiles) up to the last MFMA layer and as
// fp32 in the tail layers (<= 8 channels).  1x1 convolutions with >= 8
// outputs run on MFMA (v_mfma_f32_16x16x32_f16 where a lane holds 8 K slots,
// else 16x16x16; fp32 accumulate); depthwise 3x3, the small 1x1 layers,
// batch-norm (folded into weights/bias on the host) and GELU (LDS table) run
// on the VALU in fp32.
#include <cstddef>

#include "common.h"
#include "nn_common.h"
#include "nn_face.h"
#include "nn_face_body.h"
#include "nn_sel_body.h"
#include "sampling.h"

namespace drw {
// ------------------------------------------------------------------ kernel
// mode 0: Poisson predictor (relu head); 1: Gradient1 (norm head, faces 0,1);
// 2: Gradient2 (FaceSolverPlus, norm head, faces 2..5).
// A query whose result holds a NaN joins the FDM fallback's list (when the
// caller runs the fallback): beyond the list's capacity it keeps its NaN and
// counts as failed.
__device__ __forceinline__ void fallback_append(const FaceQueryArgs& a, int b)
{
    if (!a.fb_count) return;
    const int k = atomicAdd(a.fb_count, 1);
    if (k < a.fb_cap) a.fb_list[k] = b;
    else atomicAdd((unsigned long long*)&a.fb_stats[1], 1ull);
    atomicAdd((unsigned long long*)&a.fb_stats[2], 1ull);
}

template <int MODE, int NTH>
__global__ __launch_bounds__(NTH) void face_query_kernel(FaceQueryArgs a, FaceNetDev w)
{
    constexpr int BCS = MODE == 2 ? 32 : 16;
    __shared__ FaceSmem<BCS> s;
    const int item = blockIdx.x;
    const int b = a.full ? item / 6 : item;
    const int face = a.full ? item % 6 : a.face[b];
    const int axis = a.axis ? a.axis[b] : 2;
    if (face < 0) {  // failed face draw: a NaN result (the Poisson or the Gradient1 kernel writes it)
        const int width = MODE == 0 ? 3 : 5;
        if (MODE != 2 && threadIdx.x < width) a.out[(size_t)b * width + threadIdx.x] = __builtin_nanf("");
        if (MODE != 2 && threadIdx.x == 0 && !a.full) fallback_append(a, b);
        return;
    }
    if (MODE == 1 && !(face == 0 || face == 1)) return;
    if (MODE == 2 && !(face >= 2 && face <= 5)) return;
    if (threadIdx.x < 64) vox_desc_build(s.desc, a.layers, a.cuboids, b);
    if (threadIdx.x < 4) s.zero[threadIdx.x] = 0.f;
    for (int i = threadIdx.x; i < kGeluTabEntries; i += NTH) s.gtab[i] = w.gtab[i];
    __syncthreads();
    if constexpr (MODE == 2) face_solver_plus<NTH>(s, w, face, axis); else face_solver<NTH, MODE>(s, w, face, axis);
    __syncthreads();
    float* kern = face_output_kernel<NTH, MODE>(s, face, axis);
    if (a.full) {
        float* o = a.kern_out + (size_t)item * kNN;
        for (int p = threadIdx.x; p < kNN; p += NTH) o[p] = kern[p];
        return;
    }
    if (a.kern_out) {
        float* o = a.kern_out + (size_t)b * kNN;
        for (int p = threadIdx.x; p < kNN; p += NTH) o[p] = kern[p];
    }
    // ---- tile draw with the query's second uniform word, then locate
    if (threadIdx.x < 64) {
        uint32_t u[4];
        const uint64_t base = a.ctr_dev ? *a.ctr_dev : a.counter;
        query_uniforms(a.seed, a.qid ? a.qid[b] : base + (uint64_t)b, 0u, u);
        const int t = wave_cdf_draw<float>(kern, kNN, u[1]);
        if (threadIdx.x == 0) {
            const int width = MODE == 0 ? 3 : 5;
            float* o = a.out + (size_t)b * width;
            float r[5];
            face_result<MODE>(kern, s.desc, t, face, axis, MODE == 0 ? 0.f : a.weight[b], r);
            bool nan = false;
            for (int j = 0; j < width; j++) {
                o[j] = r[j];
                nan |= isnan(r[j]);
            }
            if (nan) fallback_append(a, b);
        }
    }
}

// ------------------------------------------------------- fused Poisson query
// LDS of the fused kernel: the selector phase (im2col offsets / GELU table,
// cube, features) and the head's activations overlay the predictor's
// activation buffers X, Y; the predictor's description, GELU table and zero
// slot sit beyond them and are filled once, at the start.  39.4 KB: four
// 256-thread workgroups per CU, as for the predictor kernel alone.
union PoissonSmem {
    FaceSmem<16> f;
    struct {
        alignas(16) int koff[27 * 16];
        half_t cube[SelSmem<4, 8>::kRegion];
        alignas(16) float feat[128];
    } sel;
    struct {
        alignas(16) half_t x[128];
        alignas(16) half_t y1[64];
        alignas(16) float logits[16];
    } head;
};
static_assert(sizeof(PoissonSmem::sel) <= offsetof(FaceSmem<16>, zero), "selector phase must not reach the predictor's tables");
static_assert(sizeof(PoissonSmem) == sizeof(FaceSmem<16>), "the fused kernel's LDS is the predictor's");

// One workgroup per query: description; PoissonSelector backbone and Linear
// head (the batched kernels' device code, so every value is the separate
// kernels'); softmax and face draw (word 0); PoissonPredictor on the face;
// output head, tile draw (word 1), locate (dnn_solver_poisson.cpp:86-111).
// FEAT_IN: the features come from the selector backbone kernel (a.feat_in),
// so the workgroup starts at the head: for large batches the backbone runs
// as its own kernel (26 KB of LDS, six workgroups per CU) and the head, the
// face draw and the predictor share one launch, without the head kernel's
// launch and its wait.
template <int NTH, bool FEAT_IN>
__global__ __launch_bounds__(NTH) void poisson_query_kernel(FaceQueryArgs a, SelNetDev sw, FaceNetDev pw)
{
    __shared__ PoissonSmem s;
    FaceSmem<16>& f = s.f;
    const int b = blockIdx.x;
    if (threadIdx.x < 64) vox_desc_build(f.desc, a.layers, a.cuboids, b);
    else if (!FEAT_IN)
        for (int i = threadIdx.x - 64; i < kSelGeluTab; i += NTH - 64) ((float2*)s.sel.koff)[i] = sw.gtab[i];
    for (int i = threadIdx.x; i < kGeluTabEntries; i += NTH) f.gtab[i] = pw.gtab[i];
    if (threadIdx.x < 4) f.zero[threadIdx.x] = 0.f;
    if constexpr (FEAT_IN) {
        for (int i = threadIdx.x; i < 128; i += NTH) s.head.x[i] = (half_t)a.feat_in[(size_t)b * 128 + i];
    } else {
        __syncthreads();
        selector_backbone_body<false, NTH>(f.desc, s.sel.cube, s.sel.koff, false, 2, sw, s.sel.feat);
        for (int i = threadIdx.x; i < 128; i += NTH) s.head.x[i] = (half_t)s.sel.feat[i];
    }
    __syncthreads();
    linear_row<128, 64, true, NTH>(s.head.x, s.head.y1, nullptr, sw.frag[3], sw.bias[3]);
    __syncthreads();
    linear_row<64, 6, false, NTH>(s.head.y1, nullptr, s.head.logits, sw.frag[4], sw.bias[4]);
    __syncthreads();
    const uint64_t qid = a.qid ? a.qid[b] : (a.ctr_dev ? *a.ctr_dev : a.counter) + (uint64_t)b;
    if (threadIdx.x == 0) {
        float p[6];
        const int face = selector_probs_draw(s.head.logits, a.seed, qid, p);
        f.ibuf[0] = face;
        if (a.probs_out)
            for (int i = 0; i < 6; i++) a.probs_out[(size_t)b * 6 + i] = p[i];
        if (a.face_out) a.face_out[b] = face;
    }
    __syncthreads();
    const int face = f.ibuf[0];
    if (face < 0) {  // no drawable face: a NaN result, for the FDM fallback
        if (threadIdx.x < 3) a.out[(size_t)b * 3 + threadIdx.x] = __builtin_nanf("");
        if (threadIdx.x == 0) fallback_append(a, b);
        return;
    }
    face_solver<NTH, 0>(f, pw, face, 2);
    __syncthreads();
    float* kern = face_output_kernel<NTH, 0>(f, face, 2);
    if (a.kern_out) {
        float* o = a.kern_out + (size_t)b * kNN;
        for (int p = threadIdx.x; p < kNN; p += NTH) o[p] = kern[p];
    }
    if (threadIdx.x < 64) {
        uint32_t u[4];
        query_uniforms(a.seed, qid, 0u, u);
        const int t = wave_cdf_draw<float>(kern, kNN, u[1]);
        if (threadIdx.x == 0) {
            float r[5];
            face_result<0>(kern, f.desc, t, face, 2, 0.f, r);
            bool nan = false;
            for (int j = 0; j < 3; j++) {
                a.out[(size_t)b * 3 + j] = r[j];
                nan |= isnan(r[j]);
            }
            if (nan) fallback_append(a, b);
        }
    }
}

int poisson_query_launch(const FaceQueryArgs& a, const SelNetDev& sel, const FaceNetDev& pred, int threads,
                         hipStream_t stream)
{
    if (a.B <= 0) return 0;
    if (a.full) return set_error(-2, "poisson_query_launch: sample mode only");
    if (threads != 256 && threads != 512) return set_error(-2, "poisson_query_launch: threads must be 256 or 512");
    if (a.feat_in) {
        if (threads == 512)
            hipLaunchKernelGGL((poisson_query_kernel<512, true>), dim3(a.B), dim3(512), 0, stream, a, sel, pred);
        else
            hipLaunchKernelGGL((poisson_query_kernel<256, true>), dim3(a.B), dim3(256), 0, stream, a, sel, pred);
    } else {
        if (threads == 512)
            hipLaunchKernelGGL((poisson_query_kernel<512, false>), dim3(a.B), dim3(512), 0, stream, a, sel, pred);
        else
            hipLaunchKernelGGL((poisson_query_kernel<256, false>), dim3(a.B), dim3(256), 0, stream, a, sel, pred);
    }
    DRW_HIP(hipGetLastError());
    return 0;
}

int face_query_launch(int mode, const FaceQueryArgs& a, const FaceNetDev& w, hipStream_t stream)
{
    const int items = a.full ? a.B * 6 : a.B;
    if (items <= 0) return 0;
    // threads per (query, face): the launch shape of this network's kernel
    const bool wide = launch_shape(mode, items) == 512;
    if (mode == 0) {
        if (wide) hipLaunchKernelGGL((face_query_kernel<0, 512>), dim3(items), dim3(512), 0, stream, a, w);
        else hipLaunchKernelGGL((face_query_kernel<0, 256>), dim3(items), dim3(256), 0, stream, a, w);
    } else if (mode == 1) {
        if (wide) hipLaunchKernelGGL((face_query_kernel<1, 512>), dim3(items), dim3(512), 0, stream, a, w);
        else hipLaunchKernelGGL((face_query_kernel<1, 256>), dim3(items), dim3(256), 0, stream, a, w);
    } else {
        if (wide) hipLaunchKernelGGL((face_query_kernel<2, 512>), dim3(items), dim3(512), 0, stream, a, w);
        else hipLaunchKernelGGL((face_query_kernel<2, 256>), dim3(items), dim3(256), 0, stream, a, w);
    }
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw
