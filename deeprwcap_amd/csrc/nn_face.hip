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
#include "common.h"
#include "nn_common.h"
#include "nn_face.h"
#include "nn_face_body.h"
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
