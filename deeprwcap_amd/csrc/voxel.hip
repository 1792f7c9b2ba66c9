// Voxelization of the layer/cuboid description and the flat (FDM-table)
// next-hop sampler.
#include "common.h"
#include "kernels.h"
#include "sampling.h"
#include "voxel_fn.h"

namespace drw {

// One workgroup per query; the 135-float description is staged in LDS.
__global__ __launch_bounds__(256) void voxelize_kernel(VoxelArgs a)
{
    __shared__ float Ly[kMaxLayers * 2], Cb[kMaxCuboids * 7], lv[kMaxLayers], cv[kMaxCuboids];
    __shared__ float smax;
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < kMaxLayers * 2) Ly[tid] = a.layers[(size_t)b * kMaxLayers * 2 + tid];
    if (tid < kMaxCuboids * 7) Cb[tid] = a.cuboids[(size_t)b * kMaxCuboids * 7 + tid];
    __syncthreads();
    if (tid == 0) {
        const float m = query_max(Ly, Cb);
        smax = m;
        if (a.maxv) a.maxv[b] = m;
    }
    __syncthreads();
    const float m = smax;
    if (tid < kMaxLayers) lv[tid] = a.normalise ? Ly[2 * tid + 1] / m : Ly[2 * tid + 1];
    if (tid < kMaxCuboids) cv[tid] = a.normalise ? Cb[7 * tid + 6] / m : Cb[7 * tid + 6];
    __syncthreads();
    const int N = a.N ? a.N[b] : a.fixed_N;
    const int NN = N * N, NNN = NN * N;
    float* out = a.out + (a.out_off ? a.out_off[b] : (int64_t)b * NNN);
    for (int v = tid; v < NNN; v += blockDim.x) {
        const int x = v % N, y = (v / N) % N, z = v / NN;
        const float xc = (2.0f * (float)x + 1.0f) / (float)N - 1.0f;
        const float yc = (2.0f * (float)y + 1.0f) / (float)N - 1.0f;
        const float zc = (2.0f * (float)z + 1.0f) / (float)N - 1.0f;
        out[v] = voxel_value(Ly, lv, Cb, cv, xc, yc, zc);
    }
}

int voxelize_launch(const VoxelArgs& a, hipStream_t stream)
{
    if (a.B <= 0) return 0;
    hipLaunchKernelGGL(voxelize_kernel, dim3(a.B), dim3(256), 0, stream, a);
    DRW_HIP(hipGetLastError());
    return 0;
}

// Flat inverse-CDF next hop over the 6 N^2 FDM kernel of each query
// (flat_draw_wave), one wave per query.
__global__ __launch_bounds__(256) void sample_flat_kernel(FlatSampleArgs a)
{
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= a.B) return;
    const uint64_t base = a.ctr_dev ? *a.ctr_dev : a.counter;
    const bool grad = a.axis != nullptr;
    flat_draw_wave<true>(a.G + a.G_off[b], a.N[b], grad, grad && a.center_eps ? a.center_eps[b] : 0.0f, a.seed,
                   a.qid ? a.qid[b] : base + (uint64_t)b, a.bad && a.bad[b], a.out + (size_t)b * (grad ? 5 : 3));
}

int sample_flat_launch(const FlatSampleArgs& a, hipStream_t stream)
{
    if (a.B <= 0) return 0;
    hipLaunchKernelGGL(sample_flat_kernel, dim3((a.B + 3) / 4), dim3(256), 0, stream, a);
    DRW_HIP(hipGetLastError());
    return 0;
}

// Centre permittivity of each raw cube (ggft_solver.cpp:649-671): odd N the
// centre cell; even N the mean of the 4 cells on the minus side of the
// gradient axis (axis 0 -> WVTX, 1 -> WVTY, 2 -> WVTZ).
__global__ void center_eps_kernel(int B, const int32_t* Nq, const int32_t* axis, const float* eps,
                                  const int64_t* eps_off, float* out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int N = Nq[b], H = N / 2, NN = N * N;
    const float* E = eps + eps_off[b];
    auto C = [&](int i, int j, int k) { return (double)E[i + j * N + k * NN]; };
    double v;
    if (N % 2) {
        v = C(H, H, H);
    } else {
        const int ax = axis ? axis[b] : 2;
        if (ax == 0) v = (C(H - 1, H, H) + C(H - 1, H - 1, H) + C(H - 1, H, H - 1) + C(H - 1, H - 1, H - 1)) / 4;
        else if (ax == 1) v = (C(H, H - 1, H) + C(H - 1, H - 1, H) + C(H, H - 1, H - 1) + C(H - 1, H - 1, H - 1)) / 4;
        else v = (C(H, H, H - 1) + C(H - 1, H - 1, H - 1) + C(H, H - 1, H - 1) + C(H - 1, H, H - 1)) / 4;
    }
    out[b] = (float)v;
}

int center_eps_launch(int B, const int32_t* N, const int32_t* axis, const float* eps, const int64_t* eps_off,
                      float* out, hipStream_t stream)
{
    if (B <= 0) return 0;
    hipLaunchKernelGGL(center_eps_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, B, N, axis, eps, eps_off, out);
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw

namespace drw {

// Per-query solve metadata for a uniform-N batch: resolution, right-hand
// side (GFT, or the WVT code of the gradient axis) and packed offsets.
__global__ void fill_meta_kernel(int B, int N, const int32_t* axis, int32_t* Nq, int32_t* code, int64_t* eps_off,
                                 int64_t* G_off, uint8_t* bad, int32_t* count)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0 && count) *count = 0;
    if (b >= B) return;
    if (bad) bad[b] = 0;
    Nq[b] = N;
    int c = kGFT;
    if (axis) {
        const int ax = axis[b];
        c = ax == 0 ? kWVTX : ax == 1 ? kWVTY : kWVTZ;
    }
    code[b] = c;
    eps_off[b] = (int64_t)b * N * N * N;
    G_off[b] = (int64_t)b * 6 * N * N;
}

int fill_meta_launch(int B, int N, const int32_t* axis, int32_t* Nq, int32_t* code, int64_t* eps_off, int64_t* G_off,
                     uint8_t* bad, int32_t* count, hipStream_t stream)
{
    if (B <= 0) return 0;
    hipLaunchKernelGGL(fill_meta_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, B, N, axis, Nq, code, eps_off,
                       G_off, bad, count);
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw

namespace drw {

// Advances a device-resident query counter (the last node of a captured step).
__global__ void advance_counter_kernel(uint64_t* c, uint64_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *c += n;
}

int advance_counter_launch(uint64_t* ctr, uint64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(advance_counter_kernel, dim3(1), dim3(64), 0, stream, ctr, n);
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw
