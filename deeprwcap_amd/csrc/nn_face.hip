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
// outputs run on v_mfma_f32_16x16x16f16 (fp32 accumulate); depthwise 3x3,
// the small 1x1 layers, batch-norm (folded into weights/bias on the host) and
// GELU run on the VALU in fp32.
#include "common.h"
#include "nn_common.h"
#include "nn_face.h"
#include "sampling.h"

namespace drw {

// -------------------------------------------------------------- layer bodies
// 1x1 conv on MFMA: in[544][CSI] (CIN_PAD channels read) -> out[544][CSO].
template <int CIN_PAD, int COUT, int COUT_PAD, int CSI, int CSO>
__device__ void pw_mfma(const half_t* in, half_t* out, const half4_t* frag, const float* bias)
{
    constexpr int KT = CIN_PAD / 16, NT = COUT_PAD / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    half4_t bw[NT][KT];
    float bb[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) {
#pragma unroll
        for (int k = 0; k < KT; k++) bw[n][k] = frag[(n * KT + k) * 64 + lane];
        bb[n] = bias[n * 16 + (lane & 15)];
    }
    const int r = lane & 15, kq = 4 * (lane >> 4);
    for (int m = wave; m < kPix / 16; m += 4) {
        half4_t a[KT];
#pragma unroll
        for (int k = 0; k < KT; k++) a[k] = *(const half4_t*)(in + (m * 16 + r) * CSI + k * 16 + kq);
#pragma unroll
        for (int n = 0; n < NT; n++) {
            float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < KT; k++) acc = mfma16(a[k], bw[n][k], acc);
            constexpr int VFULL = 16;
            const int V = COUT - n * 16 < VFULL ? COUT - n * 16 : VFULL;  // valid columns of this tile
            if (V == VFULL) {
                const int ch = n * 16 + (lane & 15);
#pragma unroll
                for (int i = 0; i < 4; i++) out[(m * 16 + kq + i) * CSO + ch] = (half_t)gelu(acc[i] + bb[n]);
            } else {
                // only V of the 16 columns hold outputs: move them across lanes
                // (output o = lane*E + e -> row o / V, column o % V, held by lane
                // col + 16*(row/4) as element row % 4) so each lane evaluates E GELUs
                const int E = V / 4;
                for (int e = 0; e < E; e++) {
                    const int o = lane * E + e, row = o / V, col = o % V;
                    const int src = col + 16 * (row >> 2), el = row & 3;
                    float v = 0.f;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const float t = __shfl(acc[i], src, 64);
                        v = el == i ? t : v;
                    }
                    out[(m * 16 + row) * CSO + n * 16 + col] = (half_t)gelu(v + bias[n * 16 + col]);
                }
            }
        }
    }
}

// Depthwise 3x3 conv (zero padding = dilation) + folded BN + GELU on the
// VALU: in [544][CSI] -> out [544][CSO], channels 0..C-1.  A thread owns one (channel pair,
// column x) and slides down the rows in steps of the dilation, so each
// output reads 3 new input pairs (the other 6 taps are reused from the
// previous step) and the column's x-border mask is computed once.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
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

// The tail layers (8 channels and fewer) keep their activations in fp32: with
// so few channels per output, fp16 rounding of their inputs would dominate
// the kernel's error budget, and their storage is small.
template <int C, int DIL, int CSI, int CSO, typename TI, typename TO>
__device__ void dw_conv(const TI* in, TO* out, const float* w, const float* bias)
{
    constexpr int CP = C / 2;
    for (int item = threadIdx.x; item < CP * kN; item += 256) {
        const int cp = item % CP, x = item / CP;
        float w0[9], w1[9];
#pragma unroll
        for (int k = 0; k < 9; k++) { w0[k] = w[(2 * cp) * 9 + k]; w1[k] = w[(2 * cp + 1) * 9 + k]; }
        const float b0 = bias[2 * cp], b1 = bias[2 * cp + 1];
        const bool okl = x - DIL >= 0, okr = x + DIL < kN;
        // three input columns (x-d, x, x+d) of one row, both channels
        auto load_row = [&](int y, float* v) {
#pragma unroll
            for (int q = 0; q < 6; q++) v[q] = 0.f;
            if (y < 0 || y >= kN) return;
            const TI* s = in + (y * kN + x) * CSI + 2 * cp;
            load_pair(s, v[2], v[3]);
            if (okl) load_pair(s - DIL * CSI, v[0], v[1]);
            if (okr) load_pair(s + DIL * CSI, v[4], v[5]);
        };
#pragma unroll 1
        for (int r0 = 0; r0 < DIL; r0++) {
            float up[6], mid[6], dn[6];
            load_row(r0 - DIL, up);
            load_row(r0, mid);
#pragma unroll 1
            for (int y = r0; y < kN; y += DIL) {
                load_row(y + DIL, dn);
                float a0 = b0, a1 = b1;
#pragma unroll
                for (int kx = 0; kx < 3; kx++) {
                    a0 += w0[kx] * up[2 * kx] + w0[3 + kx] * mid[2 * kx] + w0[6 + kx] * dn[2 * kx];
                    a1 += w1[kx] * up[2 * kx + 1] + w1[3 + kx] * mid[2 * kx + 1] + w1[6 + kx] * dn[2 * kx + 1];
                }
                store_pair(out + (y * kN + x) * CSO + 2 * cp, gelu(a0), gelu(a1));
#pragma unroll
                for (int q = 0; q < 6; q++) { up[q] = mid[q]; mid[q] = dn[q]; }
            }
        }
    }
}

// Small 1x1 conv + folded BN + GELU on the VALU (CIN <= 8, COUT <= 4), fp32
// activations.
template <int CIN, int COUT, int CSI, int CSO>
__device__ void pw_valu(const float* in, float* out, const float* w, const float* bias)
{
    for (int p = threadIdx.x; p < kNN; p += 256) {
        float v[CIN];
#pragma unroll
        for (int c = 0; c < CIN; c++) v[c] = (float)in[p * CSI + c];
#pragma unroll
        for (int o = 0; o < COUT; o++) {
            float a = bias[o];
#pragma unroll
            for (int c = 0; c < CIN; c++) a += w[o * CIN + c] * v[c];
            out[p * CSO + o] = gelu(a);
        }
    }
}

// Head 1x1 conv CIN -> 1 with bias, no activation: fp32 per pixel.
template <int CIN, int CSI>
__device__ void head_conv(const float* in, float* hout, const float* w, float b)
{
    for (int p = threadIdx.x; p < kNN; p += 256) {
        float a = b;
#pragma unroll
        for (int c = 0; c < CIN; c++) a += w[c] * in[p * CSI + c];
        hout[p] = a;
    }
}

// ------------------------------------------------------------ architectures
// Buffer A [544][32] holds the 25-channel input (+7 zero channels); layers
// alternate A <-> B.
template <int BCS>
struct FaceSmem {
    alignas(16) half_t A[kPix * 32];  // 16-byte aligned: written as half8, read as half4 / float2
    alignas(16) half_t B[kPix * BCS];
    VoxDesc desc;
    float red[8];
    int ibuf[4];
};

__device__ __forceinline__ void face_solver(FaceSmem<16>& s, const FaceNetDev& w, float* hout)
{
    pw_mfma<32, 16, 16, 32, 16>(s.A, s.B, w.frag[0], w.pw_bias[0]);
    __syncthreads();
    dw_conv<16, 1, 16, 16>(s.B, s.A, w.dw_w[0], w.dw_b[0]); __syncthreads();
    pw_mfma<16, 16, 16, 16, 16>(s.A, s.B, w.frag[1], w.pw_bias[1]); __syncthreads();
    dw_conv<16, 1, 16, 16>(s.B, s.A, w.dw_w[1], w.dw_b[1]); __syncthreads();
    pw_mfma<16, 8, 16, 16, 16>(s.A, s.B, w.frag[2], w.pw_bias[2]); __syncthreads();
    // fp32 tail: A as float[544][8], B as float[544][4]
    float* fA = (float*)s.A;
    float* fB = (float*)s.B;
    dw_conv<8, 2, 16, 8>(s.B, fA, w.dw_w[2], w.dw_b[2]); __syncthreads();
    pw_valu<8, 4, 8, 4>(fA, fB, w.pw_w[3], w.pw_bias[3]); __syncthreads();
    dw_conv<4, 3, 4, 4>(fB, fA, w.dw_w[3], w.dw_b[3]); __syncthreads();
    pw_valu<4, 2, 4, 2>(fA, fB, w.pw_w[4], w.pw_bias[4]); __syncthreads();
    head_conv<2, 2>(fB, hout, w.head_w, w.head_b);
}

__device__ __forceinline__ void face_solver_plus(FaceSmem<32>& s, const FaceNetDev& w, float* hout)
{
    pw_mfma<32, 32, 32, 32, 32>(s.A, s.B, w.frag[0], w.pw_bias[0]); __syncthreads();
    dw_conv<32, 1, 32, 32>(s.B, s.A, w.dw_w[0], w.dw_b[0]); __syncthreads();
    pw_mfma<32, 32, 32, 32, 32>(s.A, s.B, w.frag[1], w.pw_bias[1]); __syncthreads();
    dw_conv<32, 1, 32, 32>(s.B, s.A, w.dw_w[1], w.dw_b[1]); __syncthreads();
    pw_mfma<32, 24, 32, 32, 32>(s.A, s.B, w.frag[2], w.pw_bias[2]); __syncthreads();
    dw_conv<24, 1, 32, 32>(s.B, s.A, w.dw_w[2], w.dw_b[2]); __syncthreads();
    pw_mfma<32, 24, 32, 32, 32>(s.A, s.B, w.frag[3], w.pw_bias[3]); __syncthreads();
    dw_conv<24, 1, 32, 32>(s.B, s.A, w.dw_w[3], w.dw_b[3]); __syncthreads();
    pw_mfma<32, 16, 16, 32, 32>(s.A, s.B, w.frag[4], w.pw_bias[4]); __syncthreads();
    dw_conv<16, 2, 32, 32>(s.B, s.A, w.dw_w[4], w.dw_b[4]); __syncthreads();
    pw_mfma<16, 16, 16, 32, 32>(s.A, s.B, w.frag[5], w.pw_bias[5]); __syncthreads();
    dw_conv<16, 2, 32, 32>(s.B, s.A, w.dw_w[5], w.dw_b[5]); __syncthreads();
    pw_mfma<16, 8, 16, 32, 32>(s.A, s.B, w.frag[6], w.pw_bias[6]); __syncthreads();
    float* fA = (float*)s.A;
    float* fB = (float*)s.B;
    dw_conv<8, 3, 32, 8>(s.B, fA, w.dw_w[6], w.dw_b[6]); __syncthreads();
    pw_valu<8, 4, 8, 4>(fA, fB, w.pw_w[7], w.pw_bias[7]); __syncthreads();
    head_conv<4, 4>(fB, hout, w.head_w, w.head_b);
}

// ------------------------------------------------------------ input tensor
// Network input T[c][h][w] (23 channels + 2 position channels) written to
// A[p = 23h + w][c]; mode 0: rotate_poisson_f(cube); mode 1:
// rotate_gradient_f(axis_rotate(cube)); mode 2: the same, permuted so that
// channels are y (SignedFacePredictor channel_dim = 1).
__device__ void build_input(half_t* A, const VoxDesc& d, int mode, int face, int axis)
{
    // Source voxel of input element (c, h, w): every remap is a permutation
    // with flips, so along the 23 channels of a pixel exactly one source axis
    // varies (by +-1 per channel); the cuboid masks of the two fixed axes are
    // AND-ed once per pixel.
    auto src = [&](int c, int h, int w, int& z, int& y, int& x) {
        if (mode == 2) { z = h; y = c; x = w; } else { z = c; y = h; x = w; }
        if (mode == 0) {
            remap_poisson(face, z, y, x);
        } else {
            remap_gradient(face, z, y, x);
            remap_axis(axis, z, y, x);
        }
    };
    typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
    for (int it = threadIdx.x; it < kPix * 4; it += 256) {
        const int p = it >> 2, g = (it & 3) * 8;
        half8_t pk;
        if (p >= kNN) {
#pragma unroll
            for (int j = 0; j < 8; j++) pk[j] = (half_t)0.f;
        } else {
            const int h = p / kN, w = p - h * kN;
            int z0, y0, x0, z1, y1, x1;
            src(0, h, w, z0, y0, x0);
            src(1, h, w, z1, y1, x1);
            const uint16_t* vm;
            int base, dir;
            uint32_t fixedm;
            const bool zvar = z1 != z0;
            if (zvar) { vm = d.bz; base = z0; dir = z1 - z0; fixedm = (uint32_t)d.by[y0] & d.bx[x0]; }
            else if (y1 != y0) { vm = d.by; base = y0; dir = y1 - y0; fixedm = (uint32_t)d.bz[z0] & d.bx[x0]; }
            else { vm = d.bx; base = x0; dir = x1 - x0; fixedm = (uint32_t)d.bz[z0] & d.by[y0]; }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int c = g + j;
                float val = 0.f;
                if (c < kN) {
                    const int k = base + dir * c;
                    const uint32_t m = fixedm & vm[k];
                    val = m ? d.cv[__builtin_ctz(m)] : d.lz[zvar ? k : z0];
                } else if (c == kN) {
                    val = (float)w / (float)(kN - 1);
                } else if (c == kN + 1) {
                    val = (float)h / (float)(kN - 1);
                }
                pk[j] = (half_t)val;
            }
        }
        *(half8_t*)(A + p * 32 + g) = pk;
    }
}

__device__ __forceinline__ float block_sum256(float v, float* red)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const float s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

// ------------------------------------------------------------------ kernel
// mode 0: Poisson predictor (relu head); 1: Gradient1 (norm head, faces 0,1);
// 2: Gradient2 (FaceSolverPlus, norm head, faces 2..5).
template <int MODE>
__global__ __launch_bounds__(256) void face_query_kernel(FaceQueryArgs a, FaceNetDev w)
{
    constexpr int BCS = MODE == 2 ? 32 : 16;
    __shared__ FaceSmem<BCS> s;
    const int item = blockIdx.x;
    const int b = a.full ? item / 6 : item;
    const int face = a.full ? item % 6 : a.face[b];
    const int axis = a.axis ? a.axis[b] : 2;
    if (MODE == 1 && !(face == 0 || face == 1)) return;
    if (MODE == 2 && !(face >= 2 && face <= 5)) return;
    if (threadIdx.x < 64) vox_desc_build(s.desc, a.layers, a.cuboids, b);
    __syncthreads();
    build_input(s.A, s.desc, MODE == 0 ? 0 : (MODE == 1 ? 1 : 2), face, axis);
    __syncthreads();
    // head output aliases buffer A (free once the first layer has consumed it)
    float* hout = (float*)s.A;
    if constexpr (MODE == 2) face_solver_plus(s, w, hout); else face_solver(s, w, hout);
    __syncthreads();
    // ---- output head: relu + 1e-10 then /sum, or x / sum|x| (architecture.py:122-132)
    float loc = 0.f;
    for (int p = threadIdx.x; p < kNN; p += 256) {
        const float v = hout[p];
        loc += MODE == 0 ? (fmaxf(v, 0.f) + 1e-10f) : fabsf(v);
    }
    const float tot = block_sum256(loc, s.red);
    // normalised kernel -> buffer B region as fp32 [529], post-processed for gradients
    float* kern = (float*)s.B;
    for (int p = threadIdx.x; p < kNN; p += 256) {
        float v;
        if (MODE == 0) {
            v = (fmaxf(hout[p], 0.f) + 1e-10f) / tot;
        } else {
            // post_process_gradient_kernel (cuda_kernels.cu:516-578)
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
    if (a.full) {
        float* o = a.kern_out + (size_t)item * kNN;
        for (int p = threadIdx.x; p < kNN; p += 256) o[p] = kern[p];
        return;
    }
    if (a.kern_out) {
        float* o = a.kern_out + (size_t)b * kNN;
        for (int p = threadIdx.x; p < kNN; p += 256) o[p] = kern[p];
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
            if (t < 0) {
                for (int j = 0; j < width; j++) o[j] = __builtin_nanf("");
            } else {
                float pt[3];
                panel_point(kN, t, MODE == 0 ? face : axis_face_remap(face, axis), pt);
                o[0] = pt[0]; o[1] = pt[1]; o[2] = pt[2];
                if (MODE != 0) {
                    const float g = kern[t];
                    const float sg = g > 0.f ? 1.f : (g < 0.f ? -1.f : 0.f);
                    o[3] = a.weight[b] * sg;
                    // centre of the axis-rotated cube times the max (dnn_solver_gradient.cpp:260-266)
                    int z = kN / 2, y = kN / 2, x = kN / 2;
                    remap_axis(axis, z, y, x);
                    o[4] = vox_value(s.desc, z, y, x) * s.desc.maxv;
                }
            }
        }
    }
}

int face_query_launch(int mode, const FaceQueryArgs& a, const FaceNetDev& w, hipStream_t stream)
{
    const int items = a.full ? a.B * 6 : a.B;
    if (items <= 0) return 0;
    if (mode == 0) hipLaunchKernelGGL(face_query_kernel<0>, dim3(items), dim3(256), 0, stream, a, w);
    else if (mode == 1) hipLaunchKernelGGL(face_query_kernel<1>, dim3(items), dim3(256), 0, stream, a, w);
    else hipLaunchKernelGGL(face_query_kernel<2>, dim3(items), dim3(256), 0, stream, a, w);
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw
