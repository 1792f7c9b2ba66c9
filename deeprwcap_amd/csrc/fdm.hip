// Batched FDM surface Green's function on MI355X (gfx950).
//
// Solves, per cube, L y = e with the 7-point harmonic-mean operator
//   (L y)_c = sum_n w_cn (y_c - y_n) + 2 eps_c nb(c) y_c,  w = 2 eps_c eps_n/(eps_c+eps_n)
// and writes G_b = 2 eps_c y_c on the 6 N^2 boundary panels, which is what the
// reference GGFT solver computes as gft = A13 x with A = L D^-1, x = D y
// (ggft/src/ggft_solver.cpp:447-518, 552-596).  The right-hand sides follow
// update_ek (ggft_solver.cpp:137-274) for odd and even N.
//
// Design (one persistent 1024-thread workgroup per CU, one cube at a time):
//  * cells c = tid + q*1024 (q < 12 at N = 23); the x and y face couplings
//    of each cell, symmetrically Jacobi-scaled (w_cn / sqrt(d_c d_n)), live
//    in VGPRs for the whole solve; the z couplings live in LDS (the z- one of
//    cell c is the z+ one of c - N^2), next to the search direction p
//    (N^3 + 2 N^2 floats with zero guard planes, so the stencil reads need no
//    bounds checks: x/y wrap-around reads hit a zero weight); A p is parked
//    in a third LDS array across the reduction so the per-thread state stays
//    at 4 couplings + x + r per cell: the CG iteration loop runs without
//    spills at 12 cells per thread (the set-up and FP64 residual passes
//    spill a few registers to scratch).  The Chronopoulos-Gear form, one
//    fused reduction per iteration, needs x, p and A p per cell besides r
//    and spills inside the loop at 12 cells per thread, with no LDS left to
//    park a vector in;
//  * inner solver: FP32 CG on the scaled system (= Jacobi PCG), block-wide
//    reductions by wave64 butterflies + a 16-entry LDS table, summed in a
//    fixed order so every thread sees bit-identical alpha/beta;
//  * outer loop: FP64 iterative refinement.  The residual e - L y is formed
//    in FP64 with exact harmonic means, y accumulates in FP64, and the loop
//    stops at ||e - L y|| <= tol ||e|| (the reference's relative-residual
//    criterion, ggft_solver.h:123-126, 193-198);
//  * work distribution: a self-resetting atomic ticket counter hands cubes to
//    workgroups in the (optional) host-given order, e.g. largest N first.
#include "common.h"
#include "fdm.h"

namespace drw {

// Two size classes.  Cubes up to N = 23 take a 1024-thread workgroup and
// 156 KB of LDS, one per CU; cubes up to N = 16 take 256 threads and 53 KB,
// so three share a CU and each hides the others' barrier and LDS latency
// (one cube per CU leaves most of it waiting at the block reductions).
template <int T, int MAXN, int PER_CU>
struct FdmSharedT {
    static constexpr int kThreads = T, kWaves = T / 64, kMaxN = MAXN, kPerCU = PER_CU;
    static constexpr int kGuard = MAXN * MAXN, kCells = MAXN * MAXN * MAXN;
    float p[kCells + 2 * kGuard + T];  // slack covers the unused tail slots
    float wz[kCells + kGuard];         // scaled z+ couplings, zero guard plane below
    float ap[kCells];                  // A p of the current iteration (own cells only)
    float red[2][kWaves];
    double redd[kWaves];
    double rhs_val[8];
    int rhs_idx[8];
    double enorm;
    int ticket;
};
using FdmLarge = FdmSharedT<1024, kMaxN, 1>;
using FdmSmall = FdmSharedT<256, 16, 3>;  // registers capped so that three fit a CU

typedef __attribute__((address_space(3))) float lds_float;

// An LDS pointer the optimiser cannot fold into absolute addresses: all
// stencil reads then become ds_read base+immediate from one VGPR per array.
__device__ __forceinline__ lds_float* lds_base(float* p)
{
    lds_float* q = (lds_float*)p;
    asm volatile("" : "+v"(q));
    return q;
}

template <int W>
__device__ __forceinline__ float block_sum(float v, float* red)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < W; w++) s += red[w];
    return s;
}

template <int W>
__device__ __forceinline__ double block_sum_d(double v, double* red)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < W; w++) s += red[w];
    __syncthreads();  // table free for the next call
    return s;
}

// Harmonic-mean face permittivity 2ab/(a+b).  The FP32 form feeds only the
// inner solve (its rounding is corrected by the FP64 refinement); the FP64
// form uses the hardware reciprocal refined by two Newton steps (~1 ulp).
__device__ __forceinline__ float harm(float a, float b) { return __fdividef(2.0f * a * b, a + b); }
__device__ __forceinline__ double harm(double a, double b)
{
    const double t = a + b;
    double r = __builtin_amdgcn_rcp(t);
    r = fma(fma(-t, r, 1.0), r, r);
    r = fma(fma(-t, r, 1.0), r, r);
    return 2.0 * a * b * r;
}

// Right-hand side cells and values (ggft_solver.cpp:137-274); unused slots -1.
template <int N>
__device__ void fdm_rhs(const float* E, int code, int* idx, double* val)
{
    constexpr int NN = N * N, H = N / 2;
    const double h = 2.0 / N;
    auto C = [](int i, int j, int k) { return i + j * N + k * NN; };
    for (int q = 0; q < 8; q++) { idx[q] = -1; val[q] = 0.0; }
    if (N % 2) {
        const int zzz = C(H, H, H);
        if (code == kGFT) { idx[0] = zzz; val[0] = 1.0; return; }
        int p, m;
        if (code == kWVTZ) { p = C(H, H, H + 1); m = C(H, H, H - 1); }
        else if (code == kWVTX) { p = C(H + 1, H, H); m = C(H - 1, H, H); }
        else { p = C(H, H + 1, H); m = C(H, H - 1, H); }
        const double dp = (double)E[p], dz = (double)E[zzz], dm = (double)E[m];
        const double d_plus = dp + dz, d_minus = -(dz + dm);
        idx[0] = p;   val[0] = dp / d_plus / h;
        idx[1] = zzz; val[1] = dz * (1. / d_plus + 1. / d_minus) / h;
        idx[2] = m;   val[2] = dm / d_minus / h;
        return;
    }
    const int zzz = C(H, H, H), mzz = C(H - 1, H, H), zmz = C(H, H - 1, H), zzm = C(H, H, H - 1),
              mmz = C(H - 1, H - 1, H), mzm = C(H - 1, H, H - 1), zmm = C(H, H - 1, H - 1),
              mmm = C(H - 1, H - 1, H - 1);
    auto D = [&](int c) { return (double)E[c]; };
    const double dsum = D(zzz) + D(zmz) + D(zzm) + D(zmm) + D(mzz) + D(mmz) + D(mzm) + D(mmm);
    if (code == kGFT) {
        const int cells[8] = {zzz, mzz, zmz, zzm, mmz, mzm, zmm, mmm};
        for (int q = 0; q < 8; q++) { idx[q] = cells[q]; val[q] = D(cells[q]) / dsum; }
        return;
    }
    int plus[4], minus[4];
    if (code == kWVTZ) {
        const int p4[4] = {zzz, mzz, zmz, mmz}, m4[4] = {zzm, mzm, zmm, mmm};
        for (int q = 0; q < 4; q++) { plus[q] = p4[q]; minus[q] = m4[q]; }
    } else if (code == kWVTX) {
        const int p4[4] = {zzz, zmz, zzm, zmm}, m4[4] = {mzz, mmz, mzm, mmm};
        for (int q = 0; q < 4; q++) { plus[q] = p4[q]; minus[q] = m4[q]; }
    } else {
        const int p4[4] = {zzz, mzz, zzm, mzm}, m4[4] = {zmz, mmz, zmm, mmm};
        for (int q = 0; q < 4; q++) { plus[q] = p4[q]; minus[q] = m4[q]; }
    }
    const double d_minus = -(D(minus[0]) + D(minus[1]) + D(minus[2]) + D(minus[3]));
    for (int q = 0; q < 4; q++) { idx[q] = plus[q]; val[q] = D(plus[q]) / dsum / h * 2; }
    for (int q = 0; q < 4; q++) { idx[4 + q] = minus[q]; val[4 + q] = D(minus[q]) * (1. / dsum + 1. / d_minus) / h * 2; }
}

template <int N, typename SH>
__device__ void fdm_cube(const FdmBatch& a, SH& sh, int b)
{
    static_assert(N <= SH::kMaxN, "cube larger than the size class");
    constexpr int NN = N * N, NNN = N * NN;
    constexpr int T = SH::kThreads, W = SH::kWaves;
    constexpr int CPT = (NNN + T - 1) / T;
    const int tid = threadIdx.x;
    const float* E = a.eps + a.eps_off[b];
    // workspace slot of this workgroup (fdm_launch sizes the grids to fit)
    double* Y = a.y_ws + (size_t)blockIdx.x * SH::kCells;
    float* S = a.s_ws + (size_t)blockIdx.x * SH::kCells;
    const int code = a.code[b];
    float* P = sh.p + NN;  // valid for indices [-NN, NNN + NN)

    // ---- stage eps (in the A p array), clear guards and y, build the source ----
    float* WZ = sh.wz + NN;  // valid for indices [-NN, NNN)
    float* EL = sh.ap;       // eps staging during set-up and residual passes
    for (int c = tid; c < NNN; c += T) { EL[c] = E[c]; Y[c] = 0.0; }
    for (int g = tid; g < NN; g += T) { P[g - NN] = 0.f; P[NNN + g] = 0.f; WZ[g - NN] = 0.f; }
    if (tid == 0) {
        fdm_rhs<N>(E, code, sh.rhs_idx, sh.rhs_val);
        double s = 0.0;
        for (int q = 0; q < 8; q++) s += sh.rhs_val[q] * sh.rhs_val[q];
        sh.enorm = sqrt(s);
    }
    __syncthreads();

    // ---- Jacobi scaling s_c = d_c^-1/2, d_c = sum_n w_cn + 2 eps_c nb(c) -----
#pragma unroll 1
    for (int c = tid; c < NNN; c += T) {
        const int i = c % N, j = (c / N) % N, k = c / NN;
        const float ec = EL[c];
        float d = 0.f;
        d += i > 0 ? harm(ec, EL[c - 1]) : 2.0f * ec;
        d += i < N - 1 ? harm(ec, EL[c + 1]) : 2.0f * ec;
        d += j > 0 ? harm(ec, EL[c - N]) : 2.0f * ec;
        d += j < N - 1 ? harm(ec, EL[c + N]) : 2.0f * ec;
        d += k > 0 ? harm(ec, EL[c - NN]) : 2.0f * ec;
        d += k < N - 1 ? harm(ec, EL[c + NN]) : 2.0f * ec;
        const float sc = 1.0f / sqrtf(d);
        P[c] = sc;
        S[c] = sc;
    }
    __syncthreads();

    // ---- scaled couplings: x/y in VGPRs, z+ in LDS ---------------------------
    float w[4][CPT];  // x-, x+, y-, y+
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = tid + q * T;
        if (q < CPT - 1 || c < NNN) {
            const int i = c % N, j = (c / N) % N, k = c / NN;
            const float ec = EL[c], sc = P[c];
            w[0][q] = i > 0 ? harm(ec, EL[c - 1]) * sc * P[c - 1] : 0.f;
            w[1][q] = i < N - 1 ? harm(ec, EL[c + 1]) * sc * P[c + 1] : 0.f;
            w[2][q] = j > 0 ? harm(ec, EL[c - N]) * sc * P[c - N] : 0.f;
            w[3][q] = j < N - 1 ? harm(ec, EL[c + N]) * sc * P[c + N] : 0.f;
            WZ[c] = k < N - 1 ? harm(ec, EL[c + NN]) * sc * P[c + NN] : 0.f;
        } else {
#pragma unroll
            for (int d = 0; d < 4; d++) w[d][q] = 0.f;
        }
    }
    __syncthreads();

    const double enorm = sh.enorm;
    float* R32 = P;  // residual hand-off reuses the p array (p is rebuilt from it)
    double rnorm = enorm;
    int total_it = 0;
    for (int round = 0; round < a.max_rounds; round++) {
        // ---- FP64 residual r = e - L y, handed to the inner solve as FP32 S r ----
        // (one pass per cell, not unrolled, through the A p array)
        if (round > 0) {
            for (int c = tid; c < NNN; c += T) EL[c] = E[c];
            __syncthreads();
        }
        double loc = 0.0;
#pragma unroll 1
        for (int c = tid; c < NNN; c += T) {
            double rv = 0.0;
            if (round > 0) {
                const int i = c % N, j = (c / N) % N, k = c / NN;
                const double ec = EL[c], yc = Y[c];
                double acc = 0.0;
                int nb = 0;
                if (i > 0) acc += harm(ec, (double)EL[c - 1]) * (yc - Y[c - 1]); else nb++;
                if (i < N - 1) acc += harm(ec, (double)EL[c + 1]) * (yc - Y[c + 1]); else nb++;
                if (j > 0) acc += harm(ec, (double)EL[c - N]) * (yc - Y[c - N]); else nb++;
                if (j < N - 1) acc += harm(ec, (double)EL[c + N]) * (yc - Y[c + N]); else nb++;
                if (k > 0) acc += harm(ec, (double)EL[c - NN]) * (yc - Y[c - NN]); else nb++;
                if (k < N - 1) acc += harm(ec, (double)EL[c + NN]) * (yc - Y[c + NN]); else nb++;
                rv = -(acc + 2.0 * ec * nb * yc);
            }
#pragma unroll
            for (int s = 0; s < 8; s++)
                if (sh.rhs_idx[s] == c) rv += sh.rhs_val[s];
            loc += rv * rv;
            R32[c] = (float)rv * S[c];
        }
        rnorm = sqrt(block_sum_d<W>(loc, sh.redd));  // also orders the R32 writes
        if (!(rnorm > a.tol * enorm)) break;      // converged (or enorm == 0)

        // ---- FP32 CG on the scaled system, right-hand side S r / ||r|| -------
        const float inv = (float)(1.0 / rnorm);
        float xh[CPT], rh[CPT];
        float rloc = 0.f;
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int c = tid + q * T;
            xh[q] = 0.f;
            rh[q] = (q < CPT - 1 || c < NNN) ? R32[c] * inv : 0.f;
            rloc += rh[q] * rh[q];
        }
        float rr = block_sum<W>(rloc, sh.red[1]);
        double tgt = 0.3 * a.tol * enorm / rnorm;
        if (tgt < (double)a.inner_floor) tgt = a.inner_floor;
        const float thresh = (float)(tgt * tgt) * rr;
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int c = tid + q * T;
            if (q < CPT - 1 || c < NNN) P[c] = rh[q];
        }
        __syncthreads();
        // Per-thread LDS bases; element q*T + k of each is cell c = tid + q*T
        // shifted by k - N^2 (P, WZ) or by k (AP).
        lds_float* Pb = lds_base(sh.p + tid);
        lds_float* Zb = lds_base(sh.wz + tid);
        lds_float* Ab = lds_base(sh.ap + tid);
        for (int it = 0; it < a.max_inner; it++) {
            float pl = 0.f;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = tid + q * T;
                const int o = q * T + NN;  // P index of cell c relative to Pb
                if (q < CPT - 1 || c < NNN) {
                    const float pc = Pb[o];
                    float v = pc;
                    v -= w[0][q] * Pb[o - 1];
                    v -= w[1][q] * Pb[o + 1];
                    v -= w[2][q] * Pb[o - N];
                    v -= w[3][q] * Pb[o + N];
                    v -= Zb[q * T] * Pb[o - NN];
                    v -= Zb[q * T + NN] * Pb[o + NN];
                    Ab[q * T] = v;
                    pl += pc * v;
                }
            }
            const float pq = block_sum<W>(pl, sh.red[0]);
            const float alpha = rr / pq;
            float rl = 0.f;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = tid + q * T;
                if (q < CPT - 1 || c < NNN) {
                    xh[q] += alpha * Pb[q * T + NN];
                    rh[q] -= alpha * Ab[q * T];
                    rl += rh[q] * rh[q];
                }
            }
            const float rr_new = block_sum<W>(rl, sh.red[1]);
            total_it++;
            if (!(rr_new > thresh)) break;
            const float beta = rr_new / rr;
            rr = rr_new;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = tid + q * T;
                if (q < CPT - 1 || c < NNN) Pb[q * T + NN] = rh[q] + beta * Pb[q * T + NN];
            }
            __syncthreads();
        }
        // ---- y += ||r|| S x -----------------------------------------------------
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int c = tid + q * T;
            if (q < CPT - 1 || c < NNN) Y[c] += (double)S[c] * (double)xh[q] * rnorm;
        }
        __syncthreads();
    }

    // ---- G_b = 2 eps_c y_c, GGFT panel order (ggft_solver.h:65-77) ----------
    double* G = a.G + a.G_off[b];
    for (int pb = tid; pb < 6 * NN; pb += T) {
        const int f = pb / NN, t = pb - f * NN, u = t % N, v = t / N;
        int c;
        switch (f) {
        case 0: c = u + v * N; break;
        case 1: c = u + v * N + (N - 1) * NN; break;
        case 2: c = u + v * NN; break;
        case 3: c = u + (N - 1) * N + v * NN; break;
        case 4: c = u * N + v * NN; break;
        default: c = (N - 1) + u * N + v * NN; break;
        }
        G[pb] = 2.0 * (double)E[c] * Y[c];
    }
    if (tid == 0) {
        if (a.iters) a.iters[b] = total_it;
        if (a.relres) a.relres[b] = enorm > 0 ? (float)(rnorm / enorm) : 0.f;
    }
    __syncthreads();
}

__host__ __device__ constexpr bool fdm_n_ok(int n) { return n == 8 || n == 9 || n == 16 || n == 21 || n == 23; }

// One persistent kernel per size class: workgroups take cubes by atomic
// ticket (ctr[0]) and solve those of their class (supported N in
// (NLO, SH::kMaxN]), skipping the rest; the last workgroup to leave resets
// ctr for the next launch.  An unsupported N is marked (iters = -1) by
// whichever class claims it.
template <typename SH, int NLO>
__global__ __launch_bounds__(SH::kThreads, SH::kPerCU) void fdm_pcg_kernel(FdmBatch a, uint32_t* ctr)
{
    __shared__ SH sh;
    for (;;) {
        if (threadIdx.x == 0) sh.ticket = (int)atomicAdd(&ctr[0], 1u);
        __syncthreads();
        const int t = sh.ticket;
        __syncthreads();
        if (t >= a.B) break;
        const int b = a.order ? a.order[t] : t;
        const int n = a.N[b];
        const bool mine = (NLO == 0 || n > NLO) && !(n > SH::kMaxN && fdm_n_ok(n));
        if (!mine) continue;  // the other class (uniform across the workgroup)
        switch (n) {
        case 8: if constexpr (NLO < 8) fdm_cube<8>(a, sh, b); break;
        case 9: if constexpr (NLO < 9) fdm_cube<9>(a, sh, b); break;
        case 16: if constexpr (NLO < 16) fdm_cube<16>(a, sh, b); break;
        case 21: if constexpr (SH::kMaxN >= 21) fdm_cube<21>(a, sh, b); break;
        case 23: if constexpr (SH::kMaxN >= 23) fdm_cube<23>(a, sh, b); break;
        default:
            if (threadIdx.x == 0 && a.iters) a.iters[b] = -1;
            break;
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&ctr[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&ctr[0], 0u);
            atomicExch(&ctr[1], 0u);
        }
    }
}

bool fdm_supported_n(int n) { return fdm_n_ok(n); }

int fdm_grid_size(int device)
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

int fdm_launch(const FdmBatch& a, int grid, hipStream_t stream)
{
    if (a.B <= 0) return 0;
    if (grid <= 0) return set_error(-2, "fdm_launch: grid must be positive");
    // grid = CUs; the workspace holds grid * kMaxCells cells
    const bool small = a.n_hint == 0 || a.n_hint <= FdmSmall::kMaxN;
    const bool large = a.n_hint == 0 || a.n_hint > FdmSmall::kMaxN;
    if (large) {
        const int g = grid < a.B ? grid : a.B;
        hipLaunchKernelGGL((fdm_pcg_kernel<FdmLarge, FdmSmall::kMaxN>), dim3(g), dim3(FdmLarge::kThreads), 0, stream,
                           a, a.counters);
        DRW_HIP(hipGetLastError());
    }
    if (small) {
        int g = FdmSmall::kPerCU * grid;
        const int fit = (int)((size_t)grid * kMaxCells / FdmSmall::kCells);
        if (g > fit) g = fit;
        if (g > a.B) g = a.B;
        hipLaunchKernelGGL((fdm_pcg_kernel<FdmSmall, 0>), dim3(g), dim3(FdmSmall::kThreads), 0, stream, a,
                           a.counters + 2);
        DRW_HIP(hipGetLastError());
    }
    return 0;
}

}  // namespace drw
