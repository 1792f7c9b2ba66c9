// Batched FDM surface Green's function on MI355X (gfx950).
//
// Solves, per cube, L y = e with the 7-point harmonic-mean operator
//   (L y)_c = sum_n w_cn (y_c - y_n) + 2 eps_c nb(c) y_c,  w = 2 eps_c eps_n/(eps_c+eps_n)
// and writes G_b = 2 eps_c y_c on the 6 N^2 boundary panels, which is what the
// reference GGFT solver computes as gft = A13 x with A = L D^-1, x = D y
// (ggft/src/ggft_solver.cpp:447-518, 552-596).  The right-hand sides follow
// update_ek (ggft_solver.cpp:137-274) for odd and even N.
//
// Design (persistent workgroups, one cube at a time each; 1024 threads and one
// workgroup per CU up to N = 23, 256 threads and three per CU up to N = 16):
//  * a thread owns the red and the black cell of x-pairs h = tid + q*T (12
//    cells at N = 23); the x and y face couplings of each cell,
//    symmetrically Jacobi-scaled (w_cn / sqrt(d_c d_n)), live in VGPRs for
//    the whole solve; the z couplings live in LDS (the z- one of a cell is
//    the z+ one of the cell below), next to the search direction p, with
//    zero guards so the stencil reads need no bounds checks (x/y wrap-around
//    reads hit a zero weight), and the Jacobi scales s_c; x, r and p of the
//    own cells stay in registers;
//  * inner solver: FP32 CG on the red-black reduced system of the
//    Jacobi-scaled operator (half the iterations of Jacobi PCG, see
//    fdm_cube), the two colours in separate LDS arrays; block-wide
//    reductions by wave64 DPP sums + a 16-entry LDS table, summed in a
//    fixed order so every thread sees bit-identical alpha/beta (held in
//    scalar registers);
//  * outer loop: FP64 iterative refinement.  The residual e - L y is formed
//    in FP64 with exact harmonic means and the loop stops at ||e - L y|| <=
//    tol ||e|| (the reference's relative-residual criterion,
//    ggft_solver.h:123-126, 193-198).  After the first FP32 solve y = ||e||
//    S x is kept implicit (s and x in LDS, exact as a double product): the
//    FP64 check and the G panels are computed from it, so a cube that
//    converges in one solve (most of them, with the first solve aimed at
//    0.3 tol) reads its permittivities and writes its 6 N^2 outputs and
//    little else to HBM; a cube that needs a second round spills y to an
//    FP64 workspace and refines from there;
//  * work distribution: a self-resetting atomic ticket counter hands cubes to
//    workgroups in the (optional) host-given order, e.g. largest N first.
#include "common.h"
#include "fdm.h"
#include "sampling.h"
#include "voxel_fn.h"

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
    float ap[kCells];                  // eps while setting up, then the Jacobi scales s_c
    float red[2][kWaves];
    double redd[kWaves];
    double rhs_val[8];
    int rhs_idx[8];
    double enorm;
    int ticket;
};
// Both classes run textbook CG with A p in registers.  The Chronopoulos-Gear
// form (one fused reduction per iteration) needs s = A p, x and p per cell
// besides r and spills at these register budgets: N = 23 batch 16.1 ms
// against 12.5 ms, N = 16 batch 3.96 ms against 2.95 ms.
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
    return uniform_f32(s);
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
    return uniform_f64(s);
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

// One cube's solve: its permittivities, outputs, the y workspace of the
// calling workgroup and the iteration limits.
struct CubeJob {
    const float* E;     // N^3 permittivities, x fastest
    double* G;          // 6 N^2 outputs
    double* Y;          // N^3 FP64 refinement vector (workspace)
    int code;           // right-hand side
    double tol;
    float inner_floor;
    int max_inner, max_rounds;
    int32_t* iters;     // optional single outputs
    float* relres;
    int32_t* status;    // 0 converged, 1 not within max_rounds
};

template <int N, typename SH>
__device__ void fdm_cube(const CubeJob& a, SH& sh)
{
    static_assert(N <= SH::kMaxN, "cube larger than the size class");
    constexpr int NN = N * N, NNN = N * NN;
    constexpr int T = SH::kThreads, W = SH::kWaves;
    // The inner solve is CG on the red-black reduced system.  The 7-point
    // grid is two-coloured by the parity of i + j + k, and every neighbour
    // of a cell has the other colour.  With the Jacobi-scaled operator A = I - B
    // (B: the scaled couplings, red <-> black only), the black unknowns are
    // x_b = b_b + B_br x_r, and the red ones solve the Schur complement
    // (I - B_rb B_br) x_r = b_r + B_rb b_b, whose condition number is about a
    // quarter of A's: CG needs half the iterations, each one a black half
    // sweep and a red half sweep (the work of one stencil pass).  The full
    // residual of (x_r, x_b) is the reduced one on the red cells and zero on
    // the black ones, so the stopping test is unchanged.  Cells are taken in
    // pairs h = (2h, 2h + 1) along x: for odd N the parity of c = i + N j +
    // N^2 k is that of i + j + k, so pair h holds red 2h and black 2h + 1;
    // for even N a row of N / 2 pairs starts with its red cell when j + k is
    // even (par = 0) and with its black one otherwise (par = 1).  A thread's
    // slots q < CPH are the red cells of pairs h = tid + (q mod CPH) T, the
    // others the black cells of the same pairs.  For even N the row parity
    // of a thread's pairs must not depend on q (T / (N/2) rows apart, a
    // multiple of 2N: j unchanged, k by an even step); where it would, the
    // kernel keeps Jacobi PCG on c = tid + q T.
    constexpr bool RB = (N & 1) != 0 || (T % (N / 2) == 0 && (T / (N / 2)) % (2 * N) == 0);
    constexpr int CPH = ((NNN + 1) / 2 + T - 1) / T;
    constexpr int CPT = RB ? 2 * CPH : (NNN + T - 1) / T;
    // Jacobi PCG: offset of slot q's cell from the thread's first cell tid
    auto off = [](int q) __attribute__((always_inline)) { return q * T; };
    // only the last slot of each colour can lie past the last cell; the
    // other slots' tests fold away after unrolling
    auto edge = [](int q) __attribute__((always_inline)) { return RB ? (q == CPH - 1 || q == CPT - 1) : q == CPT - 1; };
    // Colour-split LDS layout of the inner solve: the red cell of pair h at
    // R[h] = p[R0 + h], the black one at K[h] = p[K0 + h], and their z+
    // couplings at ZR[h] = wz[ZR0 + h], ZK[h] = wz[ZK0 + h].  Every neighbour
    // of a colour sits at a fixed offset in the other colour's array: x at
    // -1 + par, par (red) and -par, 1 - par (black), y at -(N/2) - (N&1),
    // N/2 (red) and -(N/2), N/2 + (N&1) (black), z at -C, F (red) and -F, C
    // (black), with F = floor(N^2/2), C = ceil(N^2/2) and par = 0 for odd N;
    // so a wave's reads are consecutive words.  Zero guards of F (below R,
    // above R), C (below and above K) entries take the out-of-cube
    // neighbours, whose couplings are zero.
    constexpr int HR = (NNN + 1) / 2, HB = NNN / 2, F = NN / 2, C = (NN + 1) / 2;
    constexpr int YM = N / 2 + (N & 1), YP = N / 2;  // red y-, y+ (black: YP, YM)
    constexpr int R0 = F, K0 = 2 * F + HR + C;        // split p: [0, NNN + 2 NN)
    constexpr int ZR0 = F, ZK0 = F + HR + C;          // split wz: [0, NNN + NN)
    static_assert(!RB || K0 + HB + C <= SH::kCells + 2 * SH::kGuard, "split p exceeds its array");
    static_assert(!RB || ZK0 + HB <= SH::kCells + SH::kGuard, "split wz exceeds its array");
    // A p stays in registers and the Jacobi scales s_c in the third LDS
    // array for the whole solve (s is read before and after each inner solve
    // only, so in registers it would be spilled across the loop).
    // Opaque to the optimiser: the per-thread cell indices, LDS addresses and
    // boundary masks derive from tid, so they are recomputed for each cube
    // (a few VALU instructions) instead of being hoisted out of the
    // persistent ticket loop for every cell and cube size and spilled to
    // scratch, whose per-cube reloads (about 1 MB per workgroup) do not stay
    // in L2.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    // row parity of the thread's pairs (even N; 0 for odd N), and slot q's cell
    const int par = (N & 1) ? 0 : (((tid / (N / 2)) % N + tid / (N / 2) / N) & 1);
    auto cell = [&](int q) __attribute__((always_inline)) {
        return RB ? 2 * (tid + (q % CPH) * T) + (q >= CPH ? 1 - par : par) : tid + q * T;
    };
    const float* E = a.E;
    const int code = a.code;
    float* P = sh.p + NN;  // valid for indices [-NN, NNN + NN)
    float* SC = sh.ap;     // s_c after the set-up

    // ---- stage eps (in the A p array), clear guards, build the source -------
    float* WZ = sh.wz + NN;  // valid for indices [-NN, NNN)
    float* EL = sh.ap;       // eps staging during the set-up
    for (int c = tid; c < NNN; c += T) EL[c] = E[c];
    for (int g = tid; g < NN; g += T) { P[g - NN] = 0.f; P[NNN + g] = 0.f; WZ[g - NN] = 0.f; }
    if constexpr (RB) {
        // split wz guards: below ZR (black z- reads) and below ZK (red z- reads)
        for (int g = tid; g < F; g += T) sh.wz[ZR0 - F + g] = 0.f;
        for (int g = tid; g < C; g += T) sh.wz[ZK0 - C + g] = 0.f;
    }
    if (tid == 0) {
        fdm_rhs<N>(E, code, sh.rhs_idx, sh.rhs_val);
        double s = 0.0;
        for (int q = 0; q < 8; q++) s += sh.rhs_val[q] * sh.rhs_val[q];
        sh.enorm = sqrt(s);
    }
    __syncthreads();

    // ---- Jacobi scaling s_c = d_c^-1/2, d_c = sum_n w_cn + 2 eps_c nb(c) -----
    // (all cells in P for the coupling set-up below)
    float sc[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = cell(q);
        sc[q] = 0.f;
        if (!edge(q) || c < NNN) {
            const int i = c % N, j = (c / N) % N, k = c / NN;
            const float ec = EL[c];
            float d = 0.f;
            d += i > 0 ? harm(ec, EL[c - 1]) : 2.0f * ec;
            d += i < N - 1 ? harm(ec, EL[c + 1]) : 2.0f * ec;
            d += j > 0 ? harm(ec, EL[c - N]) : 2.0f * ec;
            d += j < N - 1 ? harm(ec, EL[c + N]) : 2.0f * ec;
            d += k > 0 ? harm(ec, EL[c - NN]) : 2.0f * ec;
            d += k < N - 1 ? harm(ec, EL[c + NN]) : 2.0f * ec;
            sc[q] = 1.0f / sqrtf(d);
            P[c] = sc[q];
        }
    }
    __syncthreads();

    // ---- scaled couplings: x/y in VGPRs, z+ in LDS ---------------------------
    float w[4][CPT];  // x-, x+, y-, y+
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = cell(q);
        if (!edge(q) || c < NNN) {
            const int i = c % N, j = (c / N) % N, k = c / NN;
            const float ec = EL[c], sc = P[c];
            w[0][q] = i > 0 ? harm(ec, EL[c - 1]) * sc * P[c - 1] : 0.f;
            w[1][q] = i < N - 1 ? harm(ec, EL[c + 1]) * sc * P[c + 1] : 0.f;
            w[2][q] = j > 0 ? harm(ec, EL[c - N]) * sc * P[c - N] : 0.f;
            w[3][q] = j < N - 1 ? harm(ec, EL[c + N]) * sc * P[c + N] : 0.f;
            const float wzc = k < N - 1 ? harm(ec, EL[c + NN]) * sc * P[c + NN] : 0.f;
            if constexpr (RB) sh.wz[(q >= CPH ? ZK0 : ZR0) + (c >> 1)] = wzc;
            else WZ[c] = wzc;
        } else {
#pragma unroll
            for (int d = 0; d < 4; d++) w[d][q] = 0.f;
        }
    }
    // Red-black: the red and the black cell of a pair are x neighbours, so
    // one coupling serves both (the red cell's, which makes the scaled
    // operator exactly symmetric in it); per pair the partner coupling wp,
    // the red cell's other x coupling wor and the black cell's wob, and the
    // y couplings in w[2], w[3]: one register per pair less than w[0..1].
    float wp[RB ? CPH : 1], wor[RB ? CPH : 1], wob[RB ? CPH : 1];
    if constexpr (RB) {
#pragma unroll
        for (int m = 0; m < CPH; m++) {
            wp[m] = par ? w[0][m] : w[1][m];
            wor[m] = par ? w[1][m] : w[0][m];
            wob[m] = par ? w[0][m + CPH] : w[1][m + CPH];
        }
    }
    __syncthreads();
    // eps staging is done: s_c moves to the third array for good
    for (int c = tid; c < NNN; c += T) SC[c] = P[c];
    __syncthreads();

    const double enorm = uniform_f64(sh.enorm);
    float* R32 = P;  // residual hand-off reuses the p array (p is rebuilt from it)
    double rnorm = enorm;
    int total_it = 0;
    bool have_y = false;  // y = 0 until the first update writes it
    bool converged = false;
    // After the first FP32 solve y = ||e|| S x exactly (a double product of
    // the FP32 s_c and x_c), so y stays implicit: s (SV) and x (XV, in the p
    // array) of every cell sit in LDS, the FP64 check reads them there and G
    // comes from them.  Only a cube that needs a second round writes y to Y.
    const float* SV = SC;
    float* XV = P;
    bool implicit = false;
    double yscale = 0.0;
    for (int round = 0; round < a.max_rounds; round++) {
        if (implicit) {
            // ---- FP64 residual norm of y = yscale S x --------------------------
            double loc = 0.0;
            const float* Ev = E;  // address arithmetic kept inside the pass (not held, and spilled, across rounds)
            asm volatile("" : "+v"(Ev));
#pragma unroll 1
            for (int c = tid; c < NNN; c += T) {
                const int i = c % N, j = (c / N) % N, k = c / NN;
                const double ec = Ev[c], uc = (double)SV[c] * (double)XV[c];
                auto u = [&](int n) { return (double)SV[n] * (double)XV[n]; };
                double acc = 0.0;
                int nb = 0;
                if (i > 0) acc += harm(ec, (double)Ev[c - 1]) * (uc - u(c - 1)); else nb++;
                if (i < N - 1) acc += harm(ec, (double)Ev[c + 1]) * (uc - u(c + 1)); else nb++;
                if (j > 0) acc += harm(ec, (double)Ev[c - N]) * (uc - u(c - N)); else nb++;
                if (j < N - 1) acc += harm(ec, (double)Ev[c + N]) * (uc - u(c + N)); else nb++;
                if (k > 0) acc += harm(ec, (double)Ev[c - NN]) * (uc - u(c - NN)); else nb++;
                if (k < N - 1) acc += harm(ec, (double)Ev[c + NN]) * (uc - u(c + NN)); else nb++;
                double rv = -yscale * (acc + 2.0 * ec * nb * uc);
#pragma unroll
                for (int s = 0; s < 8; s++)
                    if (sh.rhs_idx[s] == c) rv += sh.rhs_val[s];
                loc += rv * rv;
            }
            rnorm = sqrt(block_sum_d<W>(loc, sh.redd));  // also orders the LDS reads before reuse
            if (!(rnorm > a.tol * enorm)) {
                converged = true;
                break;
            }
            // a further round: y goes to Y, and the residual is formed from it
            double* Yv = a.Y;
            asm volatile("" : "+v"(Yv));  // addresses of the rare path stay in it
            for (int c = tid; c < NNN; c += T) Yv[c] = (double)SV[c] * (double)XV[c] * yscale;
            implicit = false;
            __syncthreads();
        }
        // ---- FP64 residual r = e - L y, handed to the inner solve as FP32 r ----
        double loc = 0.0;
        if (round == 0) {
            for (int c = tid; c < NNN; c += T) {
                double rv = 0.0;
#pragma unroll
                for (int s = 0; s < 8; s++)
                    if (sh.rhs_idx[s] == c) rv += sh.rhs_val[s];
                loc += rv * rv;
                R32[c] = (float)rv;
            }
        } else {
            const double* Yv = a.Y;
            const float* Ev = E;
            asm volatile("" : "+v"(Yv), "+v"(Ev));
#pragma unroll 1
            for (int c = tid; c < NNN; c += T) {
                const int i = c % N, j = (c / N) % N, k = c / NN;
                const double ec = Ev[c], yc = Yv[c];
                double acc = 0.0;
                int nb = 0;
                if (i > 0) acc += harm(ec, (double)Ev[c - 1]) * (yc - Yv[c - 1]); else nb++;
                if (i < N - 1) acc += harm(ec, (double)Ev[c + 1]) * (yc - Yv[c + 1]); else nb++;
                if (j > 0) acc += harm(ec, (double)Ev[c - N]) * (yc - Yv[c - N]); else nb++;
                if (j < N - 1) acc += harm(ec, (double)Ev[c + N]) * (yc - Yv[c + N]); else nb++;
                if (k > 0) acc += harm(ec, (double)Ev[c - NN]) * (yc - Yv[c - NN]); else nb++;
                if (k < N - 1) acc += harm(ec, (double)Ev[c + NN]) * (yc - Yv[c + NN]); else nb++;
                double rv = -(acc + 2.0 * ec * nb * yc);
#pragma unroll
                for (int s = 0; s < 8; s++)
                    if (sh.rhs_idx[s] == c) rv += sh.rhs_val[s];
                loc += rv * rv;
                R32[c] = (float)rv;  // scaled by s_c where the inner solve reads it
            }
        }
        rnorm = sqrt(block_sum_d<W>(loc, sh.redd));  // also orders the R32 writes
        if (!(rnorm > a.tol * enorm)) {           // converged (or enorm == 0)
            converged = true;
            break;
        }

        // ---- FP32 CG on the scaled system, right-hand side S r / ||r|| -------
        const float inv = (float)(1.0 / rnorm);
        double tgt = 0.3 * a.tol * enorm / rnorm;
        if (tgt < (double)a.inner_floor) tgt = a.inner_floor;
        const float tgt2 = (float)(tgt * tgt);
        float xh[CPT];
        {
            float rh[CPT];
            float rloc = 0.f;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = cell(q);
                xh[q] = 0.f;
                rh[q] = (!edge(q) || c < NNN) ? R32[c] * SC[c] * inv : 0.f;
                rloc += rh[q] * rh[q];
            }
            float rr = block_sum<W>(rloc, sh.red[1]);  // also: every read of R32 is done
            const float thresh = tgt2 * rr;
            if constexpr (RB) {
                // Per-thread bases of the colour-split arrays; slot q < CPH is red
                // h = tid + qT at R[qT], slot q >= CPH black h = tid + (q - CPH)T
                // at K[(q - CPH)T].
                lds_float* Rb = lds_base(sh.p + R0 + tid);
                lds_float* Kb = lds_base(sh.p + K0 + tid);
                // the other x neighbour of a red pair's cell is black pair
                // h - 1 + 2 par, of a black pair's cell red pair h + 1 - 2 par
                // (odd N: par = 0, fixed offsets from Kb, Rb)
                lds_float* Kx = (N & 1) ? Kb - 1 : lds_base(sh.p + K0 + tid - 1 + 2 * par);
                lds_float* Rx = (N & 1) ? Rb + 1 : lds_base(sh.p + R0 + tid + 1 - 2 * par);
                lds_float* Zr = lds_base(sh.wz + ZR0 + tid);
                lds_float* Zk = lds_base(sh.wz + ZK0 + tid);
                // sum over the black neighbours of red slot q of w P
                auto red_nb = [&](int q) __attribute__((always_inline)) {
                    const int j = q * T;
                    float v = wor[q] * Kx[j];
                    v += wp[q] * Kb[j];
                    v += w[2][q] * Kb[j - YM];
                    v += w[3][q] * Kb[j + YP];
                    v += Zk[j - C] * Kb[j - C];
                    v += Zr[j] * Kb[j + F];
                    return v;
                };
                // sum over the red neighbours of black slot q of w P
                auto black_nb = [&](int q) __attribute__((always_inline)) {
                    const int j = (q - CPH) * T;
                    float v = wp[q - CPH] * Rb[j];
                    v += wob[q - CPH] * Rx[j];
                    v += w[2][q] * Rb[j - YP];
                    v += w[3][q] * Rb[j + YM];
                    v += Zr[j - F] * Rb[j - F];
                    v += Zk[j] * Rb[j + C];
                    return v;
                };
                // guards of the split arrays (the p array held R32 until now)
                for (int g = tid; g < F; g += T) {
                    sh.p[R0 - F + g] = 0.f;
                    sh.p[R0 + HR + g] = 0.f;
                }
                for (int g = tid; g < C; g += T) {
                    sh.p[K0 - C + g] = 0.f;
                    sh.p[K0 + HB + g] = 0.f;
                }
                // b: red slots to R, black slots to K
#pragma unroll
                for (int q = 0; q < CPT; q++) {
                    const int c = cell(q);
                    if (!edge(q) || c < NNN) {
                        if (q < CPH) Rb[q * T] = rh[q];
                        else Kb[(q - CPH) * T] = rh[q];
                    }
                }
                __syncthreads();
                // reduced right-hand side b_r + B_rb b_b on the red slots (rh of
                // the black slots keeps b_b for the recovery below)
                float rl = 0.f;
#pragma unroll
                for (int q = 0; q < CPH; q++) {
                    const int c = cell(q);
                    if (!edge(q) || c < NNN) {
                        rh[q] += red_nb(q);
                        rl += rh[q] * rh[q];
                    }
                }
                rr = block_sum<W>(rl, sh.red[1]);  // also: every read of b is done
                float pv[CPH];
#pragma unroll
                for (int q = 0; q < CPH; q++) pv[q] = rh[q];
                if (rr > thresh) {
                    for (int it = 0; it < a.max_inner; it++) {
#pragma unroll
                        for (int q = 0; q < CPH; q++) {
                            const int c = cell(q);
                            if (!edge(q) || c < NNN) Rb[q * T] = pv[q];
                        }
                        __syncthreads();
                        // black half sweep: t_b = B_br p
#pragma unroll
                        for (int q = CPH; q < CPT; q++) {
                            const int c = cell(q);
                            if (!edge(q) || c < NNN) Kb[(q - CPH) * T] = black_nb(q);
                        }
                        __syncthreads();
                        // red half sweep: (I - B_rb B_br) p = p - B_rb t
                        float pl = 0.f;
                        float sp[CPH];
#pragma unroll
                        for (int q = 0; q < CPH; q++) {
                            const int c = cell(q);
                            sp[q] = 0.f;
                            if (!edge(q) || c < NNN) {
                                sp[q] = pv[q] - red_nb(q);
                                pl += pv[q] * sp[q];
                            }
                        }
                        const float pq = block_sum<W>(pl, sh.red[0]);
                        const float alpha = rr / pq;
                        float rl2 = 0.f;
#pragma unroll
                        for (int q = 0; q < CPH; q++) {
                            const int c = cell(q);
                            if (!edge(q) || c < NNN) {
                                xh[q] += alpha * pv[q];
                                rh[q] -= alpha * sp[q];
                                rl2 += rh[q] * rh[q];
                            }
                        }
                        const float rr_new = block_sum<W>(rl2, sh.red[1]);
                        total_it++;
                        if (!(rr_new > thresh)) break;
                        const float beta = rr_new / rr;
                        rr = rr_new;
#pragma unroll
                        for (int q = 0; q < CPH; q++) pv[q] = rh[q] + beta * pv[q];
                    }
                }
                // black unknowns x_b = b_b + B_br x_r (the last reduction above,
                // or the one before the loop, ordered every earlier read of R)
#pragma unroll
                for (int q = 0; q < CPH; q++) {
                    const int c = cell(q);
                    if (!edge(q) || c < NNN) Rb[q * T] = xh[q];
                }
                __syncthreads();
#pragma unroll
                for (int q = CPH; q < CPT; q++) {
                    const int c = cell(q);
                    if (!edge(q) || c < NNN) xh[q] = rh[q] + black_nb(q);
                }
                __syncthreads();  // every read of the split arrays done before P is rewritten
            } else {
                // Per-thread LDS bases; element off(q) + k of each is slot q's
                // cell shifted by k - N^2.
                lds_float* Pb = lds_base(sh.p + tid);
                lds_float* Zb = lds_base(sh.wz + tid);
#pragma unroll
                for (int q = 0; q < CPT; q++) {
                    const int c = cell(q);
                    if (!edge(q) || c < NNN) P[c] = rh[q];
                }
                __syncthreads();
                for (int it = 0; it < a.max_inner; it++) {
                    float pl = 0.f;
                    float apv[CPT];
#pragma unroll
                    for (int q = 0; q < CPT; q++) {
                        const int c = cell(q);
                        apv[q] = 0.f;
                        if (!edge(q) || c < NNN) {
                            const int o = off(q) + NN;  // P index of cell c relative to Pb
                            const float pc = Pb[o];
                            float v = pc;
                            v -= w[0][q] * Pb[o - 1];
                            v -= w[1][q] * Pb[o + 1];
                            v -= w[2][q] * Pb[o - N];
                            v -= w[3][q] * Pb[o + N];
                            v -= Zb[o - NN] * Pb[o - NN];
                            v -= Zb[o] * Pb[o + NN];
                            apv[q] = v;
                            pl += pc * v;
                        }
                    }
                    const float pq = block_sum<W>(pl, sh.red[0]);
                    const float alpha = rr / pq;
                    float rl = 0.f;
#pragma unroll
                    for (int q = 0; q < CPT; q++) {
                        const int c = cell(q);
                        if (!edge(q) || c < NNN) {
                            xh[q] += alpha * Pb[off(q) + NN];
                            rh[q] -= alpha * apv[q];
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
                        const int c = cell(q);
                        if (!edge(q) || c < NNN) Pb[off(q) + NN] = rh[q] + beta * Pb[off(q) + NN];
                    }
                    __syncthreads();
                }
            }
        }
        // ---- y += ||r|| S x (implicit after the first solve) -------------------
        // every thread is past its last read of its neighbours' p (the solve's
        // final reduction), so P can take x
        if (!have_y) {
            implicit = true;
            yscale = rnorm;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = cell(q);
                if (!edge(q) || c < NNN) XV[c] = xh[q];
            }
        } else {
            double* Yv = a.Y;
            asm volatile("" : "+v"(Yv));
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = cell(q);
                if (!edge(q) || c < NNN) Yv[c] += (double)SC[c] * (double)xh[q] * rnorm;
            }
        }
        have_y = true;
        __syncthreads();
    }

    // ---- G_b = 2 eps_c y_c, GGFT panel order (ggft_solver.h:65-77) ----------
    double* G = a.G;
    const float* Ev = E;
    asm volatile("" : "+v"(G), "+v"(Ev));
    if (implicit) {
        // each boundary cell's owner writes its panels (a cell on an edge or a
        // corner lies on two or three)
#pragma unroll 1
        for (int c = tid; c < NNN; c += T) {
            const int i = c % N, j = (c / N) % N, k = c / NN;
            if (i == 0 || i == N - 1 || j == 0 || j == N - 1 || k == 0 || k == N - 1) {
                const double g = 2.0 * (double)Ev[c] * ((double)SV[c] * (double)XV[c] * yscale);
                if (k == 0) G[i + j * N] = g;
                if (k == N - 1) G[NN + i + j * N] = g;
                if (j == 0) G[2 * NN + i + k * N] = g;
                if (j == N - 1) G[3 * NN + i + k * N] = g;
                if (i == 0) G[4 * NN + j + k * N] = g;
                if (i == N - 1) G[5 * NN + j + k * N] = g;
            }
        }
    } else {
        const double* Yv = a.Y;
        asm volatile("" : "+v"(Yv));
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
            G[pb] = have_y ? 2.0 * (double)Ev[c] * Yv[c] : 0.0;
        }
    }
    if (tid == 0) {
        if (a.iters) *a.iters = total_it;
        if (a.relres) *a.relres = enorm > 0 ? (float)(rnorm / enorm) : 0.f;
        if (a.status) *a.status = converged ? 0 : 1;
    }
    __syncthreads();
}

// The job of cube b of a batch, solved by the calling workgroup.
__device__ __forceinline__ CubeJob batch_job(const FdmBatch& a, int b, int cells)
{
    CubeJob j;  // the cube's pointers are the same in every thread: scalar registers
    j.E = uniform_ptr(a.eps + a.eps_off[b]);
    j.G = uniform_ptr(a.G + a.G_off[b]);
    j.Y = a.y_ws + (size_t)blockIdx.x * cells;  // fdm_launch sizes the grids to fit the workspace
    j.code = a.code[b];
    j.tol = a.tol;
    j.inner_floor = a.inner_floor;
    j.max_inner = a.max_inner;
    j.max_rounds = a.max_rounds;
    j.iters = a.iters ? a.iters + b : nullptr;
    j.relres = a.relres ? a.relres + b : nullptr;
    j.status = a.status ? a.status + b : nullptr;
    return j;
}

__host__ __device__ constexpr bool fdm_n_ok(int n) { return n == 8 || n == 9 || n == 16 || n == 21 || n == 23; }

// One persistent kernel per size class: workgroups take cubes by atomic
// ticket (ctr[0]) and solve those of their class (supported N in
// (NLO, SH::kMaxN]), skipping the rest; the last workgroup to leave resets
// ctr for the next launch.  An unsupported N is marked (iters = -1,
// status = 2) by whichever class claims it.
template <typename SH, int NLO>
__global__ __launch_bounds__(SH::kThreads, SH::kPerCU) void fdm_pcg_kernel(FdmBatch a, uint32_t* ctr)
{
    __shared__ SH sh;
    const int nB = a.B_dev ? min(a.B, *a.B_dev) : a.B;  // the same value in every thread
    for (;;) {
        if (threadIdx.x == 0) sh.ticket = (int)atomicAdd(&ctr[0], 1u);
        __syncthreads();
        const int t = sh.ticket;
        __syncthreads();
        if (t >= nB) break;
        const int b = a.order ? a.order[t] : t;
        const int n = a.N[b];
        const bool mine = (NLO == 0 || n > NLO) && !(n > SH::kMaxN && fdm_n_ok(n));
        if (!mine) continue;  // the other class (uniform across the workgroup)
        switch (n) {
        case 8: if constexpr (NLO < 8) fdm_cube<8>(batch_job(a, b, SH::kCells), sh); break;
        case 9: if constexpr (NLO < 9) fdm_cube<9>(batch_job(a, b, SH::kCells), sh); break;
        case 16: if constexpr (NLO < 16) fdm_cube<16>(batch_job(a, b, SH::kCells), sh); break;
        case 21: if constexpr (SH::kMaxN >= 21) fdm_cube<21>(batch_job(a, b, SH::kCells), sh); break;
        case 23: if constexpr (SH::kMaxN >= 23) fdm_cube<23>(batch_job(a, b, SH::kCells), sh); break;
        default:
            if (threadIdx.x == 0) {
                if (a.iters) a.iters[b] = -1;
                if (a.status) a.status[b] = 2;
            }
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

// ---------------------------------------------------------------- acceptance
// Sum and minimum of one kernel by one wave (all lanes get them).
__device__ __forceinline__ void wave_kernel_stats(const double* G, int M, double& sum, double& mn)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0, m = 1e300;
    for (int i = lane; i < M; i += 64) {
        const double g = G[i];
        s += g;  // a NaN or infinity anywhere makes the sum non-finite
        m = fmin(m, g);
    }
    sum = wave_sum(s);
    mn = wave_min(m);
}

__device__ __forceinline__ bool kernel_ok(int status, int code, double sum, double mn)
{
    if (status != 0 || !isfinite(sum)) return false;
    return code != kGFT || (fabs(sum - 1.0) <= 0.01 && mn >= 0.0);  // ggft_solver.cpp:599-625
}

// One wave per cube (first pass: all B; second pass: the listed cubes).
__global__ __launch_bounds__(256) void fdm_accept_kernel(FdmAcceptArgs a)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int n = a.list_in ? min(a.B, *a.count_in) : a.B;
    if (a.list_in && a.stats && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&a.stats[0], (unsigned long long)n);
    if (w >= n) return;
    const int b = a.list_in ? a.list_in[w] : w;
    const int N = a.N[b];
    double sum, mn;
    wave_kernel_stats(a.G + a.G_off[b], 6 * N * N, sum, mn);
    const bool ok = kernel_ok(a.status ? a.status[b] : 0, a.code[b], sum, mn);
    if ((threadIdx.x & 63) == 0 && !ok) {
        if (a.list_in) {
            if (a.bad) a.bad[b] = 1;
            if (a.stats) atomicAdd((unsigned long long*)&a.stats[1], 1ull);
        } else {
            a.list_out[atomicAdd(a.count_out, 1)] = b;
        }
    }
}

int fdm_accept_launch(const FdmAcceptArgs& a, hipStream_t stream)
{
    if (a.B <= 0) return 0;
    hipLaunchKernelGGL(fdm_accept_kernel, dim3((a.B + 3) / 4), dim3(256), 0, stream, a);
    DRW_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------- network fallback
// One workgroup: the rows of out[B][width] that hold a NaN, in list.
// Block-wide sum and minimum of one kernel (1024 threads), then acceptance.
template <int W, typename SH>
__device__ bool block_kernel_ok(const double* G, int M, int status, int code, SH& sh)
{
    double s = 0.0, m = 1e300;
    for (int i = threadIdx.x; i < M; i += SH::kThreads) {
        s += G[i];
        m = fmin(m, G[i]);
    }
    const double sum = block_sum_d<W>(s, sh.redd);
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) sh.redd[threadIdx.x >> 6] = m;
    __syncthreads();
    double mn = 1e300;
#pragma unroll
    for (int w = 0; w < W; w++) mn = fmin(mn, sh.redd[w]);
    __syncthreads();
    return kernel_ok(status, code, sum, mn);
}

// Workgroups take listed queries by ticket; each solves in its own slot of
// the eps / G / y buffers: voxelize (raw permittivities, fill_structure
// semantics), solve at tol, and on a failed acceptance once more in FP64
// mode, then draw the hop with the query's own counter.  The last workgroup
// to leave resets the ticket counter.
__global__ __launch_bounds__(FdmLarge::kThreads, 1) void nn_fallback_kernel(NnFallbackArgs a)
{
    __shared__ FdmLarge sh;
    __shared__ float Ly[kMaxLayers * 2], Cb[kMaxCuboids * 7];
    constexpr int N = kMaxN, NN = N * N, NNN = N * NN;
    const int n = min(*a.count, a.cap);  // appended by the face kernels of this batch
    const int slot = blockIdx.x;
    float* E = a.eps + (size_t)slot * kMaxCells;
    double* G = a.G + (size_t)slot * 6 * NN;
    const bool grad = a.axis != nullptr;
    for (;;) {
        if (threadIdx.x == 0) sh.ticket = (int)atomicAdd(&a.ctr[0], 1u);
        __syncthreads();
        const int t = sh.ticket;
        __syncthreads();
        if (t >= n) break;
        const int q = a.list[t];
        if (threadIdx.x < kMaxLayers * 2) Ly[threadIdx.x] = a.layers[(size_t)q * kMaxLayers * 2 + threadIdx.x];
        if (threadIdx.x < kMaxCuboids * 7) Cb[threadIdx.x] = a.cuboids[(size_t)q * kMaxCuboids * 7 + threadIdx.x];
        __syncthreads();
        float lv[kMaxLayers], cv[kMaxCuboids];
        for (int l = 0; l < kMaxLayers; l++) lv[l] = Ly[2 * l + 1];
        for (int c = 0; c < kMaxCuboids; c++) cv[c] = Cb[7 * c + 6];
        for (int v = threadIdx.x; v < NNN; v += FdmLarge::kThreads) {
            const int x = v % N, y = (v / N) % N, z = v / NN;
            E[v] = voxel_value(Ly, lv, Cb, cv, (2.0f * x + 1.0f) / N - 1.0f, (2.0f * y + 1.0f) / N - 1.0f,
                               (2.0f * z + 1.0f) / N - 1.0f);
        }
        __syncthreads();  // the slot cube is visible to the whole workgroup
        const int ax = grad ? a.axis[q] : 2;
        CubeJob j;
        j.E = E;
        j.G = G;
        j.Y = a.y_ws + (size_t)slot * kMaxCells;
        j.code = grad ? (ax == 0 ? kWVTX : ax == 1 ? kWVTY : kWVTZ) : kGFT;
        j.tol = a.tol;
        j.inner_floor = kFdmInnerFloor;
        j.max_inner = 2000;
        j.max_rounds = 12;
        j.iters = nullptr;
        j.relres = nullptr;
        j.status = &sh.ticket;  // the ticket is read above; reused as the status cell
        fdm_cube<N>(j, sh);
        bool ok = block_kernel_ok<FdmLarge::kWaves>(G, 6 * NN, sh.ticket, j.code, sh);
        if (!ok) {
            j.tol = 1e-12;
            j.max_rounds = 40;
            j.max_inner = 4000;
            fdm_cube<N>(j, sh);
            ok = block_kernel_ok<FdmLarge::kWaves>(G, 6 * NN, sh.ticket, j.code, sh);
            if (threadIdx.x == 0) {
                atomicAdd((unsigned long long*)&a.stats[0], 1ull);
                if (!ok) atomicAdd((unsigned long long*)&a.stats[1], 1ull);
            }
        }
        if (threadIdx.x < 64) {
            const uint64_t base = a.ctr_dev ? *a.ctr_dev : a.counter;
            // centre cell of the raw cube (odd N, ggft_solver.cpp:649-671)
            flat_draw_wave(G, N, grad, E[(N / 2) * (1 + N + NN)], a.seed, a.qid ? a.qid[q] : base + (uint64_t)q, !ok,
                           a.out + (size_t)q * a.width);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&a.ctr[1], 1u);
        if (done == gridDim.x - 1) {  // every workgroup has read the count: the list is free again
            atomicExch(a.count, 0);
            atomicExch(&a.ctr[0], 0u);
            atomicExch(&a.ctr[1], 0u);
        }
    }
}

int nn_fallback_launch(const NnFallbackArgs& a, hipStream_t stream)
{
    if (a.B <= 0) return 0;
    if (a.grid <= 0 || a.cap <= 0) return set_error(-2, "nn_fallback_launch: no fallback buffers");
    hipLaunchKernelGGL(nn_fallback_kernel, dim3(a.grid), dim3(FdmLarge::kThreads), 0, stream, a);
    DRW_HIP(hipGetLastError());
    return 0;
}

}  // namespace drw
