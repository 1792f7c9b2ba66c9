/*
 * CPU oracle for the deepRWCap transition-kernel path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithms on
 * the CPU so that tests/, the smoke check and the CPU-baseline leg of the
 * benchmark can check the HIP library against it.  No product code path links
 * or calls it; the library under deeprwcap_amd/ never falls back to it.
 *
 * Contents (each function cites the reference file:line it restates):
 *   - FDM surface Green's function, FP64, matrix-free 7-point operator with
 *     harmonic-mean face permittivities and Dirichlet 0 at half a cell;
 *     Jacobi-preconditioned CG (the reference uses Eigen BiCGSTAB on the
 *     D-scaled matrix A = L D^-1; the exact solution is the same)
 *   - right-hand sides GFT / WVTZ / WVTX / WVTY for odd and even N
 *   - GGFT-style dense assembly of A (tiny N only) for cross-checks
 *   - voxelization of the layer/cuboid description
 *   - face-rotation, post-process, axis pre-rotation and locate tables
 *   - Philox4x32-10 and the fixed-point inverse-CDF sampler
 *
 * Parity pins: see tests/test_oracle_*.py (closed forms, dense GGFT-form
 * solves, GGFT golden vectors at N=23 built from the reference sources,
 * invariants, torch.rot90 for the axis pre-rotation).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Grid indexing: cell c = i + j*N + k*N*N (x fastest), ggft_solver.h:28,48  */
/* ------------------------------------------------------------------------ */
static inline int cidx(int N, int i, int j, int k) { return i + j * N + k * N * N; }

/* Operator L (SPD):
 *   (L y)_c = sum_{n in nbr(c)} w_cn (y_c - y_n) + 2 eps_c nb(c) y_c,
 *   w_cn = 2 eps_c eps_n / (eps_c + eps_n)
 * GGFT stores A = L D^-1 with D = diag(2 eps)  (ggft_solver.cpp:447-518). */
static inline double wface(double a, double b) { return 2.0 * a * b / (a + b); }

void orc_fdm_apply_L(int N, const double* eps, const double* y, double* out)
{
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                int c = cidx(N, i, j, k);
                double ec = eps[c], yc = y[c], acc = 0.0;
                int nb = 0;
#define NBR(cond, n)                                              \
    if (cond) { double w = wface(ec, eps[n]); acc += w * (yc - y[n]); } \
    else { nb++; }
                NBR(i > 0, c - 1)
                NBR(i < N - 1, c + 1)
                NBR(j > 0, c - N)
                NBR(j < N - 1, c + N)
                NBR(k > 0, c - N * N)
                NBR(k < N - 1, c + N * N)
#undef NBR
                out[c] = acc + 2.0 * ec * nb * yc;
            }
}

void orc_fdm_diag(int N, const double* eps, double* d)
{
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                int c = cidx(N, i, j, k);
                double ec = eps[c], acc = 0.0;
                int nb = 0;
#define NBR(cond, n) if (cond) acc += wface(ec, eps[n]); else nb++;
                NBR(i > 0, c - 1)
                NBR(i < N - 1, c + 1)
                NBR(j > 0, c - N)
                NBR(j < N - 1, c + N)
                NBR(k > 0, c - N * N)
                NBR(k < N - 1, c + N * N)
#undef NBR
                d[c] = acc + 2.0 * ec * nb;
            }
}

/* Right-hand side e, restating ggft_solver.cpp:137-274 (update_ek).
 * code: 0 = GFT, 1 = WVTZ, 2 = WVTX, 3 = WVTY.  e must hold N^3 zeros. */
void orc_fdm_rhs(int N, const double* D, int code, double* e)
{
    const double h = 2.0 / N;
    const int H = N / 2;
    memset(e, 0, sizeof(double) * (size_t)N * N * N);
    if (N % 2) {
        int zzz = cidx(N, H, H, H);
        if (code == 0) {
            e[zzz] = 1.0;                                   /* :142-143 */
        } else {
            int p, m;                                       /* :144-177 */
            if (code == 1) { p = cidx(N, H, H, H + 1); m = cidx(N, H, H, H - 1); }
            else if (code == 2) { p = cidx(N, H + 1, H, H); m = cidx(N, H - 1, H, H); }
            else { p = cidx(N, H, H + 1, H); m = cidx(N, H, H - 1, H); }
            double d_plus = D[p] + D[zzz];
            double d_minus = -(D[zzz] + D[m]);
            e[p] = D[p] / d_plus / h;
            e[zzz] = D[zzz] * (1. / d_plus + 1. / d_minus) / h;
            e[m] = D[m] / d_minus / h;
        }
        return;
    }
    /* even N: the 8 cells around the centre vertex (:178-273) */
    int zzz = cidx(N, H, H, H), mzz = cidx(N, H - 1, H, H), zmz = cidx(N, H, H - 1, H),
        zzm = cidx(N, H, H, H - 1), mmz = cidx(N, H - 1, H - 1, H), mzm = cidx(N, H - 1, H, H - 1),
        zmm = cidx(N, H, H - 1, H - 1), mmm = cidx(N, H - 1, H - 1, H - 1);
    double dsum = D[zzz] + D[zmz] + D[zzm] + D[zmm] + D[mzz] + D[mmz] + D[mzm] + D[mmm];
    if (code == 0) {
        int cells[8] = {zzz, mzz, zmz, zzm, mmz, mzm, zmm, mmm};
        for (int q = 0; q < 8; q++) e[cells[q]] = D[cells[q]] / dsum;
        return;
    }
    int plus[4], minus[4];
    if (code == 1) {        /* WVTZ: plus = k==H, minus = k==H-1 */
        int p4[4] = {zzz, mzz, zmz, mmz}, m4[4] = {zzm, mzm, zmm, mmm};
        memcpy(plus, p4, sizeof p4); memcpy(minus, m4, sizeof m4);
    } else if (code == 2) { /* WVTX: plus = i==H, minus = i==H-1 */
        int p4[4] = {zzz, zmz, zzm, zmm}, m4[4] = {mzz, mmz, mzm, mmm};
        memcpy(plus, p4, sizeof p4); memcpy(minus, m4, sizeof m4);
    } else {                /* WVTY: plus = j==H, minus = j==H-1 */
        int p4[4] = {zzz, mzz, zzm, mzm}, m4[4] = {zmz, mmz, zmm, mmm};
        memcpy(plus, p4, sizeof p4); memcpy(minus, m4, sizeof m4);
    }
    double d_minus = -(D[minus[0]] + D[minus[1]] + D[minus[2]] + D[minus[3]]);
    for (int q = 0; q < 4; q++) e[plus[q]] = D[plus[q]] / dsum / h * 2;
    for (int q = 0; q < 4; q++) e[minus[q]] = D[minus[q]] * (1. / dsum + 1. / d_minus) / h * 2;
}

/* Boundary panel -> cell map, GGFT numbering (ggft_solver.h:65-77):
 * face 0 z-: i+jN, 1 z+: N^2+i+jN, 2 y-: 2N^2+i+kN, 3 y+: 3N^2+i+kN,
 * face 4 x-: 4N^2+j+kN, 5 x+: 5N^2+j+kN. */
void orc_panel_cells(int N, int* cell_of_panel)
{
    int NN = N * N;
    for (int b = 0; b < N; b++)
        for (int a = 0; a < N; a++) {
            int t = a + b * N;
            cell_of_panel[0 * NN + t] = cidx(N, a, b, 0);
            cell_of_panel[1 * NN + t] = cidx(N, a, b, N - 1);
            cell_of_panel[2 * NN + t] = cidx(N, a, 0, b);
            cell_of_panel[3 * NN + t] = cidx(N, a, N - 1, b);
            cell_of_panel[4 * NN + t] = cidx(N, 0, a, b);
            cell_of_panel[5 * NN + t] = cidx(N, N - 1, a, b);
        }
}

/* Jacobi-preconditioned CG on L y = e (FP64).  Returns the iteration count,
 * or -1 if maxit was reached.  relres receives ||e - L y|| / ||e||. */
int orc_fdm_cg(int N, const double* eps, const double* e, double* y, double tol, int maxit, double* relres)
{
    size_t n = (size_t)N * N * N;
    double *r = malloc(n * 8), *z = malloc(n * 8), *p = malloc(n * 8), *q = malloc(n * 8), *d = malloc(n * 8);
    orc_fdm_diag(N, eps, d);
    orc_fdm_apply_L(N, eps, y, q);
    double bnorm = 0, rz = 0;
    for (size_t c = 0; c < n; c++) {
        r[c] = e[c] - q[c];
        z[c] = r[c] / d[c];
        p[c] = z[c];
        rz += r[c] * z[c];
        bnorm += e[c] * e[c];
    }
    bnorm = sqrt(bnorm);
    int it = 0, ok = 0;
    double rn = 0;
    for (it = 0; it < maxit; it++) {
        rn = 0;
        for (size_t c = 0; c < n; c++) rn += r[c] * r[c];
        if (sqrt(rn) <= tol * bnorm) { ok = 1; break; }
        orc_fdm_apply_L(N, eps, p, q);
        double pq = 0;
        for (size_t c = 0; c < n; c++) pq += p[c] * q[c];
        double alpha = rz / pq, rz_new = 0;
        for (size_t c = 0; c < n; c++) {
            y[c] += alpha * p[c];
            r[c] -= alpha * q[c];
            z[c] = r[c] / d[c];
            rz_new += r[c] * z[c];
        }
        double beta = rz_new / rz;
        rz = rz_new;
        for (size_t c = 0; c < n; c++) p[c] = z[c] + beta * p[c];
    }
    /* true residual */
    orc_fdm_apply_L(N, eps, y, q);
    rn = 0;
    for (size_t c = 0; c < n; c++) rn += (e[c] - q[c]) * (e[c] - q[c]);
    if (relres) *relres = bnorm > 0 ? sqrt(rn) / bnorm : 0.0;
    free(r); free(z); free(p); free(q); free(d);
    return ok ? it : -1;
}

/* Full GGFT solve: G_b = 2 eps_c y_c for every boundary panel b
 * (ggft_solver.cpp:592-596, with x = D y).  Returns CG iterations (<0 on
 * non-convergence). */
int orc_fdm_solve(int N, const double* eps, int code, double tol, int maxit, double* G, double* relres)
{
    size_t n = (size_t)N * N * N;
    double* e = malloc(n * 8);
    double* y = calloc(n, 8);
    int* cells = malloc(sizeof(int) * 6 * N * N);
    orc_fdm_rhs(N, eps, code, e);
    int it = orc_fdm_cg(N, eps, e, y, tol, maxit, relres);
    orc_panel_cells(N, cells);
    for (int b = 0; b < 6 * N * N; b++) G[b] = 2.0 * eps[cells[b]] * y[cells[b]];
    free(e); free(y); free(cells);
    return it;
}

/* GGFT's own matrix A (row-major dense, n x n) for tiny N, exactly as
 * dielectric_opt assembles it (ggft_solver.cpp:447-518):
 *   A(r,c) = -eps_r/(eps_r+eps_c) for neighbours, A(c,c) = 6 - sum_n eps_c/(eps_c+eps_n). */
void orc_ggft_dense_A(int N, const double* eps, double* A)
{
    int n = N * N * N;
    memset(A, 0, sizeof(double) * (size_t)n * n);
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                int c = cidx(N, i, j, k);
                double diag = 6.0;
#define NBR(cond, m)                                               \
    if (cond) {                                                    \
        A[(size_t)(m) * n + c] = -eps[m] / (eps[m] + eps[c]);      \
        diag -= eps[c] / (eps[c] + eps[m]);                        \
    }
                NBR(i > 0, c - 1)
                NBR(i < N - 1, c + 1)
                NBR(j > 0, c - N)
                NBR(j < N - 1, c + N)
                NBR(k > 0, c - N * N)
                NBR(k < N - 1, c + N * N)
#undef NBR
                A[(size_t)c * n + c] = diag;
            }
}

/* Centre permittivity used by the FDM sampler prep (ggft_solver.cpp:649-671):
 * odd N -> centre cell; even N -> mean of the 4 cells on the minus side of the
 * gradient axis (code 1 WVTZ, 2 WVTX, 3 WVTY). */
double orc_center_eps(int N, const double* D, int code)
{
    int H = N / 2;
    if (N % 2) return D[cidx(N, H, H, H)];
    int zzm = cidx(N, H, H, H - 1), mmm = cidx(N, H - 1, H - 1, H - 1), zmm = cidx(N, H, H - 1, H - 1),
        mzm = cidx(N, H - 1, H, H - 1), mzz = cidx(N, H - 1, H, H), mmz = cidx(N, H - 1, H - 1, H),
        zmz = cidx(N, H, H - 1, H);
    if (code == 2) return (D[mzz] + D[mmz] + D[mzm] + D[mmm]) / 4;
    if (code == 3) return (D[zmz] + D[mmz] + D[zmm] + D[mmm]) / 4;
    return (D[zzm] + D[mmm] + D[zmm] + D[mzm]) / 4;
}

/* ------------------------------------------------------------------------ */
/* Voxelization, restating fill_structure_kernel (cuda_kernels.cu:21-84) and */
/* the per-query normalisation of structureConstruction (dnn_solver.cpp:274-285). */
/* layers [B,15,2] (min_z, eps), cuboids [B,15,7]; out [B,N,N,N] (z,y,x).    */
/* ------------------------------------------------------------------------ */
#define ORC_PAD (-9999.0f)
#define ORC_L 15
#define ORC_C 15

float orc_query_max(const float* layer, const float* cub)
{
    float m = layer[1];
    for (int l = 1; l < ORC_L; l++) m = fmaxf(m, layer[2 * l + 1]);
    for (int c = 0; c < ORC_C; c++) m = fmaxf(m, cub[7 * c + 6]);
    return m;
}

void orc_voxelize(int B, int N, const float* layers, const float* cuboids, int normalise, float* out, float* maxv)
{
    for (int b = 0; b < B; b++) {
        const float* Ly = layers + (size_t)b * ORC_L * 2;
        const float* Cb = cuboids + (size_t)b * ORC_C * 7;
        float mx = orc_query_max(Ly, Cb);
        if (maxv) maxv[b] = mx;
        float lv[ORC_L], cv[ORC_C];
        for (int l = 0; l < ORC_L; l++) lv[l] = normalise ? Ly[2 * l + 1] / mx : Ly[2 * l + 1];
        for (int c = 0; c < ORC_C; c++) cv[c] = normalise ? Cb[7 * c + 6] / mx : Cb[7 * c + 6];
        for (int z = 0; z < N; z++)
            for (int y = 0; y < N; y++)
                for (int x = 0; x < N; x++) {
                    const float xc = (2.0f * (float)x + 1.0f) / (float)N - 1.0f;
                    const float yc = (2.0f * (float)y + 1.0f) / (float)N - 1.0f;
                    const float zc = (2.0f * (float)z + 1.0f) / (float)N - 1.0f;
                    int li = 0;
                    for (int l = 0; l < ORC_L; l++) {
                        float mz = Ly[2 * l];
                        if (mz == ORC_PAD) break;
                        if (zc >= mz) li = l; else break;
                    }
                    float v = lv[li];
                    for (int c = 0; c < ORC_C; c++) {
                        const float* q = Cb + 7 * c;
                        if (q[0] == ORC_PAD) break;
                        if (xc >= q[0] && xc <= q[3] && yc >= q[1] && yc <= q[4] && zc >= q[2] && zc <= q[5]) {
                            v = cv[c];
                            break;
                        }
                    }
                    out[(size_t)b * N * N * N + (size_t)z * N * N + y * N + x] = v;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Index remaps.  All produce out[z][y][x] = in[src].                        */
/* ------------------------------------------------------------------------ */
/* rotate_faces_poisson_kernel (cuda_kernels.cu:299-380) */
void orc_rotate_poisson(int N, int face, const float* in, float* out)
{
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                int sz = z, sy = y, sx = x;
                switch (face) {
                case 1: sz = N - 1 - z; break;
                case 2: sz = y; sy = z; break;
                case 3: sz = y; sy = N - 1 - z; break;
                case 4: sz = y; sy = x; sx = z; break;
                case 5: sz = y; sy = x; sx = N - 1 - z; break;
                default: break;
                }
                out[(z * N + y) * N + x] = in[(sz * N + sy) * N + sx];
            }
}

/* rotate_faces_gradient_kernel (cuda_kernels.cu:411-486) */
void orc_rotate_gradient(int N, int face, const float* in, float* out)
{
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                int sz = z, sy = y, sx = x;
                switch (face) {
                case 0: sz = N - 1 - z; break;
                case 3: sy = N - 1 - y; break;
                case 4: sy = x; sx = y; break;
                case 5: sy = x; sx = N - 1 - y; break;
                default: break;
                }
                out[(z * N + y) * N + x] = in[(sz * N + sy) * N + sx];
            }
}

/* Axis pre-rotation of dnn_solver_gradient.cpp:228-238:
 * axis 0: rot90(k=-1, dims (z,x)); axis 1: rot90(k=-1, dims (z,y)); axis 2: identity. */
void orc_axis_rotate(int N, int axis, const float* in, float* out)
{
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                int sz = z, sy = y, sx = x;
                if (axis == 0) { sz = N - 1 - x; sx = z; }
                else if (axis == 1) { sz = N - 1 - y; sy = z; }
                out[(z * N + y) * N + x] = in[(sz * N + sy) * N + sx];
            }
}

/* post_process_gradient_kernel (cuda_kernels.cu:516-578) on one [N*N] face */
void orc_post_process(int N, int face, int axis, const float* in, float* out)
{
    int axX = axis == 0, axY = axis == 1;
    int f01 = face == 0 || face == 1, f45 = face == 4 || face == 5;
    int rot = axX || (axY && f45), flip = (axX && f45) || (axY && f01);
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            int sy = y, sx = x;
            if (rot && flip) { sy = x; sx = y; }
            else if (rot) { sy = x; sx = N - 1 - y; }
            else if (flip) { sy = N - 1 - y; }
            float v = in[sy * N + sx];
            out[y * N + x] = face == 0 ? -v : v;
        }
}

/* locate_index_kernel / locate_index_simple_kernel (cuda_kernels.cu:122-264).
 * axis < 0 selects the simple (no remap) variant. */
void orc_locate(int N, int tile, int face, int axis, float* out3)
{
    static const int fv[6][3] = {{1, 2, -3}, {1, 2, 3}, {1, -3, 2}, {1, 3, 2}, {-3, 1, 2}, {3, 1, 2}};
    int xid = tile % N, yid = tile / N;
    float t[3] = {(2.0f * xid + 1.0f) / N - 1.0f, (2.0f * yid + 1.0f) / N - 1.0f, 1.0f};
    int f = face;
    if (axis == 0) { if (f == 0) f = 4; else if (f == 1) f = 5; else if (f == 4) f = 1; else if (f == 5) f = 0; }
    else if (axis == 1) { if (f == 0) f = 2; else if (f == 1) f = 3; else if (f == 2) f = 1; else if (f == 3) f = 0; }
    for (int j = 0; j < 3; j++) {
        int v = fv[f][j];
        out3[j] = t[abs(v) - 1] * (v < 0 ? -1.0f : 1.0f);
    }
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Multipliers and Weyl constants as published.  */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform words for query q of a stream: counter = (q_lo, q_hi, stream, 0),
 * key = (seed_lo, seed_hi). */
void orc_query_uniforms(uint64_t seed, uint64_t q, uint32_t stream, uint32_t out[4])
{
    uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* Fixed-point inverse-CDF draw over |w[0..M)|.
 * q_i = trunc(|w_i| * 2^s) with the power of two s chosen so that the largest
 * magnitude lands in [2^23, 2^24); cumulative sums are exact integers, so the
 * result does not depend on summation order.  The drawn bin is the smallest i
 * whose cumulative sum exceeds floor(u * total / 2^32).  Returns -1 if all
 * weights are zero or the largest is not finite. */
int orc_cdf_draw_f32(const float* w, int M, uint32_t u)
{
    float m = 0.0f;
    for (int i = 0; i < M; i++) { float a = fabsf(w[i]); if (a > m) m = a; }
    if (!(m > 0.0f) || isinf(m)) return -1;
    int E;
    frexpf(m, &E);
    int s = 24 - E;
    uint64_t total = 0;
    for (int i = 0; i < M; i++) total += (uint32_t)ldexpf(fabsf(w[i]), s);
    uint64_t r = (uint64_t)(((unsigned __int128)u * total) >> 32);
    uint64_t acc = 0;
    for (int i = 0; i < M; i++) {
        acc += (uint32_t)ldexpf(fabsf(w[i]), s);
        if (acc > r) return i;
    }
    return M - 1;
}

int orc_cdf_draw_f64(const double* w, int M, uint32_t u)
{
    double m = 0.0;
    for (int i = 0; i < M; i++) { double a = fabs(w[i]); if (a > m) m = a; }
    if (!(m > 0.0) || isinf(m)) return -1;
    int E;
    frexp(m, &E);
    int s = 24 - E;
    uint64_t total = 0;
    for (int i = 0; i < M; i++) total += (uint32_t)ldexp(fabs(w[i]), s);
    uint64_t r = (uint64_t)(((unsigned __int128)u * total) >> 32);
    uint64_t acc = 0;
    for (int i = 0; i < M; i++) {
        acc += (uint32_t)ldexp(fabs(w[i]), s);
        if (acc > r) return i;
    }
    return M - 1;
}
