// Batched FDM solve (device side); see fdm.hip.
#pragma once

#include "common.h"

namespace drw {

struct FdmBatch {
    int B;                   // cubes
    const int32_t* N;        // [B] resolution of each cube
    const int32_t* code;     // [B] SolveCode
    const float* eps;        // packed cubes, eps[b] at eps + eps_off[b], x fastest
    const int64_t* eps_off;  // [B]
    double* G;               // packed outputs, 6 N^2 doubles at G + G_off[b]
    const int64_t* G_off;    // [B]
    const int32_t* order;    // [B] processing order (optional)
    double tol;              // stop at ||e - L y|| <= tol ||e||
    float inner_floor;       // smallest relative reduction one FP32 solve aims for
    int max_inner;           // FP32 CG iterations per refinement round
    int max_rounds;          // FP64 refinement rounds
    int32_t* iters;          // [B] total FP32 iterations (optional; -1 = unsupported N)
    float* relres;           // [B] final FP64 relative residual (optional)
    int32_t* status;         // [B] 0 converged, 1 not converged within max_rounds, 2 unsupported N (optional)
    const int32_t* B_dev;    // device count (optional): only tickets t < min(B, *B_dev) are solved
    double* y_ws;            // [grid * kMaxCells] workspace
    float* s_ws;             // [grid * kMaxCells] workspace
    uint32_t* counters;      // [4] zero at rest; the kernels leave them zero
    int n_hint;              // N of every cube when all are equal (one size class launches), else 0
};

// Smallest relative reduction one FP32 inner solve aims for.  At the query
// tolerance 1e-6 the first solve then targets 0.3 tol, so most cubes converge
// in one FP32 solve and one FP64 residual check instead of two refinement
// rounds (N = 23 batch: 69 iterations and 11.9 ms against 72 and 13.3 ms with
// a floor of 1e-5); FP32 CG still reaches it, its recursive residual keeps
// falling after the true one levels off, and the FP64 check catches that.
constexpr float kFdmInnerFloor = 3e-7f;

bool fdm_supported_n(int n);
int fdm_grid_size(int device);
int fdm_launch(const FdmBatch& a, int grid, hipStream_t stream);

// Acceptance of FDM kernels, the checks of ggft_solver::solve
// (ggft/src/ggft_solver.cpp:565-626): the solve converged, G is finite, and a
// Poisson kernel (GFT) has |sum G - 1| <= 0.01 and no negative panel.
//   first pass (list_in == null): every cube b < B that fails is appended to
//     list_out (count_out, which the caller has zeroed);
//   second pass (list_in, count_in): every listed cube that still fails gets
//     bad[b] = 1 and stats[1] += 1; stats[0] += count_in (cubes re-solved).
struct FdmAcceptArgs {
    int B;
    const int32_t* N;
    const int32_t* code;
    const double* G;
    const int64_t* G_off;
    const int32_t* status;
    const int32_t* list_in;   // second pass
    const int32_t* count_in;  // second pass
    int32_t* list_out;        // first pass
    int32_t* count_out;       // first pass
    uint8_t* bad;             // second pass: per query
    uint64_t* stats;          // second pass: [0] re-solved, [1] failed
};
int fdm_accept_launch(const FdmAcceptArgs& a, hipStream_t stream);

// Network-path fallback: the queries whose network result holds a NaN (a
// failed draw: non-finite or all-zero kernel or probabilities) are solved by
// the FDM at N = 23 and drawn with their own Philox counters, in one
// persistent launch that exits at once when there are none.  The face
// kernels append those queries to list (count, stats[2]; stats[1] beyond
// cap) as they write them; the fallback's workgroups take listed queries by
// ticket, voxelize the description (raw permittivities), solve at tol (FP64
// mode on a failed acceptance, as in the FDM path) and write the hop to
// out[q]; the last one to finish zeroes count.
struct NnFallbackArgs {
    int B, width, cap;           // queries in the batch, floats per result, list capacity
    float* out;                  // [B][width] network results, overwritten for listed queries
    const float* layers;         // [B][15][2]
    const float* cuboids;        // [B][15][7]
    const int32_t* axis;         // [B] (gradient) or null
    int32_t* list;               // [cap]
    int32_t* count;              // [1] appended by the face kernels, zeroed by the fallback
    uint64_t* stats;             // [0] FDM re-solves, [1] failed, [2] network fallbacks
    uint64_t seed, counter;      // Philox key and first query counter
    const uint64_t* qid;         // [B] per-query counters (optional)
    const uint64_t* ctr_dev;     // device-resident counter base (optional)
    double tol;
    int grid;                    // workgroups of the solve (slots of the buffers below)
    float* eps;                  // [grid][kMaxCells] slot cubes
    double* G;                   // [grid][6 kNN] slot kernels
    double* y_ws;                // [grid][kMaxCells]
    uint32_t* ctr;               // [2] ticket counters, zero at rest
};
int nn_fallback_launch(const NnFallbackArgs& a, hipStream_t stream);

}  // namespace drw
