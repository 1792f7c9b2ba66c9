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
    double* y_ws;            // [grid * kMaxCells] workspace
    float* s_ws;             // [grid * kMaxCells] workspace
    uint32_t* counters;      // [4] zero at rest; the kernels leave them zero
    int n_hint;              // N of every cube when all are equal (one size class launches), else 0
};

bool fdm_supported_n(int n);
int fdm_grid_size(int device);
int fdm_launch(const FdmBatch& a, int grid, hipStream_t stream);

}  // namespace drw
