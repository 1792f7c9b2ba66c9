// Launch interfaces of the voxelizer and samplers (device pointers throughout).
#pragma once

#include "common.h"

namespace drw {

struct VoxelArgs {
    int B;
    const float* layers;     // [B, 15, 2]
    const float* cuboids;    // [B, 15, 7]
    const int32_t* N;        // [B] per-query resolution, or null -> fixed_N
    int fixed_N;
    int normalise;           // divide eps by the per-query max (network input)
    float* out;              // cubes, [z][y][x]
    const int64_t* out_off;  // [B] float offsets, or null -> b * N^3
    float* maxv;             // [B] per-query max (optional)
};
int voxelize_launch(const VoxelArgs& a, hipStream_t stream);

struct FlatSampleArgs {
    int B;
    const int32_t* N;          // [B]
    const int32_t* axis;       // [B] gradient axis, or null for Poisson queries
    const double* G;           // FDM kernels, GGFT panel order
    const int64_t* G_off;      // [B]
    const float* center_eps;   // [B] (gradient only; optional)
    uint64_t seed, counter;    // Philox key and first query counter
    float* out;                // [B, 3] or [B, 5]
    const uint64_t* qid;       // [B] per-query counters (optional; else counter + b)
    const uint64_t* ctr_dev;   // device-resident counter base (optional; replaces counter)
    const uint8_t* bad;        // [B] queries whose kernel failed acceptance: NaN results (optional)
};
int sample_flat_launch(const FlatSampleArgs& a, hipStream_t stream);

int center_eps_launch(int B, const int32_t* N, const int32_t* axis, const float* eps, const int64_t* eps_off,
                      float* out, hipStream_t stream);

}  // namespace drw

namespace drw {
int advance_counter_launch(uint64_t* ctr, uint64_t n, hipStream_t stream);
// Per-query solve metadata of a uniform-N batch; also zeroes bad[B] and the
// acceptance list count (both optional).
int fill_meta_launch(int B, int N, const int32_t* axis, int32_t* Nq, int32_t* code, int64_t* eps_off, int64_t* G_off,
                     uint8_t* bad, int32_t* count, hipStream_t stream);
}  // namespace drw
