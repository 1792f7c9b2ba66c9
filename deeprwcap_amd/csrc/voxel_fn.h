// Voxel values of the layer/cuboid structure description (device functions
// shared by the voxelizer and the network-path FDM fallback).
#pragma once

#include "common.h"

namespace drw {

// Per-query maximum permittivity over all layer and cuboid entries (padded
// entries are -9999 and never win), dnn_solver.cpp:274-282.
__device__ __forceinline__ float query_max(const float* Ly, const float* Cb)
{
    float m = Ly[1];
    for (int l = 1; l < kMaxLayers; l++) m = fmaxf(m, Ly[2 * l + 1]);
    for (int c = 0; c < kMaxCuboids; c++) m = fmaxf(m, Cb[7 * c + 6]);
    return m;
}

// fill_structure_kernel semantics (cuda_kernels.cu:21-84): the last layer
// whose min_z <= z_centre (list ends at the first padded min_z), overridden by
// the first cuboid whose closed box holds the voxel centre.
__device__ __forceinline__ float voxel_value(const float* Ly, const float* lv, const float* Cb, const float* cv,
                                             float xc, float yc, float zc)
{
    int li = 0;
    for (int l = 0; l < kMaxLayers; l++) {
        const float mz = Ly[2 * l];
        if (mz == kPad) break;
        if (zc >= mz) li = l; else break;
    }
    float v = lv[li];
    for (int c = 0; c < kMaxCuboids; c++) {
        const float* q = Cb + 7 * c;
        if (q[0] == kPad) break;
        if (xc >= q[0] && xc <= q[3] && yc >= q[1] && yc <= q[4] && zc >= q[2] && zc <= q[5]) { v = cv[c]; break; }
    }
    return v;
}

}  // namespace drw
