// Limits of a query's structure description, shared by the device code and
// the host-only walker geometry.
#pragma once

namespace drw {

constexpr int kMaxLayers = 15;    // inference_cpp/src/dnn_solver.h:29
constexpr int kMaxCuboids = 15;   // inference_cpp/src/dnn_solver.h:28
constexpr float kPad = -9999.0f;  // inference_cpp/src/cuda_kernels.cu:9

}  // namespace drw
