// Network-backed transition queries (the five predictor networks at N = 23).
#pragma once

#include "common.h"
#include "nn_face.h"
#include "nn_sel.h"
#include "runtime.h"

namespace drw {

struct NetWeights;  // device-resident packed weights of one solver kind

// Loads the index-addressed model list (models.txt semantics,
// inference_cpp/src/dnn_solver.cpp:65-87): entries 0,1 for Poisson solvers,
// 2,3,4 for gradient solvers (dnn_solver_poisson.cpp:18-20,
// dnn_solver_gradient.cpp:19-22).
int net_weights_load(const char* models_txt, int kind, NetWeights** out);
void net_weights_free(NetWeights* w);

// Device views of a loaded model set: its selector, and its predictors
// (which: 0 PoissonPredictor, 1 Gradient1Predictor, 2 Gradient2Predictor).
const SelNetDev& net_selector(const NetWeights* w);
const FaceNetDev& net_face(const NetWeights* w, int which);

struct QueryArgs {
    int B;
    int full;               // 1: all six faces, kernels only (no draws)
    const float* layers;    // [B, 15, 2]
    const float* cuboids;   // [B, 15, 7]
    const int32_t* axis;    // [B] (gradient) or null (Poisson)
    float* out;             // [B, 3] or [B, 5] (sample mode)
    // optional diagnostics (device pointers, may be null)
    float* probs_out;       // [B, 6] selector face probabilities
    int32_t* face_out;      // [B] drawn face
    float* weight_out;      // [B] predicted W (gradient)
    float* kern_out;        // sample: [B, 529] kernel of the drawn face; full: [B, 6, 529]
    uint64_t seed, counter;
    const uint64_t* qid;    // [B] per-query Philox counters (optional; else counter + b)
    const uint64_t* ctr_dev;  // device-resident counter base (optional; replaces counter)
    StageTimer* timer;      // optional
    // FDM fallback of failed rows (optional: set when nn_fallback_launch follows)
    int32_t* fb_list;       // [fb_cap] failed queries, appended by the face kernels
    int32_t* fb_count;      // [1] zero at rest; the fallback kernel resets it
    int fb_cap;
    uint64_t* fb_stats;     // [1] failed beyond the list, [2] network fallbacks
};

// Enqueues the fused network query; *scratch (grown as needed, *cap bytes)
// holds per-query intermediates.
int query_nn_launch(const NetWeights* w, const QueryArgs& q, uint8_t** scratch, size_t* cap, hipStream_t stream);

}  // namespace drw
