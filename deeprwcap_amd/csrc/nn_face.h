// Device weight views and launch interface of the face-predictor kernel.
#pragma once

#include "common.h"
#include "nn_common.h"

namespace drw {

// Folded (batch-norm absorbed) weights of one FaceSolver / FaceSolverPlus.
// Layer l of the 1x1 convolutions: l = 0 is z_weight (25 -> C0), l = i + 1
// the pointwise conv of block i.  MFMA layers use frag[l] (B fragments,
// [n_tile][k_tile][64 lanes] of half4), the small VALU layers pw_w[l]
// ([cout][cin] fp32).  dw_w[i] is [cin][9] (ky-major), dw_b[i] [cin].
// frag_lo[l] holds the rounding residue W - fp16(W) of MFMA layer l: the
// 1x1 products run in split precision, W a ~= W_hi a + W_lo a.
struct FaceNetDev {
    const half4_t* frag[9];
    const half4_t* frag_lo[9];
    const float* pw_w[9];
    const float* pw_bias[9];
    const float* dw_w[8];
    const float* dw_b[8];
    const float* head_w;
    const float2* gtab;  // GELU table (gelu_tab), copied into LDS by every workgroup
    float head_b;
};

struct FaceQueryArgs {
    int B;                   // queries
    int full;                // 1: evaluate all 6 faces, write kernels only
    const float* layers;     // [B, 15, 2]
    const float* cuboids;    // [B, 15, 7]
    const int32_t* face;     // [B] selected face (sample mode)
    const int32_t* axis;     // [B] gradient axis (null for Poisson)
    const float* weight;     // [B] predicted W (gradient)
    float* out;              // [B, 3 | 5] (sample mode)
    float* kern_out;         // [items, 529] normalised (post-processed) kernels, optional in sample mode
    uint64_t seed, counter;
    const uint64_t* qid;     // [B] per-query counters (optional; else counter + b)
    const uint64_t* ctr_dev; // device-resident counter base (optional)
    int32_t* fb_list;        // failed queries for the FDM fallback (optional)
    int32_t* fb_count;
    int fb_cap;
    uint64_t* fb_stats;
    float* probs_out;        // fused Poisson query: [B, 6] face probabilities (optional)
    int32_t* face_out;       // fused Poisson query: [B] drawn face (optional)
    const float* feat_in;    // Poisson query: [B, 128] PoissonSelector features from the backbone
                             // kernel; the query kernel then runs the head and the predictor only
};

// mode 0: PoissonPredictor, 1: Gradient1Predictor, 2: Gradient2Predictor.
int face_query_launch(int mode, const FaceQueryArgs& a, const FaceNetDev& w, hipStream_t stream);

struct SelNetDev;
// The whole Poisson query in one kernel (sample mode): one workgroup per
// query runs the PoissonSelector, its Linear head, the face draw and the
// PoissonPredictor on that face, with threads = 256 or 512 per workgroup.
// With a.feat_in the selector backbone has already run (its kernel wrote
// the features) and the workgroup starts at the head.
int poisson_query_launch(const FaceQueryArgs& a, const SelNetDev& sel, const FaceNetDev& pred, int threads,
                         hipStream_t stream);

}  // namespace drw
