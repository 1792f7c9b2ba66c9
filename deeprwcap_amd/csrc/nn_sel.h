// Device weight views and launch interface of the face selectors.
#pragma once

#include "common.h"
#include "nn_common.h"

namespace drw {

// Folded weights of FaceSelector / FaceSelectorWeight.  frag[0..2]: the three
// Conv3d layers as MFMA B fragments ([n_tile][k_tile][64] half4, k = ci*27 +
// tap); frag[3..5]: the Linear layers.  bias[l] fp32 per output.
struct SelNetDev {
    const half4_t* frag[6];
    const float* bias[6];
    const half_t* conv1_w;  // first Conv3d for the direct form: per channel 14 fp16 tap pairs (conv1_pairs)
    const float2* gtab;     // GELU table of the first Conv3d (kSelGeluTab entries)
};

struct SelBackboneArgs {
    int B;
    const float* layers;   // [B, 15, 2]
    const float* cuboids;  // [B, 15, 7]
    const int32_t* axis;   // [B] gradient axis pre-rotation, or null
    float* feat;           // [B, C3 * 8]
};

struct SelHeadArgs {
    int B;
    const float* feat;     // [B, F]
    float* probs;          // [B, 6] (optional)
    float* weight;         // [B] raw W logit (FaceSelectorWeight, optional)
    int32_t* face;         // [B] drawn face (optional)
    uint64_t seed, counter;
    const uint64_t* qid;   // [B] per-query counters (optional; else counter + b)
    const uint64_t* ctr_dev;  // device-resident counter base (optional)
};

int selector_backbone_launch(bool weighted, const SelBackboneArgs& ba, const SelNetDev& w, hipStream_t stream);
int selector_head_launch(bool weighted, const SelHeadArgs& ha, const SelNetDev& w, hipStream_t stream);

}  // namespace drw
