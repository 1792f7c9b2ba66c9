// Host runtime helpers shared by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace drw {

// Device scratch owned by the library, one set per (device, stream) so that
// launches on different streams never share tickets or scratch.  Allocated on
// first use and kept until the process ends.
struct Workspace {
    int device = 0;
    int fdm_grid = 0;              // persistent FDM workgroups (= CUs)
    double* y_ws = nullptr;        // fdm_grid * kMaxCells
    float* s_ws = nullptr;         // fdm_grid * kMaxCells
    uint32_t* counters = nullptr;  // 8 self-resetting ticket counters
};

int get_workspace(hipStream_t stream, Workspace** out);

// Optional per-stage kernel timing with HIP events recorded on the launch
// stream around each stage (no synchronisation until collect()).
// Stages: 0 selector backbone, 1 selector head, 2 face predictor,
// 3 voxelize, 4 FDM solve, 5 flat sampler, 6 network fallback, 7 fused
// Poisson query.
struct StageTimer {
    static constexpr int kStages = 8;
    bool enabled = false;
    double total_ms[kStages] = {};
    int64_t count[kStages] = {};
    struct Pair {
        hipEvent_t a, b;
        int stage;
    };
    Pair open{};
    bool is_open = false;
    std::vector<Pair> pending;
    std::vector<hipEvent_t> pool;
    int begin(int stage, hipStream_t st);
    int end(hipStream_t st);
    int collect();  // waits for recorded events, accumulates totals
    void release();
};
const char* last_error();
void clear_error();

}  // namespace drw
