// Host runtime helpers shared by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
const char* last_error();
void clear_error();

}  // namespace drw
