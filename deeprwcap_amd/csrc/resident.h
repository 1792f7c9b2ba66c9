// Resident serving loop: a persistent kernel that takes N = 23 network
// transition queries (Poisson and gradient) straight from a request ring in
// pinned host memory, runs each query end to end in one workgroup (selector,
// face draw, face predictor, tile draw, locate) and writes the result back
// to the ring, with no host involvement per query or per batch.
//
// Protocol (sequence numbers s = 0, 1, 2, ...; slot = s mod kResRing):
//  * host: writes req[slot] (description, kind, axis, counter), then
//    req[slot].seq = s + 1 with release semantics;
//  * device: a rate-limited scout wave reads the seq words of the next 64
//    slots and advances ctl.published over the contiguous published prefix;
//    a free workgroup takes a ticket s = ctl.head++ and serves query s once
//    s < ctl.published: it reads the request and writes the six response
//    words, each tagged with s + 1 in its high half;
//  * host: a reaper sees the result when all six words carry the tag; the
//    slot is reused for s + kResRing only after that.
// No step needs an acquire or release at system or device scope (on this
// GPU those write back or invalidate the L2, which holds the weights): the
// ring is fine-grained host memory, read and written with system-coherent
// (cache-bypassing) accesses; the request's words are read only after its
// seq word has been seen, and each response word is single-copy atomic and
// self-identifying, so its order of arrival does not matter.  The control
// block is fine-grained device memory, so plain atomic loads of it are
// coherent across the XCDs.
// Every wait on the device is bounded: a workgroup leaves when the host's
// stop flag is seen (copied to ctl.stop by the scouts) or, as a safety net,
// after idle_ticks of the 100 MHz wall clock without its ticket being
// published.  The host stops the kernel when the loop has been idle for a
// while or FDM batches need the device, and restarts it with head rewound
// to the oldest unanswered sequence.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace drw {

constexpr uint32_t kResRing = 1u << 16;  // slots of the request / response ring

struct alignas(64) ResReq {
    float layers[kMaxLayers * 2];    // [15][2]: min_z, eps
    float cuboids[kMaxCuboids * 7];  // [15][7]: box, eps
    int32_t kind;                    // DRW_POISSON / DRW_GRADIENT
    int32_t axis;                    // gradient axis (0, 1, 2)
    uint64_t qid;                    // Philox counter of the query
    uint64_t seq;                    // sequence + 1, written last by the host
};

// Response words: w[j] = (uint64_t)(uint32_t)(sequence + 1) << 32 | payload
// j, payload 0..4 the result's float bits ([x, y, z] or [x, y, z, W sign,
// eps_centre]), payload 5 the status (0: result; 1: NaN result, no drawable
// face or tile, for the FDM fallback).
constexpr int kResWords = 6;
struct alignas(64) ResResp {
    uint64_t w[8];
};

// Device-memory control block, zeroed once before the first launch.
// Fine-grained device memory; every word is read and written with
// device-scope atomics.  The words that many workgroups touch sit in cache
// lines of their own: read-modify-writes of one line serialise.
struct ResCtl {
    alignas(128) unsigned long long head;        // next ticket (one fetch-and-add per query)
    alignas(128) unsigned long long published;   // every sequence below is published (scouts' view)
    alignas(128) unsigned long long scout_next;  // wall-clock tick before which no new scout starts
    alignas(128) unsigned long long stop;        // the host's stop flag, as last read by a scout
    alignas(128) unsigned long long served[2];   // queries served, per kind (added as workgroups leave)
    unsigned long long scouts;                   // scout polls of the host ring
    unsigned long long exits_idle;               // workgroups that left after idle_ticks (safety net)
    unsigned long long exits_stop;               // workgroups that left on the stop flag
};

struct ResidentArgs {
    ResReq* req;               // device view of the host request ring
    ResResp* resp;             // device view of the host response ring
    const uint32_t* host_stop; // device view of the host stop flag
    ResCtl* ctl;               // device memory
    uint64_t seed;             // Philox key
    uint64_t idle_ticks;       // safety net: a workgroup leaves after this many 100 MHz ticks without work
    uint64_t scout_gap;        // ticks between two scout polls of the host ring
};

struct NetWeights;

// Workgroups of the resident kernel that fit the device at once (threads per
// workgroup: 256 or 512).
int resident_grid(int device, int threads, int* grid);
// Device copy of the Poisson and gradient model sets' views for the
// resident kernel (hipFree it after the last launch has finished; the model
// sets must stay loaded as long as it is used).
int resident_nets_create(const NetWeights* poisson, const NetWeights* gradient, void** out);
// Launches the resident kernel with `grid` workgroups on `stream`.
int resident_launch(const ResidentArgs& a, const void* nets, int threads, int grid, hipStream_t stream);

}  // namespace drw
