// tx_context.hpp -- per-device state of the C ABI: stream, scratch pools, arch.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <stddef.h>
#include <stdint.h>

#include "tx_pipeline.hpp"

namespace txc {

struct DeviceContext {
    int device = -1;
    char arch[64] = {0};
    int cus = 0;
    hipStream_t stream = nullptr;  // default priority, non-blocking
    std::mutex mu;                 // guards the staging pool and host-staged calls
    uint8_t* pool = nullptr;
    size_t pool_cap = 0;

    // Binned-pipeline scratch, shared by every stream of the device.  Uses
    // are ordered through `pipe_done`: a use on a caller's stream waits for
    // it and records it afterwards; uses on the library's own `stream` (which
    // lives as long as the context) are ordered by that stream and recorded
    // only when another stream or a buffer change needs them covered
    // (`pipe_lib_pending`: the last use was on `stream` and is not recorded
    // yet) -- an event record per launch cost ~5 us between launches.
    std::mutex pipe_mu;
    uint8_t* pipe = nullptr;
    uint64_t pipe_recs = 0;
    hipEvent_t pipe_done = nullptr;
    bool pipe_lib_pending = false;

    // Device staging memory of at least `bytes` (grown geometrically); call with mu held.
    uint8_t* scratch(size_t bytes);
    // Pipeline buffers for `recs` records; call with pipe_mu held.
    bool pipe_scratch(uint64_t recs, txp::PipeBufs* out);
    // Orders a use of the pipeline scratch on `s` after every earlier use
    // (before enqueueing it) and records it for later ones (after); call
    // with pipe_mu held.
    bool pipe_before(hipStream_t s);
    bool pipe_after(hipStream_t s);
    // Makes `pipe_done` cover the library stream's unrecorded uses.
    bool pipe_cover_lib();
};

// Context of the calling thread's current HIP device, created on first use.
// Returns nullptr (and sets *err to a TX_ERR_* code when err is given) when
// no gfx950 device is usable.
DeviceContext* current_context(int* err);

// Blocks at or above this count go through the binned pipeline.
uint64_t pipeline_min_blocks();
void set_pipeline_min_blocks(uint64_t v);

}  // namespace txc
