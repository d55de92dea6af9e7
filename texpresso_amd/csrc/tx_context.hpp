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

    // Binned-pipeline scratch, shared by every stream of the device: each use
    // waits for `pipe_done` (the previous use) and records it again.
    std::mutex pipe_mu;
    uint8_t* pipe = nullptr;
    uint64_t pipe_recs = 0;
    hipEvent_t pipe_done = nullptr;

    // Device staging memory of at least `bytes` (grown geometrically); call with mu held.
    uint8_t* scratch(size_t bytes);
    // Pipeline buffers for `recs` records; call with pipe_mu held.
    bool pipe_scratch(uint64_t recs, txp::PipeBufs* out);
};

// Context of the calling thread's current HIP device, created on first use.
// Returns nullptr (and sets *err to a TX_ERR_* code when err is given) when
// no gfx950 device is usable.
DeviceContext* current_context(int* err);

// Blocks at or above this count go through the binned pipeline.
uint64_t pipeline_min_blocks();
void set_pipeline_min_blocks(uint64_t v);

}  // namespace txc
