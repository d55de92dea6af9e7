// tx_context.hpp -- per-device state of the C ABI: stream, scratch pool, arch.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <stddef.h>
#include <stdint.h>

namespace txc {

struct DeviceContext {
    int device = -1;
    char arch[64] = {0};
    hipStream_t stream = nullptr;  // default priority, non-blocking
    std::mutex mu;                 // guards the scratch pool and host-staged calls
    uint8_t* pool = nullptr;
    size_t pool_cap = 0;

    // Device scratch of at least `bytes` (grown geometrically); call with mu held.
    uint8_t* scratch(size_t bytes);
};

// Context of the calling thread's current HIP device, created on first use.
// Returns nullptr (and sets *err to a TX_ERR_* code when err is given) when
// no gfx950 device is usable.
DeviceContext* current_context(int* err);

}  // namespace txc
