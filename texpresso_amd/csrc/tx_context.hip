// tx_context.hip -- per-device contexts of the C ABI.
#include <hip/hip_runtime.h>

#include <atomic>
#include <string.h>

#include "texpresso_hip.h"
#include "tx_context.hpp"

namespace txc {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
DeviceContext* g_ctx[kMaxDevices] = {};

}  // namespace

uint8_t* DeviceContext::scratch(size_t bytes)
{
    if (bytes <= pool_cap) return pool;
    size_t cap = pool_cap ? pool_cap : ((size_t)1 << 20);
    while (cap < bytes) cap *= 2;
    if (pool) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(pool);
        pool = nullptr;
        pool_cap = 0;
    }
    if (hipMalloc(&pool, cap) != hipSuccess) {
        pool = nullptr;
        return nullptr;
    }
    // first access to fresh device memory is slow; pay it here, once
    (void)hipMemsetAsync(pool, 0, cap, stream);
    pool_cap = cap;
    return pool;
}

bool DeviceContext::pipe_scratch(uint64_t recs, txp::PipeBufs* out)
{
    if (recs > pipe_recs || !pipe) {
        uint64_t cap = pipe_recs ? pipe_recs : 4096;
        while (cap < recs) cap *= 2;
        if (pipe) {
            (void)pipe_cover_lib();
            (void)hipEventSynchronize(pipe_done);
            (void)hipFree(pipe);
            pipe = nullptr;
            pipe_recs = 0;
        }
        if (hipMalloc(&pipe, txp::PipeBufs::bytes_for(cap)) != hipSuccess) {
            pipe = nullptr;
            return false;
        }
        // touch it once (first access to fresh device memory is slow); the
        // scratch may next be used on another stream, so wait here
        if (hipMemsetAsync(pipe, 0, txp::PipeBufs::bytes_for(cap), stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            (void)hipFree(pipe);
            pipe = nullptr;
            return false;
        }
        pipe_recs = cap;
    }
    out->carve(pipe, pipe_recs);
    return true;
}

bool DeviceContext::pipe_cover_lib()
{
    if (!pipe_lib_pending) return true;
    pipe_lib_pending = false;
    return hipEventRecord(pipe_done, stream) == hipSuccess;
}

bool DeviceContext::pipe_before(hipStream_t s)
{
    // a use right after one on the same library stream is ordered by it;
    // otherwise wait for the last recorded use (a never-recorded event is
    // already complete)
    if (s == stream && pipe_lib_pending) return true;
    if (!pipe_cover_lib()) return false;
    return hipStreamWaitEvent(s, pipe_done, 0) == hipSuccess;
}

bool DeviceContext::pipe_after(hipStream_t s)
{
    if (s == stream) {
        pipe_lib_pending = true;
        return true;
    }
    return hipEventRecord(pipe_done, s) == hipSuccess;
}

namespace {
std::atomic<uint64_t> g_pipe_min_blocks{32768};
}

uint64_t pipeline_min_blocks() { return g_pipe_min_blocks.load(); }

void set_pipeline_min_blocks(uint64_t v) { g_pipe_min_blocks.store(v); }

DeviceContext* current_context(int* err)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        if (err) *err = TX_ERR_NO_DEVICE;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ctx[dev]) return g_ctx[dev];
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        if (err) *err = TX_ERR_NO_DEVICE;
        return nullptr;
    }
    // the code objects are built for gfx950 only
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        if (err) *err = TX_ERR_NO_DEVICE;
        return nullptr;
    }
    DeviceContext* c = new DeviceContext();
    c->device = dev;
    c->cus = prop.multiProcessorCount;
    strncpy(c->arch, "gfx950", sizeof c->arch - 1);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->pipe_done, hipEventDisableTiming) != hipSuccess) {
        delete c;
        if (err) *err = TX_ERR_HIP;
        return nullptr;
    }
    g_ctx[dev] = c;
    return c;
}

}  // namespace txc
