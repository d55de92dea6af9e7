// tx_context.hip -- per-device contexts of the C ABI.
#include <hip/hip_runtime.h>

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
    pool_cap = cap;
    return pool;
}

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
    strncpy(c->arch, "gfx950", sizeof c->arch - 1);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        if (err) *err = TX_ERR_HIP;
        return nullptr;
    }
    g_ctx[dev] = c;
    return c;
}

}  // namespace txc
