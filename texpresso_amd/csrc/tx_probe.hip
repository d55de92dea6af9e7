// tx_probe.hip -- microbenchmark of the plain FP32 VALU rate (separate
// multiply and add, no FMA, no packed math: the only FP32 forms the
// bit-exact encoder can use).  Used to state the encoder's compute ceiling.
// The operands are per-lane values in vector registers, as in the encoder's
// split loop: an operand in a scalar register (a kernel argument used
// directly) halves the issue rate of these instructions on this GPU.
#include <hip/hip_runtime.h>

#include "texpresso_hip.h"
#include "tx_context.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 8;

__global__ __launch_bounds__(256) void k_valu_probe(float* out, int iters, float a, float b)
{
    float x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = (float)(threadIdx.x + c) * 1e-3f;
    // lane-dependent copies of the factors keep them in vector registers
    const float av = a + (float)(threadIdx.x & 1u) * 1e-30f;
    const float bv = b + (float)(threadIdx.x & 2u) * 1e-30f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = x[c] * av;
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = x[c] + bv;
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s = s + x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" int tx_util_valu_probe_waves(int waves_per_simd, double* lane_ops_per_s)
{
    if (!lane_ops_per_s || waves_per_simd < 1 || waves_per_simd > 8) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return TX_ERR_HIP;
    // workgroups of 4 waves (one per SIMD of a CU), `waves_per_simd` of them
    // per CU; fewer waves get proportionally more iterations
    const int blocks = prop.multiProcessorCount * waves_per_simd;
    const int threads = 256, iters = 2048 * 8 / waves_per_simd;
    float* out = nullptr;
    if (hipMalloc(&out, sizeof(float) * blocks * threads) != hipSuccess) return TX_ERR_NOMEM;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(threads), 0, ctx->stream, out, iters, 0.9999999f, 1e-7f);
    (void)hipEventRecord(e0, ctx->stream);
    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(threads), 0, ctx->stream, out, iters, 0.9999999f, 1e-7f);
    (void)hipEventRecord(e1, ctx->stream);
    rc = hipEventSynchronize(e1) == hipSuccess ? TX_OK : TX_ERR_HIP;
    float ms = 0.0f;
    if (rc == TX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = TX_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc != TX_OK || ms <= 0.0f) return rc != TX_OK ? rc : TX_ERR_HIP;
    const double ops = (double)blocks * threads * iters * kUnroll * kChains * 2.0;
    *lane_ops_per_s = ops / (ms * 1e-3);
    return TX_OK;
}

extern "C" int tx_util_valu_probe(double* lane_ops_per_s) { return tx_util_valu_probe_waves(8, lane_ops_per_s); }
