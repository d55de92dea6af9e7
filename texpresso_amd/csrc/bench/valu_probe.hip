// valu_probe.hip -- benchmark helper, not part of the encoder library: a
// microbenchmark of the plain FP32 VALU rate (separate multiply and add, no
// FMA, no packed math: the only FP32 forms the bit-exact encoder can use),
// which bench.py reports beside the encoder's own rate to state its compute
// ceiling.  Built as libtx_bench_probe.so; nothing in libtexpresso_hip.so
// links it.
//
// The operands are per-lane values in vector registers, as in the encoder's
// split loop: an operand in a scalar register (a kernel argument used
// directly) halves the issue rate of these instructions on this GPU.
#include <hip/hip_runtime.h>

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

// Plain FP32 lane-operations per second on the current device with
// `waves_per_simd` (1..8) waves per SIMD.  Returns 0 on success, -1 on a bad
// argument, -2 on a HIP error.
extern "C" int txb_valu_probe(int waves_per_simd, double* lane_ops_per_s)
{
    if (!lane_ops_per_s || waves_per_simd < 1 || waves_per_simd > 8) return -1;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -2;
    // workgroups of 4 waves (one per SIMD of a CU), `waves_per_simd` of them
    // per CU; fewer waves get proportionally more iterations
    const int blocks = prop.multiProcessorCount * waves_per_simd;
    const int threads = 256, iters = 2048 * 8 / waves_per_simd;
    float* out = nullptr;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -2;
    if (hipMalloc(&out, sizeof(float) * blocks * threads) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return -2;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) ? 0 : -2;
    if (rc == 0) {
        // the first launch warms the clocks; the second is timed
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(threads), 0, s, out, iters, 0.9999999f, 1e-7f);
        (void)hipEventRecord(e0, s);
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(threads), 0, s, out, iters, 0.9999999f, 1e-7f);
        (void)hipEventRecord(e1, s);
        rc = hipEventSynchronize(e1) == hipSuccess ? 0 : -2;
    }
    float ms = 0.0f;
    if (rc == 0 && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = -2;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    (void)hipStreamDestroy(s);
    if (rc != 0 || ms <= 0.0f) return rc != 0 ? rc : -2;
    const double ops = (double)blocks * threads * iters * kUnroll * kChains * 2.0;
    *lane_ops_per_s = ops / (ms * 1e-3);
    return 0;
}
