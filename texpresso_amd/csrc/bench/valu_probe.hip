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
//
// A second probe (txb_op_rate) times fixed instruction mixes -- one
// instruction form, or two forms interleaved -- to show which forms issue
// at the plain FP32 rate and which can issue alongside one another.
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

// Op-rate mixes: eight independent chains per lane, each step one
// instruction written in inline assembly (register operands only, no memory
// access) so that the compiler neither folds nor reorders the chains.
#define TXB_OP(ins, x) asm volatile(ins : "+v"(x) : "v"(b), "v"(c))
#define TXB_MIX2(i0, i1)                                                                                   \
    TXB_OP(i0, x0); TXB_OP(i1, x1); TXB_OP(i0, x2); TXB_OP(i1, x3); TXB_OP(i0, x4); TXB_OP(i1, x5);        \
    TXB_OP(i0, x6); TXB_OP(i1, x7)

constexpr const char* kOpMixNames[] = {
    "8 v_add_f32",          "8 v_mul_f32",          "8 v_fma_f32",         "8 v_add_u32",
    "8 v_sad_u16",          "8 v_min3_u32",         "8 v_trunc_f32",       "8 v_bfe_u32",
    "8 v_mul_u32_u24",      "4 v_sad_u16 + 4 v_add_f32", "4 v_trunc_f32 + 4 v_mul_f32",
    "4 v_min3_u32 + 4 v_add_u32", "8 v_mul_f32_e64 (VOP3)", "8 v_mul_f32 clamp",
    "4 v_mul_f32 clamp + 4 v_add_f32", "8 v_med3_f32",     "8 v_max_f32",         "8 v_floor_f32",
    "8 v_fract_f32",        "8 v_cvt_u32_f32",      "8 v_cndmask_b32",     "8 v_sub_f32",
    "4 v_trunc_f32 + 4 v_add_f32", "4 v_cndmask_b32 + 4 v_mul_f32"};
constexpr int kOpMixes = (int)(sizeof(kOpMixNames) / sizeof(kOpMixNames[0]));

template <int M>
__global__ __launch_bounds__(256) void k_op_rate(uint32_t* out, int iters, uint32_t seed)
{
    const uint32_t t = threadIdx.x + blockIdx.x * 256u;
    uint32_t x0 = t ^ seed, x1 = t + 1u, x2 = t * 3u, x3 = t + 7u, x4 = t ^ 0x55u, x5 = t + 11u, x6 = t * 5u,
             x7 = t + 13u;
    const uint32_t b = (seed | 1u) + (t & 1u), c = (seed >> 3) + (t & 2u);
    for (int i = 0; i < iters; ++i) {
        if (M == 0) { TXB_MIX2("v_add_f32 %0, %0, %1", "v_add_f32 %0, %0, %1"); }
        if (M == 1) { TXB_MIX2("v_mul_f32 %0, %0, %1", "v_mul_f32 %0, %0, %1"); }
        if (M == 2) { TXB_MIX2("v_fma_f32 %0, %0, %1, %2", "v_fma_f32 %0, %0, %1, %2"); }
        if (M == 3) { TXB_MIX2("v_add_u32 %0, %0, %1", "v_add_u32 %0, %0, %1"); }
        if (M == 4) { TXB_MIX2("v_sad_u16 %0, %0, %1, %2", "v_sad_u16 %0, %0, %1, %2"); }
        if (M == 5) { TXB_MIX2("v_min3_u32 %0, %0, %1, %2", "v_min3_u32 %0, %0, %1, %2"); }
        if (M == 6) { TXB_MIX2("v_trunc_f32 %0, %0", "v_trunc_f32 %0, %0"); }
        if (M == 7) { TXB_MIX2("v_bfe_u32 %0, %0, %1, %2", "v_bfe_u32 %0, %0, %1, %2"); }
        if (M == 8) { TXB_MIX2("v_mul_u32_u24 %0, %0, %1", "v_mul_u32_u24 %0, %0, %1"); }
        if (M == 9) { TXB_MIX2("v_sad_u16 %0, %0, %1, %2", "v_add_f32 %0, %0, %1"); }
        if (M == 10) { TXB_MIX2("v_trunc_f32 %0, %0", "v_mul_f32 %0, %0, %1"); }
        if (M == 11) { TXB_MIX2("v_min3_u32 %0, %0, %1, %2", "v_add_u32 %0, %0, %1"); }
        if (M == 12) { TXB_MIX2("v_mul_f32_e64 %0, %0, %1", "v_mul_f32_e64 %0, %0, %1"); }
        if (M == 13) { TXB_MIX2("v_mul_f32_e64 %0, %0, %1 clamp", "v_mul_f32_e64 %0, %0, %1 clamp"); }
        if (M == 14) { TXB_MIX2("v_mul_f32_e64 %0, %0, %1 clamp", "v_add_f32 %0, %0, %1"); }
        if (M == 15) { TXB_MIX2("v_med3_f32 %0, %0, %1, %2", "v_med3_f32 %0, %0, %1, %2"); }
        if (M == 16) { TXB_MIX2("v_max_f32 %0, %0, %1", "v_max_f32 %0, %0, %1"); }
        if (M == 17) { TXB_MIX2("v_floor_f32 %0, %0", "v_floor_f32 %0, %0"); }
        if (M == 18) { TXB_MIX2("v_fract_f32 %0, %0", "v_fract_f32 %0, %0"); }
        if (M == 19) { TXB_MIX2("v_cvt_u32_f32 %0, %0", "v_cvt_u32_f32 %0, %0"); }
        if (M == 20) { TXB_MIX2("v_cndmask_b32 %0, %0, %1, vcc", "v_cndmask_b32 %0, %0, %1, vcc"); }
        if (M == 21) { TXB_MIX2("v_sub_f32 %0, %0, %1", "v_sub_f32 %0, %0, %1"); }
        if (M == 22) { TXB_MIX2("v_trunc_f32 %0, %0", "v_add_f32 %0, %0, %1"); }
        if (M == 23) { TXB_MIX2("v_cndmask_b32 %0, %0, %1, vcc", "v_mul_f32 %0, %0, %1"); }
    }
    out[t] = x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7;
}
#undef TXB_MIX2
#undef TXB_OP

template <int M>
void launch_op_rate(int blocks, uint32_t* out, int iters, hipStream_t s)
{
    hipLaunchKernelGGL(k_op_rate<M>, dim3(blocks), dim3(256), 0, s, out, iters, 7u);
}

typedef void (*OpLaunch)(int, uint32_t*, int, hipStream_t);
constexpr OpLaunch kOpLaunch[kOpMixes] = {
    launch_op_rate<0>, launch_op_rate<1>, launch_op_rate<2>, launch_op_rate<3>, launch_op_rate<4>,  launch_op_rate<5>,
    launch_op_rate<6>,  launch_op_rate<7>,  launch_op_rate<8>,  launch_op_rate<9>,  launch_op_rate<10>, launch_op_rate<11>,
    launch_op_rate<12>, launch_op_rate<13>, launch_op_rate<14>, launch_op_rate<15>, launch_op_rate<16>, launch_op_rate<17>,
    launch_op_rate<18>, launch_op_rate<19>, launch_op_rate<20>, launch_op_rate<21>, launch_op_rate<22>, launch_op_rate<23>};

}  // namespace

// Number of instruction mixes of txb_op_rate and the name of mix m (nullptr
// when m is out of range).
extern "C" int txb_op_mixes(void) { return kOpMixes; }
extern "C" const char* txb_op_mix_name(int m) { return m >= 0 && m < kOpMixes ? kOpMixNames[m] : nullptr; }

// Mix m at 8 waves per SIMD: VALU wave-instructions per second per SIMD
// (the loop's few scalar instructions aside).  Returns 0 on success, -1 on a
// bad argument, -2 on a HIP error.
extern "C" int txb_op_rate(int m, double* wave_instrs_per_s_per_simd)
{
    if (!wave_instrs_per_s_per_simd || m < 0 || m >= kOpMixes) return -1;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -2;
    const int blocks = prop.multiProcessorCount * 8, iters = 4096;
    uint32_t* out = nullptr;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -2;
    if (hipMalloc(&out, sizeof(uint32_t) * blocks * 256) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return -2;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) ? 0 : -2;
    if (rc == 0) {
        kOpLaunch[m](blocks, out, iters, s);  // warms the clocks
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < 4; ++r) kOpLaunch[m](blocks, out, iters, s);
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
    const double instrs = 4.0 * blocks * 4.0 /* waves per block */ * iters * 8.0;
    *wave_instrs_per_s_per_simd = instrs / (ms * 1e-3) / (prop.multiProcessorCount * 4.0);
    return 0;
}

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
