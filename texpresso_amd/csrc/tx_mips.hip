// tx_mips.hip -- device mip-chain generation (the step before the encoder
// when a caller needs every level; texpresso itself encodes one level, and
// its CLI writes level 0 only, cli/src/main.rs:148-160).
//
// Level k has size max(1, w >> k) x max(1, h >> k); each texel of level k+1
// is the rounded mean of a 2x2 box of level k with edge clamp,
// (a + b + c + d + 2) >> 2 per channel.  Integer work, HBM-bound: one thread
// per destination texel, four dword loads, one dword store.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bcn_kernels.hpp"

namespace txd {

__global__ __launch_bounds__(256) void k_mip_down(const uint32_t* __restrict__ src, uint32_t w, uint32_t h,
                                                  uint32_t* __restrict__ dst, uint32_t nw, uint32_t nh)
{
    const uint32_t x = blockIdx.x * 64u + (threadIdx.x & 63u);
    if (x >= nw) return;
    const uint32_t x0 = min(2u * x, w - 1u), x1 = min(2u * x + 1u, w - 1u);
    // rows in a grid-stride loop (the grid's y extent is capped)
    for (uint32_t y = blockIdx.y * 4u + (threadIdx.x >> 6); y < nh; y += gridDim.y * 4u) {
        const uint32_t y0 = min(2u * y, h - 1u), y1 = min(2u * y + 1u, h - 1u);
        const uint32_t a = src[(size_t)y0 * w + x0], b = src[(size_t)y0 * w + x1];
        const uint32_t c = src[(size_t)y1 * w + x0], d = src[(size_t)y1 * w + x1];
        uint32_t out = 0;
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) {
            const uint32_t s = ((a >> (8 * ch)) & 255u) + ((b >> (8 * ch)) & 255u) + ((c >> (8 * ch)) & 255u) +
                               ((d >> (8 * ch)) & 255u);
            out |= ((s + 2u) >> 2) << (8 * ch);
        }
        dst[(size_t)y * nw + x] = out;
    }
}

}  // namespace txd

namespace txk {

hipError_t launch_mip_down(const uint8_t* src, uint32_t w, uint32_t h, uint8_t* dst, hipStream_t s)
{
    const uint32_t nw = std::max(1u, w >> 1), nh = std::max(1u, h >> 1);
    if ((nw + 63u) / 64u > 0x7FFFFFFFu) return hipErrorInvalidValue;
    const dim3 grid((nw + 63u) / 64u, std::min((nh + 3u) / 4u, 16384u));
    hipLaunchKernelGGL(txd::k_mip_down, grid, dim3(256), 0, s, reinterpret_cast<const uint32_t*>(src), w, h,
                       reinterpret_cast<uint32_t*>(dst), nw, nh);
    return hipGetLastError();
}

}  // namespace txk
