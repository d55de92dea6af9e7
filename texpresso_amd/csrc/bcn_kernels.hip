// bcn_kernels.hip -- kernels and launchers of the BC1-BC5 codec (gfx950).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bcn_device.hpp"
#include "bcn_kernels.hpp"
#include "bcn_lane.hpp"

#pragma clang fp contract(off)

namespace txd {

// ------------------------------------------------------------------ encode

template <int FMT>
__device__ __forceinline__ void store_block(uint8_t* out, uint64_t b, const uint32_t (&w)[4])
{
    if (FMT == BC1 || FMT == BC4)
        *reinterpret_cast<uint2*>(out + b * 8u) = make_uint2(w[0], w[1]);
    else
        *reinterpret_cast<uint4*>(out + b * 16u) = make_uint4(w[0], w[1], w[2], w[3]);
}

// Index of this workgroup in a grid_2d launch (workgroup y * gridDim.x + x;
// below 2^32 since y < 2^16 and x <= 2^16).
__device__ __forceinline__ uint32_t group_index() { return blockIdx.y * gridDim.x + blockIdx.x; }

// Plain per-texture launch: one wave per four consecutive output blocks.
template <int FMT, int ALG>
__global__ __launch_bounds__(64) void k_encode(const uint8_t* __restrict__ rgba, uint32_t width, uint32_t height,
                                               uint32_t blocks_wide, uint64_t nblocks, uint8_t* __restrict__ out,
                                               FitParams fp)
{
    __shared__ GroupLds lds[4];
    const int tid = (int)threadIdx.x, g = tid >> 4, l = tid & 15;
    // the last row of a 2-D grid may hold groups past the texture: all their
    // blocks are not live, as in a part-filled last wave
    const uint64_t b = (uint64_t)group_index() * 4u + (uint64_t)g;
    const bool live = b < nblocks;
    const uint32_t bx = live ? (uint32_t)(b % blocks_wide) : 0u;
    const uint32_t by = live ? (uint32_t)(b / blocks_wide) : 0u;
    bool valid;
    const uint32_t px = fetch_pixel(rgba, width, height, bx, by, live, l, valid);
    uint32_t w[4];
    encode_block<FMT, ALG>(px, valid, fp, lds[g], g, l, w);
    if (live && l == 0) store_block<FMT>(out, b, w);
}

// BC4 / BC5 per-texture launch: one block per lane (alpha codebooks are
// integer work on 16 values; a lane's 16 dword loads keep many bytes in
// flight, and neighbouring lanes read neighbouring 16-byte row segments).
template <int FMT>
__global__ __launch_bounds__(256) void k_encode_alpha(const uint8_t* __restrict__ rgba, uint32_t width,
                                                      uint32_t height, uint32_t blocks_wide, uint64_t nblocks,
                                                      uint8_t* __restrict__ out)
{
    const uint64_t b = (uint64_t)group_index() * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    uint32_t px[16];
    const uint32_t valid = fetch_block_lane(rgba, width, height, (uint32_t)(b % blocks_wide),
                                            (uint32_t)(b / blocks_wide), px);
    uint32_t w[4];
    encode_alpha_lane<FMT>(px, valid, w);
    store_block<FMT>(out, b, w);
}

// RangeFit per-texture launch (BC1/BC2/BC3): one block per lane, the front of
// the pipeline's classify pass (alpha, colour set, principal axis, then the
// single-colour or range fit) without its record path.  A RangeFit block is a
// few hundred serial operations; on a 16-lane group every fold would run on
// all 16 lanes.  Single-wave workgroups (the per-wave LDS copies of the
// block pixels and weights), one per 64-block span of the texture.
template <int FMT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_encode_range_lane(
    const uint8_t* __restrict__ rgba, uint32_t width, uint32_t height, uint32_t blocks_wide, uint64_t nblocks,
    uint8_t* __restrict__ out, FitParams fp)
{
    __shared__ LaneTables LT;
    __shared__ LaneBlockLds LB;
    const int lane = (int)threadIdx.x;
    const uint64_t base = (uint64_t)group_index() * 64u;
    if (base >= nblocks) return;
    fill_lane_tables(LT, lane);
    group_sync();
    const uint64_t b = base + (uint64_t)lane;
    if (b < nblocks) {
        uint32_t px[16];
        const uint32_t valid = fetch_block_lane(rgba, width, height, (uint32_t)(b % blocks_wide),
                                                (uint32_t)(b / blocks_wide), px);
        uint32_t w[4];
        LaneSet S;
        float vx, vy, vz;
        encode_front_lane<FMT, RANGE_FIT>(px, valid, fp, LT, LB, lane, w, S, vx, vy, vz);
        store_block<FMT>(out, b, w);
    }
}

// Explicit block list (Format::compress_block_masked over many blocks): block
// b's 16 pixels are the 64 bytes at blocks + 64*b, its mask is masks[b].
template <int FMT, int ALG>
__global__ __launch_bounds__(64) void k_encode_blocks(const uint8_t* __restrict__ blocks,
                                                      const uint32_t* __restrict__ masks, uint64_t nblocks,
                                                      uint8_t* __restrict__ out, FitParams fp)
{
    __shared__ GroupLds lds[4];
    const int tid = (int)threadIdx.x, g = tid >> 4, l = tid & 15;
    const uint64_t b = (uint64_t)group_index() * 4u + (uint64_t)g;
    const bool live = b < nblocks;
    const uint32_t mask = live ? masks[b] : 0u;
    const bool valid = live && ((mask >> l) & 1u);
    const uint32_t px = valid ? reinterpret_cast<const uint32_t*>(blocks + 64 * b)[l] : 0u;
    uint32_t w[4];
    encode_block<FMT, ALG>(px, valid, fp, lds[g], g, l, w);
    if (live && l == 0) store_block<FMT>(out, b, w);
}

// ------------------------------------------------------------------ decode

// Format::decompress_block for the pixel (px, py) of `blk`,
// colourblock.rs:96-169 and alpha.rs:53-68, 258-304.
__device__ __forceinline__ uint32_t bc3_alpha_value(const uint8_t* blk, int p)
{
    const int a0 = blk[0], a1 = blk[1];
    const int grp = p >> 3, j = p & 7;
    const uint32_t v = (uint32_t)blk[2 + 3 * grp] | ((uint32_t)blk[3 + 3 * grp] << 8) |
                       ((uint32_t)blk[4 + 3 * grp] << 16);
    const int index = (int)((v >> (3 * j)) & 7u);
    if (index == 0) return (uint32_t)a0;
    if (index == 1) return (uint32_t)a1;
    if (a0 <= a1) {
        if (index == 6) return 0u;
        if (index == 7) return 255u;
        const int i = index - 1;
        return (uint32_t)(((5 - i) * a0 + i * a1) / 5);
    }
    const int i = index - 1;
    return (uint32_t)(((7 - i) * a0 + i * a1) / 7);
}

__device__ __forceinline__ uint32_t colour_value(const uint8_t* blk, bool is_bc1, int p)
{
    const uint32_t a = (uint32_t)blk[0] | ((uint32_t)blk[1] << 8);
    const uint32_t b = (uint32_t)blk[2] | ((uint32_t)blk[3] << 8);
    const int index = (blk[4 + (p >> 2)] >> (2 * (p & 3))) & 3;
    const uint32_t ra = (a >> 11) & 31u, ga = (a >> 5) & 63u, ba = a & 31u;
    const uint32_t rb = (b >> 11) & 31u, gb = (b >> 5) & 63u, bb = b & 31u;
    const uint32_t c[2][3] = {{(ra << 3) | (ra >> 2), (ga << 2) | (ga >> 4), (ba << 3) | (ba >> 2)},
                              {(rb << 3) | (rb >> 2), (gb << 2) | (gb >> 4), (bb << 3) | (bb >> 2)}};
    if (index < 2) return c[index][0] | (c[index][1] << 8) | (c[index][2] << 16) | 0xFF000000u;
    const bool three = is_bc1 && a <= b;
    uint32_t r = 0;
    for (int ch = 0; ch < 3; ++ch) {
        const uint32_t x = c[0][ch], y = c[1][ch];
        uint32_t v;
        if (three) v = index == 2 ? (x + y) / 2 : 0u;
        else v = index == 2 ? (2 * x + y) / 3 : (x + 2 * y) / 3;
        r |= v << (8 * ch);
    }
    const uint32_t alpha = (three && index == 3) ? 0u : 255u;
    return r | (alpha << 24);
}

template <int FMT>
__device__ __forceinline__ void decode_pixel(const uint8_t* __restrict__ data, uint32_t width, uint32_t x, uint32_t y,
                                             uint32_t blocks_wide, uint8_t* __restrict__ out)
{
    constexpr uint32_t bs = (FMT == BC1 || FMT == BC4) ? 8u : 16u;
    const uint8_t* blk = data + ((size_t)(y >> 2) * blocks_wide + (x >> 2)) * bs;
    const int p = (int)((y & 3u) * 4u + (x & 3u));
    uint32_t rgba;
    if (FMT == BC1 || FMT == BC2 || FMT == BC3) {
        rgba = colour_value(blk + (FMT == BC1 ? 0 : 8), FMT == BC1, p);
    } else {
        rgba = 0xFF000000u;
    }
    if (FMT == BC2) {
        const uint32_t q = (blk[p >> 1] >> (4 * (p & 1))) & 15u;
        rgba = (rgba & 0x00FFFFFFu) | ((q | (q << 4)) << 24);
    }
    if (FMT == BC3) rgba = (rgba & 0x00FFFFFFu) | (bc3_alpha_value(blk, p) << 24);
    if (FMT == BC4) {
        const uint32_t r = bc3_alpha_value(blk, p);
        rgba = r | (r << 8) | (r << 16) | 0xFF000000u;
    }
    if (FMT == BC5) {
        const uint32_t r = bc3_alpha_value(blk, p), g = bc3_alpha_value(blk + 8, p);
        rgba = r | (g << 8) | 0xFF000000u;
    }
    *reinterpret_cast<uint32_t*>(out + 4 * ((size_t)y * width + x)) = rgba;
}

// One thread per pixel of a 64 x 4 tile; grids smaller than the image walk
// its tiles grid-stride in both directions.
template <int FMT>
__global__ __launch_bounds__(256) void k_decode(const uint8_t* __restrict__ data, uint32_t width, uint32_t height,
                                                uint32_t blocks_wide, uint8_t* __restrict__ out)
{
    for (uint32_t ty = blockIdx.y; ty < (height + 3u) / 4u; ty += gridDim.y) {
        const uint32_t y = ty * 4u + (threadIdx.x >> 6);
        if (y >= height) continue;
        for (uint32_t tx = blockIdx.x; tx < (width + 63u) / 64u; tx += gridDim.x) {
            const uint32_t x = tx * 64u + (threadIdx.x & 63u);
            if (x < width) decode_pixel<FMT>(data, width, x, y, blocks_wide, out);
        }
    }
}

// The exactly specified elementary operations of the encoder, evaluated by
// the device functions the kernels use, for an exhaustive check against a
// host IEEE binary32 evaluation: the u8/255 and sqrt(count) tables the
// one-block-per-lane front fills (fill_lane_tables), sqrtf(k / 256) of the
// alpha-weighted point weights (k = 1..4096, the range of the summed
// (alpha + 1) values), and for each input value the 5- and 6-bit grid snap
// (snap, range.rs / cluster.rs) and roundf clamped to [0, 255]
// (f32_to_i32_clamped, math.rs:100-102).  tables: 256 + 17 + 4096 floats.
__global__ __launch_bounds__(64) void k_elementary(const float* __restrict__ in, uint32_t n, float* __restrict__ tables,
                                                   float* __restrict__ snap31, float* __restrict__ snap63,
                                                   int32_t* __restrict__ round255)
{
    __shared__ LaneTables T;
    const int lane = (int)threadIdx.x;
    if (blockIdx.x == 0) {
        fill_lane_tables(T, lane);
        group_sync();
        for (int i = lane; i < 256; i += 64) tables[i] = T.u8f[i];
        if (lane < 17) tables[256 + lane] = T.sqc[lane];
        for (int k = lane + 1; k <= 4096; k += 64) tables[256 + 17 + (k - 1)] = sqrtf((float)k / 256.0f);
    }
    for (uint32_t i = blockIdx.x * 64u + (uint32_t)lane; i < n; i += gridDim.x * 64u) {
        const float v = in[i];
        snap31[i] = snap(v, 31.0f, 1.0f / 31.0f);
        snap63[i] = snap(v, 63.0f, 1.0f / 63.0f);
        round255[i] = f32_to_i32_clamped(v, 255);
    }
}

}  // namespace txd

// ---------------------------------------------------------------- launchers

namespace {

template <int FMT, int ALG>
hipError_t launch_encode_t(const txk::EncodeArgs& a, hipStream_t s)
{
    const uint64_t groups = (a.nblocks + 3) / 4;
    if (groups == 0) return hipSuccess;
    if (groups > txk::kMaxGroups2d) return hipErrorInvalidValue;
    hipLaunchKernelGGL((txd::k_encode<FMT, ALG>), txk::grid_2d(groups), dim3(64), 0, s, a.rgba, a.width, a.height,
                       a.blocks_wide, a.nblocks, a.out, txd::FitParams{a.wr, a.wg, a.wb, a.weigh_alpha});
    return hipGetLastError();
}

template <int FMT>
hipError_t launch_encode_range(const txk::EncodeArgs& a, hipStream_t s)
{
    const uint64_t waves = (a.nblocks + 63) / 64;
    if (waves == 0) return hipSuccess;
    if (waves > txk::kMaxGroups2d) return hipErrorInvalidValue;
    hipLaunchKernelGGL((txd::k_encode_range_lane<FMT>), txk::grid_2d(waves), dim3(64), 0, s, a.rgba, a.width, a.height,
                       a.blocks_wide, a.nblocks, a.out, txd::FitParams{a.wr, a.wg, a.wb, a.weigh_alpha});
    return hipGetLastError();
}

// RangeFit textures of at least this many blocks take the one-block-per-lane
// kernel, smaller ones the 16-lane-group kernel: a lane-per-block wave costs
// one block's serial latency (~17 us on MI355X however few blocks it has),
// while the group kernel spreads a block over 16 lanes (measured per level
// of a 4096^2 chain, BC1: 11.6 vs 17.1 us at 4096 blocks, 18.5 vs 16.8 us
// at 16384, 166 vs 33 us at 262144).
constexpr uint64_t kRangeLaneMinBlocks = 12288;

template <int FMT>
hipError_t launch_encode_f(const txk::EncodeArgs& a, hipStream_t s)
{
    switch (a.algorithm) {
    case 0: return a.nblocks >= kRangeLaneMinBlocks ? launch_encode_range<FMT>(a, s) : launch_encode_t<FMT, 0>(a, s);
    case 1: return launch_encode_t<FMT, 1>(a, s);
    default: return launch_encode_t<FMT, 2>(a, s);
    }
}

template <int FMT, int ALG>
hipError_t launch_blocks_t(const txk::BlockListArgs& a, hipStream_t s)
{
    const uint64_t groups = (a.nblocks + 3) / 4;
    if (groups == 0) return hipSuccess;
    if (groups > txk::kMaxGroups2d) return hipErrorInvalidValue;
    hipLaunchKernelGGL((txd::k_encode_blocks<FMT, ALG>), txk::grid_2d(groups), dim3(64), 0, s, a.blocks, a.masks,
                       a.nblocks, a.out, txd::FitParams{a.wr, a.wg, a.wb, a.weigh_alpha});
    return hipGetLastError();
}

template <int FMT>
hipError_t launch_blocks_f(const txk::BlockListArgs& a, hipStream_t s)
{
    switch (a.algorithm) {
    case 0: return launch_blocks_t<FMT, 0>(a, s);
    case 1: return launch_blocks_t<FMT, 1>(a, s);
    default: return launch_blocks_t<FMT, 2>(a, s);
    }
}

template <int FMT>
hipError_t launch_decode_t(const uint8_t* data, uint32_t w, uint32_t h, uint8_t* out, hipStream_t s)
{
    if (w == 0 || h == 0) return hipSuccess;
    const uint32_t bw = (w + 3) / 4;
    const dim3 grid(txk::capped_grid(((uint64_t)w + 63) / 64, txk::kMaxDecodeTilesX),
                    txk::capped_grid(((uint64_t)h + 3) / 4, txk::kMaxDecodeTilesY));
    hipLaunchKernelGGL((txd::k_decode<FMT>), grid, dim3(256), 0, s, data, w, h, bw, out);
    return hipGetLastError();
}

}  // namespace

namespace txk {

template <int FMT>
hipError_t launch_encode_alpha(const EncodeArgs& a, hipStream_t s)
{
    const uint64_t groups = (a.nblocks + 255) / 256;
    if (groups == 0) return hipSuccess;
    if (groups > kMaxGroups2d) return hipErrorInvalidValue;
    hipLaunchKernelGGL((txd::k_encode_alpha<FMT>), grid_2d(groups), dim3(256), 0, s, a.rgba, a.width, a.height,
                       a.blocks_wide, a.nblocks, a.out);
    return hipGetLastError();
}

hipError_t launch_encode(const EncodeArgs& a, hipStream_t s)
{
    switch (a.format) {
    case 0: return launch_encode_f<0>(a, s);
    case 1: return launch_encode_f<1>(a, s);
    case 2: return launch_encode_f<2>(a, s);
    case 3: return launch_encode_alpha<3>(a, s);
    case 4: return launch_encode_alpha<4>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_encode_blocks(const BlockListArgs& a, hipStream_t s)
{
    switch (a.format) {
    case 0: return launch_blocks_f<0>(a, s);
    case 1: return launch_blocks_f<1>(a, s);
    case 2: return launch_blocks_f<2>(a, s);
    case 3: return launch_blocks_t<3, 0>(a, s);
    case 4: return launch_blocks_t<4, 0>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_elementary(const float* d_in, uint32_t n, float* d_tables, float* d_snap31, float* d_snap63,
                             int32_t* d_round255, hipStream_t s)
{
    const uint32_t groups = (uint32_t)std::min<uint64_t>(4096, std::max<uint64_t>(1, ((uint64_t)n + 63) / 64));
    hipLaunchKernelGGL(txd::k_elementary, dim3(groups), dim3(64), 0, s, d_in, n, d_tables, d_snap31, d_snap63,
                       d_round255);
    return hipGetLastError();
}

hipError_t launch_decode(int format, const uint8_t* data, uint32_t w, uint32_t h, uint8_t* out, hipStream_t s)
{
    switch (format) {
    case 0: return launch_decode_t<0>(data, w, h, out, s);
    case 1: return launch_decode_t<1>(data, w, h, out, s);
    case 2: return launch_decode_t<2>(data, w, h, out, s);
    case 3: return launch_decode_t<3>(data, w, h, out, s);
    case 4: return launch_decode_t<4>(data, w, h, out, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace txk
