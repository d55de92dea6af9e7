// bcn_kernels.hpp -- host-side launch interface of the BC1-BC5 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace txk {

struct EncodeArgs {
    int format;           // tx_format
    int algorithm;        // tx_algorithm
    const uint8_t* rgba;  // device, 4-byte aligned, width*height*4 bytes
    uint32_t width, height;
    uint32_t blocks_wide;
    uint64_t nblocks;  // output blocks; blocks past the image are fully masked
    uint8_t* out;      // device, nblocks * block_size bytes
    float wr, wg, wb;
    uint32_t weigh_alpha;
};

struct BlockListArgs {
    int format;
    int algorithm;
    const uint8_t* blocks;  // device, 64 bytes (16 RGBA8 pixels) per block
    const uint32_t* masks;  // device, one valid-pixel mask per block
    uint64_t nblocks;
    uint8_t* out;  // device, nblocks * block_size bytes
    float wr, wg, wb;
    uint32_t weigh_alpha;
};

// One texture of a persistent batch launch, as the device reads it.
struct BatchTex {
    const uint8_t* rgba;
    uint8_t* out;
    uint32_t width, height, blocks_wide, nblocks;
    uint32_t variant;  // 3*format + algorithm (BC4: 9, BC5: 10)
    uint32_t pad;
    float wr, wg, wb;
    uint32_t weigh_alpha;
};

// Counter reset + persistent kernels over chunks {texture, first block,
// block count, 0}: the first `nheavy` chunks (any format) go to the LDS
// encoder (`grid` workgroups of `waves_per_group` = 1, 2 or 4 waves), the
// next `nlight` (BC4/BC5 only) to the alpha-only launch (launch_alpha_chunks,
// `grid_light` workgroups, `alpha_lanes` lanes per block), on `s_light` when
// given (fork/join through the two events) or after the heavy kernel on `s`.
// d_counter holds two counters.
hipError_t launch_batch_raw(const BatchTex* d_tex, const uint4* d_chunks, uint32_t nheavy, uint32_t nlight,
                            uint32_t* d_counter, uint32_t grid, uint32_t grid_light, uint32_t waves_per_group,
                            uint32_t alpha_lanes, hipStream_t s, hipStream_t s_light, hipEvent_t fork,
                            hipEvent_t join);

// Blocks per chunk of an alpha-only (BC4/BC5) texture encoded one block per
// lane: the kernel runs 8 steps of 64 blocks per chunk.
constexpr uint32_t kLightChunkBlocks = 512;

// Persistent kernel over the BC4/BC5 chunks [first, first + nlight):
// alpha_lanes = 1 runs one block per lane (k_batch_light, 4-wave
// workgroups); 16 runs a 16-lane group per block (k_batch with
// `waves_per_group` waves per workgroup).  `d_counter` is its work counter
// (not reset here).
hipError_t launch_alpha_chunks(const BatchTex* d_tex, const uint4* d_chunks, uint32_t first, uint32_t nlight,
                               uint32_t* d_counter, uint32_t grid, uint32_t alpha_lanes, uint32_t waves_per_group,
                               hipStream_t s);

// Workgroups of the alpha-only launch for `nlight` chunks on `cus` CUs.
inline uint32_t alpha_grid(uint32_t cus, uint32_t nlight, uint32_t alpha_lanes, uint32_t groups_per_cu,
                           uint32_t waves_per_group)
{
    const uint64_t want = alpha_lanes == 16 ? ((uint64_t)nlight + waves_per_group - 1) / waves_per_group
                                            : ((uint64_t)nlight + 3) / 4;
    const uint64_t cap = alpha_lanes == 16 ? (uint64_t)cus * groups_per_cu : (uint64_t)cus * 8;
    return (uint32_t)(want < 1 ? 1 : want < cap ? want : cap);
}

// Counter reset (d_counter[0..1]) + the alpha-only kernel over chunks
// [first, first + nlight) on `s`.
hipError_t launch_batch_light(const BatchTex* d_tex, const uint4* d_chunks, uint32_t first, uint32_t nlight,
                              uint32_t* d_counter, uint32_t grid_light, uint32_t alpha_lanes,
                              uint32_t waves_per_group, hipStream_t s);

// Clears d_counter[0..1] on `s` (a kernel: replay-safe inside graphs).
hipError_t reset_counters(uint32_t* d_counter, hipStream_t s);

// Launch shapes that keep every grid dimension within its limit (2^32
// work-items in x, 2^16 workgroups in y) whatever the texture size.  The
// one-pass encoders take `groups` workgroups as a 2-D grid of rows of at
// most 2^16 and index workgroup y * gridDim.x + x (a grid-stride loop
// instead made the compiler keep more registers live and cost the lockstep
// kernel a wave per SIMD); the decoder caps its tile grid and walks larger
// images grid-stride.
constexpr uint64_t kGroupsPerRow = 1ull << 16;
constexpr uint64_t kMaxGroups2d = kGroupsPerRow * 65535ull;
constexpr uint64_t kMaxDecodeTilesX = 1ull << 16;   // 64 x 4 pixel tiles per grid row
constexpr uint64_t kMaxDecodeTilesY = 1ull << 14;

inline uint32_t capped_grid(uint64_t groups, uint64_t cap) { return (uint32_t)(groups < cap ? groups : cap); }

inline dim3 grid_2d(uint64_t groups)
{
    const uint64_t x = groups < kGroupsPerRow ? groups : kGroupsPerRow;
    return dim3((uint32_t)x, (uint32_t)((groups + x - 1) / x));
}

hipError_t launch_encode(const EncodeArgs& a, hipStream_t s);
hipError_t launch_encode_blocks(const BlockListArgs& a, hipStream_t s);
hipError_t launch_decode(int format, const uint8_t* data, uint32_t w, uint32_t h, uint8_t* out, hipStream_t s);
// The elementary-operation check (tx_util_elementary): d_tables holds
// 256 + 17 + 4096 floats; the three per-value outputs hold n entries each.
hipError_t launch_elementary(const float* d_in, uint32_t n, float* d_tables, float* d_snap31, float* d_snap63,
                             int32_t* d_round255, hipStream_t s);
// One mip step of the RGBA8 image src (w x h, 4-byte aligned) into dst
// (max(1, w/2) x max(1, h/2)): 2x2 box, edge clamp, (sum + 2) >> 2.
hipError_t launch_mip_down(const uint8_t* src, uint32_t w, uint32_t h, uint8_t* dst, hipStream_t s);

}  // namespace txk
