// tx_capi.hip -- C ABI of the MI355X BC1-BC5 encoder (include/texpresso_hip.h).
//
// Host-buffer entry points (tx_compress, tx_decompress, ...) stage through a
// per-device scratch pool and run synchronously on the device's stream;
// device-buffer entry points only enqueue.  There is no CPU code path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string.h>
#include <vector>

#include "bcn_kernels.hpp"
#include "texpresso_hip.h"
#include "tx_context.hpp"
#include "tx_tables.hpp"

namespace {

constexpr txt::SingleLut kHostLut = txt::make_single_lut();

bool valid_format(int f) { return f >= TX_BC1 && f <= TX_BC5; }
bool valid_params(const tx_params* p) { return p && p->algorithm >= 0 && p->algorithm <= 2; }

// 4 * width * height without overflow; false when it does not fit
bool image_bytes(size_t w, size_t h, size_t* out)
{
    if (w != 0 && h > SIZE_MAX / 4 / w) return false;
    *out = 4 * w * h;
    return true;
}

}  // namespace

// Device scratch of tx_compress_device_noalloc: the binned pipeline's
// records, keys, sorted lists and counters for TX_NOALLOC_MAX_PIPELINE_BLOCKS
// blocks.  Every GPU the library's code object is loaded on holds its own
// copy.  In host code `scratch` names a host-side stand-in, not a device
// address: the per-device address comes from hipGetSymbolAddress
// (NoallocDevice::ptr), and only that address is passed to kernels.
static_assert(txp::PipeBufs::bytes_for(TX_NOALLOC_MAX_PIPELINE_BLOCKS) == TX_NOALLOC_SCRATCH_BYTES,
              "TX_NOALLOC_SCRATCH_BYTES holds the pipeline buffers of TX_NOALLOC_MAX_PIPELINE_BLOCKS blocks");
static __device__ __align__(256) uint8_t scratch[TX_NOALLOC_SCRATCH_BYTES];

namespace {

struct NoallocDevice {
    bool ready = false;
    uint8_t* ptr = nullptr;  // this device's copy of `scratch`
    uint32_t cus = 0;
};

constexpr int kNoallocDevices = 64;
std::mutex g_noalloc_mu;
NoallocDevice g_noalloc[kNoallocDevices];

// Set-up of the calling thread's current device, once per device: its
// address of `scratch` and its compute-unit count.  Allocates nothing.
int noalloc_device(const NoallocDevice** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoallocDevices) return TX_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoallocDevice& d = g_noalloc[dev];
    if (!d.ready) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess || strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return TX_ERR_NO_DEVICE;
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return TX_ERR_HIP;
        d.ptr = static_cast<uint8_t*>(p);
        d.cus = (uint32_t)prop.multiProcessorCount;
        d.ready = true;
    }
    *out = &d;
    return TX_OK;
}

}  // namespace

extern "C" {

tx_params tx_params_default(void)
{
    tx_params p;
    p.algorithm = TX_CLUSTER_FIT;
    p.weights[0] = 0.2126f;
    p.weights[1] = 0.7152f;
    p.weights[2] = 0.0722f;
    p.weigh_colour_by_alpha = 0;
    return p;
}

size_t tx_num_blocks(size_t size) { return (size + 3) / 4; }

size_t tx_block_size(int format)
{
    if (!valid_format(format)) return 0;
    return (format == TX_BC1 || format == TX_BC4) ? 8 : 16;
}

size_t tx_compressed_size(int format, size_t width, size_t height)
{
    return tx_num_blocks(width) * tx_num_blocks(height) * tx_block_size(format);
}

const char* tx_error_string(int code)
{
    switch (code) {
    case TX_OK: return "ok";
    case TX_ERR_FORMAT: return "invalid format or algorithm";
    case TX_ERR_ZERO_WIDTH: return "width is zero";
    case TX_ERR_INPUT_SHORT: return "input buffer shorter than the image";
    case TX_ERR_OUTPUT_SHORT: return "output buffer smaller than the compressed size";
    case TX_ERR_OUTPUT_PARTIAL: return "output buffer ends inside a block";
    case TX_ERR_ALIGN: return "device pointer misaligned";
    case TX_ERR_HIP: return "HIP runtime error";
    case TX_ERR_NO_DEVICE: return "no usable gfx950 device";
    case TX_ERR_BUSY: return "queue full";
    case TX_ERR_INVALID: return "invalid argument or handle";
    case TX_ERR_NOMEM: return "device memory budget exceeded";
    case TX_ERR_TOO_LARGE: return "texture exceeds the allocation-free scratch (TX_NOALLOC_MAX_PIPELINE_BLOCKS)";
    default: return "unknown error";
    }
}

void tx_single_lut(uint8_t out[6144]) { memcpy(out, &kHostLut, sizeof kHostLut); }

uint64_t tx_set_pipeline_threshold(uint64_t min_blocks)
{
    const uint64_t old = txc::pipeline_min_blocks();
    txc::set_pipeline_min_blocks(min_blocks);
    return old;
}

const char* tx_device_arch(void)
{
    txc::DeviceContext* ctx = txc::current_context(nullptr);
    return ctx ? ctx->arch : nullptr;
}

// Validation shared by the image entry points (lib.rs:295-305 semantics).
static int check_compress(int format, size_t width, size_t height, const tx_params* params, size_t out_len,
                          size_t* in_bytes)
{
    if (!valid_format(format) || !valid_params(params)) return TX_ERR_FORMAT;
    if (width == 0) return TX_ERR_ZERO_WIDTH;
    if (width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return TX_ERR_INVALID;
    if (!image_bytes(width, height, in_bytes)) return TX_ERR_INVALID;
    const size_t bs = tx_block_size(format);
    if (out_len < tx_compressed_size(format, width, height)) return TX_ERR_OUTPUT_SHORT;
    if (out_len % bs != 0) return TX_ERR_OUTPUT_PARTIAL;
    return TX_OK;
}

static txk::EncodeArgs encode_args(int format, const tx_params* p, const uint8_t* d_in, size_t w, size_t h,
                                   uint8_t* d_out, size_t out_len)
{
    txk::EncodeArgs a;
    a.format = format;
    a.algorithm = p->algorithm;
    a.rgba = d_in;
    a.width = (uint32_t)w;
    a.height = (uint32_t)h;
    a.blocks_wide = (uint32_t)tx_num_blocks(w);
    a.nblocks = out_len / tx_block_size(format);
    a.out = d_out;
    a.wr = p->weights[0];
    a.wg = p->weights[1];
    a.wb = p->weights[2];
    a.weigh_alpha = p->weigh_colour_by_alpha ? 1u : 0u;
    return a;
}

// Blocks from which a texture goes through the binned pipeline: the set
// threshold for BC2/BC3 IterativeClusterFit, 1.25x for BC2/BC3 ClusterFit,
// 1.5x for BC1 ClusterFit and 3x for BC1 IterativeClusterFit (measured
// crossovers against the direct kernel, 256^2-1448^2: BC1's two search
// passes and single-ordering blocks gain less from independent per-block
// scheduling).
static uint64_t pipeline_threshold(int format, int algorithm)
{
    const uint64_t t = txc::pipeline_min_blocks();
    const bool bc1 = format == TX_BC1, iterative = algorithm == TX_ITERATIVE_CLUSTER_FIT;
    const uint64_t quarters = bc1 ? (iterative ? 12 : 6) : (iterative ? 4 : 5);
    return t > UINT64_MAX / quarters ? UINT64_MAX : t * quarters / 4;
}

// Whether a texture goes through the binned pipeline: ClusterFit or
// IterativeClusterFit in BC1-BC3 with at least pipeline_threshold() blocks.
static bool takes_pipeline(const txk::EncodeArgs& a)
{
    const bool cluster = a.format <= TX_BC3 && a.algorithm != TX_RANGE_FIT;
    return cluster && a.nblocks >= pipeline_threshold(a.format, a.algorithm) && a.nblocks <= 0xFFFFFFFFull;
}

// Launch shape of the binned pipeline for one texture on a device with
// `cus` compute units.
static txp::PipeLaunch pipe_launch_for(const txk::EncodeArgs& a, uint32_t cus)
{
    txp::PipeLaunch p{};
    p.d_tex = nullptr;
    p.tex0.rgba = a.rgba;
    p.tex0.out = a.out;
    p.tex0.width = a.width;
    p.tex0.height = a.height;
    p.tex0.blocks_wide = a.blocks_wide;
    p.tex0.nblocks = (uint32_t)a.nblocks;
    p.tex0.variant = 3u * (uint32_t)a.format + (uint32_t)a.algorithm;
    p.tex0.pad = 0;
    p.tex0.wr = a.wr;
    p.tex0.wg = a.wg;
    p.tex0.wb = a.wb;
    p.tex0.weigh_alpha = a.weigh_alpha;
    // persistent single-wave workgroups, 16 per CU for both passes.  Classify
    // runs one block per lane: chunks of 64 blocks fill a wave; a texture
    // with fewer than 64 blocks per SIMD is cut into chunks of at least 16
    // so that it still reaches every SIMD (one wave per SIMD with all lanes
    // busy beats several part-filled ones)
    const uint64_t cwaves = (uint64_t)cus * txp::kClassifyWavesPerCu;
    const uint64_t swaves = (uint64_t)cus * txp::kSearchWavesPerCu;
    const uint64_t simds = (uint64_t)cus * 4;
    uint64_t cb = (a.nblocks + simds - 1) / simds;
    cb = std::min<uint64_t>(64, std::max<uint64_t>(16, (cb + 15) & ~(uint64_t)15));
    p.chunk_blocks = (uint32_t)cb;
    p.nchunks = (uint32_t)((a.nblocks + cb - 1) / cb);
    p.nslots = a.nblocks;
    p.d_chunks = nullptr;
    p.classify_grid = (uint32_t)std::min<uint64_t>(cwaves, p.nchunks);
    p.search_grid = (uint32_t)std::min<uint64_t>(swaves, (a.nblocks + 3) / 4);
    p.lists = txp::lists_for((uint32_t)a.format, (uint32_t)a.algorithm);
    return p;
}

// Encodes one texture on `s`: textures that take_pipeline() go through the
// binned pipeline (classify, bin, independent per-block searches),
// everything else through the direct kernel.  Both produce the same bytes.
// With `marks` (5 events) the pipeline records them between its kernels;
// `used_pipeline` says which path ran.
static hipError_t encode_on(txc::DeviceContext* ctx, const txk::EncodeArgs& a, hipStream_t s,
                            hipEvent_t* marks = nullptr, bool* used_pipeline = nullptr)
{
    if (used_pipeline) *used_pipeline = false;
    if (!takes_pipeline(a)) return txk::launch_encode(a, s);
    if (used_pipeline) *used_pipeline = true;
    std::lock_guard<std::mutex> lock(ctx->pipe_mu);
    txp::PipeBufs bufs;
    if (!ctx->pipe_scratch(a.nblocks, &bufs)) return hipErrorOutOfMemory;
    if (!ctx->pipe_before(s)) return hipErrorUnknown;
    const hipError_t e = txp::launch_pipeline(pipe_launch_for(a, (uint32_t)ctx->cus), bufs, s, marks);
    if (!ctx->pipe_after(s)) return hipErrorUnknown;
    return e;
}

int tx_compress(int format, const uint8_t* rgba, size_t rgba_len, size_t width, size_t height,
                const tx_params* params, uint8_t* output, size_t out_len)
{
    size_t in_bytes = 0;
    int rc = check_compress(format, width, height, params, out_len, &in_bytes);
    if (rc != TX_OK) return rc;
    if (rgba_len < in_bytes) return TX_ERR_INPUT_SHORT;
    if (out_len == 0) return TX_OK;
    if ((in_bytes && !rgba) || !output) return TX_ERR_INVALID;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const size_t in_cap = (in_bytes + 255) & ~(size_t)255;
    uint8_t* d = ctx->scratch(in_cap + out_len);
    if (!d) return TX_ERR_NOMEM;
    uint8_t* d_in = d;
    uint8_t* d_out = d + in_cap;
    if (in_bytes && hipMemcpyAsync(d_in, rgba, in_bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return TX_ERR_HIP;
    if (encode_on(ctx, encode_args(format, params, d_in, width, height, d_out, out_len), ctx->stream) !=
        hipSuccess)
        return TX_ERR_HIP;
    if (hipMemcpyAsync(output, d_out, out_len, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return TX_ERR_HIP;
    return TX_OK;
}

int tx_compress_device(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                       uint8_t* d_output, size_t out_len, void* stream)
{
    size_t in_bytes = 0;
    int rc = check_compress(format, width, height, params, out_len, &in_bytes);
    if (rc != TX_OK) return rc;
    if (out_len == 0) return TX_OK;
    if ((in_bytes && !d_rgba) || !d_output) return TX_ERR_INVALID;
    if (((uintptr_t)d_rgba & 3u) || ((uintptr_t)d_output & (tx_block_size(format) - 1))) return TX_ERR_ALIGN;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (encode_on(ctx, encode_args(format, params, d_rgba, width, height, d_output, out_len), s) != hipSuccess)
        return TX_ERR_HIP;
    return TX_OK;
}

// ------------------------------------------------------------ allocation-free entry point

int tx_compress_device_noalloc(int format, const uint8_t* d_rgba, size_t width, size_t height,
                               const tx_params* params, uint8_t* d_output, size_t out_len, void* stream)
{
    size_t in_bytes = 0;
    int rc = check_compress(format, width, height, params, out_len, &in_bytes);
    if (rc != TX_OK) return rc;
    if (out_len == 0) return TX_OK;
    if ((in_bytes && !d_rgba) || !d_output) return TX_ERR_INVALID;
    if (((uintptr_t)d_rgba & 3u) || ((uintptr_t)d_output & (tx_block_size(format) - 1))) return TX_ERR_ALIGN;
    const NoallocDevice* dev = nullptr;
    rc = noalloc_device(&dev);
    if (rc != TX_OK) return rc;
    const hipStream_t s = (hipStream_t)stream;
    const txk::EncodeArgs a = encode_args(format, params, d_rgba, width, height, d_output, out_len);
    if (!takes_pipeline(a)) return txk::launch_encode(a, s) == hipSuccess ? TX_OK : TX_ERR_HIP;
    if (a.nblocks > TX_NOALLOC_MAX_PIPELINE_BLOCKS) return TX_ERR_TOO_LARGE;
    txp::PipeBufs bufs;
    bufs.carve(dev->ptr, TX_NOALLOC_MAX_PIPELINE_BLOCKS);
    return txp::launch_pipeline(pipe_launch_for(a, dev->cus), bufs, s) == hipSuccess ? TX_OK : TX_ERR_HIP;
}

// ------------------------------------------------------------ mip chains

static size_t mip_side(size_t v, uint32_t level) { return level >= 64 ? 1 : std::max<size_t>(1, v >> level); }

uint32_t tx_mip_levels(size_t width, size_t height)
{
    if (width == 0 || height == 0) return 0;
    uint32_t n = 1;
    for (size_t m = std::max(width, height); m > 1; m >>= 1) ++n;
    return n;
}

void tx_mip_dims(size_t width, size_t height, uint32_t level, size_t* w_out, size_t* h_out)
{
    if (w_out) *w_out = mip_side(width, level);
    if (h_out) *h_out = mip_side(height, level);
}

size_t tx_mip_chain_rgba_bytes(size_t width, size_t height, uint32_t levels)
{
    size_t total = 0;
    for (uint32_t k = 1; k < levels; ++k) {
        size_t b = 0;
        if (!image_bytes(mip_side(width, k), mip_side(height, k), &b) || total > SIZE_MAX - b) return SIZE_MAX;
        total += b;
    }
    return total;
}

size_t tx_mip_chain_compressed_size(int format, size_t width, size_t height, uint32_t levels)
{
    size_t total = 0;
    for (uint32_t k = 0; k < levels; ++k) total += tx_compressed_size(format, mip_side(width, k), mip_side(height, k));
    return total;
}

// Shared checks of the mip-chain entry points.
static int check_mips(size_t width, size_t height, uint32_t levels, size_t chain_len)
{
    if (width == 0) return TX_ERR_ZERO_WIDTH;
    if (height == 0 || width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return TX_ERR_INVALID;
    // the whole chain holds less than 4/3 of level 0: keep level 0 below half
    // the address space so every size computed from it fits
    size_t level0 = 0;
    if (!image_bytes(width, height, &level0) || level0 > SIZE_MAX / 2) return TX_ERR_INVALID;
    if (levels == 0 || levels > tx_mip_levels(width, height)) return TX_ERR_INVALID;
    if (chain_len < tx_mip_chain_rgba_bytes(width, height, levels)) return TX_ERR_OUTPUT_SHORT;
    return TX_OK;
}

// Levels 1..levels-1 into d_chain, each from the previous one.
static hipError_t mips_on(const uint8_t* d_rgba, size_t width, size_t height, uint32_t levels, uint8_t* d_chain,
                          hipStream_t s)
{
    const uint8_t* src = d_rgba;
    size_t off = 0;
    for (uint32_t k = 1; k < levels; ++k) {
        uint8_t* dst = d_chain + off;
        const hipError_t e = txk::launch_mip_down(src, (uint32_t)mip_side(width, k - 1),
                                                  (uint32_t)mip_side(height, k - 1), dst, s);
        if (e != hipSuccess) return e;
        src = dst;
        off += 4 * mip_side(width, k) * mip_side(height, k);
    }
    return hipSuccess;
}

// Every level of the chain (level 0 = d_rgba, then d_chain) compressed back
// to back into d_output, level 0 first.
static hipError_t compress_chain_on(txc::DeviceContext* ctx, int format, const tx_params* params,
                                    const uint8_t* d_rgba, size_t width, size_t height, uint32_t levels,
                                    const uint8_t* d_chain, uint8_t* d_output, hipStream_t s)
{
    const uint8_t* src = d_rgba;
    size_t in_off = 0, out_off = 0;
    for (uint32_t k = 0; k < levels; ++k) {
        const size_t w = mip_side(width, k), h = mip_side(height, k);
        const size_t n = tx_compressed_size(format, w, h);
        if (k > 0) {
            src = d_chain + in_off;
            in_off += 4 * w * h;
        }
        const hipError_t e = encode_on(ctx, encode_args(format, params, src, w, h, d_output + out_off, n), s);
        if (e != hipSuccess) return e;
        out_off += n;
    }
    return hipSuccess;
}

int tx_generate_mips_device(const uint8_t* d_rgba, size_t width, size_t height, uint32_t levels, uint8_t* d_chain,
                            size_t chain_len, void* stream)
{
    int rc = check_mips(width, height, levels, chain_len);
    if (rc != TX_OK) return rc;
    if (levels == 1) return TX_OK;
    if (!d_rgba || !d_chain) return TX_ERR_INVALID;
    if (((uintptr_t)d_rgba & 3u) || ((uintptr_t)d_chain & 3u)) return TX_ERR_ALIGN;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    return mips_on(d_rgba, width, height, levels, d_chain, stream ? (hipStream_t)stream : ctx->stream) == hipSuccess
               ? TX_OK
               : TX_ERR_HIP;
}

int tx_compress_mips_device(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                            uint32_t levels, uint8_t* d_chain, size_t chain_len, uint8_t* d_output, size_t out_len,
                            void* stream)
{
    if (!valid_format(format) || !valid_params(params)) return TX_ERR_FORMAT;
    int rc = check_mips(width, height, levels, chain_len);
    if (rc != TX_OK) return rc;
    if (out_len < tx_mip_chain_compressed_size(format, width, height, levels)) return TX_ERR_OUTPUT_SHORT;
    if (!d_rgba || !d_output || (levels > 1 && !d_chain)) return TX_ERR_INVALID;
    if (((uintptr_t)d_rgba & 3u) || ((uintptr_t)d_chain & 3u) || ((uintptr_t)d_output & (tx_block_size(format) - 1)))
        return TX_ERR_ALIGN;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (mips_on(d_rgba, width, height, levels, d_chain, s) != hipSuccess) return TX_ERR_HIP;
    if (compress_chain_on(ctx, format, params, d_rgba, width, height, levels, d_chain, d_output, s) != hipSuccess)
        return TX_ERR_HIP;
    return TX_OK;
}

int tx_compress_mips(int format, const uint8_t* rgba, size_t rgba_len, size_t width, size_t height,
                     const tx_params* params, uint32_t levels, uint8_t* output, size_t out_len)
{
    if (!valid_format(format) || !valid_params(params)) return TX_ERR_FORMAT;
    const size_t chain_len = tx_mip_chain_rgba_bytes(width, height, levels);
    int rc = check_mips(width, height, levels, chain_len);
    if (rc != TX_OK) return rc;
    size_t in_bytes = 0;
    if (!image_bytes(width, height, &in_bytes)) return TX_ERR_INVALID;
    if (rgba_len < in_bytes) return TX_ERR_INPUT_SHORT;
    const size_t need = tx_mip_chain_compressed_size(format, width, height, levels);
    if (out_len < need) return TX_ERR_OUTPUT_SHORT;
    if (!rgba || !output) return TX_ERR_INVALID;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    uint8_t* d = ctx->scratch(up(in_bytes) + up(chain_len) + need);
    if (!d) return TX_ERR_NOMEM;
    uint8_t* d_in = d;
    uint8_t* d_chain = d + up(in_bytes);
    uint8_t* d_out = d_chain + up(chain_len);
    if (hipMemcpyAsync(d_in, rgba, in_bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    if (mips_on(d_in, width, height, levels, d_chain, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    if (compress_chain_on(ctx, format, params, d_in, width, height, levels, d_chain, d_out, ctx->stream) !=
        hipSuccess)
        return TX_ERR_HIP;
    if (hipMemcpyAsync(output, d_out, need, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return TX_ERR_HIP;
    return TX_OK;
}

int tx_compress_block_masked(int format, const uint8_t rgba[64], uint32_t mask, const tx_params* params,
                             uint8_t* output, size_t out_len)
{
    if (!valid_format(format) || !valid_params(params)) return TX_ERR_FORMAT;
    if (!rgba || !output) return TX_ERR_INVALID;
    if (out_len < tx_block_size(format)) return TX_ERR_OUTPUT_SHORT;
    return tx_compress_blocks(format, rgba, &mask, 1, params, output);
}

int tx_compress_blocks(int format, const uint8_t* blocks, const uint32_t* masks, size_t nblocks,
                       const tx_params* params, uint8_t* output)
{
    if (!valid_format(format) || !valid_params(params)) return TX_ERR_FORMAT;
    if (nblocks == 0) return TX_OK;
    if (!blocks || !masks || !output) return TX_ERR_INVALID;
    if (nblocks > SIZE_MAX / 64) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const size_t bs = tx_block_size(format);
    const size_t pix = 64 * nblocks, msk = (4 * nblocks + 255) & ~(size_t)255;
    uint8_t* d = ctx->scratch(pix + msk + nblocks * bs);
    if (!d) return TX_ERR_NOMEM;
    uint8_t* d_pix = d;
    uint32_t* d_mask = reinterpret_cast<uint32_t*>(d + pix);
    uint8_t* d_out = d + pix + msk;
    if (hipMemcpyAsync(d_pix, blocks, pix, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    if (hipMemcpyAsync(d_mask, masks, 4 * nblocks, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return TX_ERR_HIP;
    txk::BlockListArgs a;
    a.format = format;
    a.algorithm = params->algorithm;
    a.blocks = d_pix;
    a.masks = d_mask;
    a.nblocks = nblocks;
    a.out = d_out;
    a.wr = params->weights[0];
    a.wg = params->weights[1];
    a.wb = params->weights[2];
    a.weigh_alpha = params->weigh_colour_by_alpha ? 1u : 0u;
    if (txk::launch_encode_blocks(a, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    if (hipMemcpyAsync(output, d_out, nblocks * bs, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        return TX_ERR_HIP;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return TX_ERR_HIP;
    return TX_OK;
}

int tx_decompress(int format, const uint8_t* data, size_t data_len, size_t width, size_t height, uint8_t* output,
                  size_t out_len)
{
    if (!valid_format(format)) return TX_ERR_FORMAT;
    if (width == 0) return TX_ERR_ZERO_WIDTH;
    if (width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return TX_ERR_INVALID;
    size_t px_bytes = 0;
    if (!image_bytes(width, height, &px_bytes)) return TX_ERR_INVALID;
    const size_t need = tx_compressed_size(format, width, height);
    if (data_len < need) return TX_ERR_INPUT_SHORT;
    if (out_len < px_bytes) return TX_ERR_OUTPUT_SHORT;
    if (px_bytes == 0) return TX_OK;
    if (!data || !output) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const size_t in_cap = (need + 255) & ~(size_t)255;
    uint8_t* d = ctx->scratch(in_cap + px_bytes);
    if (!d) return TX_ERR_NOMEM;
    if (hipMemcpyAsync(d, data, need, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    if (txk::launch_decode(format, d, (uint32_t)width, (uint32_t)height, d + in_cap, ctx->stream) != hipSuccess)
        return TX_ERR_HIP;
    if (hipMemcpyAsync(output, d + in_cap, px_bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        return TX_ERR_HIP;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return TX_ERR_HIP;
    return TX_OK;
}

int tx_decompress_device(int format, const uint8_t* d_data, size_t width, size_t height, uint8_t* d_output,
                         void* stream)
{
    if (!valid_format(format)) return TX_ERR_FORMAT;
    if (width == 0) return TX_ERR_ZERO_WIDTH;
    if (width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return TX_ERR_INVALID;
    if (height == 0) return TX_OK;
    if (!d_data || !d_output) return TX_ERR_INVALID;
    if ((uintptr_t)d_output & 3u) return TX_ERR_ALIGN;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (txk::launch_decode(format, d_data, (uint32_t)width, (uint32_t)height, d_output, s) != hipSuccess)
        return TX_ERR_HIP;
    return TX_OK;
}

int tx_util_malloc(void** d_ptr, size_t bytes)
{
    if (!d_ptr) return TX_ERR_INVALID;
    int rc = TX_OK;
    if (!txc::current_context(&rc)) return rc;
    if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) return TX_ERR_NOMEM;
    return TX_OK;
}

int tx_util_free(void* d_ptr)
{
    if (!d_ptr) return TX_OK;
    return hipFree(d_ptr) == hipSuccess ? TX_OK : TX_ERR_HIP;
}

int tx_util_upload(void* d_dst, const void* src, size_t bytes)
{
    if (bytes == 0) return TX_OK;
    if (!d_dst || !src) return TX_ERR_INVALID;
    return hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? TX_OK : TX_ERR_HIP;
}

int tx_util_download(void* dst, const void* d_src, size_t bytes)
{
    if (bytes == 0) return TX_OK;
    if (!dst || !d_src) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return TX_ERR_HIP;
    return hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? TX_OK : TX_ERR_HIP;
}

int tx_util_copy(void* d_dst, const void* d_src, size_t bytes)
{
    if (bytes == 0) return TX_OK;
    if (!d_dst || !d_src) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    if (hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) return TX_ERR_HIP;
    return hipStreamSynchronize(ctx->stream) == hipSuccess ? TX_OK : TX_ERR_HIP;
}

int tx_util_elementary(const float* in, size_t n, float u8f[256], float sqrt_counts[17], float sqrt_alpha[4096],
                       float* snap31, float* snap63, int32_t* round255)
{
    if (!u8f || !sqrt_counts || !sqrt_alpha || n > ((size_t)1 << 26) || (n && (!in || !snap31 || !snap63 || !round255)))
        return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    constexpr size_t kTables = 256 + 17 + 4096;
    const size_t m = n ? n : 1;
    uint8_t* d = nullptr;
    const size_t bytes = 4 * (kTables + 4 * m);
    if (hipMalloc(&d, bytes) != hipSuccess) return TX_ERR_NOMEM;
    float* d_tables = reinterpret_cast<float*>(d);
    float* d_in = d_tables + kTables;
    float* d_s31 = d_in + m;
    float* d_s63 = d_s31 + m;
    int32_t* d_r = reinterpret_cast<int32_t*>(d_s63 + m);
    rc = TX_OK;
    if (n && hipMemcpyAsync(d_in, in, 4 * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && txk::launch_elementary(d_in, (uint32_t)n, d_tables, d_s31, d_s63, d_r, ctx->stream) != hipSuccess)
        rc = TX_ERR_HIP;
    if (rc == TX_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = TX_ERR_HIP;
    float tables[kTables];
    if (rc == TX_OK && hipMemcpy(tables, d_tables, sizeof tables, hipMemcpyDeviceToHost) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && n &&
        (hipMemcpy(snap31, d_s31, 4 * n, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(snap63, d_s63, 4 * n, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(round255, d_r, 4 * n, hipMemcpyDeviceToHost) != hipSuccess))
        rc = TX_ERR_HIP;
    if (rc == TX_OK) {
        memcpy(u8f, tables, 4 * 256);
        memcpy(sqrt_counts, tables + 256, 4 * 17);
        memcpy(sqrt_alpha, tables + 256 + 17, 4 * 4096);
    }
    (void)hipFree(d);
    return rc;
}

int tx_util_time_compress(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                          uint8_t* d_output, size_t out_len, int iters, float* ms_total)
{
    if (!ms_total || iters < 1) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return TX_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return TX_ERR_HIP;
    }
    rc = hipEventRecord(e0, ctx->stream) == hipSuccess ? TX_OK : TX_ERR_HIP;
    for (int i = 0; i < iters && rc == TX_OK; ++i)
        rc = tx_compress_device(format, d_rgba, width, height, params, d_output, out_len, ctx->stream);
    if (rc == TX_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && hipEventSynchronize(e1) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && hipEventElapsedTime(ms_total, e0, e1) != hipSuccess) rc = TX_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int tx_util_time_phases(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                        uint8_t* d_output, size_t out_len, int iters, float ms_phase[4], int* used_pipeline)
{
    if (!ms_phase || !used_pipeline || iters < 1) return TX_ERR_INVALID;
    size_t in_bytes = 0;
    int rc = check_compress(format, width, height, params, out_len, &in_bytes);
    if (rc != TX_OK) return rc;
    if (!d_rgba || !d_output) return TX_ERR_INVALID;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    hipEvent_t ev[5] = {};
    for (int i = 0; i < 5 && rc == TX_OK; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) rc = TX_ERR_HIP;
    for (int k = 0; k < 4; ++k) ms_phase[k] = 0.0f;
    const txk::EncodeArgs a = encode_args(format, params, d_rgba, width, height, d_output, out_len);
    for (int it = 0; it < iters && rc == TX_OK; ++it) {
        bool pipe = false;
        if (encode_on(ctx, a, ctx->stream, ev, &pipe) != hipSuccess) rc = TX_ERR_HIP;
        *used_pipeline = pipe ? 1 : 0;
        if (rc != TX_OK || !pipe) break;
        if (hipEventSynchronize(ev[4]) != hipSuccess) rc = TX_ERR_HIP;
        for (int k = 0; k < 4 && rc == TX_OK; ++k) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) rc = TX_ERR_HIP;
            ms_phase[k] += ms;
        }
    }
    if (rc == TX_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = TX_ERR_HIP;
    for (int i = 0; i < 5; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    return rc;
}

int tx_decompress_block(int format, const uint8_t* block, size_t block_len, uint8_t rgba_out[64])
{
    if (!valid_format(format)) return TX_ERR_FORMAT;
    if (block_len < tx_block_size(format)) return TX_ERR_INPUT_SHORT;
    return tx_decompress(format, block, tx_block_size(format), 4, 4, rgba_out, 64);
}

}  // extern "C"
