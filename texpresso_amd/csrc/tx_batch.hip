// tx_batch.hip -- persistent batch encoder: many textures / mip levels of
// mixed formats in one launch, fed by a device work counter over cost-ordered
// chunks, captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "bcn_device.hpp"
#include "bcn_kernels.hpp"
#include "texpresso_hip.h"
#include "tx_context.hpp"

#pragma clang fp contract(off)

namespace txd {

using txk::BatchTex;

template <int FMT, int ALG>
__device__ __forceinline__ void run_chunk(const BatchTex& T, uint32_t first, uint32_t count, GroupLds& L, int g,
                                          int l)
{
    const FitParams fp{T.wr, T.wg, T.wb, T.weigh_alpha};
    for (uint32_t off = 0; off < count; off += 4) {
        const uint32_t b = first + off + (uint32_t)g;
        const bool live = off + (uint32_t)g < count;
        const uint32_t bx = live ? b % T.blocks_wide : 0u;
        const uint32_t by = live ? b / T.blocks_wide : 0u;
        bool valid;
        const uint32_t px = fetch_pixel(T.rgba, T.width, T.height, bx, by, live, l, valid);
        uint32_t w[4];
        encode_block<FMT, ALG>(px, valid, fp, L, g, l, w);
        if (live && l == 0) {
            if (FMT == BC1 || FMT == BC4)
                *reinterpret_cast<uint2*>(T.out + (size_t)b * 8u) = make_uint2(w[0], w[1]);
            else
                *reinterpret_cast<uint4*>(T.out + (size_t)b * 16u) = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

// Persistent kernel: every wave takes chunks {texture, first block, count},
// the first by its wave index and the rest from `counter`, until the queue
// is exhausted.  A chunk's texture fixes the format and algorithm for the
// whole wave, so the dispatch is wave-uniform.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_batch(const BatchTex* __restrict__ tex,
                                                     const uint4* __restrict__ chunks, uint32_t nchunks,
                                                     uint32_t* __restrict__ counter)
{
    __shared__ GroupLds lds[4 * WAVES];
    const int tid = (int)threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, l = lane & 15;
    GroupLds& L = lds[4 * wave + g];
    // first chunk by wave index, later ones from the counter (offset past
    // those), so the launch does not open with one atomic per wave
    const uint32_t nwaves = gridDim.x * WAVES;
    for (uint32_t c = blockIdx.x * WAVES + (uint32_t)wave; c < nchunks;) {
        const uint4 ch = chunks[c];
        const BatchTex& T = tex[ch.x];
        switch (T.variant) {
        case 0: run_chunk<BC1, RANGE_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 1: run_chunk<BC1, CLUSTER_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 2: run_chunk<BC1, ITERATIVE_CLUSTER_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 3: run_chunk<BC2, RANGE_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 4: run_chunk<BC2, CLUSTER_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 5: run_chunk<BC2, ITERATIVE_CLUSTER_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 6: run_chunk<BC3, RANGE_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 7: run_chunk<BC3, CLUSTER_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 8: run_chunk<BC3, ITERATIVE_CLUSTER_FIT>(T, ch.y, ch.z, L, g, l); break;
        case 9: run_chunk<BC4, RANGE_FIT>(T, ch.y, ch.z, L, g, l); break;
        default: run_chunk<BC5, RANGE_FIT>(T, ch.y, ch.z, L, g, l); break;
        }
        uint32_t next = 0;
        if (lane == 0) next = atomicAdd(counter, 1u);
        c = __builtin_amdgcn_readfirstlane(next) + nwaves;
    }
}

// Alpha-only chunk (BC4/BC5), one block per lane: the wave covers 64
// consecutive blocks per step and loads the next step's pixels before
// encoding the current ones.
template <int FMT>
__device__ __forceinline__ void run_chunk_light(const BatchTex& T, uint32_t first, uint32_t count, int lane)
{
    auto fetch = [&](uint32_t off, uint32_t (&px)[16]) -> uint32_t {
        const uint32_t k = off + (uint32_t)lane;
        if (k >= count) return 0u;
        const uint32_t b = first + k;
        return fetch_block_lane(T.rgba, T.width, T.height, b % T.blocks_wide, b / T.blocks_wide, px);
    };
    uint32_t px[16], nx[16];
    uint32_t valid = fetch(0, px);
    for (uint32_t off = 0; off < count; off += 64) {
        const uint32_t nvalid = fetch(off + 64, nx);
        if (off + (uint32_t)lane < count) {
            uint32_t w[4];
            encode_alpha_lane<FMT>(px, valid, w);
            const size_t b = (size_t)first + off + (uint32_t)lane;
            if (FMT == BC4)
                *reinterpret_cast<uint2*>(T.out + b * 8u) = make_uint2(w[0], w[1]);
            else
                *reinterpret_cast<uint4*>(T.out + b * 16u) = make_uint4(w[0], w[1], w[2], w[3]);
        }
#pragma unroll
        for (int p = 0; p < 16; ++p) px[p] = nx[p];
        valid = nvalid;
    }
}

// Persistent kernel for the alpha-only formats (BC4/BC5): no LDS and few
// registers, so many waves stay resident to hide memory latency.  Takes
// chunks [first, first + nchunks), the first by wave index and the rest
// through its own counter.
__global__ __launch_bounds__(256) void k_batch_light(const BatchTex* __restrict__ tex,
                                                     const uint4* __restrict__ chunks, uint32_t first,
                                                     uint32_t nchunks, uint32_t* __restrict__ counter)
{
    const int lane = (int)threadIdx.x & 63;
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < nchunks;) {
        const uint4 ch = chunks[first + c];
        const BatchTex& T = tex[ch.x];
        if (T.variant == 9u) run_chunk_light<BC4>(T, ch.y, ch.z, lane);
        else run_chunk_light<BC5>(T, ch.y, ch.z, lane);
        uint32_t next = 0;
        if (lane == 0) next = atomicAdd(counter, 1u);
        c = __builtin_amdgcn_readfirstlane(next) + nwaves;
    }
}

// Clears the work counters with a kernel (a memset node is not replay-safe
// in a graph on this platform).
__global__ void k_reset_counter(uint32_t* counter)
{
    counter[0] = 0u;
    counter[1] = 0u;
}

}  // namespace txd

namespace txk {

hipError_t reset_counters(uint32_t* d_counter, hipStream_t s)
{
    hipLaunchKernelGGL(txd::k_reset_counter, dim3(1), dim3(1), 0, s, d_counter);
    return hipGetLastError();
}

hipError_t launch_alpha_chunks(const BatchTex* d_tex, const uint4* d_chunks, uint32_t first, uint32_t nlight,
                               uint32_t* d_counter, uint32_t grid, uint32_t alpha_lanes, uint32_t waves_per_group,
                               hipStream_t s)
{
    if (nlight == 0) return hipSuccess;
    if (alpha_lanes != 16) {
        hipLaunchKernelGGL(txd::k_batch_light, dim3(grid), dim3(256), 0, s, d_tex, d_chunks, first, nlight,
                           d_counter);
        return hipGetLastError();
    }
    // 16 lanes per block: the LDS encoder's persistent kernel over the
    // alpha-only chunks (it dispatches BC4/BC5 chunks to the group code)
    switch (waves_per_group) {
    case 4:
        hipLaunchKernelGGL(txd::k_batch<4>, dim3(grid), dim3(256), 0, s, d_tex, d_chunks + first, nlight, d_counter);
        break;
    case 2:
        hipLaunchKernelGGL(txd::k_batch<2>, dim3(grid), dim3(128), 0, s, d_tex, d_chunks + first, nlight, d_counter);
        break;
    default:
        hipLaunchKernelGGL(txd::k_batch<1>, dim3(grid), dim3(64), 0, s, d_tex, d_chunks + first, nlight, d_counter);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_batch_light(const BatchTex* d_tex, const uint4* d_chunks, uint32_t first, uint32_t nlight,
                              uint32_t* d_counter, uint32_t grid_light, uint32_t alpha_lanes,
                              uint32_t waves_per_group, hipStream_t s)
{
    if (nlight == 0) return hipSuccess;
    hipLaunchKernelGGL(txd::k_reset_counter, dim3(1), dim3(1), 0, s, d_counter);
    return launch_alpha_chunks(d_tex, d_chunks, first, nlight, d_counter, grid_light, alpha_lanes, waves_per_group,
                               s);
}

hipError_t launch_batch_raw(const BatchTex* d_tex, const uint4* d_chunks, uint32_t nheavy, uint32_t nlight,
                            uint32_t* d_counter, uint32_t grid, uint32_t grid_light, uint32_t waves_per_group,
                            uint32_t alpha_lanes, hipStream_t s, hipStream_t s_light, hipEvent_t fork,
                            hipEvent_t join)
{
    hipLaunchKernelGGL(txd::k_reset_counter, dim3(1), dim3(1), 0, s, d_counter);
    hipStream_t sl = s;
    if (nlight && s_light) {
        if (hipEventRecord(fork, s) != hipSuccess || hipStreamWaitEvent(s_light, fork, 0) != hipSuccess)
            return hipErrorUnknown;
        sl = s_light;
    }
    if (nlight && launch_alpha_chunks(d_tex, d_chunks, nheavy, nlight, d_counter + 1, grid_light, alpha_lanes,
                                      waves_per_group, sl) != hipSuccess)
        return hipErrorUnknown;
    if (nheavy) {
        switch (waves_per_group) {
        case 4:
            hipLaunchKernelGGL(txd::k_batch<4>, dim3(grid), dim3(256), 0, s, d_tex, d_chunks, nheavy, d_counter);
            break;
        case 2:
            hipLaunchKernelGGL(txd::k_batch<2>, dim3(grid), dim3(128), 0, s, d_tex, d_chunks, nheavy, d_counter);
            break;
        default:
            hipLaunchKernelGGL(txd::k_batch<1>, dim3(grid), dim3(64), 0, s, d_tex, d_chunks, nheavy, d_counter);
            break;
        }
    }
    if (nlight && s_light) {
        if (hipEventRecord(join, s_light) != hipSuccess || hipStreamWaitEvent(s, join, 0) != hipSuccess)
            return hipErrorUnknown;
    }
    return hipGetLastError();
}

}  // namespace txk

struct tx_batch {
    int device = -1;
    txd::BatchTex* d_tex = nullptr;
    uint4* d_chunks = nullptr;
    uint32_t* d_counter = nullptr;
    uint32_t nchunks = 0;  // heavy chunks first, then the alpha-only ones
    uint32_t nheavy = 0;
    uint64_t nblocks = 0;
    tx_batch_tuning tuning{};
    uint32_t grid = 0, grid_light = 0;
    hipStream_t light_stream = nullptr;  // runs the alpha-only kernel beside the heavy one
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint8_t* pipe_mem = nullptr;  // binned-pipeline buffers (tuning.pipeline)
    txp::PipeBufs pipe;
    uint32_t pipe_lists = 0;  // txp::PipeLaunch::lists for the batch's textures
    uint32_t pipe_grid = 0;
    uint64_t pipe_slots = 0;  // record slots covered by the heavy chunks
    uint32_t classify_grid = 0;
    // Runs of one handle share its counters and record buffers: `mu` guards
    // the lazy graph capture, and each run's stream waits for the previous
    // run (`done`), so runs on different streams execute one after another.
    std::mutex mu;
    hipEvent_t done = nullptr;
    bool ran = false;
};

namespace {

bool valid_texture(const tx_texture& t)
{
    if (t.format < TX_BC1 || t.format > TX_BC5) return false;
    if (t.params.algorithm < 0 || t.params.algorithm > 2) return false;
    if (t.width == 0 || t.width > 0x7FFFFFFFu || t.height > 0x7FFFFFFFu) return false;
    if (t.height > 0 && (!t.d_rgba || !t.d_output)) return false;
    return true;
}

// Relative per-block cost used to order chunks (most expensive first).
uint32_t cost_class(const tx_texture& t)
{
    if (t.format == TX_BC4 || t.format == TX_BC5) return 0;
    if (t.params.algorithm == TX_RANGE_FIT) return 1;
    if (t.params.algorithm == TX_CLUSTER_FIT) return t.format == TX_BC1 ? 3 : 2;
    return 4;
}

hipError_t launch_batch(tx_batch* b, hipStream_t s)
{
    if (!b->tuning.pipeline || b->nheavy == 0)
        return txk::launch_batch_raw(b->d_tex, b->d_chunks, b->nheavy, b->nchunks - b->nheavy, b->d_counter,
                                     b->grid, b->grid_light, b->tuning.waves_per_group, b->tuning.lanes_per_block,
                                     s, b->light_stream, b->fork, b->join);
    // alpha-only chunks on the side stream, the binned pipeline on `s`
    const uint32_t nlight = b->nchunks - b->nheavy;
    hipLaunchKernelGGL(txd::k_reset_counter, dim3(1), dim3(1), 0, s, b->d_counter);
    hipStream_t sl = s;
    if (nlight && b->light_stream) {
        if (hipEventRecord(b->fork, s) != hipSuccess || hipStreamWaitEvent(b->light_stream, b->fork, 0) != hipSuccess)
            return hipErrorUnknown;
        sl = b->light_stream;
    }
    if (nlight && txk::launch_alpha_chunks(b->d_tex, b->d_chunks, b->nheavy, nlight, b->d_counter + 1, b->grid_light,
                                           b->tuning.lanes_per_block, b->tuning.waves_per_group, sl) != hipSuccess)
        return hipErrorUnknown;
    txp::PipeLaunch p{};
    p.d_tex = b->d_tex;
    p.d_chunks = b->d_chunks;
    p.nchunks = b->nheavy;
    p.nslots = b->pipe_slots;
    p.chunk_blocks = b->tuning.chunk_blocks;
    p.classify_grid = std::min(b->classify_grid, b->nheavy);
    p.search_grid = b->pipe_grid;
    p.lists = b->pipe_lists;
    const hipError_t e = txp::launch_pipeline(p, b->pipe, s);
    if (e != hipSuccess) return e;
    if (nlight && b->light_stream) {
        if (hipEventRecord(b->join, b->light_stream) != hipSuccess || hipStreamWaitEvent(s, b->join, 0) != hipSuccess)
            return hipErrorUnknown;
    }
    return hipGetLastError();
}

void free_batch(tx_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->light_stream) (void)hipStreamDestroy(b->light_stream);
    if (b->pipe_mem) (void)hipFree(b->pipe_mem);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->d_tex) (void)hipFree(b->d_tex);
    if (b->d_chunks) (void)hipFree(b->d_chunks);
    if (b->d_counter) (void)hipFree(b->d_counter);
    delete b;
}

}  // namespace

extern "C" {

tx_batch_tuning tx_batch_tuning_default(void)
{
    tx_batch_tuning t;
    t.chunk_blocks = 64;
    t.waves_per_group = 1;
    t.groups_per_cu = 12;
    t.use_graph = 1;
    t.pipeline = 0;
    t.lanes_per_block = 1;
    return t;
}

int tx_batch_create(tx_batch** out, const tx_texture* textures, size_t count, const tx_batch_tuning* tuning)
{
    if (!out) return TX_ERR_INVALID;
    *out = nullptr;
    if (count && !textures) return TX_ERR_INVALID;
    if (count > 0xFFFFFFFFu) return TX_ERR_INVALID;
    tx_batch_tuning t = tuning ? *tuning : tx_batch_tuning_default();
    if (t.chunk_blocks == 0 || t.chunk_blocks % 4 != 0) return TX_ERR_INVALID;
    if (t.waves_per_group != 1 && t.waves_per_group != 2 && t.waves_per_group != 4) return TX_ERR_INVALID;
    if (t.groups_per_cu == 0 || t.groups_per_cu > 64 || t.pipeline > 1) return TX_ERR_INVALID;
    if (t.lanes_per_block == 0) t.lanes_per_block = 1;
    if (t.lanes_per_block != 1 && t.lanes_per_block != 16) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;

    std::vector<txd::BatchTex> tex(count);
    std::vector<uint4> chunks;
    std::vector<uint32_t> cls;
    uint64_t nblocks = 0;
    for (size_t i = 0; i < count; ++i) {
        const tx_texture& s = textures[i];
        if (!valid_texture(s)) return TX_ERR_INVALID;
        const size_t bs = tx_block_size(s.format);
        if (((uintptr_t)s.d_rgba & 3u) || ((uintptr_t)s.d_output & (bs - 1))) return TX_ERR_ALIGN;
        txd::BatchTex& d = tex[i];
        d.rgba = s.d_rgba;
        d.out = s.d_output;
        d.width = s.width;
        d.height = s.height;
        d.blocks_wide = (s.width + 3) / 4;
        const uint64_t nb = (uint64_t)d.blocks_wide * ((s.height + 3) / 4);
        if (nb > 0xFFFFFFFFull) return TX_ERR_INVALID;
        d.nblocks = (uint32_t)nb;
        d.variant = (s.format >= TX_BC4) ? 9u + (uint32_t)(s.format - TX_BC4)
                                         : 3u * (uint32_t)s.format + (uint32_t)s.params.algorithm;
        d.pad = 0;
        d.wr = s.params.weights[0];
        d.wg = s.params.weights[1];
        d.wb = s.params.weights[2];
        d.weigh_alpha = s.params.weigh_colour_by_alpha ? 1u : 0u;
        // alpha-only textures one block per lane take larger chunks
        const uint32_t cb = (cost_class(s) == 0 && t.lanes_per_block == 1) ? txk::kLightChunkBlocks : t.chunk_blocks;
        for (uint32_t f = 0; f < d.nblocks; f += cb) {
            chunks.push_back(make_uint4((uint32_t)i, f, std::min(cb, d.nblocks - f), 0u));
            cls.push_back(cost_class(s));
        }
        nblocks += nb;
    }
    // most expensive chunks first (longest-processing-time order); stable
    // so each texture's chunks stay in block order
    std::vector<size_t> order(chunks.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cls[a] > cls[b]; });
    // heavy chunks also get their first record slot of the binned pipeline
    std::vector<uint4> sorted(chunks.size());
    uint64_t slots = 0;
    for (size_t i = 0; i < order.size(); ++i) {
        sorted[i] = chunks[order[i]];
        if (cls[order[i]] > 0) {
            sorted[i].w = (uint32_t)slots;
            slots += sorted[i].z;
        }
    }
    if (slots > 0xFFFFFFFFull) return TX_ERR_INVALID;

    tx_batch* b = new (std::nothrow) tx_batch();
    if (!b) return TX_ERR_NOMEM;
    b->device = ctx->device;
    b->tuning = t;
    b->nchunks = (uint32_t)sorted.size();
    b->nheavy = 0;
    for (size_t i = 0; i < order.size(); ++i) b->nheavy += cls[order[i]] > 0 ? 1u : 0u;
    b->nblocks = nblocks;
    hipDeviceProp_t prop;
    if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
        hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) {
        free_batch(b);
        return TX_ERR_HIP;
    }
    // persistent workgroups: the tuned count, but never more than there are chunks
    b->grid = (uint32_t)std::min<uint64_t>((uint64_t)prop.multiProcessorCount * t.groups_per_cu,
                                           std::max<uint64_t>(1, (b->nheavy + t.waves_per_group - 1) /
                                                                     t.waves_per_group));
    // alpha-only kernel: one block per lane in 4-wave workgroups (up to 8 per
    // CU), or 16 lanes per block in the tuned workgroups
    const uint32_t nlight = b->nchunks - b->nheavy;
    b->grid_light = txk::alpha_grid((uint32_t)prop.multiProcessorCount, nlight, t.lanes_per_block, t.groups_per_cu,
                                    t.waves_per_group);
    if (nlight && b->nheavy) {
        if (hipStreamCreateWithFlags(&b->light_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess) {
            free_batch(b);
            return TX_ERR_HIP;
        }
    }
    if (t.pipeline && b->nheavy) {
        const uint64_t heavy_blocks = slots;
        b->pipe_slots = slots;
        for (size_t i = 0; i < count; ++i)
            if (tex[i].variant < 9u) b->pipe_lists |= txp::lists_for(tex[i].variant / 3u, tex[i].variant % 3u);
        b->pipe_grid = (uint32_t)std::min<uint64_t>((uint64_t)prop.multiProcessorCount * txp::kSearchWavesPerCu,
                                                    std::max<uint64_t>(1, (heavy_blocks + 3) / 4));
        b->classify_grid = (uint32_t)prop.multiProcessorCount * txp::kClassifyWavesPerCu;
        if (hipMalloc(&b->pipe_mem, txp::PipeBufs::bytes_for(std::max<uint64_t>(1, heavy_blocks))) != hipSuccess) {
            free_batch(b);
            return TX_ERR_NOMEM;
        }
        b->pipe.carve(b->pipe_mem, std::max<uint64_t>(1, heavy_blocks));
    }
    if (hipMalloc(&b->d_tex, sizeof(txd::BatchTex) * std::max<size_t>(1, tex.size())) != hipSuccess ||
        hipMalloc(&b->d_chunks, sizeof(uint4) * std::max<size_t>(1, sorted.size())) != hipSuccess ||
        hipMalloc(&b->d_counter, 256) != hipSuccess) {
        free_batch(b);
        return TX_ERR_NOMEM;
    }
    if ((tex.size() && hipMemcpy(b->d_tex, tex.data(), sizeof(txd::BatchTex) * tex.size(),
                                 hipMemcpyHostToDevice) != hipSuccess) ||
        (sorted.size() && hipMemcpy(b->d_chunks, sorted.data(), sizeof(uint4) * sorted.size(),
                                    hipMemcpyHostToDevice) != hipSuccess)) {
        free_batch(b);
        return TX_ERR_HIP;
    }
    *out = b;
    return TX_OK;
}

int tx_batch_run(tx_batch* b, void* stream)
{
    if (!b) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    if (ctx->device != b->device) return TX_ERR_INVALID;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (b->nchunks == 0) return TX_OK;
    std::lock_guard<std::mutex> lock(b->mu);
    if (b->ran && hipStreamWaitEvent(s, b->done, 0) != hipSuccess) return TX_ERR_HIP;
    if (!b->tuning.use_graph) {
        if (launch_batch(b, s) != hipSuccess) return TX_ERR_HIP;
    } else {
        if (!b->exec) {
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return TX_ERR_HIP;
            const hipError_t le = launch_batch(b, s);
            const hipError_t ce = hipStreamEndCapture(s, &b->graph);
            if (le != hipSuccess || ce != hipSuccess) {
                if (b->graph) (void)hipGraphDestroy(b->graph);
                b->graph = nullptr;
                return TX_ERR_HIP;
            }
            if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
                b->exec = nullptr;
                return TX_ERR_HIP;
            }
        }
        if (hipGraphLaunch(b->exec, s) != hipSuccess) return TX_ERR_HIP;
    }
    if (hipEventRecord(b->done, s) != hipSuccess) return TX_ERR_HIP;
    b->ran = true;
    return TX_OK;
}

int tx_batch_run_timed(tx_batch* b, int iters, float* ms_total)
{
    if (!b || !ms_total || iters < 1) return TX_ERR_INVALID;
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
    for (int i = 0; i < iters && rc == TX_OK; ++i) rc = tx_batch_run(b, ctx->stream);
    if (rc == TX_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && hipEventSynchronize(e1) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && hipEventElapsedTime(ms_total, e0, e1) != hipSuccess) rc = TX_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int tx_batch_destroy(tx_batch* b)
{
    if (!b) return TX_OK;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    {
        // the last run, on whichever stream it was enqueued
        std::lock_guard<std::mutex> lock(b->mu);
        if (b->ran) (void)hipEventSynchronize(b->done);
    }
    free_batch(b);
    return TX_OK;
}

int tx_batch_info(const tx_batch* b, uint64_t* chunks, uint64_t* blocks)
{
    if (!b) return TX_ERR_INVALID;
    if (chunks) *chunks = b->nchunks;
    if (blocks) *blocks = b->nblocks;
    return TX_OK;
}

int tx_batch_autotune(const tx_texture* textures, size_t count, const tx_batch_tuning* candidates, size_t ncand,
                      int iters, tx_batch_tuning* best, float* best_ms)
{
    if (!best || !best_ms || iters < 1) return TX_ERR_INVALID;
    std::vector<tx_batch_tuning> grid;
    bool alpha_only = false;  // BC4/BC5 textures present: lanes per block is tuned too
    for (size_t i = 0; i < count && textures; ++i) alpha_only = alpha_only || textures[i].format >= TX_BC4;
    if (candidates && ncand) {
        grid.assign(candidates, candidates + ncand);
    } else {
        const uint32_t chunk[] = {16, 32, 64, 128, 256};
        const uint32_t wpg[] = {1, 2, 4};
        const uint32_t waves_per_cu[] = {4, 8, 12};
        for (uint32_t c : chunk) {
            // binned pipeline: its grids are fixed, only the classify chunk varies
            tx_batch_tuning p = tx_batch_tuning_default();
            p.chunk_blocks = c;
            p.pipeline = 1;
            grid.push_back(p);
            for (uint32_t w : wpg)
                for (uint32_t v : waves_per_cu) {
                    tx_batch_tuning t = tx_batch_tuning_default();
                    t.chunk_blocks = c;
                    t.waves_per_group = w;
                    t.groups_per_cu = std::max(1u, v / w);
                    t.use_graph = 1;
                    t.pipeline = 0;
                    grid.push_back(t);
                }
        }
        if (alpha_only) {
            const size_t n = grid.size();
            for (size_t i = 0; i < n; ++i) {
                tx_batch_tuning t = grid[i];
                t.lanes_per_block = 16;
                grid.push_back(t);
            }
        }
    }
    float best_t = -1.0f;
    for (const tx_batch_tuning& t : grid) {
        tx_batch* b = nullptr;
        int rc = tx_batch_create(&b, textures, count, &t);
        if (rc != TX_OK) return rc;
        float ms = 0.0f;
        rc = tx_batch_run_timed(b, 1, &ms);  // warm-up (and graph capture)
        if (rc == TX_OK) rc = tx_batch_run_timed(b, iters, &ms);
        tx_batch_destroy(b);
        if (rc != TX_OK) return rc;
        ms /= (float)iters;
        if (best_t < 0.0f || ms < best_t) {
            best_t = ms;
            *best = t;
        }
    }
    *best_ms = best_t;
    return TX_OK;
}

int tx_util_time_per_texture(const tx_texture* textures, size_t count, int iters, float* ms_total)
{
    if (!ms_total || iters < 1 || (count && !textures)) return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    for (size_t i = 0; i < count; ++i)
        if (!valid_texture(textures[i])) return TX_ERR_INVALID;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return TX_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return TX_ERR_HIP;
    }
    rc = hipEventRecord(e0, ctx->stream) == hipSuccess ? TX_OK : TX_ERR_HIP;
    for (int it = 0; it < iters && rc == TX_OK; ++it) {
        for (size_t i = 0; i < count && rc == TX_OK; ++i) {
            const tx_texture& t = textures[i];
            rc = tx_compress_device(t.format, t.d_rgba, t.width, t.height, &t.params, t.d_output,
                                    tx_compressed_size(t.format, t.width, t.height), ctx->stream);
        }
    }
    if (rc == TX_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && hipEventSynchronize(e1) != hipSuccess) rc = TX_ERR_HIP;
    if (rc == TX_OK && hipEventElapsedTime(ms_total, e0, e1) != hipSuccess) rc = TX_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // extern "C"
