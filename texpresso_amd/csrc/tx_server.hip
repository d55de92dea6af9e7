// tx_server.hip -- continuous-batching compression server.
//
// One scheduler thread per server.  Each turn it (1) retires finished
// launches (event query), (2) admits pending jobs in arrival order while their
// source and output fit in the device arena, and (3) folds all newly
// admitted jobs into one persistent-kernel launch on a free stream of the
// ring.  Launches on different streams overlap on the device, so work that
// arrives while the GPU is busy joins the next launch instead of waiting for
// the previous one to drain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <time.h>
#include <unordered_map>
#include <vector>

#include "bcn_kernels.hpp"
#include "texpresso_hip.h"
#include "tx_context.hpp"

namespace {

uint64_t now_ns()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

// First-fit allocator over one device allocation (offsets, 256-byte aligned).
class Arena {
public:
    bool init(size_t bytes)
    {
        size_ = (bytes + 255) & ~(size_t)255;
        if (hipMalloc(&base_, size_) != hipSuccess) {
            base_ = nullptr;
            return false;
        }
        free_[0] = size_;
        return true;
    }
    void release()
    {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
    }
    uint8_t* base() const { return base_; }
    size_t size() const { return size_; }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        }
        return false;
    }
    void free(size_t off, size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    uint8_t* base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;
};

struct Job {
    tx_job spec{};
    tx_job_info info{};
    size_t in_off = 0, in_bytes = 0, out_off = 0, out_bytes = 0;
    bool has_memory = false;
    bool release_when_done = false;  // release requested while running or being fetched
    uint32_t fetching = 0;           // tx_server_fetch copies in progress from the job's output
};

struct Launch {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    txk::BatchTex* h_tex = nullptr;  // pinned
    uint4* h_chunks = nullptr;       // pinned
    txk::BatchTex* d_tex = nullptr;
    uint4* d_chunks = nullptr;
    uint32_t* d_counter = nullptr;   // [0..1] colour kernel, [2..3] alpha-only kernel
    size_t chunk_cap = 0;
    uint8_t* pipe_mem = nullptr;     // binned-pipeline records of this slot (grown on demand)
    txp::PipeBufs pipe;
    // chunk tables outgrown while serving, freed when the server is destroyed:
    // hipFree and hipHostFree wait for the whole device, which would stall
    // the launches of the other slots
    std::vector<void*> retired_host, retired_dev;
    std::vector<uint64_t> jobs;
    bool busy = false;
};

// Largest number of BC1-BC3 blocks a launch sends through the binned
// pipeline (records: 336 bytes each, so 1.4 GB per launch slot at most);
// larger launches use the lockstep persistent kernel.
constexpr uint64_t kServerPipeMaxBlocks = 4u << 20;

}  // namespace

struct tx_server {
    tx_server_config cfg{};
    int device = -1;
    uint32_t grid = 0;
    uint32_t cus = 0;
    Arena arena;
    std::vector<Launch> launches;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Job> jobs;
    uint64_t next_id = 1;
    uint32_t next_launch = 1;
    bool stop = false;
    std::thread worker;
    hipStream_t fetch_stream = nullptr;
    std::mutex fetch_mu;  // serialises copies on fetch_stream

    void run();
    bool start_launch(Launch& L, std::vector<uint64_t>& ids, std::unique_lock<std::mutex>& lock);
    void free_job_memory(Job& j)
    {
        if (j.has_memory) {
            arena.free(j.in_off, j.in_bytes);
            arena.free(j.out_off, j.out_bytes);
            j.has_memory = false;
        }
    }
    // Frees and forgets a finished job whose release was requested, once no
    // fetch reads its output any more; call with `mu` held.
    void release_if_requested(std::unordered_map<uint64_t, Job>::iterator it)
    {
        Job& j = it->second;
        if (!j.release_when_done || j.fetching) return;
        free_job_memory(j);
        jobs.erase(it);
        cv_work.notify_one();
    }
};

namespace {

bool grow_chunks(Launch& L, size_t need)
{
    if (need <= L.chunk_cap) return true;
    size_t cap = std::max<size_t>(need, L.chunk_cap * 2);
    if (L.h_chunks) L.retired_host.push_back(L.h_chunks);
    if (L.d_chunks) L.retired_dev.push_back(L.d_chunks);
    L.h_chunks = nullptr;
    L.d_chunks = nullptr;
    L.chunk_cap = 0;
    if (hipHostMalloc(&L.h_chunks, cap * sizeof(uint4)) != hipSuccess) return false;
    if (hipMalloc(&L.d_chunks, cap * sizeof(uint4)) != hipSuccess) return false;
    L.chunk_cap = cap;
    return true;
}

uint32_t variant_of(const tx_job& s)
{
    return s.format >= TX_BC4 ? 9u + (uint32_t)(s.format - TX_BC4) : 3u * (uint32_t)s.format + (uint32_t)s.params.algorithm;
}

// Record capacity of a slot's pipeline scratch: at least `blocks`.  The slot
// is idle (its previous launch has retired), so the old buffer can go.
bool grow_pipe(Launch& L, uint64_t blocks)
{
    if (L.pipe_mem && L.pipe.capacity >= blocks) return true;
    const uint64_t cap = std::min<uint64_t>(kServerPipeMaxBlocks, std::max<uint64_t>(blocks, 2 * L.pipe.capacity));
    if (L.pipe_mem) (void)hipFree(L.pipe_mem);
    L.pipe_mem = nullptr;
    L.pipe = txp::PipeBufs{};
    if (hipMalloc(&L.pipe_mem, txp::PipeBufs::bytes_for(cap)) != hipSuccess) {
        L.pipe_mem = nullptr;
        return false;
    }
    L.pipe.carve(L.pipe_mem, cap);
    return true;
}

}  // namespace

// Called with the lock held; drops it while enqueueing device work.
bool tx_server::start_launch(Launch& L, std::vector<uint64_t>& ids, std::unique_lock<std::mutex>& lock)
{
    // chunks small enough that every resident wave gets work even when the
    // launch holds a single small job (latency), capped by the tuned size
    uint64_t colour_blocks = 0;  // BC1-BC3 (LDS encoder or pipeline); BC4/BC5 go to the alpha-only kernel
    uint32_t lists = 0;  // pipeline lists the launch's records can go to
    for (uint64_t id : ids) {
        const Job& j = jobs[id];
        if (j.spec.format <= TX_BC3)
            colour_blocks += (uint64_t)((j.spec.width + 3) / 4) * ((j.spec.height + 3) / 4);
        lists |= txp::lists_for((uint32_t)j.spec.format, (uint32_t)j.spec.params.algorithm);
    }
    const uint64_t waves = (uint64_t)grid * cfg.tuning.waves_per_group;
    uint64_t want = (colour_blocks + waves - 1) / std::max<uint64_t>(1, waves);
    want = std::max<uint64_t>(4, (want + 3) & ~(uint64_t)3);
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(cfg.tuning.chunk_blocks, want);
    // BC4/BC5 jobs: one block per lane in chunks of 512, or 16 lanes per block
    // in the colour chunks' size
    const uint32_t alpha_lanes = cfg.tuning.lanes_per_block == 16 ? 16u : 1u;
    const uint32_t light_chunk = alpha_lanes == 16 ? chunk : txk::kLightChunkBlocks;
    const bool pipeline = cfg.tuning.pipeline && colour_blocks >= txc::pipeline_min_blocks() &&
                          colour_blocks <= kServerPipeMaxBlocks;
    size_t nchunks = 0;
    std::vector<tx_job> specs;
    std::vector<size_t> in_offs, out_offs;
    for (uint64_t id : ids) {
        Job& j = jobs[id];
        const uint64_t nb = (uint64_t)((j.spec.width + 3) / 4) * ((j.spec.height + 3) / 4);
        const uint32_t cb = j.spec.format >= TX_BC4 ? light_chunk : chunk;
        nchunks += (size_t)((nb + cb - 1) / cb);
        specs.push_back(j.spec);
        in_offs.push_back(j.in_off);
        out_offs.push_back(j.out_off);
        j.info.status = TX_JOB_RUNNING;
        j.info.t_launch_ns = now_ns();
    }
    const uint32_t launch_id = next_launch++;
    for (uint64_t id : ids) {
        jobs[id].info.launch_id = launch_id;
        jobs[id].info.launch_jobs = (uint32_t)ids.size();
    }
    L.jobs = ids;
    L.busy = true;
    lock.unlock();

    bool ok = grow_chunks(L, std::max<size_t>(1, nchunks));
    // without room for the records the launch falls back to the lockstep kernel
    const bool use_pipe = ok && pipeline && grow_pipe(L, colour_blocks);
    uint8_t* base = arena.base();
    size_t c = 0;
    // expensive jobs first so the cheap ones fill the tail
    std::vector<size_t> order(specs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return (specs[a].format >= TX_BC4) < (specs[b].format >= TX_BC4); });
    for (size_t k = 0; ok && k < specs.size(); ++k) {
        const tx_job& s = specs[k];
        const size_t in_bytes = (size_t)s.width * s.height * 4;
        if (in_bytes && hipMemcpyAsync(base + in_offs[k], s.rgba, in_bytes, hipMemcpyHostToDevice, L.stream) !=
                            hipSuccess)
            ok = false;
        txk::BatchTex& t = L.h_tex[k];
        t.rgba = base + in_offs[k];
        t.out = base + out_offs[k];
        t.width = s.width;
        t.height = s.height;
        t.blocks_wide = (s.width + 3) / 4;
        t.nblocks = t.blocks_wide * ((s.height + 3) / 4);
        t.variant = variant_of(s);
        t.pad = 0;
        t.wr = s.params.weights[0];
        t.wg = s.params.weights[1];
        t.wb = s.params.weights[2];
        t.weigh_alpha = s.params.weigh_colour_by_alpha ? 1u : 0u;
    }
    // colour chunks first (each with its first record slot for the pipeline),
    // then the alpha-only chunks
    uint32_t nheavy = 0, slot = 0;
    for (size_t q = 0; ok && q < order.size(); ++q) {
        const size_t k = order[q];
        const txk::BatchTex& t = L.h_tex[k];
        const bool light = t.variant >= 9u;
        const uint32_t cb = light ? light_chunk : chunk;
        for (uint32_t f = 0; f < t.nblocks; f += cb) {
            const uint32_t n = std::min(cb, t.nblocks - f);
            L.h_chunks[c++] = make_uint4((uint32_t)k, f, n, light ? 0u : slot);
            if (!light) {
                slot += n;
                ++nheavy;
            }
        }
    }
    const uint32_t nlight = (uint32_t)c - nheavy;
    if (ok && hipMemcpyAsync(L.d_tex, L.h_tex, sizeof(txk::BatchTex) * specs.size(), hipMemcpyHostToDevice,
                             L.stream) != hipSuccess)
        ok = false;
    if (ok && c && hipMemcpyAsync(L.d_chunks, L.h_chunks, sizeof(uint4) * c, hipMemcpyHostToDevice, L.stream) !=
                       hipSuccess)
        ok = false;
    // no more persistent workgroups than chunks
    const uint32_t launch_grid = (uint32_t)std::min<uint64_t>(
        grid, std::max<uint64_t>(1, (nheavy + cfg.tuning.waves_per_group - 1) / cfg.tuning.waves_per_group));
    const uint32_t light_grid = txk::alpha_grid(cus, nlight, alpha_lanes, cfg.tuning.groups_per_cu,
                                                cfg.tuning.waves_per_group);
    if (ok && use_pipe) {
        if (txk::launch_batch_light(L.d_tex, L.d_chunks, nheavy, nlight, L.d_counter + 2, light_grid, alpha_lanes,
                                    cfg.tuning.waves_per_group, L.stream) !=
            hipSuccess)
            ok = false;
        txp::PipeLaunch p{};
        p.d_tex = L.d_tex;
        p.d_chunks = L.d_chunks;
        p.nchunks = nheavy;
        p.nslots = slot;
        p.chunk_blocks = chunk;
        p.classify_grid = std::min<uint32_t>(cus * txp::kClassifyWavesPerCu, nheavy);
        p.search_grid = (uint32_t)std::min<uint64_t>((uint64_t)cus * txp::kSearchWavesPerCu, (slot + 3) / 4);
        p.lists = lists;
        if (ok && txp::launch_pipeline(p, L.pipe, L.stream) != hipSuccess) ok = false;
    } else if (ok && c &&
               txk::launch_batch_raw(L.d_tex, L.d_chunks, nheavy, nlight, L.d_counter, launch_grid, light_grid,
                                     cfg.tuning.waves_per_group, alpha_lanes, L.stream, nullptr, nullptr,
                                     nullptr) != hipSuccess) {
        ok = false;
    }
    if (ok && hipEventRecord(L.done, L.stream) != hipSuccess) ok = false;
    lock.lock();
    if (!ok) {
        for (uint64_t id : ids) {
            auto it = jobs.find(id);
            if (it == jobs.end()) continue;
            it->second.info.status = TX_ERR_HIP;
            it->second.info.t_done_ns = now_ns();
            // a release requested while the lock was dropped frees it now
            release_if_requested(it);
        }
        (void)hipStreamSynchronize(L.stream);
        L.busy = false;
        L.jobs.clear();
        cv_done.notify_all();
    }
    return ok;
}

void tx_server::run()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
        bool progressed = false;
        bool inflight = false;
        // (1) retire finished launches
        for (Launch& L : launches) {
            if (!L.busy) continue;
            const hipError_t q = hipEventQuery(L.done);
            if (q == hipErrorNotReady) {
                inflight = true;
                continue;
            }
            const uint64_t t = now_ns();
            for (uint64_t id : L.jobs) {
                auto it = jobs.find(id);
                if (it == jobs.end()) continue;
                it->second.info.status = q == hipSuccess ? TX_JOB_DONE : TX_ERR_HIP;
                it->second.info.t_done_ns = t;
                release_if_requested(it);
            }
            L.jobs.clear();
            L.busy = false;
            progressed = true;
            cv_done.notify_all();
        }
        if (stop && !inflight && pending.empty()) break;
        // (2) admit in arrival order while memory allows, (3) launch them
        Launch* slot = nullptr;
        for (Launch& L : launches)
            if (!L.busy) {
                slot = &L;
                break;
            }
        if (slot && !pending.empty()) {
            std::vector<uint64_t> ids;
            while (!pending.empty() && ids.size() < cfg.max_jobs_per_launch) {
                Job& j = jobs[pending.front()];
                if (!arena.alloc(j.in_bytes, &j.in_off)) break;
                if (!arena.alloc(j.out_bytes, &j.out_off)) {
                    arena.free(j.in_off, j.in_bytes);
                    break;
                }
                j.has_memory = true;
                j.info.status = TX_JOB_ADMITTED;
                j.info.t_admit_ns = now_ns();
                j.info.d_output = arena.base() + j.out_off;
                ids.push_back(pending.front());
                pending.pop_front();
            }
            if (!ids.empty()) {
                start_launch(*slot, ids, lock);
                progressed = true;
                inflight = true;
            }
        }
        if (!progressed) {
            if (inflight) cv_work.wait_for(lock, std::chrono::microseconds(20));
            else cv_work.wait_for(lock, std::chrono::milliseconds(5));
        }
    }
}

extern "C" {

tx_server_config tx_server_config_default(void)
{
    tx_server_config c;
    c.hbm_budget = (size_t)8 << 30;
    c.max_inflight = 3;
    c.max_jobs_per_launch = 64;
    c.max_pending = 4096;
    c.tuning = tx_batch_tuning_default();
    c.tuning.use_graph = 0;
    c.tuning.pipeline = 1;  // launches with many ClusterFit blocks use the binned pipeline
    return c;
}

int tx_server_create(tx_server** out, const tx_server_config* config)
{
    if (!out) return TX_ERR_INVALID;
    *out = nullptr;
    tx_server_config cfg = config ? *config : tx_server_config_default();
    if (cfg.max_inflight < 1 || cfg.max_inflight > 8 || cfg.max_jobs_per_launch < 1 ||
        cfg.max_jobs_per_launch > 4096 || cfg.max_pending < 1 || cfg.hbm_budget < 4096)
        return TX_ERR_INVALID;
    if (cfg.tuning.chunk_blocks == 0 || cfg.tuning.chunk_blocks % 4 ||
        (cfg.tuning.waves_per_group != 1 && cfg.tuning.waves_per_group != 2 && cfg.tuning.waves_per_group != 4) ||
        cfg.tuning.groups_per_cu == 0 || cfg.tuning.groups_per_cu > 64 || cfg.tuning.pipeline > 1 ||
        (cfg.tuning.lanes_per_block != 0 && cfg.tuning.lanes_per_block != 1 && cfg.tuning.lanes_per_block != 16))
        return TX_ERR_INVALID;
    int rc = TX_OK;
    txc::DeviceContext* ctx = txc::current_context(&rc);
    if (!ctx) return rc;
    tx_server* s = new (std::nothrow) tx_server();
    if (!s) return TX_ERR_NOMEM;
    s->cfg = cfg;
    s->device = ctx->device;
    hipDeviceProp_t prop;
    bool ok = hipGetDeviceProperties(&prop, ctx->device) == hipSuccess;
    s->grid = ok ? (uint32_t)prop.multiProcessorCount * cfg.tuning.groups_per_cu : 0;
    s->cus = ok ? (uint32_t)prop.multiProcessorCount : 0;
    ok = ok && s->arena.init(cfg.hbm_budget);
    ok = ok && hipStreamCreateWithFlags(&s->fetch_stream, hipStreamNonBlocking) == hipSuccess;
    s->launches.resize(cfg.max_inflight);
    for (Launch& L : s->launches) {
        ok = ok && hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&L.done, hipEventDisableTiming) == hipSuccess;
        ok = ok && hipHostMalloc(&L.h_tex, sizeof(txk::BatchTex) * cfg.max_jobs_per_launch) == hipSuccess;
        ok = ok && hipMalloc(&L.d_tex, sizeof(txk::BatchTex) * cfg.max_jobs_per_launch) == hipSuccess;
        ok = ok && hipMalloc(&L.d_counter, 256) == hipSuccess;
        // chunk table for a launch of 2^20 colour blocks (one 4096^2 texture)
        // at the tuned chunk size, or one chunk per resident wave for small
        // launches, without growing while serving
        ok = ok && grow_chunks(L, (size_t)s->grid * cfg.tuning.waves_per_group +
                                      ((size_t)1 << 20) / cfg.tuning.chunk_blocks +
                                      2 * (size_t)cfg.max_jobs_per_launch);
        // record scratch for the largest launch the arena can hold (its
        // inputs take 64 bytes per block), allocated once: a later free would
        // wait for the whole device in the middle of serving
        if (cfg.tuning.pipeline)
            ok = ok && grow_pipe(L, std::min<uint64_t>(kServerPipeMaxBlocks, std::max<uint64_t>(1, cfg.hbm_budget / 64)));
    }
    // touch the arena and the record scratch once now: the first access to
    // fresh device memory costs milliseconds, which would otherwise land on
    // the first jobs of each launch slot
    ok = ok && hipMemsetAsync(s->arena.base(), 0, s->arena.size(), s->fetch_stream) == hipSuccess;
    for (Launch& L : s->launches)
        if (ok && L.pipe_mem)
            ok = hipMemsetAsync(L.pipe_mem, 0, txp::PipeBufs::bytes_for(L.pipe.capacity), s->fetch_stream) ==
                 hipSuccess;
    ok = ok && hipStreamSynchronize(s->fetch_stream) == hipSuccess;
    // a stream gets its hardware queue at its first dispatch: do that now
    // (a counter reset) rather than in the first job of each slot
    for (Launch& L : s->launches)
        if (ok) ok = txk::reset_counters(L.d_counter, L.stream) == hipSuccess &&
                     hipStreamSynchronize(L.stream) == hipSuccess;
    // the runtime stages copies from pageable host memory in 1 MiB pieces
    // and sets that path up on a stream's first large copy: a first 2048^2
    // job spent ~0.4 ms between each of its 16 upload pieces (14 ms instead
    // of 1.8 ms from launch to done), and fetches ran at 1.8 ms per 4 MiB
    // instead of 0.6 ms.  One 16 MiB copy each way now, through every slot
    // stream (uploads) and the fetch stream (downloads), moves that here.
    {
        const size_t n = std::min<size_t>((size_t)16 << 20, s->arena.size());
        std::vector<uint8_t> host(n, 0);
        for (Launch& L : s->launches)
            if (ok && n) ok = hipMemcpyAsync(s->arena.base(), host.data(), n, hipMemcpyHostToDevice, L.stream) ==
                             hipSuccess &&
                         hipStreamSynchronize(L.stream) == hipSuccess;
        if (ok && n) ok = hipMemcpyAsync(host.data(), s->arena.base(), n, hipMemcpyDeviceToHost, s->fetch_stream) ==
                         hipSuccess &&
                     hipStreamSynchronize(s->fetch_stream) == hipSuccess;
    }
    if (!ok) {
        tx_server_destroy(s);
        return TX_ERR_NOMEM;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return TX_OK;
}

int tx_server_submit(tx_server* s, const tx_job* job, uint64_t* id)
{
    if (!s || !job || !id) return TX_ERR_INVALID;
    if (job->format < TX_BC1 || job->format > TX_BC5 || job->params.algorithm < 0 || job->params.algorithm > 2)
        return TX_ERR_FORMAT;
    if (job->width == 0 || job->height == 0 || job->width > 0x7FFFFFFFu || job->height > 0x7FFFFFFFu || !job->rgba)
        return TX_ERR_INVALID;
    Job j;
    j.spec = *job;
    j.in_bytes = (size_t)job->width * job->height * 4;
    j.out_bytes = tx_compressed_size(job->format, job->width, job->height);
    if (((j.in_bytes + 255) & ~(size_t)255) + ((j.out_bytes + 255) & ~(size_t)255) > s->arena.size())
        return TX_ERR_NOMEM;
    j.info.status = TX_JOB_PENDING;
    j.info.out_len = j.out_bytes;
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stop) return TX_ERR_INVALID;
    if (s->pending.size() >= s->cfg.max_pending) return TX_ERR_BUSY;
    j.info.t_submit_ns = now_ns();
    const uint64_t nid = s->next_id++;
    s->jobs.emplace(nid, j);
    s->pending.push_back(nid);
    *id = nid;
    s->cv_work.notify_one();
    return TX_OK;
}

int tx_server_query(tx_server* s, uint64_t id, tx_job_info* info)
{
    if (!s || !info) return TX_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return TX_ERR_INVALID;
    *info = it->second.info;
    return TX_OK;
}

int tx_server_wait(tx_server* s, uint64_t id, int timeout_ms, tx_job_info* info)
{
    if (!s) return TX_ERR_INVALID;
    std::unique_lock<std::mutex> lock(s->mu);
    auto finished = [&] {
        auto it = s->jobs.find(id);
        return it == s->jobs.end() || it->second.info.status == TX_JOB_DONE || it->second.info.status < 0;
    };
    if (timeout_ms < 0) s->cv_done.wait(lock, finished);
    else if (!s->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), finished)) return TX_ERR_BUSY;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return TX_ERR_INVALID;
    if (info) *info = it->second.info;
    return it->second.info.status < 0 ? it->second.info.status : TX_OK;
}

int tx_server_fetch(tx_server* s, uint64_t id, uint8_t* out, size_t out_len)
{
    if (!s || !out) return TX_ERR_INVALID;
    const uint8_t* src = nullptr;
    size_t n = 0;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end() || it->second.info.status != TX_JOB_DONE || it->second.release_when_done)
            return TX_ERR_INVALID;
        src = it->second.info.d_output;
        n = it->second.info.out_len;
        if (out_len < n) return TX_ERR_OUTPUT_SHORT;
        // the output region stays allocated until this copy ends, even if the
        // job is released meanwhile (the release is deferred to the end)
        ++it->second.fetching;
    }
    (void)hipSetDevice(s->device);
    // one copy at a time on the fetch stream, so a failing copy's error is its own
    int rc = TX_OK;
    {
        std::lock_guard<std::mutex> copy_lock(s->fetch_mu);
        if (hipMemcpyAsync(out, src, n, hipMemcpyDeviceToHost, s->fetch_stream) != hipSuccess ||
            hipStreamSynchronize(s->fetch_stream) != hipSuccess)
            rc = TX_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it != s->jobs.end()) {
        --it->second.fetching;
        s->release_if_requested(it);
    }
    return rc;
}

int tx_server_release(tx_server* s, uint64_t id)
{
    if (!s) return TX_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return TX_ERR_INVALID;
    Job& j = it->second;
    if (j.info.status == TX_JOB_DONE || j.info.status < 0) {
        // freed now, or when the fetches still copying its output end
        j.release_when_done = true;
        s->release_if_requested(it);
        return TX_OK;
    }
    if (j.info.status == TX_JOB_PENDING) {
        s->pending.erase(std::remove(s->pending.begin(), s->pending.end(), id), s->pending.end());
        s->jobs.erase(it);
        return TX_OK;
    }
    j.release_when_done = true;
    return TX_OK;
}

int tx_server_destroy(tx_server* s)
{
    if (!s) return TX_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        // jobs never admitted are cancelled
        for (uint64_t id : s->pending) {
            Job& j = s->jobs[id];
            j.info.status = TX_ERR_INVALID;
            j.info.t_done_ns = now_ns();
        }
        s->pending.clear();
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device >= 0 ? s->device : 0);
    for (Launch& L : s->launches) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
        if (L.h_tex) (void)hipHostFree(L.h_tex);
        if (L.h_chunks) (void)hipHostFree(L.h_chunks);
        if (L.d_tex) (void)hipFree(L.d_tex);
        if (L.d_chunks) (void)hipFree(L.d_chunks);
        if (L.d_counter) (void)hipFree(L.d_counter);
        if (L.pipe_mem) (void)hipFree(L.pipe_mem);
        for (void* p : L.retired_host) (void)hipHostFree(p);
        for (void* p : L.retired_dev) (void)hipFree(p);
    }
    if (s->fetch_stream) (void)hipStreamDestroy(s->fetch_stream);
    s->arena.release();
    delete s;
    return TX_OK;
}

}  // extern "C"
