// tx_pipeline.hip -- binned ClusterFit pipeline: classify, bin, search.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bcn_device.hpp"
#include "bcn_lane.hpp"
#include "tx_pipeline.hpp"

#pragma clang fp contract(off)

namespace txp {

void PipeBufs::carve(uint8_t* base, uint64_t cap)
{
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = 0;
    recs = reinterpret_cast<ClusterRec*>(base + off);
    off += up(sizeof(ClusterRec) * cap);
    keys = reinterpret_cast<uint32_t*>(base + off);
    off += up(4 * cap);
    for (int k = 0; k < kLists; ++k) {
        list[k] = reinterpret_cast<uint32_t*>(base + off);
        off += up(4 * cap);
    }
    ctl = reinterpret_cast<PipeCtl*>(base + off);
    capacity = cap;
}

}  // namespace txp

namespace txd {

using txk::BatchTex;
using txp::ClusterRec;
using txp::PipeCtl;

using txp::kLists;
using txp::kListA;
using txp::kListB;
using txp::kListC;
using txp::kSearchWaves;
constexpr uint32_t key_bit(int list) { return 1u << (5 + list); }

__global__ void k_pipe_reset(PipeCtl* ctl)
{
    uint32_t* w = reinterpret_cast<uint32_t*>(ctl);
    for (uint32_t i = threadIdx.x; i < sizeof(PipeCtl) / 4; i += blockDim.x) w[i] = 0u;
}

__device__ __forceinline__ const BatchTex& tex_at(const BatchTex* tex, const BatchTex& tex0, uint32_t i)
{
    return tex ? tex[i] : tex0;
}

// Calls f(g, m) once for every value g of `group` among the lanes with `in`
// set, m = the ballot of those lanes.  Uniform over the wave.
template <class F>
__device__ __forceinline__ void for_each_group(bool in, uint32_t group, F&& f)
{
    uint64_t rem = __ballot(in);
    while (rem) {
        const int src = (int)__builtin_ctzll(rem);
        const uint32_t g = (uint32_t)__builtin_amdgcn_readlane((int)group, src);
        const uint64_t m = __ballot(in && group == g);
        rem &= ~m;
        f(g, m);
    }
}

// Counts a wave of keys (lane slot `slot`) into the count copies: one ballot
// per (copy, colour count) present among the lanes of each list the texture
// variant feeds, and one atomic add by its first lane.
template <int FMT, int ALG>
__device__ __forceinline__ void count_keys(PipeCtl* __restrict__ ctl, uint32_t key, uint32_t slot, int lane)
{
    const uint32_t group = ((slot / txp::kBinSeg) % txp::kHistCopies) << 5 | (key & 31u);
#pragma unroll
    for (int k = 0; k < kLists; ++k) {
        if (ALG == RANGE_FIT) continue;
        if (ALG == CLUSTER_FIT && k != kListC) continue;
        if (ALG == ITERATIVE_CLUSTER_FIT && (k == kListC || (k == kListA && FMT != BC1))) continue;
        for_each_group((key & key_bit(k)) != 0u, group, [&](uint32_t g, uint64_t m) {
            if (lane == (int)__builtin_ctzll(m)) atomicAdd(&ctl->part[g >> 5][k][g & 31u], (uint32_t)__popcll(m));
        });
    }
}

// ClusterFit block on one lane: the rest of its record in slot r (the points
// were written by the front's centroid fold), and its key.
template <int FMT, int ALG>
__device__ __forceinline__ uint32_t write_record_lane(ClusterRec* __restrict__ recs, uint32_t r,
                                                      const LaneSet& S, const uint32_t (&out)[4],
                                                      float vx, float vy, float vz, const FitParams& fp,
                                                      uint32_t tex_idx, uint32_t block)
{
    ClusterRec& R = recs[r];  // points: written by the front's centroid fold
    const bool need4 = !(FMT == BC1 && S.transparent);
    const uint32_t info = (uint32_t)S.n | ((S.transparent ? 1u : 0u) << 8) | ((uint32_t)FMT << 12) |
                          ((ALG == ITERATIVE_CLUSTER_FIT ? 8u : 1u) << 16);
    uint4* tail = reinterpret_cast<uint4*>(R.remap);
    tail[0] = make_uint4(S.remap[0], S.remap[1], S.remap[2], S.remap[3]);
    tail[1] = make_uint4(out[0], out[1], __float_as_uint(vx), __float_as_uint(vy));
    tail[2] = make_uint4(__float_as_uint(vz), __float_as_uint(FLT_MAX), 0u, 0u);
    tail[3] = make_uint4(__float_as_uint(fp.wr), __float_as_uint(fp.wg), __float_as_uint(fp.wb), tex_idx);
    tail[4] = make_uint4(block, info, 0u, 0u);
    // ClusterFit: list C (all of the block's searches in one pass);
    // IterativeClusterFit: the 3-colour search for BC1 (list A) and the
    // 4-colour one unless transparent (list B)
    uint32_t key = (uint32_t)S.n;
    if (ALG == CLUSTER_FIT) {
        key |= key_bit(kListC);
    } else {
        if (FMT == BC1) key |= key_bit(kListA);
        if (need4) key |= key_bit(kListB);
    }
    return key;
}

// Blocks [first, first + count) of one texture, one per lane, 64 at a time.
// Single-colour and range-fit blocks are finished here; ClusterFit blocks
// become records in slots [slot0, slot0 + count), counted in ctl->part.
template <int FMT, int ALG>
__device__ __forceinline__ void classify_chunk(const BatchTex& T, uint32_t tex_idx, uint32_t first, uint32_t count,
                                               uint32_t slot0, const LaneTables& LT, LaneBlockLds& LB, int lane,
                                               ClusterRec* recs, uint32_t* keys, PipeCtl* ctl)
{
    const FitParams fp{T.wr, T.wg, T.wb, T.weigh_alpha};
    for (uint32_t off = 0; off < count; off += 64) {
        const uint32_t k = off + (uint32_t)lane;
        uint32_t key = 0;
        if (k < count) {
            const uint32_t b = first + k;
            uint32_t px[16];
            const uint32_t valid = fetch_block_lane(T.rgba, T.width, T.height, b % T.blocks_wide,
                                                    b / T.blocks_wide, px);
            uint32_t w[4];
            LaneSet S;
            float vx, vy, vz;
            const int kind = encode_front_lane<FMT, ALG>(px, valid, fp, LT, LB, lane, w, S, vx, vy, vz,
                                                         recs[slot0 + k].pts);
            if (kind == KIND_CLUSTER) {
                key = write_record_lane<FMT, ALG>(recs, slot0 + k, S, w, vx, vy, vz, fp, tex_idx, b);
            } else if (FMT == BC1) {
                *reinterpret_cast<uint2*>(T.out + (size_t)b * 8u) = make_uint2(w[0], w[1]);
            } else {
                *reinterpret_cast<uint4*>(T.out + (size_t)b * 16u) = make_uint4(w[0], w[1], w[2], w[3]);
            }
            keys[slot0 + k] = key;
        }
        count_keys<FMT, ALG>(ctl, key, slot0 + k, lane);
    }
}

// Classify over the heavy chunks (BC1/BC2/BC3 textures), each workgroup
// taking chunks blockIdx.x, blockIdx.x + gridDim.x, ...  Chunk c covers
// blocks [first, first + count) of one texture and record slots
// [slot0, slot0 + count).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_pipe_classify(const BatchTex* __restrict__ tex, BatchTex tex0,
                                                      const uint4* __restrict__ chunks, uint32_t nchunks,
                                                      uint32_t chunk_blocks, ClusterRec* __restrict__ recs,
                                                      uint32_t* __restrict__ keys, PipeCtl* __restrict__ ctl)
{
    __shared__ LaneTables LT;
    __shared__ LaneBlockLds LB;
    const int lane = (int)threadIdx.x;
    fill_lane_tables(LT, lane);
    group_sync();
    // chunks hold similar work, so a plain grid-stride loop balances them
    // (a shared work counter would serialise thousands of atomics at launch)
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        uint4 ch;
        if (chunks) {
            ch = chunks[c];
        } else {
            const uint32_t f = c * chunk_blocks;
            ch = make_uint4(0u, f, min(chunk_blocks, tex0.nblocks - f), f);
        }
        const BatchTex& T = tex_at(tex, tex0, ch.x);
        switch (T.variant) {
        case 0: classify_chunk<BC1, RANGE_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        case 1: classify_chunk<BC1, CLUSTER_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        case 2: classify_chunk<BC1, ITERATIVE_CLUSTER_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        case 3: classify_chunk<BC2, RANGE_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        case 4: classify_chunk<BC2, CLUSTER_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        case 5: classify_chunk<BC2, ITERATIVE_CLUSTER_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        case 6: classify_chunk<BC3, RANGE_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        case 7: classify_chunk<BC3, CLUSTER_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        default: classify_chunk<BC3, ITERATIVE_CLUSTER_FIT>(T, ch.x, ch.y, ch.z, ch.w, LT, LB, lane, recs, keys, ctl); break;
        }
    }
}

// Counting sort of the records by colour count, 16 first, one workgroup
// per slot segment s (kBinSeg slots).  Bin (list k, colour count v) starts
// after bins v+1..16 of its list and holds the records of count copy 0, then
// copy 1, ...; the workgroup sums the copies into the totals (workgroup 0
// also stores them in ctl->hist for the search passes) and into the start
// of its own copy c = s % kHistCopies, counts its keys per bin in LDS,
// reserves one range per non-empty bin with one atomic add on copy c's
// cursor, then places its records inside those ranges (order within a bin
// is free: it only steers which records share a wave).
__global__ __launch_bounds__(256) void k_pipe_bin(const uint32_t* __restrict__ keys, uint64_t nslots,
                                                  PipeCtl* __restrict__ ctl, uint32_t* __restrict__ list_a,
                                                  uint32_t* __restrict__ list_b, uint32_t* __restrict__ list_c)
{
    constexpr uint32_t kCounts = kLists * 17;
    __shared__ uint32_t cnt[kLists][17], base[kLists][17], tot[kLists][17], pre[kLists][17];
    uint32_t* const lists[kLists] = {list_a, list_b, list_c};
    const uint32_t t = threadIdx.x;
    const uint32_t copy = blockIdx.x % txp::kHistCopies;
    if (t < 17)
        for (int k = 0; k < kLists; ++k) cnt[k][t] = tot[k][t] = pre[k][t] = 0u;
    __syncthreads();
    // totals and the start of this copy: 4 threads per count, each summing
    // every 4th copy
    if (t < 4 * kCounts) {
        const uint32_t km = t % kCounts;
        uint32_t sum = 0, below = 0;
        for (uint32_t c = t / kCounts; c < txp::kHistCopies; c += 4) {
            const uint32_t v = (&ctl->part[c][0][0])[km];
            sum += v;
            below += c < copy ? v : 0u;
        }
        if (sum) atomicAdd(&(&tot[0][0])[km], sum);
        if (below) atomicAdd(&(&pre[0][0])[km], below);
    }
    const uint64_t s0 = (uint64_t)blockIdx.x * txp::kBinSeg;
    const uint64_t s1 = s0 + txp::kBinSeg < nslots ? s0 + txp::kBinSeg : nslots;
    for (uint64_t r = s0 + t; r < s1; r += blockDim.x) {
        const uint32_t key = keys[r];
        for (int k = 0; k < kLists; ++k)
            if (key & key_bit(k)) atomicAdd(&cnt[k][key & 31u], 1u);
    }
    __syncthreads();
    if (blockIdx.x == 0 && t < kCounts) (&ctl->hist[0][0])[t] = (&tot[0][0])[t];
    if (t < 17) {
        for (int k = 0; k < kLists; ++k) {
            uint32_t o = pre[k][t];
            for (int m = 16; m > (int)t; --m) o += tot[k][m];
            base[k][t] = cnt[k][t] ? o + atomicAdd(&ctl->pos[copy][k][t], cnt[k][t]) : 0u;
            cnt[k][t] = 0u;
        }
    }
    __syncthreads();
    for (uint64_t r = s0 + t; r < s1; r += blockDim.x) {
        const uint32_t key = keys[r];
        const uint32_t n = key & 31u;
        for (int k = 0; k < kLists; ++k)
            if (key & key_bit(k)) lists[k][base[k][n] + atomicAdd(&cnt[k][n], 1u)] = (uint32_t)r;
    }
}

// The record fields a finished block's output needs, read with the rest of
// the record when the group takes it (not behind a load at emit time).
struct RecOut {
    uint32_t tex, block, alpha0, alpha1;
};

// Result of a finished block: the output block, or (NC = 3 for an opaque BC1
// block) the 3-colour result kept in the record for the 4-colour pass.
template <int NC>
__device__ __forceinline__ void emit_block(const BatchTex* __restrict__ tex, const BatchTex& tex0, ClusterRec& R,
                                           const RecOut& ro, uint32_t info, const GroupLds& L, int l, int n,
                                           int remap, float lb_err, float thr, uint32_t lb_code, int lb_it,
                                           const float (&sa)[3], const float (&sb)[3])
{
    const uint32_t fmt = (info >> 12) & 15u;
    uint32_t lo = 0, hi = 0;
    const bool won = lb_err < thr;
    if (won) {
        write_colour(NC == 4, sa[0], sa[1], sa[2], sb[0], sb[1], sb[2], split_index<NC>(L, n, remap, lb_code, lb_it),
                     l, lo, hi);
    } else if (NC == 4 && fmt == BC1) {
        lo = R.fit3[0];
        hi = R.fit3[1];
    }
    if (l == 0) {
        const BatchTex& T = tex_at(tex, tex0, ro.tex);
        if (NC == 3) {
            if ((info >> 8) & 1u) {
                *reinterpret_cast<uint2*>(T.out + (size_t)ro.block * 8u) = make_uint2(lo, hi);
            } else {
                R.fit3_err = won ? lb_err : thr;
                R.fit3[0] = lo;
                R.fit3[1] = hi;
            }
        } else if (fmt == BC1) {
            *reinterpret_cast<uint2*>(T.out + (size_t)ro.block * 8u) = make_uint2(lo, hi);
        } else {
            *reinterpret_cast<uint4*>(T.out + (size_t)ro.block * 16u) = make_uint4(ro.alpha0, ro.alpha1, lo, hi);
        }
    }
}

// Lane 0's value of v (the lane that took a wave's list positions).
__device__ __forceinline__ uint32_t lane0(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, 0); }

// Takes `nwant` (1..4) list positions for one wave from its workgroup's run
// [next, end), packed as (end << 32) | next in LDS and shared by the
// workgroup's 16 groups.  Executed by one lane.  A wave that finds the run
// empty takes the refill lock, reserves the next `refill` positions from
// the shared cursor and installs them; the other waves wait for the lock
// to drop (its holder waits on nothing).  The positions come in up to
// nwant consecutive pieces and are returned in order in pos[0..nwant).
// Positions past the list's end are handed out like any other; their
// groups go idle.
__device__ __forceinline__ void take_positions(uint64_t* run, uint32_t* refilling, uint32_t* __restrict__ cursor,
                                               uint32_t cursor_base, uint32_t refill, uint32_t nwant,
                                               uint32_t (&pos)[4])
{
    uint32_t got = 0;
    while (got < nwant) {
        uint64_t w = __hip_atomic_load(run, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t next = (uint32_t)w, end = (uint32_t)(w >> 32);
        if (next < end) {
            const uint32_t k = min(nwant - got, end - next);
            if (__hip_atomic_compare_exchange_strong(run, &w, ((uint64_t)end << 32) | (next + k), __ATOMIC_ACQ_REL,
                                                     __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) {
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i)
                    if (i >= got && i < got + k) pos[i] = next + (i - got);
                got += k;
            }
            continue;
        }
        uint32_t unlocked = 0u;
        if (__hip_atomic_compare_exchange_strong(refilling, &unlocked, 1u, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP)) {
            // only the lock holder installs a run, and no taker changes an
            // empty one, so a run still empty here stays empty until the store
            w = __hip_atomic_load(run, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            if ((uint32_t)w >= (uint32_t)(w >> 32)) {
                const uint32_t fresh = atomicAdd(cursor, refill) + cursor_base;
                __hip_atomic_store(run, ((uint64_t)(fresh + refill) << 32) | fresh, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            __hip_atomic_store(refilling, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// Persistent search: each 16-lane group pulls records and runs the block's
// ordering iterations independently (cluster.rs:174-249 / 301-389).  NC = 3:
// BC1's compress3, whose result goes to the record (or, for a BC1 block with
// transparent pixels, straight to the output).  NC = 4: compress4 from the
// 3-colour result (BC1) or from nothing (BC2/BC3), then the output block.
//
// Each step has two phases.  Ordering phase: groups without a pending search
// build their next ordering; an ordering equal to an earlier one ends that
// block at once (cluster.rs:176-178, 303-305), and the group pulls its next record and
// tries again, so no group enters the search phase empty-handed while records
// remain.  Search phase: every group with a fresh ordering evaluates its
// candidate splits; the candidate loop runs as long as the largest colour
// count among them needs, and the records are sorted by colour count.
//
// A workgroup is kSearchWaves waves (4 groups each) sharing one run of list
// positions: a group takes one position at a time from it, so near the end
// of the list no group holds records that an idle group could have taken,
// while the shared cursor still sees one atomic per 16 positions.  The waves
// are otherwise independent (no workgroup barrier after the set-up).
template <int NC>
__global__ __launch_bounds__(64 * kSearchWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_pipe_search(
    const BatchTex* __restrict__ tex, BatchTex tex0, ClusterRec* __restrict__ recs, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ hist, uint32_t* __restrict__ cursor)
{
    __shared__ GroupLds lds[4 * kSearchWaves];
    __shared__ uint64_t run;
    __shared__ uint32_t refilling;
    const int lane = (int)(threadIdx.x & 63u), g = lane >> 4, l = lane & 15;
    GroupLds& L = lds[4 * (threadIdx.x >> 6) + g];
    uint32_t total = 0;
    for (int m = 0; m < 17; ++m) total += hist[m];
    // refills of kRefill positions: at most one per group of the workgroup
    // and at most the records per wave, so a short list still spreads over
    // every workgroup; the first run of each workgroup is fixed by its index
    // (no atomic burst at launch) and refills come from the shared cursor,
    // offset past those runs
    const uint32_t kRefill = min(16u, max(4u, total / (kSearchWaves * gridDim.x)));
    const uint32_t kFirst = kSearchWaves * kRefill;
    const uint32_t cursor_base = gridDim.x * kFirst;
    if (threadIdx.x == 0) {
        const uint32_t first = blockIdx.x * kFirst;
        run = ((uint64_t)(first + kFirst) << 32) | first;
        refilling = 0u;
    }
    __syncthreads();

    bool idle = false, need = true, ready = false;
    uint32_t r = 0, info = 0, lb_code = 0;
    RecOut ro{0u, 0u, 0u, 0u};
    int n = 0, iters = 1, it = 0, lb_it = 0, remap = -1;
    float thr = FLT_MAX, lb_err = FLT_MAX;
    float sa[3] = {0.0f, 0.0f, 0.0f}, sb[3] = {0.0f, 0.0f, 0.0f};
    float axx = 0.0f, axy = 0.0f, axz = 0.0f;
    FitParams fp{0.0f, 0.0f, 0.0f, 0u};
    for (;;) {
        // ---- ordering phase
        for (;;) {
            // groups without a block take the next record
            if (wave_any(need && !idle)) {
                const bool take = need && !idle;
                const uint64_t want = __ballot(take && l == 0);
                const uint32_t nwant = (uint32_t)__popcll(want);
                const uint32_t rank = (uint32_t)__popcll(want & ((1ull << (uint32_t)(g * 16)) - 1ull));
                uint32_t pos[4] = {0u, 0u, 0u, 0u};
                if (lane == 0) take_positions(&run, &refilling, cursor, cursor_base, kRefill, nwant, pos);
                const uint32_t p0 = lane0(pos[0]), p1 = lane0(pos[1]), p2 = lane0(pos[2]), p3 = lane0(pos[3]);
                const uint32_t p = rank == 0u ? p0 : rank == 1u ? p1 : rank == 2u ? p2 : p3;
                if (take) {
                    if (p >= total) {
                        idle = true;
                    } else {
                        r = list[p];
                        const ClusterRec& R = recs[r];
                        L.pts[l] = R.pts[l];
                        info = R.info;
                        ro = RecOut{R.tex, R.block, R.alpha[0], R.alpha[1]};
                        n = (int)(info & 31u);
                        iters = (int)((info >> 16) & 15u);
                        const uint32_t byte = (R.remap[l >> 2] >> (8 * (l & 3))) & 255u;
                        remap = byte == 255u ? -1 : (int)byte;
                        fp = FitParams{R.w[0], R.w[1], R.w[2], 0u};
                        axx = R.axis[0];
                        axy = R.axis[1];
                        axz = R.axis[2];
                        thr = NC == 4 ? R.fit3_err : FLT_MAX;
                        lb_err = thr;
                        lb_code = 0;
                        lb_it = 0;
                        it = 0;
                        sa[0] = sa[1] = sa[2] = 0.0f;
                        sb[0] = sb[1] = sb[2] = 0.0f;
                        need = false;
                    }
                }
                group_sync();
            }
            const bool todo = !idle && !ready;
            if (!wave_any(todo)) break;
            const bool unique = ordering_step(L, l, n, todo, it, axx, axy, axz);
            ready = ready || unique;
            const bool repeated = todo && !unique;
            if (wave_any(repeated)) {
                if (repeated) {
                    emit_block<NC>(tex, tex0, recs[r], ro, info, L, l, n, remap, lb_err, thr, lb_code, lb_it, sa, sb);
                    need = true;
                }
            }
        }
        if (!wave_any(ready)) break;
        // ---- search phase
        float be;
        uint32_t bc;
        search_step<NC>(L, l, n, ready, fp, be, bc);
        bool fin = false;
        if (ready) {
            ready = false;
            if (be < lb_err) {
                lb_err = be;
                lb_code = bc;
                lb_it = it;
                split_endpoints<NC>(L, n, bc, fp, sa, sb);
            }
            if (lb_it != it || it + 1 >= iters) {
                fin = true;
            } else {
                axx = sb[0] - sa[0];
                axy = sb[1] - sa[1];
                axz = sb[2] - sa[2];
                ++it;
            }
        }
        if (wave_any(fin)) {
            if (fin) {
                emit_block<NC>(tex, tex0, recs[r], ro, info, L, l, n, remap, lb_err, thr, lb_code, lb_it, sa, sb);
                need = true;
            }
        }
    }
}

// ClusterFit records (colourfit.rs:44-60 with a single ordering per search,
// cluster.rs:152-417): the four groups of a wave pull a record each and run
// the block's searches on it side by side, exactly as the direct kernel
// does -- compress3 and then, unless the block is transparent, compress4 for
// BC1; compress4 for BC2/BC3.  With one ordering per search there is nothing
// to gain from letting the groups run independently, and for BC1 one pull
// serves both searches.  Records are sorted by colour count, so the four
// candidate loops have the same length.  As in k_pipe_search, the
// workgroup's kSearchWaves waves take their positions, four per step, from
// one shared run, so at the end of the list no wave holds steps that an
// idle wave could have run.
__global__ __launch_bounds__(64 * kSearchWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_pipe_cluster(
    const BatchTex* __restrict__ tex, BatchTex tex0, const ClusterRec* __restrict__ recs,
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ hist, uint32_t* __restrict__ cursor)
{
    __shared__ GroupLds lds[4 * kSearchWaves];
    __shared__ uint64_t run;
    __shared__ uint32_t refilling;
    const int lane = (int)(threadIdx.x & 63u), g = lane >> 4, l = lane & 15;
    GroupLds& L = lds[4 * (threadIdx.x >> 6) + g];
    uint32_t total = 0;
    for (int m = 0; m < 17; ++m) total += hist[m];
    // refills of up to four steps for one wave (16 positions), at most the
    // records per wave; the first run of each workgroup is fixed by its
    // index, refills come from the shared cursor, offset past those runs
    const uint32_t kRefill = 4u * min(4u, max(1u, total / (4u * kSearchWaves * gridDim.x)));
    const uint32_t kFirst = kSearchWaves * kRefill;
    const uint32_t cursor_base = gridDim.x * kFirst;
    if (threadIdx.x == 0) {
        const uint32_t first = blockIdx.x * kFirst;
        run = ((uint64_t)(first + kFirst) << 32) | first;
        refilling = 0u;
    }
    __syncthreads();
    for (;;) {
        uint32_t pos[4] = {0u, 0u, 0u, 0u};
        if (lane == 0) take_positions(&run, &refilling, cursor, cursor_base, kRefill, 4u, pos);
        const uint32_t p0 = lane0(pos[0]), p1 = lane0(pos[1]), p2 = lane0(pos[2]), p3 = lane0(pos[3]);
        const uint32_t p = g == 0 ? p0 : g == 1 ? p1 : g == 2 ? p2 : p3;
        const bool has = p < total;
        if (!wave_any(has)) break;
        ColourSet cs{0, -1, false};
        FitParams fp{0.0f, 0.0f, 0.0f, 0u};
        float axx = 0.0f, axy = 0.0f, axz = 0.0f;
        uint32_t tex_idx = 0, block = 0, fmt = BC1, alo = 0, ahi = 0;
        if (has) {
            const ClusterRec& R = recs[list[p]];
            L.pts[l] = R.pts[l];
            const uint32_t info = R.info;
            fmt = (info >> 12) & 15u;
            alo = R.alpha[0];
            ahi = R.alpha[1];
            cs.n = (int)(info & 31u);
            cs.transparent = ((info >> 8) & 1u) != 0u;
            const uint32_t byte = (R.remap[l >> 2] >> (8 * (l & 3))) & 255u;
            cs.remap = byte == 255u ? -1 : (int)byte;
            fp = FitParams{R.w[0], R.w[1], R.w[2], 0u};
            axx = R.axis[0];
            axy = R.axis[1];
            axz = R.axis[2];
            tex_idx = R.tex;
            block = R.block;
        }
        group_sync();
        float best = FLT_MAX;
        uint32_t lo = 0, hi = 0;  // all-zero block if no candidate wins (colourfit.rs:44-60)
        run_cluster<3, 1>(L, cs, l, has && fmt == BC1, axx, axy, axz, fp, best, lo, hi);
        run_cluster<4, 1>(L, cs, l, has && !(fmt == BC1 && cs.transparent), axx, axy, axz, fp, best, lo, hi,
                          has && fmt == BC1);
        if (has && l == 0) {
            const BatchTex& T = tex_at(tex, tex0, tex_idx);
            if (fmt == BC1)
                *reinterpret_cast<uint2*>(T.out + (size_t)block * 8u) = make_uint2(lo, hi);
            else
                *reinterpret_cast<uint4*>(T.out + (size_t)block * 16u) = make_uint4(alo, ahi, lo, hi);
        }
    }
}

}  // namespace txd

namespace txp {

hipError_t launch_pipeline(const PipeLaunch& p, const PipeBufs& b, hipStream_t s, hipEvent_t* marks)
{
    if (p.nslots > b.capacity || p.nslots > 0xFFFFFFFFull) return hipErrorInvalidValue;
    auto mark = [&](int i) { return marks ? hipEventRecord(marks[i], s) : hipSuccess; };
    hipLaunchKernelGGL(txd::k_pipe_reset, dim3(1), dim3(256), 0, s, b.ctl);
    if (p.nchunks == 0 || p.nslots == 0) {
        for (int i = 0; i < 5; ++i)
            if (mark(i) != hipSuccess) return hipErrorUnknown;
        return hipGetLastError();
    }
    if (mark(0) != hipSuccess) return hipErrorUnknown;
    hipLaunchKernelGGL(txd::k_pipe_classify, dim3(p.classify_grid), dim3(64), 0, s, p.d_tex, p.tex0, p.d_chunks,
                       p.nchunks, p.chunk_blocks, b.recs, b.keys, b.ctl);
    if (mark(1) != hipSuccess) return hipErrorUnknown;
    const uint32_t bin_grid = (uint32_t)((p.nslots + kBinSeg - 1) / kBinSeg);
    const uint32_t search_wgs = (p.search_grid + kSearchWaves - 1) / kSearchWaves;
    hipLaunchKernelGGL(txd::k_pipe_bin, dim3(bin_grid), dim3(256), 0, s, b.keys, p.nslots, b.ctl, b.list[kListA],
                       b.list[kListB], b.list[kListC]);
    if (mark(2) != hipSuccess) return hipErrorUnknown;
    if (p.lists & (1u << kListC))
        hipLaunchKernelGGL(txd::k_pipe_cluster, dim3(search_wgs), dim3(64 * kSearchWaves), 0, s, p.d_tex, p.tex0, b.recs,
                           b.list[kListC], b.ctl->hist[kListC], &b.ctl->cursor[kListC]);
    if (p.lists & (1u << kListA))
        hipLaunchKernelGGL(txd::k_pipe_search<3>, dim3(search_wgs), dim3(64 * kSearchWaves), 0, s, p.d_tex, p.tex0, b.recs,
                           b.list[kListA], b.ctl->hist[kListA], &b.ctl->cursor[kListA]);
    if (mark(3) != hipSuccess) return hipErrorUnknown;
    if (p.lists & (1u << kListB))
        hipLaunchKernelGGL(txd::k_pipe_search<4>, dim3(search_wgs), dim3(64 * kSearchWaves), 0, s, p.d_tex, p.tex0, b.recs,
                           b.list[kListB], b.ctl->hist[kListB], &b.ctl->cursor[kListB]);
    if (mark(4) != hipSuccess) return hipErrorUnknown;
    return hipGetLastError();
}

}  // namespace txp
