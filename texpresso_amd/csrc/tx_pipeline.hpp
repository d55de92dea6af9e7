// tx_pipeline.hpp -- the binned ClusterFit pipeline (host-side interface).
//
//   reset -> classify -> bin -> search<3> (BC1) -> search<4>
//
// classify: alpha, colour set and fit choice for every block, one block per
//   lane; single-colour and range-fit blocks are finished here, ClusterFit
//   blocks become a ClusterRec and are counted by distinct-colour count n.
// bin: counting sort of the records by n (16 first) into one list per pass
//   (list A: BC1 IterativeClusterFit records for the 3-colour search; list
//   B: records that need the 4-colour search -- BC1 IterativeClusterFit
//   without transparency, BC2/BC3 IterativeClusterFit; list C: ClusterFit
//   records of every format, which get all their searches in one pass).
// search<NC>: persistent waves; each 16-lane group pulls the next record
//   from its list and runs that block's ordering iterations on its own, so
//   blocks with one ordering and blocks with several share a wave without
//   waiting for each other, and neighbouring records have the same n (the
//   same number of candidate rounds).
// cluster: ClusterFit (a single ordering per search): each group pulls a
//   record and runs its searches on it (BC1: 3-colour then 4-colour).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "bcn_kernels.hpp"

namespace txp {

// One ClusterFit block between the passes (336 bytes).
struct __align__(16) ClusterRec {
    float4 pts[16];     // distinct colours: x, y, z, sqrt(weight)
    uint32_t remap[4];  // one byte per pixel: point index, 255 = excluded
    uint32_t alpha[2];  // alpha half of the block (BC2/BC3)
    float axis[3];      // principal axis
    float fit3_err;     // 3-colour result: ClusterFit::best_error after compress3
    uint32_t fit3[2];   // 3-colour result: ClusterFit::best_compressed after compress3
    float w[3];         // Params::weights
    uint32_t tex;       // texture index
    uint32_t block;     // block index within the texture
    uint32_t info;      // n | transparent << 8 | format << 12 | iterations << 16
    uint32_t pad[2];
};
static_assert(sizeof(ClusterRec) == 336, "record layout");
// the classify pass stores the tail as five 16-byte words starting at remap
static_assert(offsetof(ClusterRec, remap) == 256 && offsetof(ClusterRec, alpha) == 272 &&
                  offsetof(ClusterRec, axis) == 280 && offsetof(ClusterRec, fit3_err) == 292 &&
                  offsetof(ClusterRec, fit3) == 296 && offsetof(ClusterRec, w) == 304 &&
                  offsetof(ClusterRec, tex) == 316 && offsetof(ClusterRec, block) == 320 &&
                  offsetof(ClusterRec, info) == 324,
              "record tail layout");

// Counters of one pipeline run (cleared by the reset kernel).  Every update
// is aggregated (per workgroup or per batch of records) so that no counter
// takes one global atomic per block.
enum { kListA = 0, kListB = 1, kListC = 2, kLists = 3 };

// The bin pass sorts kBinSeg consecutive record slots per workgroup.
constexpr uint32_t kBinSeg = 1024;

// Bin counts and bin cursors are kept in kHistCopies copies: the records of
// slot segment s (slots [s * kBinSeg, (s + 1) * kBinSeg)) are counted in copy
// s % kHistCopies, and bin (list k, colour count v) is laid out copy after
// copy.  Device-scope atomic adds on one word serialise; with a single copy,
// thousands of classify workgroups finishing together, nearly all counting
// 16-colour blocks, queued on the same word (2048x2048: ~27 us per list of
// a ~70 us pass), and every bin workgroup on the same cursor.  With copies,
// no word takes more than a few dozen adds.
constexpr uint32_t kHistCopies = 64;

struct PipeCtl {
    uint32_t cursor[kLists];
    uint32_t hist[kLists][17];                // records per colour count (totals, written by the bin pass)
    uint32_t part[kHistCopies][kLists][17];   // records per colour count, copy by copy (classify pass)
    uint32_t pos[kHistCopies][kLists][17];    // records placed per colour count, copy by copy (bin pass)
};

struct PipeBufs {
    ClusterRec* recs = nullptr;          // capacity records; slot = the block's position among the chunks
    uint32_t* keys = nullptr;            // capacity: n | in list L << (5 + L); 0 = no record
    uint32_t* list[kLists] = {nullptr};  // capacity each: record slots sorted by n, 16 first
    PipeCtl* ctl = nullptr;
    uint64_t capacity = 0;

    // bytes of device memory for `cap` records (one allocation, 256-aligned parts)
    static constexpr size_t bytes_for(uint64_t cap)
    {
        return align256(sizeof(ClusterRec) * cap) + (1 + kLists) * align256(4 * cap) + align256(sizeof(PipeCtl));
    }
    static constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
    // carves the parts out of one allocation of bytes_for(cap) bytes
    void carve(uint8_t* base, uint64_t cap);
};

// Resident waves per CU: classify (single-wave workgroups, one block per
// lane) is limited by registers and LDS (8 KiB of pixel and weight copies
// per workgroup), the searches by LDS and registers (4 x 2512 B per wave);
// both fit 4 waves per SIMD.  The 16-colour searches run kSearchWaves waves
// per workgroup (k_pipe_search), ClusterFit's single wave (k_pipe_cluster).
constexpr uint32_t kClassifyWavesPerCu = 16;
constexpr uint32_t kSearchWavesPerCu = 16;
constexpr uint32_t kSearchWaves = 4;  // waves per search workgroup (sharing one run of list positions)

struct PipeLaunch {
    const txk::BatchTex* d_tex;  // texture descriptors (null: only `tex0`)
    txk::BatchTex tex0;          // the texture when d_tex is null
    const uint4* d_chunks;       // heavy chunks (texture, first block, count, first record slot);
                                 // null: implicit chunks of tex0
    uint32_t nchunks;            // heavy chunks
    uint32_t chunk_blocks;       // implicit chunk size
    uint64_t nslots;             // record slots the chunks cover (<= capacity)
    uint32_t classify_grid;      // persistent 64-thread workgroups of classify
    uint32_t search_grid;        // persistent 64-thread workgroups of each search
    uint32_t lists;              // bit L set: records may go to list L (its search is launched)
};

// The lists a texture's records can go to: ClusterFit -> C; BC1
// IterativeClusterFit -> A and B; BC2/BC3 IterativeClusterFit -> B; RangeFit
// -> none (classify finishes those blocks).
inline uint32_t lists_for(uint32_t format, uint32_t algorithm)
{
    if (algorithm == 0 || format > 2) return 0u;
    if (algorithm == 1) return 1u << kListC;
    return format == 0 ? (1u << kListA) | (1u << kListB) : 1u << kListB;
}

// Launches reset, classify, bin and the searches on `s`.  With `marks`
// (5 events), marks[0] is recorded before classify, marks[1] after it,
// marks[2] after bin, marks[3] after the ClusterFit pass and the 3-colour
// search (recorded even when there are none) and marks[4] after the
// 4-colour search.
hipError_t launch_pipeline(const PipeLaunch& p, const PipeBufs& b, hipStream_t s, hipEvent_t* marks = nullptr);

}  // namespace txp
