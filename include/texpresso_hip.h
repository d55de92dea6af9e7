/*
 * texpresso_hip.h -- C ABI of the MI355X (gfx950) BC1-BC5 encoder.
 *
 * The entry points mirror texpresso 2.0.2's public Rust API one to one
 * (lib/src/lib.rs); every declaration cites the reference item it replaces.
 * All compute runs in hand-written HIP kernels; there is no CPU fallback: an
 * entry point that needs the GPU returns TX_ERR_NO_DEVICE when none is
 * usable.  Output bytes are identical to the reference encoder's.
 *
 * Ownership: the caller owns every buffer; nothing is retained after a call
 * returns, except the library's per-device context (streams, scratch pools)
 * and the opaque batch/server handles, which the caller destroys.
 * Thread safety: every entry point may be called from any thread.  Calls
 * with host buffers run synchronously and are serialised per device on that
 * device's context lock.  Device-pointer calls only enqueue; where two calls
 * share library-owned device memory (the binned pipeline's intermediate
 * buffer, one batch handle's counters and records) the library orders their
 * kernels with events, so they run one after another whatever streams they
 * were issued on.
 * Errors: negative return codes (see TX_ERR_*).  Where the reference panics
 * (short input, short output, zero width) these return an error instead and
 * leave the output unspecified.
 */
#ifndef TEXPRESSO_HIP_H
#define TEXPRESSO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Format, lib/src/lib.rs:39-46 (same order, same meaning). */
typedef enum { TX_BC1 = 0, TX_BC2 = 1, TX_BC3 = 2, TX_BC4 = 3, TX_BC5 = 4 } tx_format;

/* Algorithm, lib/src/lib.rs:49-59; default ClusterFit (lib/src/lib.rs:61-65). */
typedef enum { TX_RANGE_FIT = 0, TX_CLUSTER_FIT = 1, TX_ITERATIVE_CLUSTER_FIT = 2 } tx_algorithm;

/* Params, lib/src/lib.rs:76-90. */
typedef struct {
    int32_t algorithm;             /* tx_algorithm */
    float weights[3];              /* ColourWeights: RGB channel weights */
    uint8_t weigh_colour_by_alpha; /* bool */
} tx_params;

/* Error codes (the reference panics in these cases). */
#define TX_OK 0
#define TX_ERR_FORMAT (-1)         /* format or algorithm out of range */
#define TX_ERR_ZERO_WIDTH (-2)     /* width == 0: lib/src/lib.rs:301 chunks_mut(0) panics */
#define TX_ERR_INPUT_SHORT (-3)    /* input slice shorter than the image: lib/src/lib.rs:324 */
#define TX_ERR_OUTPUT_SHORT (-4)   /* output smaller than compressed_size: lib/src/lib.rs:295 */
#define TX_ERR_OUTPUT_PARTIAL (-5) /* output ends inside a block: lib/src/lib.rs:198-214 slice panics */
#define TX_ERR_ALIGN (-6)          /* device pointer not 4-byte aligned */
#define TX_ERR_HIP (-7)            /* HIP runtime error */
#define TX_ERR_NO_DEVICE (-8)      /* no usable gfx950 device */
#define TX_ERR_BUSY (-9)           /* serving queue full / admission refused */
#define TX_ERR_INVALID (-10)       /* bad handle or argument */
#define TX_ERR_NOMEM (-11)         /* device memory budget exceeded */
#define TX_ERR_TOO_LARGE (-12)     /* texture exceeds the fixed scratch of tx_compress_device_noalloc */

/* Params::default(), lib/src/lib.rs:92-100: ClusterFit, perceptual weights
 * [0.2126, 0.7152, 0.0722] (lib/src/lib.rs:74), no alpha weighting. */
tx_params tx_params_default(void);

/* num_blocks, lib/src/lib.rs:103-105. */
size_t tx_num_blocks(size_t size);

/* Format::block_size, lib/src/lib.rs:159-168 (0 for an invalid format). */
size_t tx_block_size(int format);

/* Format::compressed_size, lib/src/lib.rs:175-179 (0 for an invalid format). */
size_t tx_compressed_size(int format, size_t width, size_t height);

/* Format::compress, lib/src/lib.rs:287-335.  Host buffers; synchronous.
 * Like the reference, every whole block of `output` is written: blocks past
 * the image (an oversized buffer) are encoded as fully masked blocks
 * (lib/src/lib.rs:300-305).  out_len must be a multiple of the block size. */
int tx_compress(int format, const uint8_t* rgba, size_t rgba_len, size_t width, size_t height,
                const tx_params* params, uint8_t* output, size_t out_len);

/* Format::decompress, lib/src/lib.rs:124-156.  Host buffers; synchronous.
 * Writes the width*height*4 RGBA8 pixels. */
int tx_decompress(int format, const uint8_t* data, size_t data_len, size_t width, size_t height, uint8_t* output,
                  size_t out_len);

/* Format::compress_block_masked, lib/src/lib.rs:188-234.  `rgba` holds 16
 * RGBA8 pixels in row-major 4x4 order; bit i of `mask` enables pixel i. */
int tx_compress_block_masked(int format, const uint8_t rgba[64], uint32_t mask, const tx_params* params,
                             uint8_t* output, size_t out_len);

/* Many Format::compress_block_masked calls in one launch (lib/src/lib.rs:188-234
 * applied to each block): block b is the 64 bytes at blocks + 64*b, its mask
 * is masks[b], its encoding goes to output + b*block_size.  Host buffers;
 * synchronous.  No reference counterpart as a single call; it is how a
 * caller with its own tiling reaches the per-block encoder without one
 * launch per block. */
int tx_compress_blocks(int format, const uint8_t* blocks, const uint32_t* masks, size_t nblocks,
                       const tx_params* params, uint8_t* output);

/* Format::decompress_block, lib/src/lib.rs:240-277. */
int tx_decompress_block(int format, const uint8_t* block, size_t block_len, uint8_t rgba_out[64]);

/* Device-resident variants of Format::compress / Format::decompress: all
 * pointers are device memory (4-byte aligned), `stream` is a hipStream_t
 * (NULL = the library's stream for the current device).  The call enqueues
 * the kernel and returns without synchronising.  Textures that take the
 * binned ClusterFit pipeline (tx_set_pipeline_threshold) share one
 * intermediate buffer per device; the library makes each such call's stream
 * wait for the previous one's kernels, so these calls run one after another
 * even when issued on different streams. */
int tx_compress_device(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                       uint8_t* d_output, size_t out_len, void* stream);
int tx_decompress_device(int format, const uint8_t* d_data, size_t width, size_t height, uint8_t* d_output,
                         void* stream);

/* Allocation-free variant of tx_compress_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and output bytes as tx_compress_device, except:
 *  - no hipMalloc, hipFree or other allocating call is made, and no stream
 *    or event is created; `stream` NULL means the null stream;
 *  - the binned ClusterFit pipeline's intermediate buffer (records, keys,
 *    sorted lists, counters) is a fixed device array inside the library,
 *    TX_NOALLOC_SCRATCH_BYTES on every GPU the library runs on, held until
 *    the process ends, instead of the grown per-device allocation; a
 *    texture that takes the pipeline (tx_set_pipeline_threshold) may have at
 *    most TX_NOALLOC_MAX_PIPELINE_BLOCKS blocks, larger ones return
 *    TX_ERR_TOO_LARGE before anything is enqueued.  Textures below the
 *    threshold, RangeFit and BC4/BC5 use no scratch and have no such limit.
 * All calls on one GPU share that GPU's array and the library does not order
 * them: calls that may overlap on one GPU (issued on different streams, or
 * from threads with no ordering between them) race and corrupt each other's
 * output.  Calls on one stream are ordered by the stream.  The array keeps
 * the last call's contents; every call clears the counters it needs. */
#define TX_NOALLOC_MAX_PIPELINE_BLOCKS 262144u /* a 2048 x 2048 texture */
#define TX_NOALLOC_SCRATCH_BYTES 92301056u     /* 88 MiB: 352 bytes per block plus the counters */
int tx_compress_device_noalloc(int format, const uint8_t* d_rgba, size_t width, size_t height,
                               const tx_params* params, uint8_t* d_output, size_t out_len, void* stream);

/* ---------------------------------------------------------------------
 * Batched encoding (no single reference counterpart: it runs many
 * Format::compress calls, lib/src/lib.rs:287-335, as one persistent launch).
 * Every texture (or mip level) is a tx_texture with its own format and
 * params; all are split into chunks of consecutive blocks, ordered by
 * estimated cost, and consumed by a persistent kernel through one device
 * work counter.  The launch (counter reset + kernel) is captured once in a
 * hipGraph and replayed.  Outputs are byte-identical to per-texture calls.
 */
typedef struct {
    int32_t format;       /* tx_format */
    tx_params params;
    const uint8_t* d_rgba; /* device, width*height*4 bytes, 4-byte aligned */
    uint8_t* d_output;     /* device, tx_compressed_size bytes, aligned to the block size */
    uint32_t width, height;
} tx_texture;

typedef struct {
    uint32_t chunk_blocks;     /* blocks per work-queue chunk (multiple of 4) */
    uint32_t waves_per_group;  /* waves per workgroup: 1, 2 or 4 (4 blocks per wave) */
    uint32_t groups_per_cu;    /* persistent workgroups per compute unit */
    uint32_t use_graph;        /* 1: capture the launch in a hipGraph and replay it */
    uint32_t pipeline;         /* 1: ClusterFit blocks through the binned pipeline (classify, sort by
                                  colour count, independent per-block searches); 0: one persistent
                                  kernel with the four blocks of a wave in lockstep */
    uint32_t lanes_per_block;  /* lanes that encode one BC4/BC5 block: 1 (0 = default: a lane runs the
                                  block's codebook fits alone, chunks of 512 blocks) or 16 (a 16-lane
                                  group per block, min/max and error sums across the group, chunks of
                                  chunk_blocks).  BC1-BC3 blocks use 16 lanes in the lockstep kernel
                                  and the searches, one in the pipeline's classify pass */
} tx_batch_tuning;

typedef struct tx_batch tx_batch;

/* Default tuning (used when `tuning` is NULL). */
tx_batch_tuning tx_batch_tuning_default(void);
/* Validates the textures, builds and uploads the chunk queue. */
int tx_batch_create(tx_batch** batch, const tx_texture* textures, size_t count, const tx_batch_tuning* tuning);
/* Enqueues one pass over every texture on `stream` (NULL = library stream).
 * Passes of one handle run one after another even when issued on different
 * streams or threads (each waits for the previous one). */
int tx_batch_run(tx_batch* batch, void* stream);
/* Enqueues the pass and waits for it; returns the elapsed ms (HIP events). */
int tx_batch_run_timed(tx_batch* batch, int iters, float* ms_total);
int tx_batch_destroy(tx_batch* batch);
/* Number of chunks and blocks in the batch. */
int tx_batch_info(const tx_batch* batch, uint64_t* chunks, uint64_t* blocks);

/* Autotuner: times every tuning of the candidate grid on these textures and
 * returns the fastest (ms = mean per pass).  `candidates`/`ncand` may be
 * NULL/0 for the built-in grid. */
int tx_batch_autotune(const tx_texture* textures, size_t count, const tx_batch_tuning* candidates, size_t ncand,
                      int iters, tx_batch_tuning* best, float* best_ms);

/* Baseline for the batch path: one plain launch per texture on the library
 * stream, `iters` passes, elapsed ms (HIP events). */
int tx_util_time_per_texture(const tx_texture* textures, size_t count, int iters, float* ms_total);

/* ---------------------------------------------------------------------
 * Serving (no reference counterpart).  A scheduler thread admits submitted
 * jobs in arrival order while their source + output fit in the device memory
 * budget, uploads the sources into a device arena, and folds every admitted
 * job into the next persistent-kernel launch on a ring of streams, so jobs
 * arriving while others run are batched into launches that overlap on the
 * device.  Outputs stay device-resident until fetched and released.  Each
 * job's bytes are identical to a plain tx_compress of the same texture.
 */
typedef struct {
    size_t hbm_budget;             /* device bytes for job sources + outputs */
    uint32_t max_inflight;         /* concurrent launches (streams), 1..8 */
    uint32_t max_jobs_per_launch;  /* jobs folded into one launch */
    uint32_t max_pending;          /* submissions waiting for admission */
    tx_batch_tuning tuning;        /* chunking and grid of each launch; with tuning.pipeline = 1
                                      (default) launches of >= tx_set_pipeline_threshold BC1-BC3
                                      blocks (and <= 4M) use the binned pipeline, whose records
                                      (336 B per block for min(4M, hbm_budget/64) blocks per
                                      in-flight launch) are allocated when the server is created */
} tx_server_config;

typedef struct {
    int32_t format; /* tx_format */
    tx_params params;
    const uint8_t* rgba; /* host RGBA8, width*height*4; valid until the job is done */
    uint32_t width, height;
} tx_job;

/* Job states */
#define TX_JOB_PENDING 0
#define TX_JOB_ADMITTED 1
#define TX_JOB_RUNNING 2
#define TX_JOB_DONE 3

typedef struct {
    int32_t status;                                       /* TX_JOB_* or a negative error */
    uint64_t t_submit_ns, t_admit_ns, t_launch_ns, t_done_ns; /* CLOCK_MONOTONIC */
    const uint8_t* d_output;                              /* device-resident result */
    size_t out_len;
    uint32_t launch_id;   /* the launch that ran the job */
    uint32_t launch_jobs; /* jobs batched into that launch */
} tx_job_info;

typedef struct tx_server tx_server;

tx_server_config tx_server_config_default(void);
int tx_server_create(tx_server** server, const tx_server_config* config);
/* TX_ERR_BUSY when max_pending submissions wait; TX_ERR_NOMEM when the job
 * alone exceeds the budget; TX_ERR_FORMAT/INVALID for a malformed job. */
int tx_server_submit(tx_server* server, const tx_job* job, uint64_t* job_id);
/* Waits up to timeout_ms (< 0: forever) for the job; TX_ERR_BUSY on timeout. */
int tx_server_wait(tx_server* server, uint64_t job_id, int timeout_ms, tx_job_info* info);
int tx_server_query(tx_server* server, uint64_t job_id, tx_job_info* info);
/* Copies a finished job's output to host memory. */
int tx_server_fetch(tx_server* server, uint64_t job_id, uint8_t* output, size_t out_len);
/* Frees the job's device memory (after completion). */
int tx_server_release(tx_server* server, uint64_t job_id);
/* Stops admitting, waits for running launches, frees everything. */
int tx_server_destroy(tx_server* server);

/* Host copy of the four single-colour lookup tables compiled into the device
 * code (lib/src/colourfit/single_lut.rs), 6 bytes per entry
 * (start0,end0,err0,start1,end1,err1), order 5_3, 6_3, 5_4, 6_4. */
void tx_single_lut(uint8_t out[6144]);

/* Utilities (no reference counterpart): device memory on the current device
 * and kernel timing with HIP events on the library's stream, so callers and
 * benchmarks need no other GPU runtime binding. */
int tx_util_malloc(void** d_ptr, size_t bytes);
int tx_util_free(void* d_ptr);
int tx_util_upload(void* d_dst, const void* src, size_t bytes);
int tx_util_download(void* dst, const void* d_src, size_t bytes);
/* Device-to-device copy on the library stream (synchronous). */
int tx_util_copy(void* d_dst, const void* d_src, size_t bytes);
/* The encoder's exactly specified elementary operations as the device
 * evaluates them, for an exhaustive check against a host IEEE binary32
 * evaluation: u8f[u] = u / 255 (the point coordinates of colourset.rs:60-66),
 * sqrt_counts[c] = sqrtf(c) (point weights of c duplicate pixels),
 * sqrt_alpha[k-1] = sqrtf(k / 256) for k = 1..4096 (alpha-weighted point
 * weights, colourset.rs:103-110), and for each of the n values in[i] (n <=
 * 2^26): snap31[i] / snap63[i] = the 5-/6-bit grid snap of range.rs and
 * cluster.rs (clamp to [0, 1], truncf(g * v + 0.5) * (1/g)) and round255[i] =
 * roundf(in[i]) clamped to [0, 255], NaN to 0 (math.rs:100-102).  Synchronous;
 * host pointers. */
int tx_util_elementary(const float* in, size_t n, float u8f[256], float sqrt_counts[17], float sqrt_alpha[4096],
                       float* snap31, float* snap63, int32_t* round255);
/* Enqueues `iters` tx_compress_device launches between two events on the
 * library stream, waits, and stores the elapsed milliseconds. */
int tx_util_time_compress(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                          uint8_t* d_output, size_t out_len, int iters, float* ms_total);
/* Per-phase timing of the binned ClusterFit pipeline: runs tx_compress_device
 * `iters` times on the library stream with events between its kernels and
 * adds up, in ms_phase[0..3], classify, bin, 3-colour search and 4-colour
 * search milliseconds.  *used_pipeline = 0 (and nothing is timed) when the
 * texture goes through the direct kernel (below tx_set_pipeline_threshold,
 * RangeFit, BC4/BC5). */
int tx_util_time_phases(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                        uint8_t* d_output, size_t out_len, int iters, float ms_phase[4], int* used_pipeline);

/* BC2/BC3 IterativeClusterFit textures with at least `min_blocks` blocks are
 * encoded by the binned pipeline (classify, counting sort by colour count,
 * per-block independent searches), smaller ones by the single direct kernel;
 * BC2/BC3 ClusterFit use 1.25x, BC1 ClusterFit 1.5x and BC1
 * IterativeClusterFit 3x this threshold (the measured crossovers).  Both
 * paths give identical bytes.  0 = always the pipeline, UINT64_MAX = never.
 * Process-wide; returns the previous value (default 32768). */
uint64_t tx_set_pipeline_threshold(uint64_t min_blocks);

/* Human-readable text for an error code. */
const char* tx_error_string(int code);

/* Name of the GPU architecture the library drives ("gfx950"), or NULL when
 * no device is usable. */
const char* tx_device_arch(void);

/* ---- Mip chains: the step before the encoder.  No counterpart in texpresso
 * itself (it encodes the level it is given; its CLI writes level 0 only,
 * cli/src/main.rs:148-160).  Level k is max(1, w >> k) x max(1, h >> k); each
 * texel of level k+1 is (a + b + c + d + 2) >> 2 per channel over the 2x2 box
 * of level k, edge-clamped. */

/* Levels of a full chain down to 1x1 (0 when a side is 0). */
uint32_t tx_mip_levels(size_t width, size_t height);
/* Size of level `level`. */
void tx_mip_dims(size_t width, size_t height, uint32_t level, size_t* width_out, size_t* height_out);
/* RGBA8 bytes of levels 1..levels-1 stored back to back. */
size_t tx_mip_chain_rgba_bytes(size_t width, size_t height, uint32_t levels);
/* Compressed bytes of levels 0..levels-1 stored back to back (the data
 * layout of a DDS file with mip levels). */
size_t tx_mip_chain_compressed_size(int format, size_t width, size_t height, uint32_t levels);
/* Device: levels 1..levels-1 of d_rgba into d_chain (4-byte aligned,
 * chain_len >= tx_mip_chain_rgba_bytes).  Enqueued on `stream` (NULL: the
 * library stream). */
int tx_generate_mips_device(const uint8_t* d_rgba, size_t width, size_t height, uint32_t levels, uint8_t* d_chain,
                            size_t chain_len, void* stream);
/* Device: generates the chain in d_chain, then compresses every level
 * (level 0 = d_rgba) back to back into d_output, level 0 first; each level
 * is encoded exactly like tx_compress_device would encode it alone. */
int tx_compress_mips_device(int format, const uint8_t* d_rgba, size_t width, size_t height, const tx_params* params,
                            uint32_t levels, uint8_t* d_chain, size_t chain_len, uint8_t* d_output, size_t out_len,
                            void* stream);
/* Host buffers: the same, synchronous. */
int tx_compress_mips(int format, const uint8_t* rgba, size_t rgba_len, size_t width, size_t height,
                     const tx_params* params, uint32_t levels, uint8_t* output, size_t out_len);

/* ---- DDS container: the step after the encoder, replacing the ddsfile calls
 * of texpresso's CLI (write: cli/src/main.rs:143-164, DXGI mapping 220-228;
 * read: 167-196, 230-248).  Header = "DDS " + DDS_HEADER + DDS_HEADER_DXT10
 * (148 bytes): Texture2D, DXGI BC1-3 UNORM_SRGB / BC4-5 UNORM, alpha mode
 * premultiplied for BC1 and straight otherwise; level data follows. */
size_t tx_dds_header_size(void);
int tx_dds_write_header(int format, size_t width, size_t height, uint32_t levels, uint8_t* out, size_t out_len);
/* Accepts the five DXGI formats above with a Texture2D dimension, or a
 * legacy DXT1/DXT3/DXT5 FourCC header; anything else is TX_ERR_FORMAT. */
int tx_dds_parse_header(const uint8_t* data, size_t len, int* format, size_t* width, size_t* height,
                        uint32_t* levels, size_t* data_offset);

/* ---- PNG image I/O of the command-line tool (host only), replacing the png
 * crate calls of texpresso's CLI: read = cli/src/image/png.rs:59-99 (EXPAND |
 * STRIP_16, then grey / grey-alpha / RGB / RGBA to RGBA8), write =
 * cli/src/image/png.rs:101-111 (RGBA8, 8 bits per sample).  Buffers belong
 * to the caller: tx_png_info gives the size, tx_png_decode fills 4*w*h
 * bytes; tx_png_encode_bound gives a sufficient output size (0 if the image
 * is too large) and tx_png_encode stores the file length in *out_len.
 * Malformed files (signature, CRC, chunk order, zlib stream, filter type)
 * return TX_ERR_FORMAT or TX_ERR_INPUT_SHORT. */
int tx_png_info(const uint8_t* data, size_t len, size_t* width, size_t* height);
int tx_png_decode(const uint8_t* data, size_t len, uint8_t* rgba, size_t rgba_len);
size_t tx_png_encode_bound(size_t width, size_t height);
int tx_png_encode(const uint8_t* rgba, size_t width, size_t height, uint8_t* out, size_t out_cap, size_t* out_len);

#ifdef __cplusplus
}
#endif
#endif
