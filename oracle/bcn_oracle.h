/*
 * bcn_oracle.h -- CPU restatement of texpresso 2.0.2's BC1-BC5 codec.
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); never part of
 * the product library.  Enum values follow lib/src/lib.rs:39-59.
 */
#ifndef BCN_ORACLE_H
#define BCN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TX_BC1 = 0, TX_BC2 = 1, TX_BC3 = 2, TX_BC4 = 3, TX_BC5 = 4 };
enum { TX_RANGE_FIT = 0, TX_CLUSTER_FIT = 1, TX_ITERATIVE_CLUSTER_FIT = 2 };

/* lib.rs:76-90 */
typedef struct {
    int32_t algorithm;
    float weights[3];
    uint8_t weigh_colour_by_alpha;
} ora_params;

/* per-block statistics used to count algorithmic work */
typedef struct {
    int32_t count;       /* distinct colours in the colour set */
    int32_t transparent; /* BC1 block with a pixel of alpha < 128 */
    int32_t fit;         /* 0 single colour, 1 range fit, 2 cluster fit */
    int32_t searched3;   /* orderings searched by compress3 */
    int32_t searched4;   /* orderings searched by compress4 */
} ora_block_stats;

/* winner of one ClusterFit fit (compress3 or compress4) */
typedef struct {
    int32_t won;              /* the fit produced the final block of its call */
    int32_t iteration;        /* ordering index of the winner */
    int32_t i, j, k;          /* split (k unused for compress3) */
    float error;              /* weighted error without the constant term */
    float a_raw[3], b_raw[3]; /* least-squares endpoints before clamping */
    float a[3], b[3];         /* clamped, grid-snapped endpoints */
    uint8_t order[16];        /* the winning ordering */
} ora_fit_trace;

/* internals of the colour fit of one block (for closed-form checks) */
typedef struct {
    int32_t count;
    int32_t transparent;
    float points[16][3];
    float weights[16]; /* square-rooted */
    float cov[6];
    float principle[3];
    ora_fit_trace fit3, fit4;
} ora_cluster_trace;

int ora_trace_block(int fmt, const uint8_t rgba[64], uint32_t mask, const ora_params* p, uint8_t* out,
                    ora_cluster_trace* t);

int ora_compress_block_masked(int fmt, const uint8_t rgba[64], uint32_t mask, const ora_params* p, uint8_t* out,
                              ora_block_stats* st);
int ora_compress(int fmt, const uint8_t* rgba, size_t rgba_len, size_t w, size_t h, const ora_params* p,
                 uint8_t* out, size_t out_len, int nthreads, ora_block_stats* stats);
int ora_compress_blocks(int fmt, const uint8_t* blocks, const uint32_t* masks, size_t n, const ora_params* p,
                        uint8_t* out, int nthreads);
void ora_decompress_block(int fmt, const uint8_t* block, uint8_t out[64]);
int ora_decompress(int fmt, const uint8_t* data, size_t data_len, size_t w, size_t h, uint8_t* out, size_t out_len);
void ora_single_lut(uint8_t out[4 * 256 * 6]);
size_t ora_block_size(int fmt);
size_t ora_num_blocks(size_t n);

#ifdef __cplusplus
}
#endif
#endif
