/*
 * bcn_oracle.c -- CPU restatement of texpresso 2.0.2's BC1-BC5 codec.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * encoder and the "port" CPU baseline timed by bench.py.  It is never linked
 * into, loaded by, or called from the product library; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function follows the reference operation by operation (file:line
 * citations are relative to the texpresso source tree).  The build flags
 * (-O2 -ffp-contract=off -fno-fast-math, x86-64 SSE scalar) keep every float
 * operation a single correctly rounded IEEE-754 binary32 operation, as Rust
 * does.  Loops keep the reference's incremental form (part0/part1/part2
 * accumulators) so that this file stays a literal restatement; the GPU path
 * uses a different (prefix-table) formulation and is checked against this one.
 *
 * What pins each part (the tests under tests/ do the pinning):
 *   - colourblock write3/write4, pack_565, alpha BC2/BC3/BC4/BC5 encoders,
 *     ClusterFit/IterativeClusterFit/RangeFit on 3-colour blocks with uniform
 *     weights, and all decoders: the reference's known-answer vectors
 *     (lib/src/test_data.rs), replayed from tests/golden/kat.json.
 *   - single-colour lookup tables: regenerated here and compared byte for
 *     byte with the reference tables (tests/golden/single_lut.bin) and their
 *     SHA-256.
 *   - the BC3 alpha encoder: every constant block and every two-value block
 *     (all 32,896 pairs, three layouts) decodes exactly; hand-derived vectors
 *     for fix_range at both edges of 0..255 and for blocks decided by the
 *     7-point codebook's slots 0-1 holding min5/max5 (alpha.rs:238-239).
 *   - SingleColourFit: every grey and 1,500 random colours predicted from the
 *     reference's tables (index 0 unless source 1 is strictly better; BC1's
 *     4-colour fit only when strictly better), decoded pixels exactly the
 *     tables' error away.
 *   - RangeFit: float64 principal axis, first extreme points, integer grid
 *     snap, nearest code under the colour weights and 3-/4-colour choice on
 *     4,500 blocks of 2-16 colours under perceptual, uniform and skewed
 *     weights, masks and BC1 transparency (blocks within a rounding margin
 *     of a tie skipped: 257 of them).
 *   - planted mistakes: tests/test_oracle_pins.py builds copies of this file
 *     with one plausible mistake each (squared range weights, tie rules of
 *     the single-colour and alpha choices, min7/max7 in the 7-point slots,
 *     fix_range's upper clamp) and requires the pins above to fail on them.
 *   - covariance, power iteration, least-squares endpoints, error formula and
 *     candidate search of ClusterFit: float64 closed forms (weighted
 *     least squares, eigenvectors, exhaustive candidate minimum) within a
 *     stated tolerance, and stable-sort order of the projections.
 *   - IterativeClusterFit's orderings: each ordering after the first is the
 *     stable sort along end - start of the previous ordering's best split
 *     (quantised, binary32), and a first-iteration winner means the second
 *     ordering repeated or did not improve -- float64 replay of the searches
 *     up to the third ordering; a flipped or unquantised axis fails it.
 *   - alpha-weighted colour sets: point weights sqrt(sum (alpha + 1) / 256)
 *     bit-exact, their covariance and their first-iteration least-squares
 *     endpoints and error, float64 closed forms.
 *   - parity unpinned at the bit level: the binary32 operation order of
 *     ClusterFit/RangeFit for n > 3 distinct colours, perceptual weights and
 *     iterations from the fourth on has no reference vector; the float64
 *     checks above pin every decision that is not within rounding of a tie,
 *     and near-ties (including the first-extreme rule of the range scan on
 *     equal projections) rest on faithful restatement.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "bcn_oracle.h"

/* ------------------------------------------------------------------ */
/* L0 math: lib/src/math.rs, lib/src/math/vec3.rs, lib/src/math/vec4.rs */

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } v4;

static v3 mk3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v4 mk4(float x, float y, float z, float w) { v4 r = {x, y, z, w}; return r; }

/* vec3.rs:51-53: (x*x' + y*y') + z*z' */
static float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

static v4 add4(v4 a, v4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static v4 sub4(v4 a, v4 b) { return mk4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static v4 mul4(v4 a, v4 b) { return mk4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
static v4 splat(float s) { return mk4(s, s, s, s); }
/* vec4.rs:94-96: 1/x per lane */
static v4 rcp4(v4 a) { return mk4(1.0f / a.x, 1.0f / a.y, 1.0f / a.z, 1.0f / a.w); }
/* vec4.rs:76-92: f32::max / f32::min (NaN operand yields the other) */
static v4 max4(v4 a, v4 b) { return mk4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w)); }
static v4 min4(v4 a, v4 b) { return mk4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), fminf(a.w, b.w)); }
/* vec4.rs:102-109: libm::truncf per lane */
static v4 trunc4(v4 a) { return mk4(truncf(a.x), truncf(a.y), truncf(a.z), truncf(a.w)); }
/* vec4.rs:98-100 */
static int any_less4(v4 a, v4 b) { return a.x < b.x || a.y < b.y || a.z < b.z || a.w < b.w; }

/* math.rs:100-102: roundf(a).max(0).min(limit) as i32 (NaN -> 0) */
static int f32_to_i32_clamped(float a, int limit)
{
    float r = fminf(fmaxf(roundf(a), 0.0f), (float)limit);
    return (int)r;
}

/* math.rs:44-73 */
static void weighted_covariance(const v3* pts, const float* wts, int n, float cov[6])
{
    float total = 0.0f;
    for (int i = 0; i < n; ++i) total = total + wts[i];
    v3 c = mk3(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < n; ++i) {
        v3 pw = mk3(pts[i].x * wts[i], pts[i].y * wts[i], pts[i].z * wts[i]);
        c = mk3(c.x + pw.x, c.y + pw.y, c.z + pw.z);
    }
    if (total > FLT_EPSILON) c = mk3(c.x / total, c.y / total, c.z / total);
    for (int k = 0; k < 6; ++k) cov[k] = 0.0f;
    for (int i = 0; i < n; ++i) {
        v3 a = mk3(pts[i].x - c.x, pts[i].y - c.y, pts[i].z - c.z);
        v3 b = mk3(a.x * wts[i], a.y * wts[i], a.z * wts[i]);
        cov[0] += a.x * b.x;
        cov[1] += a.x * b.y;
        cov[2] += a.x * b.z;
        cov[3] += a.y * b.y;
        cov[4] += a.y * b.z;
        cov[5] += a.z * b.z;
    }
}

/* math.rs:75-97: eight power-iteration steps, v = w * (1/max(w)) */
static v3 principle_component(const float c[6])
{
    v4 row0 = mk4(c[0], c[1], c[2], 0.0f);
    v4 row1 = mk4(c[1], c[3], c[4], 0.0f);
    v4 row2 = mk4(c[2], c[4], c[5], 0.0f);
    v4 v = splat(1.0f);
    for (int it = 0; it < 8; ++it) {
        v4 w = mul4(row0, splat(v.x));
        w = add4(mul4(row1, splat(v.y)), w);
        w = add4(mul4(row2, splat(v.z)), w);
        float a = fmaxf(w.x, fmaxf(w.y, w.z));
        v = mul4(w, rcp4(splat(a)));
    }
    return mk3(v.x, v.y, v.z);
}

/* ------------------------------------------------------------------ */
/* colourset.rs:26-142 */

typedef struct {
    int count;
    v3 points[16];
    float weights[16];
    int8_t remap[16];
    int transparent;
} colourset;

static void colourset_new(colourset* s, const uint8_t rgba[16][4], uint32_t mask, int fmt, int alpha_weighted)
{
    memset(s, 0, sizeof *s);
    for (int i = 0; i < 16; ++i) {
        if ((mask & (1u << i)) == 0) { s->remap[i] = -1; continue; }
        if (fmt == TX_BC1 && rgba[i][3] < 128) { s->remap[i] = -1; s->transparent = 1; continue; }
        for (int j = 0; j < 16; ++j) {
            if (j == i) {
                float x = (float)rgba[i][0] / 255.0f;
                float y = (float)rgba[i][1] / 255.0f;
                float z = (float)rgba[i][2] / 255.0f;
                float w = (float)((int)rgba[i][3] + 1) / 256.0f;
                s->points[s->count] = mk3(x, y, z);
                s->weights[s->count] = alpha_weighted ? w : 1.0f;
                s->remap[i] = (int8_t)s->count;
                s->count += 1;
                break;
            }
            int dup = ((mask & (1u << j)) != 0) && rgba[i][0] == rgba[j][0] && rgba[i][1] == rgba[j][1] &&
                      rgba[i][2] == rgba[j][2] && (fmt != TX_BC1 || rgba[j][3] >= 128);
            if (dup) {
                int idx = s->remap[j];
                float w = (float)((int)rgba[i][3] + 1) / 256.0f;
                s->weights[idx] += alpha_weighted ? w : 1.0f;
                s->remap[i] = (int8_t)idx;
                break;
            }
        }
    }
    for (int k = 0; k < 16; ++k) s->weights[k] = sqrtf(s->weights[k]);
}

/* colourset.rs:130-141 */
static void remap_indices(const colourset* s, const uint8_t src[16], uint8_t dst[16])
{
    for (int i = 0; i < 16; ++i) dst[i] = (s->remap[i] == -1) ? 3 : src[(int)s->remap[i]];
}

/* ------------------------------------------------------------------ */
/* colourblock.rs:27-94 */

static uint16_t pack_565(v3 c)
{
    int r = f32_to_i32_clamped(31.0f * c.x, 31);
    int g = f32_to_i32_clamped(63.0f * c.y, 63);
    int b = f32_to_i32_clamped(31.0f * c.z, 31);
    return (uint16_t)((r << 11) | (g << 5) | b);
}

static void write_block(uint16_t a, uint16_t b, const uint8_t idx[16], uint8_t* blk)
{
    blk[0] = (uint8_t)(a & 0xFF); blk[1] = (uint8_t)(a >> 8);
    blk[2] = (uint8_t)(b & 0xFF); blk[3] = (uint8_t)(b >> 8);
    for (int i = 0; i < 4; ++i)
        blk[4 + i] = (uint8_t)(((idx[4 * i + 3] & 3) << 6) | ((idx[4 * i + 2] & 3) << 4) |
                               ((idx[4 * i + 1] & 3) << 2) | (idx[4 * i] & 3));
}

static void write3(v3 start, v3 end, const uint8_t idx[16], uint8_t* blk)
{
    uint16_t a = pack_565(start), b = pack_565(end);
    uint8_t r[16];
    memcpy(r, idx, 16);
    if (a > b) {
        uint16_t t = a; a = b; b = t;
        for (int i = 0; i < 16; ++i) r[i] = (r[i] == 0) ? 1 : (r[i] == 1) ? 0 : r[i];
    }
    write_block(a, b, r, blk);
}

static void write4(v3 start, v3 end, const uint8_t idx[16], uint8_t* blk)
{
    uint16_t a = pack_565(start), b = pack_565(end);
    uint8_t r[16] = {0};
    if (a < b) {
        uint16_t t = a; a = b; b = t;
        for (int i = 0; i < 16; ++i) r[i] = (uint8_t)((idx[i] ^ 1) & 3);
    } else if (a > b) {
        memcpy(r, idx, 16);
    }
    write_block(a, b, r, blk);
}

/* ------------------------------------------------------------------ */
/* single_lut.rs: regenerated from the BC1 endpoint expansion.
 * entry[t] = (sources[0], sources[1]) where sources[0] reaches palette code 0
 * and sources[1] reaches the first interpolated code (index 2). */

typedef struct { uint8_t start, end, error; } source_block;
typedef struct { source_block s[2]; } single_lookup;

static single_lookup g_lut[4][256]; /* 5_3, 6_3, 5_4, 6_4 */
static pthread_once_t g_lut_once = PTHREAD_ONCE_INIT;

static void build_one_lut(int bits, int colours, single_lookup out[256])
{
    /* per palette index: start, end, err, valid */
    int st[4][256], en[4][256], er[4][256], ok[4][256];
    memset(ok, 0, sizeof ok);
    int levels = 1 << bits;
    for (int v1 = 0; v1 < levels; ++v1) {
        for (int v2 = 0; v2 < levels; ++v2) {
            int a = (v1 << (8 - bits)) | (v1 >> (2 * bits - 8));
            int b = (v2 << (8 - bits)) | (v2 >> (2 * bits - 8));
            int codes[4];
            codes[0] = a; codes[1] = b;
            if (colours == 3) { codes[2] = (a + b) / 2; codes[3] = 0; }
            else { codes[2] = (2 * a + b) / 3; codes[3] = (a + 2 * b) / 3; }
            for (int idx = 0; idx < colours; ++idx) {
                int t = codes[idx];
                if (!ok[idx][t]) { ok[idx][t] = 1; st[idx][t] = v1; en[idx][t] = v2; er[idx][t] = 0; }
            }
        }
    }
    for (int changed = 1; changed;) {
        changed = 0;
        for (int idx = 0; idx < colours; ++idx) {
            for (int t = 0; t < 256; ++t) {
                int nbr[2] = {t + 1, t - 1};
                for (int q = 0; q < 2; ++q) {
                    int nt = nbr[q];
                    if (nt < 0 || nt > 255 || !ok[idx][nt]) continue;
                    if (!ok[idx][t] || er[idx][nt] + 1 < er[idx][t]) {
                        ok[idx][t] = 1;
                        st[idx][t] = st[idx][nt]; en[idx][t] = en[idx][nt]; er[idx][t] = er[idx][nt] + 1;
                        changed = 1;
                    }
                }
            }
        }
    }
    for (int t = 0; t < 256; ++t) {
        out[t].s[0].start = (uint8_t)st[0][t]; out[t].s[0].end = (uint8_t)en[0][t]; out[t].s[0].error = (uint8_t)er[0][t];
        out[t].s[1].start = (uint8_t)st[2][t]; out[t].s[1].end = (uint8_t)en[2][t]; out[t].s[1].error = (uint8_t)er[2][t];
    }
}

static void build_luts(void)
{
    build_one_lut(5, 3, g_lut[0]);
    build_one_lut(6, 3, g_lut[1]);
    build_one_lut(5, 4, g_lut[2]);
    build_one_lut(6, 4, g_lut[3]);
}

void ora_single_lut(uint8_t out[4 * 256 * 6])
{
    pthread_once(&g_lut_once, build_luts);
    for (int l = 0; l < 4; ++l)
        for (int t = 0; t < 256; ++t) {
            uint8_t* p = out + (l * 256 + t) * 6;
            p[0] = g_lut[l][t].s[0].start; p[1] = g_lut[l][t].s[0].end; p[2] = g_lut[l][t].s[0].error;
            p[3] = g_lut[l][t].s[1].start; p[4] = g_lut[l][t].s[1].end; p[5] = g_lut[l][t].s[1].error;
        }
}

/* ------------------------------------------------------------------ */
/* single.rs:44-165 */

typedef struct {
    const colourset* set;
    v3 start, end;
    uint8_t index;
    uint32_t error, best_error;
    uint8_t best[8];
} single_fit;

static void single_compute_endpoints(single_fit* f, const single_lookup* lut[3])
{
    int colour[3] = {
        f32_to_i32_clamped(f->set->points[0].x * 255.0f, 255),
        f32_to_i32_clamped(f->set->points[0].y * 255.0f, 255),
        f32_to_i32_clamped(f->set->points[0].z * 255.0f, 255),
    };
    f->error = UINT32_MAX;
    for (int index = 0; index < 2; ++index) {
        uint32_t error = 0;
        const source_block* src[3];
        for (int ch = 0; ch < 3; ++ch) {
            src[ch] = &lut[ch][colour[ch]].s[index];
            uint32_t d = src[ch]->error;
            error += d * d;
        }
        if (error < f->error) {
            f->start = mk3((float)src[0]->start / 31.0f, (float)src[1]->start / 63.0f, (float)src[2]->start / 31.0f);
            f->end = mk3((float)src[0]->end / 31.0f, (float)src[1]->end / 63.0f, (float)src[2]->end / 31.0f);
            f->index = (uint8_t)(2 * index);
            f->error = error;
        }
    }
}

static void single_compress3(single_fit* f)
{
    const single_lookup* lut[3] = {g_lut[0], g_lut[1], g_lut[0]};
    single_compute_endpoints(f, lut);
    if (f->error < f->best_error) {
        uint8_t src[16], idx[16];
        memset(src, f->index, 16);
        remap_indices(f->set, src, idx);
        write3(f->start, f->end, idx, f->best);
        f->best_error = f->error;
    }
}

static void single_compress4(single_fit* f)
{
    const single_lookup* lut[3] = {g_lut[2], g_lut[3], g_lut[2]};
    single_compute_endpoints(f, lut);
    if (f->error < f->best_error) {
        uint8_t src[16], idx[16];
        memset(src, f->index, 16);
        remap_indices(f->set, src, idx);
        write4(f->start, f->end, idx, f->best);
        f->best_error = f->error;
    }
}

/* ------------------------------------------------------------------ */
/* range.rs:44-192 */

typedef struct {
    const colourset* set;
    v3 weights, start, end;
    uint8_t indices[16];
    float best_error;
    uint8_t best[8];
} range_fit;

static v3 clamp_grid3(v3 p)
{
    /* one.min(zero.max(p)), then (grid*p + half).truncate() * gridrcp */
    p = mk3(fminf(1.0f, fmaxf(0.0f, p.x)), fminf(1.0f, fmaxf(0.0f, p.y)), fminf(1.0f, fmaxf(0.0f, p.z)));
    return mk3(truncf(31.0f * p.x + 0.5f) * (1.0f / 31.0f), truncf(63.0f * p.y + 0.5f) * (1.0f / 63.0f),
               truncf(31.0f * p.z + 0.5f) * (1.0f / 31.0f));
}

static void range_init(range_fit* f, const colourset* s, const float weights[3])
{
    memset(f, 0, sizeof *f);
    f->set = s;
    f->weights = mk3(weights[0], weights[1], weights[2]);
    f->best_error = FLT_MAX;
    float cov[6];
    weighted_covariance(s->points, s->weights, s->count, cov);
    v3 principle = principle_component(cov);
    v3 start = mk3(0, 0, 0), end = mk3(0, 0, 0);
    if (s->count > 0) {
        start = s->points[0];
        end = start;
        float mn = dot3(start, principle), mx = mn;
        for (int i = 1; i < s->count; ++i) {
            float d = dot3(s->points[i], principle);
            if (d < mn) { start = s->points[i]; mn = d; }
            else if (d > mx) { end = s->points[i]; mx = d; }
        }
    }
    f->start = clamp_grid3(start);
    f->end = clamp_grid3(end);
}

static int range_helper(range_fit* f, const v3* codes, int ncodes)
{
    const colourset* s = f->set;
    uint8_t closest[16] = {0};
    float error = 0.0f;
    for (int i = 0; i < s->count; ++i) {
        float dist = FLT_MAX;
        int idx = 0;
        for (int j = 0; j < ncodes; ++j) {
            v3 d = mk3(s->points[i].x - codes[j].x, s->points[i].y - codes[j].y, s->points[i].z - codes[j].z);
            v3 e = mk3(f->weights.x * d.x, f->weights.y * d.y, f->weights.z * d.z);
            float dd = e.x * e.x + e.y * e.y + e.z * e.z;
            if (dd < dist) { dist = dd; idx = j; }
        }
        closest[i] = (uint8_t)idx;
        error += dist;
    }
    if (error < f->best_error) {
        remap_indices(s, closest, f->indices);
        f->best_error = error;
        return 1;
    }
    return 0;
}

static void range_compress3(range_fit* f)
{
    v3 s = f->start, e = f->end;
    v3 codes[3] = {s, e, mk3(s.x * 0.5f + e.x * 0.5f, s.y * 0.5f + e.y * 0.5f, s.z * 0.5f + e.z * 0.5f)};
    if (range_helper(f, codes, 3)) write3(f->start, f->end, f->indices, f->best);
}

static void range_compress4(range_fit* f)
{
    v3 s = f->start, e = f->end;
    const float t2 = 2.0f / 3.0f, t1 = 1.0f / 3.0f;
    v3 codes[4] = {s, e, mk3(s.x * t2 + e.x * t1, s.y * t2 + e.y * t1, s.z * t2 + e.z * t1),
                   mk3(s.x * t1 + e.x * t2, s.y * t1 + e.y * t2, s.z * t1 + e.z * t2)};
    if (range_helper(f, codes, 4)) write4(f->start, f->end, f->indices, f->best);
}

/* ------------------------------------------------------------------ */
/* cluster.rs:33-418 */

typedef struct {
    const colourset* set;
    int fmt;
    v4 weights;
    int num_iterations;
    v3 principle;
    uint8_t order[8][16];
    v4 pw[16];
    v4 xsum;
    v4 best_error;
    uint8_t best[8];
    int searched3, searched4; /* orderings searched (statistics only) */
    ora_cluster_trace* trace;  /* optional: records each fit's winner */
} cluster_fit;

/* records the winner of one fit (3 or 4 colours) into the trace */
static void trace_fit(cluster_fit* f, int nc, int it, int i, int j, int k, float err, v4 ua, v4 ub, v4 a, v4 b)
{
    if (!f->trace) return;
    ora_fit_trace* t = nc == 3 ? &f->trace->fit3 : &f->trace->fit4;
    t->won = 1;
    t->iteration = it;
    t->i = i; t->j = j; t->k = k;
    t->error = err;
    t->a_raw[0] = ua.x; t->a_raw[1] = ua.y; t->a_raw[2] = ua.z;
    t->b_raw[0] = ub.x; t->b_raw[1] = ub.y; t->b_raw[2] = ub.z;
    t->a[0] = a.x; t->a[1] = a.y; t->a[2] = a.z;
    t->b[0] = b.x; t->b[1] = b.y; t->b[2] = b.z;
    memcpy(t->order, f->order[it], 16);
}

static void cluster_init(cluster_fit* f, const colourset* s, int fmt, const float w[3], int iterate)
{
    memset(f, 0, sizeof *f);
    f->set = s;
    f->fmt = fmt;
    f->weights = mk4(w[0], w[1], w[2], 1.0f);
    f->num_iterations = iterate ? 8 : 1;
    f->best_error = splat(FLT_MAX);
    float cov[6];
    weighted_covariance(s->points, s->weights, s->count, cov);
    f->principle = principle_component(cov);
}

/* cluster.rs:90-97: non-finite keys tie with each other and sort last */
static int key_less(float a, float b)
{
    int fa = isfinite(a), fb = isfinite(b);
    if (!fa && !fb) return 0;
    if (!fa) return 0;
    if (!fb) return 1;
    return a < b;
}

/* cluster.rs:78-136.  Rust's sort_unstable_by on 16 elements is an
 * insertion sort (len <= 20), which is stable; restated as such. */
static int construct_ordering(cluster_fit* f, v3 axis, int it)
{
    int count = f->set->count;
    int idx[16];
    float key[16];
    for (int i = 0; i < 16; ++i) { idx[i] = 0; key[i] = FLT_MAX; }
    for (int i = 0; i < count; ++i) { idx[i] = i; key[i] = dot3(f->set->points[i], axis); }
    for (int i = 1; i < 16; ++i) {
        int ti = idx[i];
        float tk = key[i];
        int j = i;
        while (j > 0 && key_less(tk, key[j - 1])) { idx[j] = idx[j - 1]; key[j] = key[j - 1]; --j; }
        idx[j] = ti; key[j] = tk;
    }
    for (int m = 0; m < 16; ++m) f->order[it][m] = (uint8_t)idx[m];
    for (int p = 0; p < it; ++p)
        if (memcmp(f->order[it], f->order[p], 16) == 0) return 0;
    f->xsum = splat(0.0f);
    for (int i = 0; i < count; ++i) {
        int j = f->order[it][i];
        v4 p = mk4(f->set->points[j].x, f->set->points[j].y, f->set->points[j].z, 1.0f);
        v4 x = mul4(p, splat(f->set->weights[j]));
        f->pw[i] = x;
        f->xsum = add4(f->xsum, x);
    }
    return 1;
}

static void cluster_compress3(cluster_fit* f)
{
    int count = f->set->count;
    const v4 two = splat(2.0f), one = splat(1.0f), zero = splat(0.0f), half = splat(0.5f);
    const v4 half_half2 = mk4(0.5f, 0.5f, 0.5f, 0.25f);
    const v4 grid = mk4(31.0f, 63.0f, 31.0f, 0.0f);
    const v4 gridrcp = mk4(1.0f / 31.0f, 1.0f / 63.0f, 1.0f / 31.0f, 0.0f);

    v4 best_start = zero, best_end = zero, best_error = f->best_error, best_ua = zero, best_ub = zero;
    int best_iteration = 0, best_i = 0, best_j = 0;
    v3 axis = f->principle;

    for (int it = 0; it < f->num_iterations; ++it) {
        if (!construct_ordering(f, axis, it)) break;
        f->searched3++;
        v4 part0 = zero;
        for (int i = 0; i < count; ++i) {
            v4 part1 = (i == 0) ? f->pw[0] : zero;
            int jmin = (i == 0) ? 1 : i;
            for (int j = jmin; j <= count; ++j) {
                v4 part2 = sub4(sub4(f->xsum, part1), part0);
                v4 alphax = add4(mul4(part1, half_half2), part0);
                v4 alpha2 = splat(alphax.w);
                v4 betax = add4(mul4(part1, half_half2), part2);
                v4 beta2 = splat(betax.w);
                v4 alphabeta = splat(mul4(part1, half_half2).w);
                v4 factor = rcp4(sub4(mul4(alpha2, beta2), mul4(alphabeta, alphabeta)));
                v4 a = mul4(sub4(mul4(alphax, beta2), mul4(betax, alphabeta)), factor);
                v4 b = mul4(sub4(mul4(betax, alpha2), mul4(alphax, alphabeta)), factor);
                v4 ua = a, ub = b;
                a = min4(one, max4(zero, a));
                b = min4(one, max4(zero, b));
                a = mul4(trunc4(add4(mul4(grid, a), half)), gridrcp);
                b = mul4(trunc4(add4(mul4(grid, b), half)), gridrcp);
                v4 e1 = add4(mul4(mul4(a, a), alpha2), mul4(mul4(b, b), beta2));
                v4 e2 = sub4(mul4(mul4(a, b), alphabeta), mul4(a, alphax));
                v4 e3 = sub4(e2, mul4(b, betax));
                v4 e4 = add4(mul4(two, e3), e1);
                v4 e5 = mul4(e4, f->weights);
                v4 error = add4(add4(splat(e5.x), splat(e5.y)), splat(e5.z));
                if (any_less4(error, best_error)) {
                    best_start = a; best_end = b; best_i = i; best_j = j;
                    best_error = error; best_iteration = it;
                    best_ua = ua; best_ub = ub;
                }
                if (j < count) part1 = add4(part1, f->pw[j]);
            }
            part0 = add4(part0, f->pw[i]);
        }
        if (best_iteration != it) break;
        axis = mk3(best_end.x - best_start.x, best_end.y - best_start.y, best_end.z - best_start.z);
    }

    if (any_less4(best_error, f->best_error)) {
        const uint8_t* order = f->order[best_iteration];
        uint8_t unordered[16] = {0}, idx[16];
        for (int m = best_i; m < best_j; ++m) unordered[order[m]] = 2;
        for (int m = best_j; m < count; ++m) unordered[order[m]] = 1;
        remap_indices(f->set, unordered, idx);
        write3(mk3(best_start.x, best_start.y, best_start.z), mk3(best_end.x, best_end.y, best_end.z), idx, f->best);
        f->best_error = best_error;
        trace_fit(f, 3, best_iteration, best_i, best_j, 0, best_error.x, best_ua, best_ub, best_start, best_end);
    }
}

static void cluster_compress4(cluster_fit* f)
{
    int count = f->set->count;
    const v4 two = splat(2.0f), one = splat(1.0f), zero = splat(0.0f), half = splat(0.5f);
    const v4 c13 = mk4(1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 9.0f);
    const v4 c23 = mk4(2.0f / 3.0f, 2.0f / 3.0f, 2.0f / 3.0f, 4.0f / 9.0f);
    const v4 twoninths = splat(2.0f / 9.0f);
    const v4 grid = mk4(31.0f, 63.0f, 31.0f, 0.0f);
    const v4 gridrcp = mk4(1.0f / 31.0f, 1.0f / 63.0f, 1.0f / 31.0f, 0.0f);

    v4 best_start = zero, best_end = zero, best_error = f->best_error, best_ua = zero, best_ub = zero;
    int best_iteration = 0, best_i = 0, best_j = 0, best_k = 0;
    v3 axis = f->principle;

    for (int it = 0; it < f->num_iterations; ++it) {
        if (!construct_ordering(f, axis, it)) break;
        f->searched4++;
        v4 part0 = zero;
        for (int i = 0; i < count; ++i) {
            v4 part1 = zero;
            for (int j = i; j <= count; ++j) {
                v4 part2 = (j == 0) ? f->pw[0] : zero;
                int kmin = (j == 0) ? 1 : j;
                for (int k = kmin; k <= count; ++k) {
                    v4 part3 = sub4(sub4(sub4(f->xsum, part2), part1), part0);
                    v4 alphax = add4(mul4(part2, c13), add4(mul4(part1, c23), part0));
                    v4 alpha2 = splat(alphax.w);
                    v4 betax = add4(mul4(part1, c13), add4(mul4(part2, c23), part3));
                    v4 beta2 = splat(betax.w);
                    v4 alphabeta = mul4(twoninths, splat(add4(part1, part2).w));
                    v4 factor = rcp4(sub4(mul4(alpha2, beta2), mul4(alphabeta, alphabeta)));
                    v4 a = mul4(sub4(mul4(alphax, beta2), mul4(betax, alphabeta)), factor);
                    v4 b = mul4(sub4(mul4(betax, alpha2), mul4(alphax, alphabeta)), factor);
                    v4 ua = a, ub = b;
                    a = min4(one, max4(zero, a));
                    b = min4(one, max4(zero, b));
                    a = mul4(trunc4(add4(mul4(grid, a), half)), gridrcp);
                    b = mul4(trunc4(add4(mul4(grid, b), half)), gridrcp);
                    v4 e1 = add4(mul4(mul4(a, a), alpha2), mul4(mul4(b, b), beta2));
                    v4 e2 = sub4(mul4(mul4(a, b), alphabeta), mul4(a, alphax));
                    v4 e3 = sub4(e2, mul4(b, betax));
                    v4 e4 = add4(mul4(two, e3), e1);
                    v4 e5 = mul4(e4, f->weights);
                    v4 error = add4(add4(splat(e5.x), splat(e5.y)), splat(e5.z));
                    if (any_less4(error, best_error)) {
                        best_start = a; best_end = b; best_i = i; best_j = j; best_k = k;
                        best_error = error; best_iteration = it;
                        best_ua = ua; best_ub = ub;
                    }
                    if (k < count) part2 = add4(part2, f->pw[k]);
                }
                if (j < count) part1 = add4(part1, f->pw[j]);
            }
            part0 = add4(part0, f->pw[i]);
        }
        if (best_iteration != it) break;
        axis = mk3(best_end.x - best_start.x, best_end.y - best_start.y, best_end.z - best_start.z);
    }

    if (any_less4(best_error, f->best_error)) {
        const uint8_t* order = f->order[best_iteration];
        uint8_t unordered[16] = {0}, idx[16];
        for (int m = best_i; m < best_j; ++m) unordered[order[m]] = 2;
        for (int m = best_j; m < count; ++m) unordered[order[m]] = 3;
        for (int m = best_k; m < count; ++m) unordered[order[m]] = 1;
        remap_indices(f->set, unordered, idx);
        write4(mk3(best_start.x, best_start.y, best_start.z), mk3(best_end.x, best_end.y, best_end.z), idx, f->best);
        f->best_error = best_error;
        trace_fit(f, 4, best_iteration, best_i, best_j, best_k, best_error.x, best_ua, best_ub, best_start, best_end);
    }
}

/* ------------------------------------------------------------------ */
/* alpha.rs:27-256 */

static void compress_bc2(const uint8_t rgba[16][4], uint32_t mask, uint8_t* blk)
{
    for (int i = 0; i < 8; ++i) {
        float a1 = (float)rgba[2 * i][3] * (15.0f / 255.0f);
        float a2 = (float)rgba[2 * i + 1][3] * (15.0f / 255.0f);
        uint8_t q1 = (uint8_t)f32_to_i32_clamped(a1, 15);
        uint8_t q2 = (uint8_t)f32_to_i32_clamped(a2, 15);
        if ((mask & (1u << (2 * i))) == 0) q1 = 0;
        if ((mask & (1u << (2 * i + 1))) == 0) q2 = 0;
        blk[i] = (uint8_t)(q1 | (q2 << 4));
    }
}

static void fix_range(int* mn, int* mx, int steps)
{
    if (*mx - *mn < steps) *mx = (*mn + steps < 255) ? *mn + steps : 255;
    if (*mx - *mn < steps) *mn = (*mx - steps > 0) ? *mx - steps : 0;
}

static uint32_t fit_codes(const uint8_t rgba[16][4], int ch, uint32_t mask, const int codes[8], uint8_t idx[16])
{
    uint32_t err = 0;
    for (int i = 0; i < 16; ++i) {
        if ((mask & (1u << i)) == 0) { idx[i] = 0; continue; }
        int v = rgba[i][ch];
        uint32_t least = UINT32_MAX;
        int index = 0;
        for (int j = 0; j < 8; ++j) {
            int d = v - codes[j];
            uint32_t dd = (uint32_t)(d * d);
            if (dd < least) { least = dd; index = j; }
        }
        idx[i] = (uint8_t)index;
        err += least;
    }
    return err;
}

static void write_alpha_block(int a0, int a1, const uint8_t idx[16], uint8_t* blk)
{
    blk[0] = (uint8_t)a0;
    blk[1] = (uint8_t)a1;
    for (int i = 0; i < 2; ++i) {
        uint32_t v = 0;
        for (int j = 0; j < 8; ++j) v |= (uint32_t)idx[8 * i + j] << (3 * j);
        for (int j = 0; j < 3; ++j) blk[2 + 3 * i + j] = (uint8_t)((v >> (8 * j)) & 0xFF);
    }
}

static void write_alpha_block5(int a0, int a1, const uint8_t idx[16], uint8_t* blk)
{
    if (a0 > a1) {
        uint8_t s[16];
        for (int i = 0; i < 16; ++i) {
            int x = idx[i];
            s[i] = (uint8_t)((x == 0) ? 1 : (x == 1) ? 0 : (x >= 2 && x <= 5) ? 7 - x : x);
        }
        write_alpha_block(a1, a0, s, blk);
    } else {
        write_alpha_block(a0, a1, idx, blk);
    }
}

static void write_alpha_block7(int a0, int a1, const uint8_t idx[16], uint8_t* blk)
{
    if (a0 < a1) {
        uint8_t s[16];
        for (int i = 0; i < 16; ++i) {
            int x = idx[i];
            s[i] = (uint8_t)((x == 0) ? 1 : (x == 1) ? 0 : 9 - x);
        }
        write_alpha_block(a1, a0, s, blk);
    } else {
        write_alpha_block(a0, a1, idx, blk);
    }
}

static void compress_bc3_alpha(const uint8_t rgba[16][4], int ch, uint32_t mask, uint8_t* blk)
{
    int min5 = 255, max5 = 0, min7 = 255, max7 = 0;
    for (int i = 0; i < 16; ++i) {
        if ((mask & (1u << i)) == 0) continue;
        int v = rgba[i][ch];
        if (v < min7) min7 = v;
        if (v > max7) max7 = v;
        if (v != 0 && v < min5) min5 = v;
        if (v != 255 && v > max5) max5 = v;
    }
    if (min5 > max5) min5 = max5;
    if (min7 > max7) min7 = max7;
    fix_range(&min5, &max5, 5);
    fix_range(&min7, &max7, 7);
    int codes5[8], codes7[8];
    codes5[0] = min5; codes5[1] = max5;
    for (int i = 1; i < 5; ++i) codes5[1 + i] = ((5 - i) * min5 + i * max5) / 5;
    codes5[6] = 0; codes5[7] = 255;
    /* alpha.rs:238-239: the 7-point codebook keeps min5/max5 in slots 0 and 1 */
    codes7[0] = min5; codes7[1] = max5;
    for (int i = 1; i < 7; ++i) codes7[1 + i] = ((7 - i) * min7 + i * max7) / 7;
    uint8_t idx5[16], idx7[16];
    uint32_t err5 = fit_codes(rgba, ch, mask, codes5, idx5);
    uint32_t err7 = fit_codes(rgba, ch, mask, codes7, idx7);
    if (err5 <= err7) write_alpha_block5(min5, max5, idx5, blk);
    else write_alpha_block7(min7, max7, idx7, blk);
}

/* ------------------------------------------------------------------ */
/* lib.rs:188-234 */

static int compress_block_impl(int fmt, const uint8_t rgba_flat[64], uint32_t mask, const ora_params* p,
                               uint8_t* out, ora_block_stats* st, ora_cluster_trace* tr);

int ora_compress_block_masked(int fmt, const uint8_t rgba_flat[64], uint32_t mask, const ora_params* p,
                              uint8_t* out, ora_block_stats* st)
{
    return compress_block_impl(fmt, rgba_flat, mask, p, out, st, NULL);
}

int ora_trace_block(int fmt, const uint8_t rgba[64], uint32_t mask, const ora_params* p, uint8_t* out,
                    ora_cluster_trace* t)
{
    memset(t, 0, sizeof *t);
    return compress_block_impl(fmt, rgba, mask, p, out, NULL, t);
}

static int compress_block_impl(int fmt, const uint8_t rgba_flat[64], uint32_t mask, const ora_params* p,
                               uint8_t* out, ora_block_stats* st, ora_cluster_trace* tr)
{
    if (fmt < TX_BC1 || fmt > TX_BC5 || p->algorithm < 0 || p->algorithm > 2) return -1;
    pthread_once(&g_lut_once, build_luts);
    const uint8_t(*rgba)[4] = (const uint8_t(*)[4])rgba_flat;
    if (st) memset(st, 0, sizeof *st);
    switch (fmt) {
    case TX_BC2: compress_bc2(rgba, mask, out); break;
    case TX_BC3: compress_bc3_alpha(rgba, 3, mask, out); break;
    case TX_BC4: compress_bc3_alpha(rgba, 0, mask, out); break;
    case TX_BC5: compress_bc3_alpha(rgba, 0, mask, out); compress_bc3_alpha(rgba, 1, mask, out + 8); break;
    default: break;
    }
    if (fmt == TX_BC1 || fmt == TX_BC2 || fmt == TX_BC3) {
        colourset set;
        colourset_new(&set, rgba, mask, fmt, p->weigh_colour_by_alpha);
        uint8_t* cb = out + (fmt == TX_BC1 ? 0 : 8);
        int is_bc1 = fmt == TX_BC1;
        if (st) { st->count = set.count; st->transparent = set.transparent; }
        if (tr) {
            tr->count = set.count;
            tr->transparent = set.transparent;
            for (int q = 0; q < 16; ++q) {
                tr->points[q][0] = set.points[q].x; tr->points[q][1] = set.points[q].y;
                tr->points[q][2] = set.points[q].z; tr->weights[q] = set.weights[q];
            }
            weighted_covariance(set.points, set.weights, set.count, tr->cov);
            v3 pc = principle_component(tr->cov);
            tr->principle[0] = pc.x; tr->principle[1] = pc.y; tr->principle[2] = pc.z;
        }
        if (set.count == 1) {
            single_fit f;
            memset(&f, 0, sizeof f);
            f.set = &set;
            f.error = UINT32_MAX;
            f.best_error = UINT32_MAX;
            if (is_bc1) { single_compress3(&f); if (!set.transparent) single_compress4(&f); }
            else single_compress4(&f);
            memcpy(cb, f.best, 8);
            if (st) st->fit = 0;
        } else if (p->algorithm == TX_RANGE_FIT || set.count == 0) {
            range_fit f;
            range_init(&f, &set, p->weights);
            if (is_bc1) { range_compress3(&f); if (!set.transparent) range_compress4(&f); }
            else range_compress4(&f);
            memcpy(cb, f.best, 8);
            if (st) st->fit = 1;
        } else {
            cluster_fit f;
            cluster_init(&f, &set, fmt, p->weights, p->algorithm == TX_ITERATIVE_CLUSTER_FIT);
            f.trace = tr;
            if (is_bc1) { cluster_compress3(&f); if (!set.transparent) cluster_compress4(&f); }
            else cluster_compress4(&f);
            memcpy(cb, f.best, 8);
            if (st) { st->fit = 2; st->searched3 = f.searched3; st->searched4 = f.searched4; }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* lib.rs:287-335, rows split over threads like rayon's par_chunks_mut */

size_t ora_block_size(int fmt) { return (fmt == TX_BC1 || fmt == TX_BC4) ? 8 : 16; }

typedef struct {
    int fmt;
    const uint8_t* blocks;
    const uint32_t* masks;
    const ora_params* p;
    uint8_t* out;
    size_t begin, end;
} blocks_job;

static void* blocks_thread(void* arg)
{
    const blocks_job* j = (const blocks_job*)arg;
    size_t bs = ora_block_size(j->fmt);
    for (size_t b = j->begin; b < j->end; ++b)
        ora_compress_block_masked(j->fmt, j->blocks + 64 * b, j->masks[b], j->p, j->out + b * bs, NULL);
    return NULL;
}

/* many independent compress_block_masked calls, split over threads */
int ora_compress_blocks(int fmt, const uint8_t* blocks, const uint32_t* masks, size_t n, const ora_params* p,
                        uint8_t* out, int nthreads)
{
    if (fmt < TX_BC1 || fmt > TX_BC5) return -1;
    pthread_once(&g_lut_once, build_luts);
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > n) nthreads = n ? (int)n : 1;
    blocks_job* jobs = (blocks_job*)calloc((size_t)nthreads, sizeof *jobs);
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof *th);
    if (!jobs || !th) { free(jobs); free(th); return -6; }
    for (int t = 0; t < nthreads; ++t) {
        blocks_job* j = &jobs[t];
        j->fmt = fmt; j->blocks = blocks; j->masks = masks; j->p = p; j->out = out;
        j->begin = n * (size_t)t / (size_t)nthreads;
        j->end = n * (size_t)(t + 1) / (size_t)nthreads;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, blocks_thread, &jobs[t]);
    blocks_thread(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
    return 0;
}
size_t ora_num_blocks(size_t n) { return (n + 3) / 4; }

typedef struct {
    int fmt;
    const uint8_t* rgba;
    size_t w, h, blocks_wide, nblocks;
    const ora_params* p;
    uint8_t* out;
    ora_block_stats* st;
    size_t row_begin, row_end;
} compress_job;

static void compress_rows(const compress_job* j)
{
    size_t bs = ora_block_size(j->fmt);
    uint8_t px[64];
    for (size_t y = j->row_begin; y < j->row_end; ++y) {
        memset(px, 0, sizeof px);
        for (size_t x = 0; x < j->blocks_wide; ++x) {
            size_t b = y * j->blocks_wide + x;
            if (b >= j->nblocks) return;
            uint32_t mask = 0;
            for (int py = 0; py < 4; ++py)
                for (int pxx = 0; pxx < 4; ++pxx) {
                    int i = 4 * py + pxx;
                    size_t sx = 4 * x + pxx, sy = 4 * y + py;
                    if (sx < j->w && sy < j->h) {
                        memcpy(px + 4 * i, j->rgba + 4 * (j->w * sy + sx), 4);
                        mask |= 1u << i;
                    }
                }
            ora_compress_block_masked(j->fmt, px, mask, j->p, j->out + b * bs, j->st ? j->st + b : NULL);
        }
    }
}

static void* compress_thread(void* arg) { compress_rows((const compress_job*)arg); return NULL; }

int ora_compress(int fmt, const uint8_t* rgba, size_t rgba_len, size_t w, size_t h, const ora_params* p,
                 uint8_t* out, size_t out_len, int nthreads, ora_block_stats* stats)
{
    if (fmt < TX_BC1 || fmt > TX_BC5) return -1;
    if (w == 0) return -2;
    if (rgba_len < 4 * w * h) return -3;
    size_t bs = ora_block_size(fmt);
    size_t bw = ora_num_blocks(w);
    if (out_len < bw * ora_num_blocks(h) * bs) return -4;
    if (out_len % bs != 0) return -5;
    pthread_once(&g_lut_once, build_luts);
    size_t nblocks = out_len / bs;
    size_t rows = (nblocks + bw - 1) / bw;
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > rows) nthreads = rows ? (int)rows : 1;
    compress_job* jobs = (compress_job*)calloc((size_t)nthreads, sizeof *jobs);
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof *th);
    if (!jobs || !th) { free(jobs); free(th); return -6; }
    for (int t = 0; t < nthreads; ++t) {
        compress_job* j = &jobs[t];
        j->fmt = fmt; j->rgba = rgba; j->w = w; j->h = h; j->blocks_wide = bw; j->nblocks = nblocks;
        j->p = p; j->out = out; j->st = stats;
        j->row_begin = rows * (size_t)t / (size_t)nthreads;
        j->row_end = rows * (size_t)(t + 1) / (size_t)nthreads;
    }
    if (nthreads == 1) {
        compress_rows(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, compress_thread, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    free(jobs);
    free(th);
    return 0;
}

/* ------------------------------------------------------------------ */
/* decoders: colourblock.rs:96-169, alpha.rs:53-68 and 258-304, lib.rs:124-156, 240-277 */

static void unpack_565(const uint8_t* p, uint8_t out[4])
{
    unsigned v = (unsigned)p[0] | ((unsigned)p[1] << 8);
    unsigned r = (v >> 11) & 0x1F, g = (v >> 5) & 0x3F, b = v & 0x1F;
    out[0] = (uint8_t)((r << 3) | (r >> 2));
    out[1] = (uint8_t)((g << 2) | (g >> 4));
    out[2] = (uint8_t)((b << 3) | (b >> 2));
    out[3] = 255;
}

static void decompress_colour(const uint8_t* bytes, int is_bc1, uint8_t rgba[16][4])
{
    uint8_t codes[16];
    unsigned a = (unsigned)bytes[0] | ((unsigned)bytes[1] << 8);
    unsigned b = (unsigned)bytes[2] | ((unsigned)bytes[3] << 8);
    unpack_565(bytes, codes);
    unpack_565(bytes + 2, codes + 4);
    for (int i = 0; i < 4; ++i) {
        unsigned c = codes[i], d = codes[4 + i];
        if (is_bc1 && a <= b) { codes[8 + i] = (uint8_t)((c + d) / 2); codes[12 + i] = 0; }
        else { codes[8 + i] = (uint8_t)((2 * c + d) / 3); codes[12 + i] = (uint8_t)((c + 2 * d) / 3); }
    }
    codes[8 + 3] = 255;
    codes[12 + 3] = (is_bc1 && a <= b) ? 0 : 255;
    for (int i = 0; i < 16; ++i) {
        int ind = (bytes[4 + i / 4] >> (2 * (i % 4))) & 3;
        memcpy(rgba[i], codes + 4 * ind, 4);
    }
}

static void decompress_bc2_alpha(uint8_t rgba[16][4], const uint8_t* bytes)
{
    for (int i = 0; i < 8; ++i) {
        uint8_t q = bytes[i], lo = q & 0x0F, hi = q & 0xF0;
        rgba[2 * i][3] = (uint8_t)(lo | (lo << 4));
        rgba[2 * i + 1][3] = (uint8_t)(hi | (hi >> 4));
    }
}

static void decompress_bc3_alpha(uint8_t rgba[16][4], int ch, const uint8_t* bytes)
{
    int a0 = bytes[0], a1 = bytes[1];
    int codes[8];
    codes[0] = a0; codes[1] = a1;
    if (a0 <= a1) {
        for (int i = 1; i < 5; ++i) codes[1 + i] = ((5 - i) * a0 + i * a1) / 5;
        codes[6] = 0; codes[7] = 255;
    } else {
        for (int i = 1; i < 7; ++i) codes[1 + i] = ((7 - i) * a0 + i * a1) / 7;
    }
    for (int i = 0; i < 2; ++i) {
        int v = bytes[2 + 3 * i] | (bytes[3 + 3 * i] << 8) | (bytes[4 + 3 * i] << 16);
        for (int j = 0; j < 8; ++j) rgba[8 * i + j][ch] = (uint8_t)codes[(v >> (3 * j)) & 7];
    }
}

void ora_decompress_block(int fmt, const uint8_t* block, uint8_t out_flat[64])
{
    uint8_t(*rgba)[4] = (uint8_t(*)[4])out_flat;
    if (fmt == TX_BC1 || fmt == TX_BC2 || fmt == TX_BC3) {
        decompress_colour(block + (fmt == TX_BC1 ? 0 : 8), fmt == TX_BC1, rgba);
    } else {
        for (int i = 0; i < 16; ++i) { rgba[i][0] = 0; rgba[i][1] = 0; rgba[i][2] = 0; rgba[i][3] = 255; }
    }
    switch (fmt) {
    case TX_BC2: decompress_bc2_alpha(rgba, block); break;
    case TX_BC3: decompress_bc3_alpha(rgba, 3, block); break;
    case TX_BC4:
        decompress_bc3_alpha(rgba, 0, block);
        for (int i = 0; i < 16; ++i) { rgba[i][1] = rgba[i][0]; rgba[i][2] = rgba[i][0]; }
        break;
    case TX_BC5: decompress_bc3_alpha(rgba, 0, block); decompress_bc3_alpha(rgba, 1, block + 8); break;
    default: break;
    }
}

int ora_decompress(int fmt, const uint8_t* data, size_t data_len, size_t w, size_t h, uint8_t* out, size_t out_len)
{
    if (fmt < TX_BC1 || fmt > TX_BC5) return -1;
    if (w == 0) return -2;
    size_t bs = ora_block_size(fmt), bw = ora_num_blocks(w);
    /* lib.rs:129: output is walked in bands of 4 pixel rows (width*16 bytes) */
    size_t band = w * 16;
    size_t bands = (out_len + band - 1) / band;
    if (out_len < 4 * w * h) return -4;
    for (size_t y = 0; y < bands; ++y) {
        for (size_t x = 0; x < bw; ++x) {
            size_t bidx = (x + y * bw) * bs;
            if (bidx + bs > data_len) return -3;
            uint8_t rgba[64];
            ora_decompress_block(fmt, data + bidx, rgba);
            for (int py = 0; py < 4; ++py)
                for (int px = 0; px < 4; ++px) {
                    size_t sx = 4 * x + px, sy = (size_t)py;
                    if (sx < w && 4 * y + sy < h) memcpy(out + y * band + 4 * (sx + sy * w), rgba + 4 * (px + py * 4), 4);
                }
        }
    }
    return 0;
}
