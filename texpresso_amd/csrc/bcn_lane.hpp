// bcn_lane.hpp -- front half of a BC1/BC2/BC3 block with the whole block on
// one lane: alpha, colour set, principal axis, and the single-colour and
// range fits (lib.rs:188-230, colourset.rs:35-112, math.rs:44-97,
// single.rs:100-164, range.rs:44-192).
//
// The 16-lane-group front (encode_front) spreads a block over 16 lanes, which
// suits the colour-set comparisons but repeats every serial fold (centroid,
// covariance, eight power-iteration steps, the range fit's point loops) on
// all 16 lanes.  Here each lane owns one block, so a wave runs 64 blocks and
// each serial step is done once per block.  The arithmetic is the same
// expression for expression: points are visited in point order (first
// occurrence order, i.e. increasing pixel index), so every fold matches.
#pragma once
#include "bcn_device.hpp"

namespace txd {

// Exact values the front needs, filled once per workgroup with the same
// expressions the group front evaluates: u8 / 255 for the colour channels and
// sqrtf(count) for unweighted point weights.
struct LaneTables {
    float u8f[256];
    float sqc[17];
};

__device__ __forceinline__ void fill_lane_tables(LaneTables& T, int lane)
{
    for (int i = lane; i < 256; i += 64) T.u8f[i] = (float)i / 255.0f;
    if (lane < 17) T.sqc[lane] = sqrtf((float)lane);
}

// Per-wave LDS of the lane front: the pixels and point weights of every
// lane's block, pixel-major with the lane fastest (each access by the wave is
// 64 consecutive words, free of bank conflicts).  Loops over pixels read
// them here by index instead of holding 16-wide arrays in registers.
struct LaneBlockLds {
    uint32_t px[16][64];
    float w[16][64];
};

// The colour set of one block on one lane (colourset.rs:35-112).  Pixels are
// kept in pixel order; a point is the first included pixel of its colour.
struct LaneSet {
    LaneBlockLds* L;
    int lane;
    uint32_t first;     // bit p: pixel p opens a point
    uint32_t inc;       // bit p: pixel p is in the set
    uint32_t remap[4];  // one byte per pixel: point index, 255 = excluded
    int n;
    bool transparent;

    __device__ uint32_t pix(int p) const { return L->px[p][lane]; }
    __device__ float wt(int p) const { return L->w[p][lane]; }  // sqrt(weight), 0 off the points
};

// Colour of a pixel as a point (x, y, z in [0, 1]), from the u8 / 255 table.
// (An arithmetic form, c * (1/255) corrected by one fma step from the exact
// residual, is correctly rounded for all 256 bytes but measured no faster:
// its fma and conversion issue at half the rate of plain adds.)
__device__ __forceinline__ void xyz(const LaneTables& T, uint32_t rgb, float& x, float& y, float& z)
{
    x = T.u8f[rgb & 255u];
    y = T.u8f[(rgb >> 8) & 255u];
    z = T.u8f[(rgb >> 16) & 255u];
}

// Pixel p's colour as a comparison key: its RGB when included, otherwise a
// value no colour has (so it matches nothing).
__device__ __forceinline__ uint32_t colour_key(uint32_t px, uint32_t inc, int p)
{
    return ((inc >> p) & 1u) ? (px & 0xFFFFFFu) : 0x1000000u + (uint32_t)p;
}

// Bit m set iff key == keys[m], m = 0..15.  Each compare writes VCC and one
// add-with-carry shifts the mask left and brings the compare in at bit 0
// (mask + mask + carry), keys visited from 15 down to 0: two instructions
// per key, where compare, select and or take three and more.  The carry
// hand-off is the one of a 64-bit add (v_add_co / v_addc_co), which needs no
// wait states.  One asm block, so nothing is scheduled between the pairs.
__device__ __forceinline__ uint32_t match_mask16(uint32_t key, const uint32_t (&keys)[16])
{
    uint32_t mask = 0u;
    asm("v_cmp_eq_u32_e32 vcc, %1, %17\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %16\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %15\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %14\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %13\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %12\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %11\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %10\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %9\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %8\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %7\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %6\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %5\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %4\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %3\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
        "v_cmp_eq_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
        : "+v"(mask)
        : "v"(key), "v"(keys[0]), "v"(keys[1]), "v"(keys[2]), "v"(keys[3]), "v"(keys[4]), "v"(keys[5]),
          "v"(keys[6]), "v"(keys[7]), "v"(keys[8]), "v"(keys[9]), "v"(keys[10]), "v"(keys[11]), "v"(keys[12]),
          "v"(keys[13]), "v"(keys[14]), "v"(keys[15])
        : "vcc");
    return mask;
}

// Batcher's odd-even merge sort of 16 values, ascending: 63 compare-exchange
// steps, straight-line once unrolled (the network sorts every 0/1 input, so
// it sorts every input).
__device__ __forceinline__ void sort16(uint32_t (&k)[16])
{
#pragma unroll
    for (int p = 1; p < 16; p <<= 1)
#pragma unroll
        for (int d = p; d >= 1; d >>= 1)
#pragma unroll
            for (int j = d % p; j + d < 16; j += 2 * d)
#pragma unroll
                for (int i = 0; i < d; ++i)
                    if (i + j + d < 16 && (i + j) / (2 * p) == (i + j + d) / (2 * p)) {
                        const uint32_t a = k[i + j], b = k[i + j + d];
                        k[i + j] = min(a, b);
                        k[i + j + d] = max(a, b);
                    }
}

// colourset.rs:35-112 on one lane, unweighted points (weigh_colour_by_alpha
// off): the 16 keys (colour << 4 | pixel; an excluded pixel's colour is a
// value no colour has) sorted put each colour's pixels in one run, lowest
// pixel first.  The run's first pixel opens the colour's point when it is in
// the set (it is the lowest pixel of that colour); every pixel of the run has
// it as its lowest same-colour pixel; the point's weight is sqrtf(run
// length).  Same sets, remaps and weights as the pairwise compares of the
// weighted form below.
template <int FMT>
__device__ __forceinline__ void colour_set_lane_sorted(const uint32_t (&px)[16], uint32_t inc, const LaneTables& T,
                                                       LaneBlockLds& L, int lane, LaneSet& S)
{
    uint32_t k[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) k[p] = (colour_key(px[p], inc, p) << 4) | (uint32_t)p;
    sort16(k);
    uint32_t first = 0, head = 0, cnt = 0;
    uint64_t lowest = 0;  // nibble p: the lowest pixel of pixel p's colour
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const uint32_t i = k[s] & 15u;
        const bool start = s == 0 || (k[s] >> 4) != (k[s - 1] >> 4);
        head = start ? i : head;
        cnt = start ? 1u : cnt + 1u;
        const bool in = (inc >> i) & 1u;
        first |= (start && in ? 1u : 0u) << i;
        lowest |= (uint64_t)head << (4 * i);
        // weights: the run's first pixel holds sqrt(pixels so far), so its
        // run's last pixel leaves the full count; every other pixel 0, and an
        // excluded pixel (a run of its own) 0
        L.w[i][lane] = start && in ? T.sqc[1] : 0.0f;
        if (!start) L.w[head][lane] = T.sqc[cnt];
    }
    uint64_t remap = 0;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const uint32_t f = (uint32_t)(lowest >> (4 * p)) & 15u;
        const uint64_t r = ((inc >> p) & 1u) ? (uint64_t)__popc(first & ((1u << f) - 1u)) : 255u;
        remap |= r << (8 * (p & 7));
        if (p == 7) {
            S.remap[0] = (uint32_t)remap;
            S.remap[1] = (uint32_t)(remap >> 32);
            remap = 0;
        }
    }
    S.remap[2] = (uint32_t)remap;
    S.remap[3] = (uint32_t)(remap >> 32);
    S.first = first;
    S.n = __popc(first);
}

// colourset.rs:35-112 on one lane.  For each pixel p, same = the pixels with
// p's colour (p included): p opens a point when none of them is lower; its
// point index counts the points opened below its lowest same-colour pixel;
// the point's weight is the pixel count, or the sum of (alpha + 1) / 256
// over those pixels (exact in binary32), then sqrtf.  Unweighted points take
// the sorted form above.
template <int FMT>
__device__ __forceinline__ void colour_set_lane(const uint32_t (&px)[16], uint32_t valid, const FitParams& fp,
                                                const LaneTables& T, LaneBlockLds& L, int lane, LaneSet& S)
{
    S.L = &L;
    S.lane = lane;
    uint32_t inc = 0, transp = 0;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const uint32_t a = px[p] >> 24;
        const bool v = (valid >> p) & 1u;
        inc |= (v && !(FMT == BC1 && a < 128u) ? 1u : 0u) << p;
        transp |= (v && a < 128u ? 1u : 0u) << p;
        L.px[p][lane] = px[p];
    }
    S.inc = inc;
    S.transparent = FMT == BC1 && transp != 0;
    if (!fp.weigh_alpha) {
        colour_set_lane_sorted<FMT>(px, inc, T, L, lane, S);
        return;
    }
    uint32_t first = 0;
    uint64_t remap = 0;
    uint32_t keys[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) keys[m] = colour_key(px[m], inc, m);
    // the pixel loop stays rolled (pixel p comes from LDS) and the 16
    // compares against the register copies are unrolled
#pragma unroll 1
    for (int p = 0; p < 16; ++p) {
        const uint32_t kp = colour_key(L.px[p][lane], inc, p);
        uint32_t same = 0, asum = 0;
        if (fp.weigh_alpha) {
            // e = 1 when the keys match: keys are below 2^25, so for
            // distinct keys (difference - 1) never has bit 31 set
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const uint32_t e = ((kp ^ keys[m]) - 1u) >> 31;
                same |= e << m;
                asum += e * ((px[m] >> 24) + 1u);
            }
        } else {
            same = match_mask16(kp, keys);
        }
        const bool in = (inc >> p) & 1u;
        const bool opens = in && (same & ((1u << p) - 1u)) == 0u;
        first |= (opens ? 1u : 0u) << p;
        const uint32_t f = (uint32_t)__builtin_ctz(same);  // lowest pixel of p's colour
        const uint64_t r = in ? (uint64_t)__popc(first & ((1u << f) - 1u)) : 255u;
        remap |= r << (8 * (p & 7));
        if (p == 7) {
            S.remap[0] = (uint32_t)remap;
            S.remap[1] = (uint32_t)(remap >> 32);
            remap = 0;
        }
        float w = 0.0f;
        if (opens) w = fp.weigh_alpha ? sqrtf((float)asum / 256.0f) : T.sqc[__popc(same)];
        L.w[p][lane] = w;
    }
    S.remap[2] = (uint32_t)remap;
    S.remap[3] = (uint32_t)(remap >> 32);
    S.first = first;
    S.n = __popc(first);
}

// math.rs:44-97 over the points of a lane's set (same folds as
// principal_axis).  The folds run over all 16 pixels: a pixel that opens no
// point has w = 0, so its terms are +0 or -0 and adding them leaves the
// (never negative-zero) accumulators unchanged.  With `pts`, the first fold
// also stores each point as (x, y, z, weight) at its point index there (the
// ClusterFit record of the block).
__device__ __forceinline__ void principal_axis_lane(const LaneSet& S, const LaneTables& T, float& vx, float& vy,
                                                    float& vz, float4* __restrict__ pts = nullptr)
{
    float total = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
#pragma unroll 4
    for (int p = 0; p < 16; ++p) {
        float x, y, z;
        xyz(T, S.pix(p), x, y, z);
        const float w = S.wt(p);
        if (pts && ((S.first >> p) & 1u)) pts[__popc(S.first & ((1u << p) - 1u))] = make_float4(x, y, z, w);
        total = total + w;
        cx = cx + x * w;
        cy = cy + y * w;
        cz = cz + z * w;
    }
    if (total > FLT_EPSILON) {
        cx = cx / total;
        cy = cy / total;
        cz = cz / total;
    }
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f, c4 = 0.0f, c5 = 0.0f;
#pragma unroll 4
    for (int p = 0; p < 16; ++p) {
        float x, y, z;
        xyz(T, S.pix(p), x, y, z);
        const float w = S.wt(p);
        const float ax = x - cx, ay = y - cy, az = z - cz;
        const float bx = ax * w, by = ay * w, bz = az * w;
        c0 = c0 + ax * bx;
        c1 = c1 + ax * by;
        c2 = c2 + ax * bz;
        c3 = c3 + ay * by;
        c4 = c4 + ay * bz;
        c5 = c5 + az * bz;
    }
    vx = 1.0f; vy = 1.0f; vz = 1.0f;
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        float wx = c0 * vx, wy = c1 * vx, wz = c2 * vx;
        wx = c1 * vy + wx; wy = c3 * vy + wy; wz = c4 * vy + wz;
        wx = c2 * vz + wx; wy = c4 * vz + wy; wz = c5 * vz + wz;
        const float a = fmaxf(wx, fmaxf(wy, wz));
        const float r = 1.0f / a;
        vx = wx * r; vy = wy * r; vz = wz * r;
    }
}

// colourblock.rs:58-94 with all 16 indices on one lane (pixel p in bits
// 2p..2p+1 of `idx`).
__device__ __forceinline__ void write_colour_lane(bool four, float sx, float sy, float sz, float ex, float ey, float ez,
                                                  uint32_t idx, uint32_t& lo, uint32_t& hi)
{
    uint32_t a = pack_565(sx, sy, sz), b = pack_565(ex, ey, ez);
    if (!four) {
        if (a > b) {  // write3: a <= b, swap palette slots 0 and 1
            const uint32_t t = a; a = b; b = t;
            idx ^= ~(idx >> 1) & 0x55555555u;  // 0 <-> 1 where the high bit is clear
        }
    } else {
        if (a < b) {  // write4: a >= b, index ^ 1
            const uint32_t t = a; a = b; b = t;
            idx ^= 0x55555555u;
        } else if (a == b) {
            idx = 0u;
        }
    }
    lo = a | (b << 16);
    hi = idx;
}

// Two index bits per pixel: `in` for pixels of the set, 3 for the others.
__device__ __forceinline__ uint32_t spread_index(uint32_t inc, uint32_t in)
{
    uint32_t m2 = 0;
#pragma unroll
    for (int p = 0; p < 16; ++p) m2 |= ((inc >> p) & 1u ? 3u : 0u) << (2 * p);
    return ((in * 0x55555555u) & m2) | (0xFFFFFFFFu & ~m2);
}

// Colour of point 0 (the lowest pixel that opens a point).
__device__ __forceinline__ void point0_lane(const LaneSet& S, const LaneTables& T, float& qx, float& qy, float& qz)
{
    xyz(T, S.pix(S.first ? __builtin_ctz(S.first) : 0), qx, qy, qz);
}

// single.rs:100-164 (same steps as single_fit)
template <int FMT>
__device__ __forceinline__ void single_fit_lane(const LaneSet& S, const LaneTables& T, uint32_t& lo, uint32_t& hi)
{
    float qx, qy, qz;
    point0_lane(S, T, qx, qy, qz);
    const int cr = f32_to_i32_clamped(qx * 255.0f, 255);
    const int cg = f32_to_i32_clamped(qy * 255.0f, 255);
    const int cb = f32_to_i32_clamped(qz * 255.0f, 255);
    uint32_t best = 0xFFFFFFFFu;
    float s[3], e[3];
    uint32_t index = 0;
    if (FMT == BC1) {
        const uint32_t err = single_endpoints(0, 1, cr, cg, cb, s, e, index);
        if (err < best) {
            write_colour_lane(false, s[0], s[1], s[2], e[0], e[1], e[2], spread_index(S.inc, index), lo, hi);
            best = err;
        }
    }
    if (FMT != BC1 || !S.transparent) {
        const uint32_t err = single_endpoints(2, 3, cr, cg, cb, s, e, index);
        if (err < best)
            write_colour_lane(true, s[0], s[1], s[2], e[0], e[1], e[2], spread_index(S.inc, index), lo, hi);
    }
}

// range.rs:44-192 (same steps as range_fit): the index of every included
// pixel is that of its colour, which is its point's.
template <int FMT>
__device__ __forceinline__ void range_fit_lane(const LaneSet& S, const LaneTables& T, const FitParams& fp, float px,
                                               float py, float pz, uint32_t& lo, uint32_t& hi)
{
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
    if (S.n > 0) {
        float q0x, q0y, q0z;
        point0_lane(S, T, q0x, q0y, q0z);
        s0 = q0x; s1 = q0y; s2 = q0z;
        e0 = s0; e1 = s1; e2 = s2;
        float mn = q0x * px + q0y * py + q0z * pz;
        float mx = mn;
        const uint32_t rest = S.first & (S.first - 1u);  // points 1..n-1
#pragma unroll 2
        for (int p = 0; p < 16; ++p) {
            if ((rest >> p) & 1u) {
                float x, y, z;
                xyz(T, S.pix(p), x, y, z);
                const float d = x * px + y * py + z * pz;
                if (d < mn) {
                    s0 = x; s1 = y; s2 = z; mn = d;
                } else if (d > mx) {
                    e0 = x; e1 = y; e2 = z; mx = d;
                }
            }
        }
    }
    s0 = snap(s0, 31.0f, 1.0f / 31.0f); s1 = snap(s1, 63.0f, 1.0f / 63.0f); s2 = snap(s2, 31.0f, 1.0f / 31.0f);
    e0 = snap(e0, 31.0f, 1.0f / 31.0f); e1 = snap(e1, 63.0f, 1.0f / 63.0f); e2 = snap(e2, 31.0f, 1.0f / 31.0f);

    // range.rs:104-141 for the 3-colour pass (BC1) and the 4-colour pass
    // (unless BC1 with transparent pixels) in one loop over the pixels: the
    // two codebooks share codes 0 and 1 (the snapped start and end), so each
    // pixel's distances to them are computed once and serve both passes.
    // Each pass keeps its own first strict minimum per pixel (codes in
    // index order) and its own error folded in point order, exactly as two
    // separate passes would.
    const bool do3 = FMT == BC1;
    const bool do4 = !(FMT == BC1 && S.transparent);
    const float t2 = 2.0f / 3.0f, t1 = 1.0f / 3.0f;
    const float h0 = s0 * 0.5f + e0 * 0.5f, h1 = s1 * 0.5f + e1 * 0.5f, h2 = s2 * 0.5f + e2 * 0.5f;
    const float f0 = s0 * t2 + e0 * t1, f1 = s1 * t2 + e1 * t1, f2 = s2 * t2 + e2 * t1;
    const float g0 = s0 * t1 + e0 * t2, g1 = s1 * t1 + e1 * t2, g2 = s2 * t1 + e2 * t2;
    auto dist_to = [&](float x, float y, float z, float cx, float cy, float cz) {
        const float dx = fp.wr * (x - cx);
        const float dy = fp.wg * (y - cy);
        const float dz = fp.wb * (z - cz);
        return dx * dx + dy * dy + dz * dz;
    };
    float error3 = 0.0f, error4 = 0.0f;
    uint32_t idx3 = 0, idx4 = 0;
#pragma unroll 2
    for (int p = 0; p < 16; ++p) {
        float x, y, z;
        xyz(T, S.pix(p), x, y, z);
        const float d0 = dist_to(x, y, z, s0, s1, s2);
        const float d1 = dist_to(x, y, z, e0, e1, e2);
        // codes 0 and 1: the first strict minimum starts from FLT_MAX
        float m01 = FLT_MAX;
        uint32_t k01 = 0;
        if (d0 < m01) m01 = d0;
        if (d1 < m01) {
            m01 = d1;
            k01 = 1;
        }
        const bool in = (S.inc >> p) & 1u, pt = (S.first >> p) & 1u;
        if (do3) {
            float dist = m01;
            uint32_t k = k01;
            const float d2 = dist_to(x, y, z, h0, h1, h2);
            if (d2 < dist) {
                dist = d2;
                k = 2;
            }
            idx3 |= (in ? k : 3u) << (2 * p);
            if (pt) error3 += dist;
        }
        if (do4) {
            float dist = m01;
            uint32_t k = k01;
            const float d2 = dist_to(x, y, z, f0, f1, f2);
            if (d2 < dist) {
                dist = d2;
                k = 2;
            }
            const float d3 = dist_to(x, y, z, g0, g1, g2);
            if (d3 < dist) {
                dist = d3;
                k = 3;
            }
            idx4 |= (in ? k : 3u) << (2 * p);
            if (pt) error4 += dist;
        }
    }
    float best = FLT_MAX;
    if (do3 && error3 < best) {
        write_colour_lane(false, s0, s1, s2, e0, e1, e2, idx3, lo, hi);
        best = error3;
    }
    if (do4 && error4 < best) write_colour_lane(true, s0, s1, s2, e0, e1, e2, idx4, lo, hi);
}

// BC2 alpha of one lane's block (alpha.rs:27-51, same steps as bc2_alpha).
__device__ __forceinline__ void bc2_alpha_lane(const uint32_t (&px)[16], uint32_t valid, uint32_t& lo, uint32_t& hi)
{
    lo = hi = 0u;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const float af = (float)(px[p] >> 24) * (15.0f / 255.0f);
        uint32_t q = (uint32_t)f32_to_i32_clamped(af, 15);
        if (!((valid >> p) & 1u)) q = 0;
        if (p < 8) lo |= q << (4 * p);
        else hi |= q << (4 * (p - 8));
    }
}

// Front of one block on one lane: the alpha half in out[0..1] (BC2/BC3),
// the colour set in S, and for single-colour and range-fit blocks the
// finished colour half in out.  Returns the fit kind; for KIND_CLUSTER the
// principal axis comes back in (vx, vy, vz).
template <int FMT, int ALG>
__device__ __forceinline__ int encode_front_lane(const uint32_t (&px)[16], uint32_t valid, const FitParams& fp,
                                                 const LaneTables& T, LaneBlockLds& L, int lane, uint32_t (&out)[4],
                                                 LaneSet& S, float& vx, float& vy, float& vz,
                                                 float4* __restrict__ cluster_pts = nullptr)
{
    out[0] = out[1] = out[2] = out[3] = 0;
    colour_set_lane<FMT>(px, valid, fp, T, L, lane, S);
    if (FMT == BC2) bc2_alpha_lane(px, valid, out[0], out[1]);
    // BC3 alpha from the pixel copies in LDS, loops rolled (the unrolled
    // 16 x 8 codebook search would need many more registers)
    if (FMT == BC3) {
        auto alpha = [&](int p) { return (int)(S.pix(p) >> 24); };
        if (valid == 0xFFFFu) bc3_alpha_full<1>(alpha, out[0], out[1]);
        else bc3_alpha_core<1>(alpha, valid, out[0], out[1]);
    }
    const int kind = S.n == 1 ? KIND_SINGLE : (ALG == RANGE_FIT || S.n == 0) ? KIND_RANGE : KIND_CLUSTER;
    uint32_t lo = 0, hi = 0;
    vx = vy = vz = 0.0f;
    if (kind == KIND_SINGLE) single_fit_lane<FMT>(S, T, lo, hi);
    if (kind != KIND_SINGLE) principal_axis_lane(S, T, vx, vy, vz, kind == KIND_CLUSTER ? cluster_pts : nullptr);
    if (kind == KIND_RANGE) range_fit_lane<FMT>(S, T, fp, vx, vy, vz, lo, hi);
    set_colour<FMT>(out, lo, hi);
    return kind;
}

}  // namespace txd
