// bcn_device.hpp -- per-block BC1-BC5 encoder for gfx950 (CDNA4).
//
// Execution model.  A wave64 encodes four 4x4 blocks at once: each block is
// owned by a 16-lane group (one DPP row), lane l of the group owns pixel l
// (l = 4*py + px, the reference's mask bit).  Per-pixel work (colour-set
// deduplication, alpha codebook fits, index assembly and packing) runs one
// pixel per lane; per-block sequential folds (centroid, covariance, power
// iteration) are computed redundantly by the 16 lanes from LDS, which costs
// the same issue slots as computing them on one lane.  The ClusterFit split
// search -- nearly all of the arithmetic -- is spread over the 16 lanes: the
// splits of a block are read from a candidate list and evaluated
// independently from a prefix table S[a][b] built once per ordering, then
// reduced with DPP row rotations to the reference's first strict minimum.
//
// Bit-exactness.  Every float operation is the reference's binary32 operation
// in the reference's order (lib/src/colourfit/cluster.rs, range.rs,
// single.rs, colourset.rs, math.rs).  The build disables FP contraction
// (-ffp-contract=off plus the pragma below) and packed FP32 math, and keeps
// HIP's correctly rounded division and square root.  The prefix table holds
// exactly the values of the reference's incremental part0/part1/part2
// accumulators: S[a][b] = ((0 + pw[a]) + pw[a+1]) + ... + pw[b-1].
//
// No MFMA: nothing here is a dense contraction.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "tx_tables.hpp"

#pragma clang fp contract(off)

namespace txd {

enum { BC1 = 0, BC2 = 1, BC3 = 2, BC4 = 3, BC5 = 4 };
enum { RANGE_FIT = 0, CLUSTER_FIT = 1, ITERATIVE_CLUSTER_FIT = 2 };
enum { KIND_SINGLE = 0, KIND_RANGE = 1, KIND_CLUSTER = 2 };

struct FitParams {
    float wr, wg, wb;      // Params::weights
    uint32_t weigh_alpha;  // Params::weigh_colour_by_alpha
};

// LDS of the front half of a block (colour set, principal axis, single and
// range fits): the pixels and the distinct colours, 320 bytes.  The front
// functions take any LDS type with `pts` and the put_pix / pix accessors.
struct __align__(16) FrontLds {
    float4 pts[16];       // distinct colours: x, y, z, sqrt(weight)
    uint32_t pixels[16];  // the block's RGBA8 pixels

    __device__ void put_pix(int l, uint32_t px) { pixels[l] = px; }
    __device__ uint32_t pix(int m) const { return pixels[m]; }
};

// Per-group LDS (one per 4x4 block in flight), 2512 bytes, so that 16
// single-wave workgroups (64 groups) fit in a CU's 160 KiB.  The first 80
// bytes of the prefix table are also scratch while they hold nothing live:
// the block's pixels at the start of a block, then, while an ordering is
// built, the projection keys (bytes 0..63) and the ordering as bytes
// (64..79).  Every table entry there is rewritten before it is read again.
struct __align__(16) GroupLds {
    float4 S[txt::kSEntries];  // prefix sums of the weighted ordered points (txt::soff layout)
    float4 pts[16];            // distinct colours: x, y, z, sqrt(weight)
    uint2 ord[8];              // orderings searched so far, 16 four-bit indices each

    __device__ float* keys() { return reinterpret_cast<float*>(S); }
    __device__ const float* keys() const { return reinterpret_cast<const float*>(S); }
    __device__ uint8_t* order_bytes() { return reinterpret_cast<uint8_t*>(S + 4); }
    // pixels kept as float bit patterns (loads and stores only) in the key area
    __device__ void put_pix(int l, uint32_t px) { keys()[l] = __uint_as_float(px); }
    __device__ uint32_t pix(int m) const { return __float_as_uint(keys()[m]); }
};
static_assert(sizeof(GroupLds) == 2512, "group LDS size");

// Device copies of the constexpr tables (one per translation unit: the code
// objects are built without relocatable device code).
static __constant__ txt::SingleLut c_lut = txt::make_single_lut();
static __constant__ txt::CandTables c_cand = txt::make_cand_tables();
static __constant__ txt::UnitTables c_unit = txt::make_unit_tables();

// Word 3 of a raw (untyped) buffer resource on CDNA: 32-bit data format.
constexpr int kBufferFlags = 0x00020000;

// ---------------------------------------------------------------- helpers

template <int N>
__device__ __forceinline__ uint32_t ror16(uint32_t v)
{
    // DPP row_ror:N -- rotate within each row of 16 lanes
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x120 + N, 0xF, 0xF, false);
}
template <int N>
__device__ __forceinline__ float ror16f(float v)
{
    return __builtin_bit_cast(float, ror16<N>(__builtin_bit_cast(uint32_t, v)));
}

__device__ __forceinline__ uint32_t grp_or(uint32_t v)
{
    v |= ror16<8>(v);
    v |= ror16<4>(v);
    v |= ror16<2>(v);
    v |= ror16<1>(v);
    return v;
}
__device__ __forceinline__ uint32_t grp_add(uint32_t v)
{
    v += ror16<8>(v);
    v += ror16<4>(v);
    v += ror16<2>(v);
    v += ror16<1>(v);
    return v;
}
__device__ __forceinline__ uint32_t grp_min(uint32_t v)
{
    v = min(v, ror16<8>(v));
    v = min(v, ror16<4>(v));
    v = min(v, ror16<2>(v));
    v = min(v, ror16<1>(v));
    return v;
}
__device__ __forceinline__ uint32_t grp_max(uint32_t v)
{
    v = max(v, ror16<8>(v));
    v = max(v, ror16<4>(v));
    v = max(v, ror16<2>(v));
    v = max(v, ror16<1>(v));
    return v;
}

// (error, code) argmin over the group; ties go to the smaller code, whose
// high bits are the split's loop-order rank.
template <int N>
__device__ __forceinline__ void argmin_step(float& e, uint32_t& c)
{
    const float e2 = ror16f<N>(e);
    const uint32_t c2 = ror16<N>(c);
    const bool take = (e2 < e) || (e2 == e && c2 < c);
    e = take ? e2 : e;
    c = take ? c2 : c;
}
__device__ __forceinline__ void grp_argmin(float& e, uint32_t& c)
{
    argmin_step<8>(e, c);
    argmin_step<4>(e, c);
    argmin_step<2>(e, c);
    argmin_step<1>(e, c);
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0; }

// math.rs:100-102: roundf(a).max(0).min(limit) as i32, NaN -> 0
__device__ __forceinline__ int f32_to_i32_clamped(float a, int limit)
{
    return (int)fminf(fmaxf(roundf(a), 0.0f), (float)limit);
}

// colourblock.rs:27-34
__device__ __forceinline__ uint32_t pack_565(float x, float y, float z)
{
    const uint32_t r = (uint32_t)f32_to_i32_clamped(31.0f * x, 31);
    const uint32_t g = (uint32_t)f32_to_i32_clamped(63.0f * y, 63);
    const uint32_t b = (uint32_t)f32_to_i32_clamped(31.0f * z, 31);
    return (r << 11) | (g << 5) | b;
}

// colourblock.rs:58-94.  `idx` is this lane's pixel index; the group's
// packed block comes back in (lo, hi) on every lane of the group.
__device__ __forceinline__ void write_colour(bool four, float sx, float sy, float sz, float ex, float ey, float ez,
                                             uint32_t idx, int l, uint32_t& lo, uint32_t& hi)
{
    uint32_t a = pack_565(sx, sy, sz), b = pack_565(ex, ey, ez);
    if (!four) {
        if (a > b) {  // write3: a <= b, swap palette slots 0 and 1
            const uint32_t t = a; a = b; b = t;
            idx = idx == 0 ? 1u : (idx == 1 ? 0u : idx);
        }
    } else {
        if (a < b) {  // write4: a >= b, index ^ 1
            const uint32_t t = a; a = b; b = t;
            idx = (idx ^ 1u) & 3u;
        } else if (a == b) {
            idx = 0;
        }
    }
    lo = a | (b << 16);
    hi = grp_or((idx & 3u) << (2 * l));
}

// ------------------------------------------------------------ alpha paths

// alpha.rs:27-51
__device__ __forceinline__ void bc2_alpha(uint32_t px, bool valid, int l, uint32_t& lo, uint32_t& hi)
{
    const float af = (float)(px >> 24) * (15.0f / 255.0f);
    uint32_t q = (uint32_t)f32_to_i32_clamped(af, 15);
    if (!valid) q = 0;
    lo = grp_or(l < 8 ? q << (4 * l) : 0u);
    hi = grp_or(l >= 8 ? q << (4 * (l - 8)) : 0u);
}

// alpha.rs:70-77
__device__ __forceinline__ void fix_range(int& mn, int& mx, int steps)
{
    if (mx - mn < steps) mx = min(mn + steps, 255);
    if (mx - mn < steps) mn = max(mx - steps, 0);
}

// alpha.rs:79-118 for one pixel: the first strict minimum of (v - code)^2
// over the 8 codes, and that squared distance.  The square is monotone in
// |v - code|, so the same code wins the smallest key 8|v - code| + j: the
// distance orders the keys and, at equal distance, the lower index.  Values
// and codes are below 256, so v_sad_u16 forms each key from 8v and 8*code in
// one instruction.
__device__ __forceinline__ uint32_t fit8(int v, const int (&codes)[8], uint32_t& least)
{
    const uint32_t v8 = (uint32_t)v << 3;
    uint32_t key = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < 8; ++j) key = min(key, __builtin_amdgcn_sad_u16(v8, (uint32_t)codes[j] << 3, (uint32_t)j));
    const uint32_t d = (key >> 3) & 255u;  // |v - code| <= 255
    least = d * d;
    return key & 7u;
}

// alpha.rs:187-256 (BC3 alpha, BC4, BC5 channels)
__device__ __forceinline__ void bc3_alpha(int v, bool valid, int l, uint32_t& lo, uint32_t& hi)
{
    const uint32_t uv = (uint32_t)v;
    int min7 = (int)grp_min(valid ? uv : 255u);
    int max7 = (int)grp_max(valid ? uv : 0u);
    int min5 = (int)grp_min(valid && v != 0 ? uv : 255u);
    int max5 = (int)grp_max(valid && v != 255 ? uv : 0u);
    if (min5 > max5) min5 = max5;
    if (min7 > max7) min7 = max7;
    fix_range(min5, max5, 5);
    fix_range(min7, max7, 7);
    int codes5[8], codes7[8];
    codes5[0] = min5;
    codes5[1] = max5;
#pragma unroll
    for (int i = 1; i < 5; ++i) codes5[1 + i] = ((5 - i) * min5 + i * max5) / 5;
    codes5[6] = 0;
    codes5[7] = 255;
    // alpha.rs:238-239: slots 0 and 1 of the 7-point codebook hold min5/max5
    codes7[0] = min5;
    codes7[1] = max5;
#pragma unroll
    for (int i = 1; i < 7; ++i) codes7[1 + i] = ((7 - i) * min7 + i * max7) / 7;
    uint32_t l5, l7;
    uint32_t i5 = fit8(v, codes5, l5);
    uint32_t i7 = fit8(v, codes7, l7);
    if (!valid) {
        i5 = 0; i7 = 0; l5 = 0; l7 = 0;
    }
    const uint32_t err5 = grp_add(l5), err7 = grp_add(l7);
    int a0, a1;
    uint32_t idx;
    if (err5 <= err7) {  // write_alpha_block5, alpha.rs:146-165
        if (min5 > max5) {
            a0 = max5; a1 = min5;
            idx = i5 == 0 ? 1u : i5 == 1 ? 0u : (i5 >= 2 && i5 <= 5) ? 7u - i5 : i5;
        } else {
            a0 = min5; a1 = max5; idx = i5;
        }
    } else {  // write_alpha_block7, alpha.rs:167-185
        if (min7 < max7) {
            a0 = max7; a1 = min7;
            idx = i7 == 0 ? 1u : i7 == 1 ? 0u : 9u - i7;
        } else {
            a0 = min7; a1 = max7; idx = i7;
        }
    }
    // alpha.rs:120-144: two 24-bit little-endian groups of eight 3-bit indices
    const uint32_t A = grp_or(l < 8 ? idx << (3 * l) : 0u);
    const uint32_t B = grp_or(l >= 8 ? idx << (3 * (l - 8)) : 0u);
    lo = (uint32_t)a0 | ((uint32_t)a1 << 8) | ((A & 0xFFFFu) << 16);
    hi = (A >> 16) | (B << 8);
}

// ----------------------------------------------------------- colour fits

struct ColourSet {
    int n;            // ColourSet::count
    int remap;        // this pixel's point index, -1 when excluded
    bool transparent; // BC1 block with alpha < 128 somewhere
};

// lib/src/math.rs:44-97: weighted covariance of the n points in LDS, then
// eight power-iteration steps.  Folds run in point order on every lane.
template <class LDS>
__device__ __forceinline__ void principal_axis(const LDS& L, int n, float& vx, float& vy, float& vz)
{
    float total = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        if (p < n) {
            const float4 q = L.pts[p];
            total = total + q.w;
            cx = cx + q.x * q.w;
            cy = cy + q.y * q.w;
            cz = cz + q.z * q.w;
        }
    }
    if (total > FLT_EPSILON) {
        cx = cx / total;
        cy = cy / total;
        cz = cz / total;
    }
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f, c4 = 0.0f, c5 = 0.0f;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        if (p < n) {
            const float4 q = L.pts[p];
            const float ax = q.x - cx, ay = q.y - cy, az = q.z - cz;
            const float bx = ax * q.w, by = ay * q.w, bz = az * q.w;
            c0 = c0 + ax * bx;
            c1 = c1 + ax * by;
            c2 = c2 + ax * bz;
            c3 = c3 + ay * by;
            c4 = c4 + ay * bz;
            c5 = c5 + az * bz;
        }
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

// single.rs:58-98 for one codebook; returns the winning error
__device__ __forceinline__ uint32_t single_endpoints(int tr, int tg, int cr, int cg, int cb, float (&s)[3],
                                                     float (&e)[3], uint32_t& index)
{
    uint32_t best = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint8_t* r = c_lut.v[tr][cr] + 3 * k;
        const uint8_t* g = c_lut.v[tg][cg] + 3 * k;
        const uint8_t* b = c_lut.v[tr][cb] + 3 * k;
        const uint32_t er = r[2], eg = g[2], eb = b[2];
        const uint32_t err = er * er + eg * eg + eb * eb;
        if (err < best) {
            s[0] = (float)r[0] / 31.0f; s[1] = (float)g[0] / 63.0f; s[2] = (float)b[0] / 31.0f;
            e[0] = (float)r[1] / 31.0f; e[1] = (float)g[1] / 63.0f; e[2] = (float)b[1] / 31.0f;
            index = (uint32_t)(2 * k);
            best = err;
        }
    }
    return best;
}

// single.rs:100-164
template <int FMT, class LDS>
__device__ __forceinline__ void single_fit(const LDS& L, const ColourSet& cs, int l, uint32_t& lo, uint32_t& hi)
{
    const float4 q = L.pts[0];
    const int cr = f32_to_i32_clamped(q.x * 255.0f, 255);
    const int cg = f32_to_i32_clamped(q.y * 255.0f, 255);
    const int cb = f32_to_i32_clamped(q.z * 255.0f, 255);
    uint32_t best = 0xFFFFFFFFu;
    float s[3], e[3];
    uint32_t index = 0;
    if (FMT == BC1) {
        const uint32_t err = single_endpoints(0, 1, cr, cg, cb, s, e, index);
        if (err < best) {
            write_colour(false, s[0], s[1], s[2], e[0], e[1], e[2], cs.remap < 0 ? 3u : index, l, lo, hi);
            best = err;
        }
    }
    if (FMT != BC1 || !cs.transparent) {
        const uint32_t err = single_endpoints(2, 3, cr, cg, cb, s, e, index);
        if (err < best) write_colour(true, s[0], s[1], s[2], e[0], e[1], e[2], cs.remap < 0 ? 3u : index, l, lo, hi);
    }
}

// range.rs: clamp to [0,1] then snap to the 565 grid
__device__ __forceinline__ float snap(float v, float grid, float rcp)
{
    v = fminf(1.0f, fmaxf(0.0f, v));
    return truncf(grid * v + 0.5f) * rcp;
}

// range.rs:44-192
template <int FMT, class LDS>
__device__ __forceinline__ void range_fit(const LDS& L, const ColourSet& cs, int l, const FitParams& fp,
                                          float px, float py, float pz, uint32_t& lo, uint32_t& hi)
{
    const int n = cs.n;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
    if (n > 0) {
        const float4 q0 = L.pts[0];
        s0 = q0.x; s1 = q0.y; s2 = q0.z;
        e0 = s0; e1 = s1; e2 = s2;
        float mn = q0.x * px + q0.y * py + q0.z * pz;
        float mx = mn;
        for (int p = 1; p < n; ++p) {
            const float4 q = L.pts[p];
            const float d = q.x * px + q.y * py + q.z * pz;
            if (d < mn) {
                s0 = q.x; s1 = q.y; s2 = q.z; mn = d;
            } else if (d > mx) {
                e0 = q.x; e1 = q.y; e2 = q.z; mx = d;
            }
        }
    }
    s0 = snap(s0, 31.0f, 1.0f / 31.0f); s1 = snap(s1, 63.0f, 1.0f / 63.0f); s2 = snap(s2, 31.0f, 1.0f / 31.0f);
    e0 = snap(e0, 31.0f, 1.0f / 31.0f); e1 = snap(e1, 63.0f, 1.0f / 63.0f); e2 = snap(e2, 31.0f, 1.0f / 31.0f);

    // range.rs:104-141: closest code per point, error folded in point order,
    // for the 3-colour pass (BC1) and the 4-colour pass (unless BC1 with
    // transparent pixels) in one loop: both codebooks start with the snapped
    // start and end, so a point's distances to those two serve both passes.
    const bool do3 = FMT == BC1;
    const bool do4 = !(FMT == BC1 && cs.transparent);
    const float t2 = 2.0f / 3.0f, t1 = 1.0f / 3.0f;
    const float h0 = s0 * 0.5f + e0 * 0.5f, h1 = s1 * 0.5f + e1 * 0.5f, h2 = s2 * 0.5f + e2 * 0.5f;
    const float f0 = s0 * t2 + e0 * t1, f1 = s1 * t2 + e1 * t1, f2 = s2 * t2 + e2 * t1;
    const float g0 = s0 * t1 + e0 * t2, g1 = s1 * t1 + e1 * t2, g2 = s2 * t1 + e2 * t2;
    auto dist_to = [&](const float4& q, float cx, float cy, float cz) {
        const float dx = fp.wr * (q.x - cx);
        const float dy = fp.wg * (q.y - cy);
        const float dz = fp.wb * (q.z - cz);
        return dx * dx + dy * dy + dz * dz;
    };
    float error3 = 0.0f, error4 = 0.0f;
    uint32_t mine3 = 0, mine4 = 0;
    for (int p = 0; p < n; ++p) {
        const float4 q = L.pts[p];
        const float d0 = dist_to(q, s0, s1, s2);
        const float d1 = dist_to(q, e0, e1, e2);
        float m01 = FLT_MAX;  // first strict minimum of codes 0 and 1
        uint32_t k01 = 0;
        if (d0 < m01) m01 = d0;
        if (d1 < m01) {
            m01 = d1;
            k01 = 1;
        }
        if (do3) {
            float dist = m01;
            uint32_t k = k01;
            const float d2 = dist_to(q, h0, h1, h2);
            if (d2 < dist) {
                dist = d2;
                k = 2;
            }
            if (p == cs.remap) mine3 = k;
            error3 += dist;
        }
        if (do4) {
            float dist = m01;
            uint32_t k = k01;
            const float d2 = dist_to(q, f0, f1, f2);
            if (d2 < dist) {
                dist = d2;
                k = 2;
            }
            const float d3 = dist_to(q, g0, g1, g2);
            if (d3 < dist) {
                dist = d3;
                k = 3;
            }
            if (p == cs.remap) mine4 = k;
            error4 += dist;
        }
    }
    float best = FLT_MAX;
    if (do3 && error3 < best) {
        write_colour(false, s0, s1, s2, e0, e1, e2, cs.remap < 0 ? 3u : mine3, l, lo, hi);
        best = error3;
    }
    if (do4 && error4 < best) write_colour(true, s0, s1, s2, e0, e1, e2, cs.remap < 0 ? 3u : mine4, l, lo, hi);
}

// The three prefix sums a split reads: part0 = S[0][i], part1, part2.
struct SplitSums {
    float4 p0, p1, p2;
};

template <int NC>
__device__ __forceinline__ SplitSums load_split(const GroupLds& L, uint32_t code)
{
    SplitSums s;
    s.p0 = L.S[(code >> 16) & 15u];
    s.p1 = L.S[code & 255u];
    if (NC == 4) s.p2 = L.S[(code >> 8) & 255u];
    else s.p2 = s.p1;
    return s;
}

// One split of the ClusterFit search (cluster.rs:189-230 for compress3,
// 320-364 for compress4).  Returns the weighted error; the quantised
// endpoints come back through a and b.  With UNIT the weight terms
// (alpha2_sum, beta2_sum, alphabeta_sum, factor) come from `u`, the
// unit-weight table entry of this candidate, which holds exactly the values
// the expressions below produce when every point weight is 1.
template <int NC, bool UNIT = false>
__device__ __forceinline__ float eval_split(const SplitSums& s, const float4& xs, const FitParams& fp, float (&a)[3],
                                            float (&b)[3], const txt::UnitSplit& u = txt::UnitSplit{})
{
    const float4 p0 = s.p0;
    const float4 p1 = s.p1;
    float ax[3], bx[3], a2, b2, ab;
    if (NC == 4) {
        const float4 p2 = s.p2;
        // part3 = xsum - part2 - part1 - part0
        const float p3x = ((xs.x - p2.x) - p1.x) - p0.x;
        const float p3y = ((xs.y - p2.y) - p1.y) - p0.y;
        const float p3z = ((xs.z - p2.z) - p1.z) - p0.z;
        // alphax = part2*(1/3,1/3,1/3,1/9) + (part1*(2/3,2/3,2/3,4/9) + part0)
        // betax  = part1*(1/3,..,1/9) + (part2*(2/3,..,4/9) + part3)
        // The binary32 constants satisfy 2/3 = 2 * (1/3) and 4/9 = 4 * (1/9)
        // exactly (same significand), and scaling by a power of two commutes
        // with rounding for the normal values summed here, so
        // p * (2/3) + q == fma(2, p * (1/3), q) bit for bit: one product per
        // part serves both alphax and betax.
        const float t13 = 1.0f / 3.0f, n19 = 1.0f / 9.0f;
        static_assert(2.0f * (1.0f / 3.0f) == 2.0f / 3.0f && 4.0f * (1.0f / 9.0f) == 4.0f / 9.0f,
                      "2/3 and 4/9 are exact multiples of 1/3 and 1/9");
        const float m1x = p1.x * t13, m1y = p1.y * t13, m1z = p1.z * t13;
        const float m2x = p2.x * t13, m2y = p2.y * t13, m2z = p2.z * t13;
        ax[0] = m2x + __builtin_fmaf(2.0f, m1x, p0.x);
        ax[1] = m2y + __builtin_fmaf(2.0f, m1y, p0.y);
        ax[2] = m2z + __builtin_fmaf(2.0f, m1z, p0.z);
        bx[0] = m1x + __builtin_fmaf(2.0f, m2x, p3x);
        bx[1] = m1y + __builtin_fmaf(2.0f, m2y, p3y);
        bx[2] = m1z + __builtin_fmaf(2.0f, m2z, p3z);
        if (UNIT) {
            a2 = u.a2;
            b2 = u.b2;
            ab = u.ab;
        } else {
            const float p3w = ((xs.w - p2.w) - p1.w) - p0.w;
            const float w1 = p1.w * n19, w2 = p2.w * n19;
            a2 = w2 + __builtin_fmaf(4.0f, w1, p0.w);
            b2 = w1 + __builtin_fmaf(4.0f, w2, p3w);
            ab = (2.0f / 9.0f) * (p1.w + p2.w);
        }
    } else {
        // part2 = xsum - part1 - part0; alphax = part1*(.5,.5,.5) + part0,
        // betax = part1*(.5,.5,.5) + part2.  Halving is exact for the normal
        // prefix sums, so the fused form rounds once exactly as the
        // product-then-add does.
        const float p2x = (xs.x - p1.x) - p0.x;
        const float p2y = (xs.y - p1.y) - p0.y;
        const float p2z = (xs.z - p1.z) - p0.z;
        ax[0] = __builtin_fmaf(0.5f, p1.x, p0.x);
        ax[1] = __builtin_fmaf(0.5f, p1.y, p0.y);
        ax[2] = __builtin_fmaf(0.5f, p1.z, p0.z);
        bx[0] = __builtin_fmaf(0.5f, p1.x, p2x);
        bx[1] = __builtin_fmaf(0.5f, p1.y, p2y);
        bx[2] = __builtin_fmaf(0.5f, p1.z, p2z);
        if (UNIT) {
            a2 = u.a2;
            b2 = u.b2;
            ab = u.ab;
        } else {
            const float p2w = (xs.w - p1.w) - p0.w;
            const float hw = p1.w * 0.25f;
            a2 = hw + p0.w;
            b2 = hw + p2w;
            ab = hw;
        }
    }
    const float factor = UNIT ? u.factor : 1.0f / ((a2 * b2) - ab * ab);
    const float grid[3] = {31.0f, 63.0f, 31.0f};
    const float rcp[3] = {1.0f / 31.0f, 1.0f / 63.0f, 1.0f / 31.0f};
    const float wts[3] = {fp.wr, fp.wg, fp.wb};
    float e5[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float av = ((ax[c] * b2) - bx[c] * ab) * factor;
        float bv = ((bx[c] * a2) - ax[c] * ab) * factor;
        av = snap(av, grid[c], rcp[c]);
        bv = snap(bv, grid[c], rcp[c]);
        a[c] = av;
        b[c] = bv;
        const float e1 = (av * av) * a2 + ((bv * bv) * b2);
        const float e2 = ((av * bv) * ab) - av * ax[c];
        const float e3 = e2 - bv * bx[c];
        // 2 * e3 is exact (a power-of-two scaling of a small finite value),
        // so the fused form rounds once, exactly like the product-then-add
        const float e4 = __builtin_fmaf(2.0f, e3, e1);
        e5[c] = e4 * wts[c];
    }
    return (e5[0] + e5[1]) + e5[2];
}

// LDS hand-off between the lanes of one group.  Groups never share LDS and a
// wave's LDS operations execute in order, so a wavefront-scope fence (which
// stops the compiler from reordering LDS accesses across it) is all that is
// needed; no workgroup barrier, so waves of one workgroup stay independent
// (the persistent batch kernel gives them different chunks).
__device__ __forceinline__ void group_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// construct_ordering (cluster.rs:78-136) for the ordering slot `it` of every
// group with `active` set, then the weighted points and the prefix table of
// that ordering.  Executed by the whole wave.  Returns false when the new
// ordering repeats one of slots 0..it-1 (the search then stops, and the
// prefix table is left untouched).
__device__ __forceinline__ bool ordering_step(GroupLds& L, int l, int n, bool active, int it, float axx, float axy,
                                              float axz)
{
    float key = FLT_MAX;
    uint32_t idx = 0;
    if (l < n) {
        const float4 q = L.pts[l];
        key = (q.x * axx + q.y * axy) + q.z * axz;
        idx = (uint32_t)l;
    }
    if (!__builtin_isfinite(key)) key = __builtin_huge_valf();  // non-finite keys tie, after all finite
    float* keys = L.keys();
    if (active) keys[l] = key;
    group_sync();
    // stable rank: entries with a smaller key, plus equal keys at lower index
    int rank = 0;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const float km = keys[m];
        rank += (km < key) ? 1 : 0;
        rank += (m < l && km == key) ? 1 : 0;
    }
    uint8_t* bytes = L.order_bytes();
    if (active) bytes[rank] = (uint8_t)idx;
    group_sync();
    // the ordering as 16 four-bit indices (position q in bits 4q..4q+3)
    bool unique = active;
    uint2 cur = make_uint2(0u, 0u);
    if (active) {
        const uint4 b = *reinterpret_cast<const uint4*>(bytes);
        auto nib = [](uint32_t w) {
            return (w & 0xFu) | ((w >> 4) & 0xF0u) | ((w >> 8) & 0xF00u) | ((w >> 12) & 0xF000u);
        };
        cur = make_uint2(nib(b.x) | (nib(b.y) << 16), nib(b.z) | (nib(b.w) << 16));
        for (int p = 0; p < it; ++p) {
            const uint2 o = L.ord[p];
            if (o.x == cur.x && o.y == cur.y) unique = false;
        }
        if (l == 0) L.ord[it] = cur;
    }
    // S[a][b] = ((0 + pw[a]) + pw[a+1]) + ... + pw[b-1] with the weighted
    // points pw in order.  S[a][a+1] = 0 + pw[a] goes in first (it equals
    // pw[a]: every component is a product of non-negative finite values), then
    // each lane extends its row a = l from the S[m][m+1] of the others.
    if (unique && l < n) {
        const uint32_t o = ((l < 8 ? cur.x : cur.y) >> (4 * (l & 7))) & 15u;
        const float4 q = L.pts[o];
        L.S[txt::soff(l, l + 1)] =
            make_float4(0.0f + q.x * q.w, 0.0f + q.y * q.w, 0.0f + q.z * q.w, 0.0f + 1.0f * q.w);
    }
    if (unique && l == 0) L.S[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // the shared empty range
    group_sync();
    if (unique && l < n) {
        float4 s = L.S[txt::soff(l, l + 1)];
        // the diagonal entries this row adds, five at a time ahead of their
        // dependent adds (one LDS round trip per five instead of per entry;
        // the loop writes only off-diagonal entries, and reads past the set
        // are clamped into the table and never used)
#pragma unroll
        for (int t0 = 0; t0 < 15; t0 += 5) {
            float4 w[5];
#pragma unroll
            for (int u = 0; u < 5; ++u) {
                const int bb = min(l + 2 + t0 + u, 16);
                w[u] = L.S[txt::soff(bb - 1, bb)];
            }
#pragma unroll
            for (int u = 0; u < 5; ++u) {
                const int bb = l + 2 + t0 + u;
                if (bb <= n) {
                    s.x = s.x + w[u].x; s.y = s.y + w[u].y; s.z = s.z + w[u].z; s.w = s.w + w[u].w;
                    L.S[txt::soff(l, bb)] = s;
                }
            }
        }
    }
    group_sync();
    return unique;
}

// Largest value of v over the four 16-lane groups of the wave (v uniform
// within each group), as a scalar.
__device__ __forceinline__ uint32_t wave_max_of_groups(uint32_t v)
{
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    return max(max(a, b), max(c, d));
}

// Candidate loop of one split search (see search_step).
template <int NC, bool UNIT>
__device__ __forceinline__ void search_loop(const GroupLds& L, int l, uint32_t mylen, const float4& xs,
                                            const FitParams& fp, float& be, uint32_t& bc)
{
    // Trip q covers list entries 32q .. 32q + 31: lane l evaluates entry
    // 32q + l (A) and 32q + 16 + l (B), each while 32q is below its limit.
    // Loop control and the trip's offsets are scalar: the code list is read
    // through a buffer resource with the lane's offset in a vector register
    // and the trip's in a scalar one (hardware bounds checks: a read past
    // the end returns 0), the 16-byte unit entries from a scalar base plus
    // the lane's offset.  Codes are read two half-trips ahead and prefix
    // sums one half-trip ahead.  The kept minimum is the smallest (error,
    // code), codes ordered by loop-order rank, so evaluation order does not
    // matter; the three comparisons are combined without branches.  The
    // lists are padded past the longest trip: reads need no guard, and
    // entries at or past `mylen` are loaded but never evaluated.
    constexpr int kStored = NC == 4 ? txt::kCand4Stored : txt::kCand3Stored;
    const __amdgpu_buffer_rsrc_t rt =
        __builtin_amdgcn_make_buffer_rsrc((void*)(NC == 4 ? c_cand.t4 : c_cand.t3), 0, kStored * 4, kBufferFlags);
    const char* ubase = reinterpret_cast<const char*>(NC == 4 ? c_unit.u4 : c_unit.u3);
    const uint32_t vo = 4u * (uint32_t)l, uo = 16u * (uint32_t)l;
    auto code_at = [&](uint32_t soff) { return __builtin_amdgcn_raw_buffer_load_b32(rt, vo, soff, 0); };
    auto unit_at = [&](const char* base) { return *reinterpret_cast<const txt::UnitSplit*>(base + uo); };
    const uint32_t limA = mylen > (uint32_t)l ? mylen - (uint32_t)l : 0u;
    const uint32_t limB = mylen > (uint32_t)l + 16u ? mylen - (uint32_t)l - 16u : 0u;
    const uint32_t wlen = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_max_of_groups(mylen));
    be = __builtin_huge_valf();
    bc = 0xFFFFFFFFu;
    auto keep = [&](float err, uint32_t c) {
        const bool take = (err < be) | ((err == be) & (c < bc));
        be = take ? err : be;
        bc = take ? c : bc;
    };
    uint32_t cA = code_at(0);
    uint32_t cB = code_at(64);
    SplitSums sA = load_split<NC>(L, cA);
    txt::UnitSplit uA{};
    if (UNIT) uA = unit_at(ubase);
    for (uint32_t base = 0; base < wlen; base += 32) {
        const uint32_t so = 4u * base;
        const char* ub = ubase + 16u * base;
        const uint32_t cA2 = code_at(so + 128);
        const SplitSums sB = load_split<NC>(L, cB);
        txt::UnitSplit uB{};
        if (UNIT) uB = unit_at(ub + 256);
        if (base < limA) {
            float a[3], b[3];
            keep(eval_split<NC, UNIT>(sA, xs, fp, a, b, uA), cA);
        }
        const uint32_t cB2 = code_at(so + 192);
        sA = load_split<NC>(L, cA2);
        if (UNIT) uA = unit_at(ub + 512);
        if (base < limB) {
            float a[3], b[3];
            keep(eval_split<NC, UNIT>(sB, xs, fp, a, b, uB), cB);
        }
        cA = cA2;
        cB = cB2;
    }
}

// The split search over the current prefix table: every group with `run`
// set evaluates its `len` candidates (16 lanes, list entries two rounds and
// prefix sums one round ahead) and returns the first strict minimum
// (error, code) on all of its lanes.  Executed by the whole wave.  When
// every searching group holds 16 points of weight exactly 1, the weight
// terms come from the unit-weight table (same values, fewer operations).
template <int NC>
__device__ __forceinline__ void search_step(const GroupLds& L, int l, int n, bool run, const FitParams& fp,
                                            float& be, uint32_t& bc)
{
    const uint32_t mylen = run ? (NC == 4 ? c_cand.len4[n] : c_cand.len3[n]) : 0u;
    const float4 xs = L.S[n];  // xsum_wsum = S[0][n]
    const bool unit = n == 16 && grp_min(L.pts[l].w == 1.0f ? 1u : 0u) == 1u;
    if (!wave_any(run && !unit))
        search_loop<NC, true>(L, l, mylen, xs, fp, be, bc);
    else
        search_loop<NC, false>(L, l, mylen, xs, fp, be, bc);
    grp_argmin(be, bc);
}

// Quantised endpoints of split `code` on the current prefix table.
template <int NC>
__device__ __forceinline__ void split_endpoints(const GroupLds& L, int n, uint32_t code, const FitParams& fp,
                                                float (&sa)[3], float (&sb)[3])
{
    eval_split<NC>(load_split<NC>(L, code), L.S[n], fp, sa, sb);
}

// Index of this lane's pixel for the winning split (cluster.rs:252-273 /
// 391-416 followed by remap_indices): the three overwriting loops leave a
// point with the value of the range holding its largest position q >= i.
template <int NC>
__device__ __forceinline__ uint32_t split_index(const GroupLds& L, int n, int remap, uint32_t code, int it)
{
    if (remap < 0) return 3u;
    const int i = (int)((code >> 16) & 15u);
    const uint32_t o1 = code & 255u, o2 = (code >> 8) & 255u;
    const int j = o1 ? (int)c_cand.off_b[o1] : i;
    const int k = NC == 4 ? (o2 ? (int)c_cand.off_b[o2] : j) : 0;
    const uint2 ord = L.ord[it];
    int qmax = -1;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int pq = (int)(((q < 8 ? ord.x : ord.y) >> (4 * (q & 7))) & 15u);
        if (q >= i && q < n && pq == remap) qmax = q;
    }
    if (qmax < 0) return 0u;
    if (qmax < j) return 2u;
    if (NC == 4 && qmax < k) return 3u;
    return 1u;
}

// cluster.rs:152-274 / 276-417 with all groups of the wave in lockstep.
// Executed by the whole wave; `act` marks the groups running ClusterFit.
// `best_err`, `lo`, `hi` carry ClusterFit::best_error / best_compressed
// across compress3 -> compress4.  `table_ready` (single-ordering searches
// only): the group's LDS already holds this axis's ordering and prefix
// table -- compress3 of ClusterFit built them, and its search, endpoints and
// index assembly only read them -- so compress4 does not build them again.
template <int NC, int ITERS>
__device__ __forceinline__ void run_cluster(GroupLds& L, const ColourSet& cs, int l, bool act, float px, float py,
                                            float pz, const FitParams& fp, float& best_err, uint32_t& lo,
                                            uint32_t& hi, bool table_ready = false)
{
    static_assert(ITERS >= 1, "at least one ordering");
    const int n = cs.n;
    const float thr = best_err;
    float lb_err = thr;
    uint32_t lb_code = 0;
    int lb_it = 0;
    float sa[3] = {0.0f, 0.0f, 0.0f}, sb[3] = {0.0f, 0.0f, 0.0f};  // best_start / best_end
    float axx = px, axy = py, axz = pz;
    bool done = !act;

    for (int it = 0; it < ITERS; ++it) {
        if (!wave_any(!done)) break;
        const bool build = !done && !(ITERS == 1 && table_ready);
        if (wave_any(build)) {
            // an ordering built at it = 0 is always unique; only later ones can repeat
            const bool unique = ordering_step(L, l, n, build, it, axx, axy, axz);
            if (build && !unique) done = true;
        }
        float be;
        uint32_t bc;
        search_step<NC>(L, l, n, !done, fp, be, bc);
        // iteration control, cluster.rs:222-249
        if (!done) {
            if (be < lb_err) {
                lb_err = be;
                lb_code = bc;
                lb_it = it;
                split_endpoints<NC>(L, n, bc, fp, sa, sb);
            }
            if (lb_it != it) {
                done = true;
            } else {
                axx = sb[0] - sa[0];
                axy = sb[1] - sa[1];
                axz = sb[2] - sa[2];
            }
        }
    }

    if (act && lb_err < thr) {
        uint32_t nlo, nhi;
        write_colour(NC == 4, sa[0], sa[1], sa[2], sb[0], sb[1], sb[2], split_index<NC>(L, n, cs.remap, lb_code, lb_it),
                     l, nlo, nhi);
        lo = nlo;
        hi = nhi;
        best_err = lb_err;
    }
}

// alpha.rs:187-256 for a whole block on one lane: the same arithmetic as
// bc3_alpha, with v(p) the value of pixel p and `valid` the mask bits.  U is
// the unroll factor of the pixel loops (16 for values in registers; a rolled
// loop over values read from memory keeps register use down).
template <int U, class Src>
__device__ __forceinline__ void bc3_alpha_core(const Src& v, uint32_t valid, uint32_t& lo, uint32_t& hi)
{
    int min7 = 255, max7 = 0, min5 = 255, max5 = 0;
#pragma unroll U
    for (int p = 0; p < 16; ++p) {
        if ((valid >> p) & 1u) {
            const int vp = v(p);
            min7 = min(min7, vp);
            max7 = max(max7, vp);
            if (vp != 0) min5 = min(min5, vp);
            if (vp != 255) max5 = max(max5, vp);
        }
    }
    if (min5 > max5) min5 = max5;
    if (min7 > max7) min7 = max7;
    fix_range(min5, max5, 5);
    fix_range(min7, max7, 7);
    int codes5[8], codes7[8];
    codes5[0] = min5;
    codes5[1] = max5;
#pragma unroll
    for (int i = 1; i < 5; ++i) codes5[1 + i] = ((5 - i) * min5 + i * max5) / 5;
    codes5[6] = 0;
    codes5[7] = 255;
    codes7[0] = min5;
    codes7[1] = max5;
#pragma unroll
    for (int i = 1; i < 7; ++i) codes7[1 + i] = ((7 - i) * min7 + i * max7) / 7;
    // 3-bit indices of both codebooks, pixel p at bits 3p; masked pixels 0
    uint64_t i5 = 0, i7 = 0;
    uint32_t err5 = 0, err7 = 0;
#pragma unroll U
    for (int p = 0; p < 16; ++p) {
        if ((valid >> p) & 1u) {
            const int vp = v(p);
            uint32_t l5, l7;
            i5 |= (uint64_t)fit8(vp, codes5, l5) << (3 * p);
            i7 |= (uint64_t)fit8(vp, codes7, l7) << (3 * p);
            err5 += l5;
            err7 += l7;
        }
    }
    int a0, a1;
    uint64_t idx = 0;
    if (err5 <= err7) {  // write_alpha_block5, alpha.rs:146-165
        if (min5 > max5) {
            a0 = max5; a1 = min5;
#pragma unroll U
            for (int p = 0; p < 16; ++p) {
                const uint32_t i = (uint32_t)(i5 >> (3 * p)) & 7u;
                const uint32_t o = i == 0 ? 1u : i == 1 ? 0u : (i >= 2 && i <= 5) ? 7u - i : i;
                idx |= (uint64_t)o << (3 * p);
            }
        } else {
            a0 = min5; a1 = max5; idx = i5;
        }
    } else {  // write_alpha_block7, alpha.rs:167-185
        if (min7 < max7) {
            a0 = max7; a1 = min7;
#pragma unroll U
            for (int p = 0; p < 16; ++p) {
                const uint32_t i = (uint32_t)(i7 >> (3 * p)) & 7u;
                const uint32_t o = i == 0 ? 1u : i == 1 ? 0u : 9u - i;
                idx |= (uint64_t)o << (3 * p);
            }
        } else {
            a0 = min7; a1 = max7; idx = i7;
        }
    }
    // alpha.rs:120-144: the 48 index bits follow the two endpoint bytes
    lo = (uint32_t)a0 | ((uint32_t)a1 << 8) | ((uint32_t)(idx & 0xFFFFu) << 16);
    hi = (uint32_t)(idx >> 16);
}

// alpha.rs:187-256 for a block whose 16 pixels are all valid (every block
// of an image but those of its last partial row and column), branch-free:
//  - min5 is the minimum over the non-zero values and max5 the maximum over
//    the values below 255 (255 and 0 when there are none, alpha.rs:192-218):
//    (v + 255) & 255 maps 0 to 255 and v > 0 to v - 1, so min5 is one more
//    than the minimum of those keys (capped at 255); (v + 1) & 255 maps 255
//    to 0 and v < 255 to v + 1, so max5 is one less than their maximum
//    (floored at 0).
//  - after fix_range, max5 >= min5 + 5 and max7 >= min7 + 7, so
//    write_alpha_block5 never swaps and write_alpha_block7 always does
//    (alpha.rs:146-185): the 5-point block is (min5, max5, indices) and the
//    7-point block (max7, min7, indices with 0 <-> 1 and x -> 9 - x).
//  - indices are kept 8 to a word (3 bits each, the layout of
//    write_alpha_block's two 24-bit groups, alpha.rs:120-144), and the
//    7-point remap runs on all 8 fields of a word at once.
// Same bytes as bc3_alpha_core with valid = 0xFFFF.
__device__ __forceinline__ uint32_t remap7_fields(uint32_t x)
{
    // per 3-bit field: x < 2 -> x ^ 1, else 9 - x = (x ^ 7) + 2 (no carry out
    // of the field, since x ^ 7 <= 5 there)
    constexpr uint32_t kOnes = 0x249249u;  // bit 0 of each of the 8 fields
    const uint32_t ge2 = ((x >> 1) | (x >> 2)) & kOnes;
    const uint32_t hi_mask = ge2 * 7u;
    return (((x ^ 0xFFFFFFu) + 2u * ge2) & hi_mask) | ((x ^ kOnes) & ~hi_mask & 0xFFFFFFu);
}

template <int U, class Src>
__device__ __forceinline__ void bc3_alpha_full(const Src& v, uint32_t& lo, uint32_t& hi)
{
    uint32_t mn7 = 255u, mx7 = 0u, mn5k = 255u, mx5k = 0u;
#pragma unroll U
    for (int p = 0; p < 16; ++p) {
        const uint32_t x = (uint32_t)v(p);
        mn7 = min(mn7, x);
        mx7 = max(mx7, x);
        mn5k = min(mn5k, (x + 255u) & 255u);
        mx5k = max(mx5k, (x + 1u) & 255u);
    }
    int min5 = (int)min(mn5k + 1u, 255u);
    int max5 = max((int)mx5k - 1, 0);
    int min7 = (int)mn7, max7 = (int)mx7;
    if (min5 > max5) min5 = max5;
    fix_range(min5, max5, 5);
    fix_range(min7, max7, 7);
    int codes5[8], codes7[8];
    codes5[0] = min5;
    codes5[1] = max5;
#pragma unroll
    for (int i = 1; i < 5; ++i) codes5[1 + i] = ((5 - i) * min5 + i * max5) / 5;
    codes5[6] = 0;
    codes5[7] = 255;
    codes7[0] = min5;
    codes7[1] = max5;
#pragma unroll
    for (int i = 1; i < 7; ++i) codes7[1 + i] = ((7 - i) * min7 + i * max7) / 7;
    // per pixel and book: the key 8|v - code| + j of the winning code (fit8's
    // keys; slots 0-1 are min5/max5 in both books, so their keys are shared),
    // its distance squared into the error, its slot into the index word
    uint32_t c5[8], c7[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        c5[j] = (uint32_t)codes5[j] << 3;
        c7[j] = (uint32_t)codes7[j] << 3;
    }
    uint32_t i5[2] = {0u, 0u}, i7[2] = {0u, 0u};
    uint32_t err5 = 0, err7 = 0;
#pragma unroll U
    for (int p = 0; p < 16; ++p) {
        const uint32_t v8 = (uint32_t)v(p) << 3;
        const uint32_t k01 = min(__builtin_amdgcn_sad_u16(v8, c5[0], 0u), __builtin_amdgcn_sad_u16(v8, c5[1], 1u));
        uint32_t k5 = k01, k7 = k01;
#pragma unroll
        for (int j = 2; j < 8; ++j) {
            k5 = min(k5, __builtin_amdgcn_sad_u16(v8, c5[j], (uint32_t)j));
            k7 = min(k7, __builtin_amdgcn_sad_u16(v8, c7[j], (uint32_t)j));
        }
        const uint32_t d5 = __builtin_amdgcn_ubfe(k5, 3, 8), d7 = __builtin_amdgcn_ubfe(k7, 3, 8);
        err5 = __umul24(d5, d5) + err5;
        err7 = __umul24(d7, d7) + err7;
        i5[p >> 3] |= (k5 & 7u) << (3 * (p & 7));
        i7[p >> 3] |= (k7 & 7u) << (3 * (p & 7));
    }
    uint32_t a0, a1, g0, g1;
    if (err5 <= err7) {
        a0 = (uint32_t)min5; a1 = (uint32_t)max5; g0 = i5[0]; g1 = i5[1];
    } else {
        a0 = (uint32_t)max7; a1 = (uint32_t)min7; g0 = remap7_fields(i7[0]); g1 = remap7_fields(i7[1]);
    }
    // bytes: a0, a1, then the two 24-bit index groups little-endian
    lo = a0 | (a1 << 8) | (g0 << 16);
    hi = (g0 >> 16) | (g1 << 8);
}

// The BC4 / BC5 channel form: the 16 values in registers.
__device__ __forceinline__ void bc3_alpha_lane(const int (&v)[16], uint32_t valid, uint32_t& lo, uint32_t& hi)
{
    if (valid == 0xFFFFu) bc3_alpha_full<16>([&](int p) { return v[p]; }, lo, hi);
    else bc3_alpha_core<16>([&](int p) { return v[p]; }, valid, lo, hi);
}

// One block per lane: the 16 pixels of block (bx, by) as 16 dword loads
// (pixels outside the image are masked out, lib.rs:309-330).
__device__ __forceinline__ uint32_t fetch_block_lane(const uint8_t* __restrict__ rgba, uint32_t width, uint32_t height,
                                                     uint32_t bx, uint32_t by, uint32_t (&px)[16])
{
    uint32_t valid = 0;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const uint32_t sx = bx * 4u + (uint32_t)(p & 3), sy = by * 4u + (uint32_t)(p >> 2);
        const bool ok = sx < width && sy < height;
        px[p] = ok ? *reinterpret_cast<const uint32_t*>(rgba + 4 * ((size_t)sy * width + sx)) : 0u;
        valid |= (ok ? 1u : 0u) << p;
    }
    return valid;
}

// BC4 (channel 0) or BC5 (channels 0 and 1) block from one lane's pixels.
template <int FMT>
__device__ __forceinline__ void encode_alpha_lane(const uint32_t (&px)[16], uint32_t valid, uint32_t (&out)[4])
{
    int v[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) v[p] = (int)(px[p] & 255u);
    bc3_alpha_lane(v, valid, out[0], out[1]);
    out[2] = out[3] = 0u;
    if (FMT == BC5) {
#pragma unroll
        for (int p = 0; p < 16; ++p) v[p] = (int)((px[p] >> 8) & 255u);
        bc3_alpha_lane(v, valid, out[2], out[3]);
    }
}

// lib.rs:309-330: lane l of the group fetches pixel l of block (bx, by);
// pixels outside the image are masked out.
__device__ __forceinline__ uint32_t fetch_pixel(const uint8_t* __restrict__ rgba, uint32_t width, uint32_t height,
                                                uint32_t bx, uint32_t by, bool live, int l, bool& valid)
{
    const uint32_t sx = bx * 4u + (uint32_t)(l & 3), sy = by * 4u + (uint32_t)(l >> 2);
    valid = live && sx < width && sy < height;
    return valid ? *reinterpret_cast<const uint32_t*>(rgba + 4 * ((size_t)sy * width + sx)) : 0u;
}

// BC4 / BC5 blocks: alpha codebooks only (lib.rs:200-204), no LDS.
template <int FMT>
__device__ __forceinline__ void encode_alpha_only(uint32_t px, bool valid, int l, uint32_t (&out)[4])
{
    out[0] = out[1] = out[2] = out[3] = 0;
    bc3_alpha((int)(px & 255u), valid, l, out[0], out[1]);
    if (FMT == BC5) bc3_alpha((int)((px >> 8) & 255u), valid, l, out[2], out[3]);
}

// The colour half of a BC1/BC2/BC3 block: words 0-1 for BC1, 2-3 otherwise.
template <int FMT>
__device__ __forceinline__ void set_colour(uint32_t (&out)[4], uint32_t lo, uint32_t hi)
{
    if (FMT == BC1) {
        out[0] = lo;
        out[1] = hi;
    } else {
        out[2] = lo;
        out[3] = hi;
    }
}

// Front half of compress_block_masked (lib.rs:188-230) for BC1/BC2/BC3,
// one 4x4 block per 16-lane group (lane l holds pixel l, `px` RGBA8
// little-endian, and its mask bit `valid`): alpha, colour set, fit choice,
// and the finished colour half for the single-colour and range fits.
// Returns the fit kind; for KIND_CLUSTER the principal axis comes back in
// (vx, vy, vz) and the colour half is left to the caller.  Every lane of the
// wave must call this (LDS hand-offs inside).
template <int FMT, int ALG, class LDS>
__device__ __forceinline__ int encode_front(uint32_t px, bool valid, const FitParams& fp, LDS& L, int g, int l,
                                            uint32_t (&out)[4], ColourSet& cs, float& vx, float& vy, float& vz)
{
    const int gshift = 16 * g;
    out[0] = out[1] = out[2] = out[3] = 0;
    if (FMT == BC2) bc2_alpha(px, valid, l, out[0], out[1]);
    if (FMT == BC3) bc3_alpha((int)(px >> 24), valid, l, out[0], out[1]);

    // ---- ColourSet::new, colourset.rs:35-112
    const uint32_t rgb = px & 0xFFFFFFu, alpha = px >> 24;
    const bool inc = valid && !(FMT == BC1 && alpha < 128u);
    const uint32_t incmask = (uint32_t)(__ballot(inc) >> gshift) & 0xFFFFu;
    cs.transparent = FMT == BC1 && ((uint32_t)(__ballot(valid && alpha < 128u) >> gshift) & 0xFFFFu) != 0;
    L.put_pix(l, px);
    group_sync();
    uint32_t same = 0, asum = 0;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const uint32_t pm = L.pix(m);
        const bool s = ((incmask >> m) & 1u) && ((pm & 0xFFFFFFu) == rgb);
        same |= (s ? 1u : 0u) << m;
        asum += s ? (pm >> 24) + 1u : 0u;
    }
    const uint32_t lower = (1u << l) - 1u;
    const uint32_t dup = same & lower;
    const bool is_new = inc && dup == 0;
    const uint32_t newmask = (uint32_t)(__ballot(is_new) >> gshift) & 0xFFFFu;
    cs.n = __popc(newmask);
    cs.remap = -1;
    if (inc) cs.remap = is_new ? __popc(newmask & lower) : __popc(newmask & ((1u << (__ffs(dup) - 1)) - 1u));
    if (is_new) {
        // weights sum exactly (multiples of 1/256 below 16), then sqrtf
        const float w = fp.weigh_alpha ? (float)asum / 256.0f : (float)__popc(same);
        L.pts[cs.remap] = make_float4((float)(rgb & 255u) / 255.0f, (float)((rgb >> 8) & 255u) / 255.0f,
                                      (float)((rgb >> 16) & 255u) / 255.0f, sqrtf(w));
    }
    group_sync();

    // ---- fit selection, lib.rs:217-230
    const int kind = cs.n == 1 ? KIND_SINGLE
                     : (ALG == RANGE_FIT || cs.n == 0) ? KIND_RANGE : KIND_CLUSTER;
    uint32_t lo = 0, hi = 0;
    if (kind == KIND_SINGLE) single_fit<FMT>(L, cs, l, lo, hi);
    vx = 0.0f;
    vy = 0.0f;
    vz = 0.0f;
    if (kind != KIND_SINGLE) principal_axis(L, cs.n, vx, vy, vz);
    if (kind == KIND_RANGE) range_fit<FMT>(L, cs, l, fp, vx, vy, vz, lo, hi);
    set_colour<FMT>(out, lo, hi);
    return kind;
}

// Encodes one 4x4 block per 16-lane group (lib.rs:188-234): lane l holds
// pixel l (`px`, RGBA8 little-endian) and its mask bit (`valid`).  Every lane
// of the wave must call this (LDS hand-offs inside); groups without a block
// pass valid = false and discard the result.  Returns the block words in
// out[0..3] (8-byte formats use out[0..1]).  The four groups of a wave run
// their ClusterFit searches in lockstep.
template <int FMT, int ALG>
__device__ __forceinline__ void encode_block(uint32_t px, bool valid, const FitParams& fp, GroupLds& L, int g, int l,
                                             uint32_t (&out)[4])
{
    if (FMT == BC4 || FMT == BC5) {
        encode_alpha_only<FMT>(px, valid, l, out);
        return;
    }
    ColourSet cs;
    float vx, vy, vz;
    const int kind = encode_front<FMT, ALG>(px, valid, fp, L, g, l, out, cs, vx, vy, vz);
    if (ALG != RANGE_FIT) {
        // wave-uniform region: run_cluster contains LDS barriers
        constexpr int ITERS = ALG == ITERATIVE_CLUSTER_FIT ? 8 : 1;
        const bool act = kind == KIND_CLUSTER;
        if (wave_any(act)) {
            float best_err = FLT_MAX;
            uint32_t clo = 0, chi = 0;
            if (FMT == BC1) run_cluster<3, ITERS>(L, cs, l, act, vx, vy, vz, fp, best_err, clo, chi);
            run_cluster<4, ITERS>(L, cs, l, act && !(FMT == BC1 && cs.transparent), vx, vy, vz, fp, best_err, clo,
                                  chi, FMT == BC1 && act);
            if (act) set_colour<FMT>(out, clo, chi);
        }
    }
}

}  // namespace txd
