// tx_tables.hpp -- compile-time tables for the BC1-BC5 encoder.
//
// Both tables are computed by constexpr functions, so the device copy
// (__constant__) and the host copy are the same bytes by construction.
//
// 1. Single-colour lookup (replaces lib/src/colourfit/single_lut.rs:52-1086).
//    For a channel value t in 0..255 the table gives, per palette slot s in
//    {0, 2}, the quantised endpoints (start, end) and the distance `error`
//    between t and the closest value that slot can reproduce.  The tables are
//    regenerated from the BC1 endpoint expansion: every (start, end) pair is
//    expanded to 8 bits, each palette slot records the first pair that hits a
//    value exactly, and the remaining values are filled by repeated nearest
//    neighbour relaxation (t+1 before t-1), adding 1 to the error per step.
//
// 2. ClusterFit candidate lists (the split loops of
//    lib/src/colourfit/cluster.rs:182-236 and 309-380).  compress4 searches
//    the splits (i, j, k) with 0 <= i < n, i <= j <= n, max(j,1) <= k <= n;
//    compress3 the splits (i, j) with 0 <= i < n, max(i,1) <= j <= n.  A
//    split belongs to the search for n distinct colours iff
//    max(i+1, last index) <= n, so listing the n = 16 splits sorted by that
//    bound (stable in loop order) makes the list for every n a prefix of
//    length len[n].  Each entry packs the prefix-table offsets the split
//    reads and its rank in the reference's loop order, which breaks error
//    ties exactly like the reference's "first strict minimum wins":
//      bits  0..7   soff(i, j)  (compress3: part1 offset; compress4: part1)
//      bits  8..15  soff(j, k)  (compress4 only: part2 offset)
//      bits 16..19  i           (part0 = S[0][i] lives at offset i)
//      bits 20..29  loop-order rank
//    The prefix table S[a][b], a <= b <= 16, is stored compactly: every
//    empty range S[a][a] (all zero) shares offset 0, row 0 (S[0][b], b >= 1)
//    sits at offsets 1..16 so that part0 = S[0][i] is at offset i, and rows
//    a = 1..15 (b = a+1..16) follow: 137 entries in all.
#pragma once
#include <stdint.h>

namespace txt {

constexpr int soff(int a, int b)
{
    return a == b ? 0 : a == 0 ? b : 17 + 16 * (a - 1) - (((a - 1) * a) >> 1) + (b - a - 1);
}
constexpr int kSEntries = 137;  // soff(15, 16) + 1
constexpr int kCand4 = 967;       // compress4 splits for 16 colours
constexpr int kCand3 = 151;       // compress3 splits for 16 colours
// Stored lengths: the candidate loop reads up to 48 entries past its last
// 32-candidate trip, so the lists are padded (code 0 = prefix entries S[0],
// unit entries zero) and every read is in bounds without a guard.
constexpr int kCand4Stored = 1024;
constexpr int kCand3Stored = 192;

struct SingleLut {
    uint8_t v[4][256][6];  // [5_3, 6_3, 5_4, 6_4][value][start0,end0,err0,start1,end1,err1]
};

struct CandTables {
    uint32_t t4[kCand4Stored];
    uint32_t t3[kCand3Stored];
    uint16_t len4[17];
    uint16_t len3[17];
    uint8_t off_b[kSEntries + 3];  // b of the entry at each offset (0 for the shared zero)
};

constexpr void build_lut_one(int bits, int colours, uint8_t (&out)[256][6])
{
    int st[3][256] = {}, en[3][256] = {}, er[3][256] = {};
    bool ok[3][256] = {};
    // slots 0 and 2 of the palette are the ones the encoder uses; slot 1 is
    // tracked only to keep the iteration identical in shape
    const int levels = 1 << bits;
    for (int v1 = 0; v1 < levels; ++v1) {
        for (int v2 = 0; v2 < levels; ++v2) {
            const int a = (v1 << (8 - bits)) | (v1 >> (2 * bits - 8));
            const int b = (v2 << (8 - bits)) | (v2 >> (2 * bits - 8));
            const int c2 = colours == 3 ? (a + b) / 2 : (2 * a + b) / 3;
            const int codes[3] = {a, b, c2};
            for (int s = 0; s < 3; ++s) {
                const int t = codes[s];
                if (!ok[s][t]) {
                    ok[s][t] = true;
                    st[s][t] = v1;
                    en[s][t] = v2;
                    er[s][t] = 0;
                }
            }
        }
    }
    for (bool changed = true; changed;) {
        changed = false;
        for (int s = 0; s < 3; ++s) {
            for (int t = 0; t < 256; ++t) {
                for (int q = 0; q < 2; ++q) {
                    const int nt = q == 0 ? t + 1 : t - 1;
                    if (nt < 0 || nt > 255 || !ok[s][nt]) continue;
                    if (!ok[s][t] || er[s][nt] + 1 < er[s][t]) {
                        ok[s][t] = true;
                        st[s][t] = st[s][nt];
                        en[s][t] = en[s][nt];
                        er[s][t] = er[s][nt] + 1;
                        changed = true;
                    }
                }
            }
        }
    }
    for (int t = 0; t < 256; ++t) {
        out[t][0] = (uint8_t)st[0][t];
        out[t][1] = (uint8_t)en[0][t];
        out[t][2] = (uint8_t)er[0][t];
        out[t][3] = (uint8_t)st[2][t];
        out[t][4] = (uint8_t)en[2][t];
        out[t][5] = (uint8_t)er[2][t];
    }
}

constexpr SingleLut make_single_lut()
{
    SingleLut L{};
    build_lut_one(5, 3, L.v[0]);
    build_lut_one(6, 3, L.v[1]);
    build_lut_one(5, 4, L.v[2]);
    build_lut_one(6, 4, L.v[3]);
    return L;
}

constexpr CandTables make_cand_tables()
{
    CandTables T{};
    // compress4: enumerate in loop order, then bucket by the bound m
    {
        uint32_t lex[kCand4] = {};
        int bound[kCand4] = {};
        int c = 0;
        for (int i = 0; i < 16; ++i)
            for (int j = i; j <= 16; ++j)
                for (int k = (j > 1 ? j : 1); k <= 16; ++k) {
                    lex[c] = (uint32_t)soff(i, j) | ((uint32_t)soff(j, k) << 8) | ((uint32_t)i << 16) |
                             ((uint32_t)c << 20);
                    bound[c] = (i + 1 > k) ? i + 1 : k;
                    ++c;
                }
        int pos = 0;
        for (int m = 0; m <= 16; ++m) {
            for (int e = 0; e < kCand4; ++e)
                if (bound[e] == m) T.t4[pos++] = lex[e];
            T.len4[m] = (uint16_t)pos;
        }
    }
    {
        uint32_t lex[kCand3] = {};
        int bound[kCand3] = {};
        int c = 0;
        for (int i = 0; i < 16; ++i)
            for (int j = (i > 1 ? i : 1); j <= 16; ++j) {
                lex[c] = (uint32_t)soff(i, j) | ((uint32_t)i << 16) | ((uint32_t)c << 20);
                bound[c] = (i + 1 > j) ? i + 1 : j;
                ++c;
            }
        int pos = 0;
        for (int m = 0; m <= 16; ++m) {
            for (int e = 0; e < kCand3; ++e)
                if (bound[e] == m) T.t3[pos++] = lex[e];
            T.len3[m] = (uint16_t)pos;
        }
    }
    for (int a = 0; a <= 16; ++a)
        for (int b = a + 1; b <= 16; ++b) T.off_b[soff(a, b)] = (uint8_t)b;
    return T;
}

// j (or k) of a packed candidate from the offset of S[i][j] (S[j][k]): an
// empty range (offset 0) ends where it starts.
constexpr int split_end(const CandTables& T, uint32_t off, int start)
{
    return off == 0 ? start : (int)T.off_b[off];
}

// 3. Unit-weight split constants.  When a block has 16 distinct colours and
//    every point weight is exactly 1 (no duplicates, no alpha weighting),
//    the weight components of the split sums are exact integers fixed by
//    the split alone: part0.w = i, part1.w = j - i, part2.w = k - j,
//    xsum.w = 16.  alpha2_sum, beta2_sum, alphabeta_sum and the factor of
//    cluster.rs:193-202 / 325-335 are then the same numbers for every such
//    block.  They are evaluated here with the kernel's own expressions
//    (constant evaluation is IEEE binary32 with round-to-nearest-even and no
//    contraction, like the device code), one entry per candidate in the
//    order of CandTables, so the device reads bit-identical values instead
//    of recomputing them.
struct UnitSplit {
    float a2, b2, ab, factor;
};

struct UnitTables {
    UnitSplit u4[kCand4Stored];
    UnitSplit u3[kCand3Stored];
};

constexpr UnitSplit unit_split4(int i, int j, int k)
{
    const float p0w = (float)i, p1w = (float)(j - i), p2w = (float)(k - j), xsw = 16.0f;
    const float p3w = ((xsw - p2w) - p1w) - p0w;
    const float n19 = 1.0f / 9.0f, n49 = 4.0f / 9.0f;
    const float a2 = p2w * n19 + (p1w * n49 + p0w);
    const float b2 = p1w * n19 + (p2w * n49 + p3w);
    const float ab = (2.0f / 9.0f) * (p1w + p2w);
    return UnitSplit{a2, b2, ab, 1.0f / ((a2 * b2) - ab * ab)};
}

constexpr UnitSplit unit_split3(int i, int j)
{
    const float p0w = (float)i, p1w = (float)(j - i), xsw = 16.0f;
    const float p2w = (xsw - p1w) - p0w;
    const float hw = p1w * 0.25f;
    const float a2 = hw + p0w, b2 = hw + p2w, ab = hw;
    return UnitSplit{a2, b2, ab, 1.0f / ((a2 * b2) - ab * ab)};
}

// (i, j, k) of a packed candidate: i from bits 16..19, j and k from the
// prefix-table offsets soff(i, j) and soff(j, k).

constexpr UnitTables make_unit_tables()
{
    const CandTables C = make_cand_tables();
    UnitTables U{};
    for (int e = 0; e < kCand4; ++e) {
        const uint32_t code = C.t4[e];
        const int i = (int)((code >> 16) & 15u);
        const int j = split_end(C, code & 255u, i);
        const int k = split_end(C, (code >> 8) & 255u, j);
        U.u4[e] = unit_split4(i, j, k);
    }
    for (int e = 0; e < kCand3; ++e) {
        const uint32_t code = C.t3[e];
        const int i = (int)((code >> 16) & 15u);
        U.u3[e] = unit_split3(i, split_end(C, code & 255u, i));
    }
    return U;
}

static_assert(soff(0, 16) == 16 && soff(1, 2) == 17 && soff(15, 16) == kSEntries - 1, "compact prefix table");

}  // namespace txt
