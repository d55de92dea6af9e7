// PNG image I/O of the command-line tool, host code only.
//
// Reading follows what texpresso's CLI gets from the png crate 0.17 with
// Transformations::EXPAND | STRIP_16 followed by its per-colour-type
// conversion to RGBA8 (cli/src/image/png.rs:59-99): palette images are
// expanded through PLTE (and tRNS to alpha), grey of depth 1/2/4 is scaled
// to 8 bits (x255, x85, x17), a tRNS colour key becomes alpha 0 on matching
// pixels (compared at the image's own depth) and 255 elsewhere, 16-bit
// samples keep their high byte, and grey / grey-alpha / RGB become RGBA with
// the grey value replicated and alpha 255 where absent.  Writing produces an
// RGBA8, non-interlaced PNG (cli/src/image/png.rs:101-111); its pixels are
// what counts, the compressed stream need not match the png crate's bytes.
//
// The decoder is written from the PNG specification (ISO/IEC 15948):
// chunk layout with CRC-32, zlib-compressed IDAT, the five per-row filters
// and Adam7 interlacing.
#include "texpresso_hip.h"

#include <zlib.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr uint8_t kSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};
// most pixels a decode may produce: bounds allocations on hostile input
constexpr uint64_t kMaxPixels = 1ull << 32;

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put32(std::vector<uint8_t>& v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

struct Image {
    uint32_t width = 0, height = 0;
    uint8_t depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte, trns, idat;
    bool have_trns = false;
};

int channels(uint8_t ctype)
{
    switch (ctype) {
    case 0: return 1;
    case 2: return 3;
    case 3: return 1;
    case 4: return 2;
    case 6: return 4;
    default: return 0;
    }
}

bool depth_ok(uint8_t ctype, uint8_t depth)
{
    switch (ctype) {
    case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
    case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
    case 2:
    case 4:
    case 6: return depth == 8 || depth == 16;
    default: return false;
    }
}

// Chunks up to IEND: header fields, palette, transparency and the
// concatenated IDAT payload.  Every chunk's CRC is checked.
int parse(const uint8_t* d, size_t n, Image& im)
{
    if (!d || n < 8 || memcmp(d, kSignature, 8) != 0) return TX_ERR_FORMAT;
    size_t off = 8;
    bool have_ihdr = false, have_iend = false, have_idat = false;
    while (off + 12 <= n) {
        const uint32_t len = be32(d + off);
        if (len > 0x7FFFFFFFu || (uint64_t)len + 12 > n - off) return TX_ERR_INPUT_SHORT;
        const uint8_t* type = d + off + 4;
        const uint8_t* body = d + off + 8;
        const uint32_t crc = be32(body + len);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc) return TX_ERR_FORMAT;
        off += (size_t)len + 12;
        if (!memcmp(type, "IHDR", 4)) {
            if (have_ihdr || len != 13) return TX_ERR_FORMAT;
            im.width = be32(body);
            im.height = be32(body + 4);
            im.depth = body[8];
            im.ctype = body[9];
            im.interlace = body[12];
            if (im.width == 0 || im.height == 0 || im.width > 0x7FFFFFFFu || im.height > 0x7FFFFFFFu)
                return TX_ERR_FORMAT;
            if (!depth_ok(im.ctype, im.depth) || body[10] != 0 || body[11] != 0 || im.interlace > 1)
                return TX_ERR_FORMAT;
            have_ihdr = true;
            continue;
        }
        if (!have_ihdr) return TX_ERR_FORMAT;  // IHDR must come first
        if (!memcmp(type, "PLTE", 4)) {
            if (len % 3 != 0 || len == 0 || len > 768) return TX_ERR_FORMAT;
            im.plte.assign(body, body + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            if (im.ctype == 0 && len != 2) return TX_ERR_FORMAT;
            if (im.ctype == 2 && len != 6) return TX_ERR_FORMAT;
            if (im.ctype == 3 && len > 256) return TX_ERR_FORMAT;
            if (im.ctype == 4 || im.ctype == 6) return TX_ERR_FORMAT;
            im.trns.assign(body, body + len);
            im.have_trns = true;
        } else if (!memcmp(type, "IDAT", 4)) {
            im.idat.insert(im.idat.end(), body, body + len);
            have_idat = true;
        } else if (!memcmp(type, "IEND", 4)) {
            have_iend = true;
            break;
        } else if (!(type[0] & 0x20)) {
            return TX_ERR_FORMAT;  // unknown critical chunk
        }
    }
    if (!have_ihdr || !have_idat || !have_iend) return TX_ERR_INPUT_SHORT;
    if (im.ctype == 3 && im.plte.empty()) return TX_ERR_FORMAT;
    if ((uint64_t)im.width * im.height > kMaxPixels) return TX_ERR_INVALID;
    return TX_OK;
}

struct Pass {
    uint32_t x0, y0, dx, dy;
};
constexpr Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                            {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};

uint32_t pass_extent(uint32_t size, uint32_t start, uint32_t step) { return size > start ? (size - start + step - 1) / step : 0; }

uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = p > a ? p - a : a - p;
    const int pb = p > b ? p - b : b - p;
    const int pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

// Undo one row's filter in place; `prev` is the previous unfiltered row of
// the same pass (all zero for the first).
bool unfilter(uint8_t type, uint8_t* row, const uint8_t* prev, size_t len, size_t bpp)
{
    switch (type) {
    case 0: return true;
    case 1:
        for (size_t i = bpp; i < len; ++i) row[i] = (uint8_t)(row[i] + row[i - bpp]);
        return true;
    case 2:
        for (size_t i = 0; i < len; ++i) row[i] = (uint8_t)(row[i] + prev[i]);
        return true;
    case 3:
        for (size_t i = 0; i < len; ++i) {
            const int left = i >= bpp ? row[i - bpp] : 0;
            row[i] = (uint8_t)(row[i] + ((left + prev[i]) >> 1));
        }
        return true;
    case 4:
        for (size_t i = 0; i < len; ++i) {
            const int left = i >= bpp ? row[i - bpp] : 0;
            const int upleft = i >= bpp ? prev[i - bpp] : 0;
            row[i] = (uint8_t)(row[i] + paeth(left, prev[i], upleft));
        }
        return true;
    default: return false;
    }
}

// Sample c of pixel x in an unfiltered row, at the image's own depth.
uint32_t sample(const uint8_t* row, uint32_t x, int c, int nch, int depth)
{
    const uint64_t idx = (uint64_t)x * nch + c;
    if (depth == 8) return row[idx];
    if (depth == 16) return (uint32_t)row[2 * idx] << 8 | row[2 * idx + 1];
    const uint64_t bit = idx * depth;
    const int shift = 8 - depth - (int)(bit & 7);
    return (row[bit >> 3] >> shift) & ((1u << depth) - 1u);
}

// 8-bit value of a sample: low depths scaled to the full range, 16-bit
// reduced to its high byte.
uint8_t to8(uint32_t s, int depth)
{
    switch (depth) {
    case 1: return (uint8_t)(s * 255u);
    case 2: return (uint8_t)(s * 85u);
    case 4: return (uint8_t)(s * 17u);
    case 16: return (uint8_t)(s >> 8);
    default: return (uint8_t)s;
    }
}

void put_pixel(const Image& im, const uint8_t* row, uint32_t x, uint8_t* px)
{
    const int nch = channels(im.ctype), depth = im.depth;
    switch (im.ctype) {
    case 0: {
        const uint32_t g = sample(row, x, 0, 1, depth);
        const uint8_t v = to8(g, depth);
        px[0] = px[1] = px[2] = v;
        px[3] = (im.have_trns && g == ((uint32_t)im.trns[0] << 8 | im.trns[1])) ? 0 : 255;
        break;
    }
    case 2: {
        uint32_t s[3];
        bool key = im.have_trns;
        for (int c = 0; c < 3; ++c) {
            s[c] = sample(row, x, c, 3, depth);
            px[c] = to8(s[c], depth);
            if (key && s[c] != ((uint32_t)im.trns[2 * c] << 8 | im.trns[2 * c + 1])) key = false;
        }
        px[3] = key ? 0 : 255;
        break;
    }
    case 3: {
        const uint32_t i = sample(row, x, 0, 1, depth);
        // an index past the palette reads as opaque black
        if (3 * i + 2 < im.plte.size()) {
            px[0] = im.plte[3 * i];
            px[1] = im.plte[3 * i + 1];
            px[2] = im.plte[3 * i + 2];
        } else {
            px[0] = px[1] = px[2] = 0;
        }
        px[3] = i < im.trns.size() ? im.trns[i] : 255;
        break;
    }
    case 4: {
        const uint8_t v = to8(sample(row, x, 0, nch, depth), depth);
        px[0] = px[1] = px[2] = v;
        px[3] = to8(sample(row, x, 1, nch, depth), depth);
        break;
    }
    default:
        for (int c = 0; c < 4; ++c) px[c] = to8(sample(row, x, c, nch, depth), depth);
        break;
    }
}

int decode(const Image& im, uint8_t* rgba)
{
    const int nch = channels(im.ctype);
    const uint64_t bits_pp = (uint64_t)nch * im.depth;
    const size_t bpp = bits_pp >= 8 ? (size_t)(bits_pp / 8) : 1;
    const int npass = im.interlace ? 7 : 1;
    uint64_t raw_len = 0;
    for (int p = 0; p < npass; ++p) {
        const Pass ps = im.interlace ? kAdam7[p] : Pass{0, 0, 1, 1};
        const uint64_t pw = pass_extent(im.width, ps.x0, ps.dx), ph = pass_extent(im.height, ps.y0, ps.dy);
        if (pw && ph) raw_len += ph * (1 + (pw * bits_pp + 7) / 8);
    }
    if (raw_len > (1ull << 36)) return TX_ERR_INVALID;
    std::vector<uint8_t> raw((size_t)raw_len);
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return TX_ERR_NOMEM;
    size_t in_off = 0, out_off = 0;
    int zr = Z_OK;
    while (out_off < raw.size() && zr == Z_OK) {
        const size_t in_chunk = std::min<size_t>(im.idat.size() - in_off, 1u << 30);
        const size_t out_chunk = std::min<size_t>(raw.size() - out_off, 1u << 30);
        zs.next_in = const_cast<uint8_t*>(im.idat.data() + in_off);
        zs.avail_in = (uInt)in_chunk;
        zs.next_out = raw.data() + out_off;
        zs.avail_out = (uInt)out_chunk;
        zr = inflate(&zs, Z_NO_FLUSH);
        in_off += in_chunk - zs.avail_in;
        out_off += out_chunk - zs.avail_out;
        if (zr == Z_BUF_ERROR && zs.avail_in == 0 && in_off == im.idat.size()) break;
        if (zr == Z_BUF_ERROR) zr = Z_OK;
    }
    inflateEnd(&zs);
    if (zr != Z_OK && zr != Z_STREAM_END) return TX_ERR_FORMAT;
    if (out_off < raw.size()) return TX_ERR_INPUT_SHORT;
    size_t pos = 0;
    for (int p = 0; p < npass; ++p) {
        const Pass ps = im.interlace ? kAdam7[p] : Pass{0, 0, 1, 1};
        const uint32_t pw = pass_extent(im.width, ps.x0, ps.dx), ph = pass_extent(im.height, ps.y0, ps.dy);
        if (!pw || !ph) continue;
        const size_t len = (size_t)(((uint64_t)pw * bits_pp + 7) / 8);
        std::vector<uint8_t> zero(len, 0);
        const uint8_t* prev = zero.data();
        for (uint32_t y = 0; y < ph; ++y) {
            uint8_t* row = raw.data() + pos + 1;
            if (!unfilter(raw[pos], row, prev, len, bpp)) return TX_ERR_FORMAT;
            const size_t oy = (size_t)ps.y0 + (size_t)y * ps.dy;
            for (uint32_t x = 0; x < pw; ++x) {
                const size_t ox = (size_t)ps.x0 + (size_t)x * ps.dx;
                put_pixel(im, row, x, rgba + 4 * (oy * im.width + ox));
            }
            prev = row;
            pos += len + 1;
        }
    }
    return TX_OK;
}

void chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* body, size_t len)
{
    put32(out, (uint32_t)len);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    if (len) out.insert(out.end(), body, body + len);
    put32(out, (uint32_t)crc32(crc32(0L, Z_NULL, 0), out.data() + start, (uInt)(len + 4)));
}

}  // namespace

extern "C" {

int tx_png_info(const uint8_t* data, size_t len, size_t* width, size_t* height)
{
    Image im;
    const int rc = parse(data, len, im);
    if (rc != TX_OK) return rc;
    if (width) *width = im.width;
    if (height) *height = im.height;
    return TX_OK;
}

int tx_png_decode(const uint8_t* data, size_t len, uint8_t* rgba, size_t rgba_len)
{
    Image im;
    const int rc = parse(data, len, im);
    if (rc != TX_OK) return rc;
    if (rgba_len < 4ull * im.width * im.height) return TX_ERR_OUTPUT_SHORT;
    if (!rgba) return TX_ERR_INVALID;
    return decode(im, rgba);
}

size_t tx_png_encode_bound(size_t width, size_t height)
{
    if (width == 0 || height == 0 || width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return 0;
    const uint64_t raw = (uint64_t)height * (1 + 4ull * width);
    if (raw > (1ull << 36)) return 0;
    // deflate bound, then signature, IHDR, IDAT framing (one chunk per
    // 2^31 - 1 bytes) and IEND
    const uint64_t z = compressBound((uLong)raw);
    return (size_t)(z + 8 + 25 + 12 * (z / 0x7FFFFFFFull + 1) + 12);
}

int tx_png_encode(const uint8_t* rgba, size_t width, size_t height, uint8_t* out, size_t out_cap, size_t* out_len)
{
    if (width == 0 || height == 0 || width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return TX_ERR_INVALID;
    if (!rgba || !out || !out_len) return TX_ERR_INVALID;
    const size_t stride = 4 * width;
    const uint64_t raw_len = (uint64_t)height * (1 + stride);
    if (raw_len > (1ull << 36)) return TX_ERR_INVALID;
    // every row with filter 0 (none)
    std::vector<uint8_t> raw((size_t)raw_len);
    for (size_t y = 0; y < height; ++y) {
        raw[y * (1 + stride)] = 0;
        memcpy(&raw[y * (1 + stride) + 1], rgba + y * stride, stride);
    }
    std::vector<uint8_t> z(compressBound((uLong)raw.size()));
    uLongf zlen = (uLongf)z.size();
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return TX_ERR_NOMEM;
    std::vector<uint8_t> png(kSignature, kSignature + 8);
    uint8_t ihdr[13];
    const uint32_t w = (uint32_t)width, h = (uint32_t)height;
    for (int i = 0; i < 4; ++i) {
        ihdr[i] = (uint8_t)(w >> (24 - 8 * i));
        ihdr[4 + i] = (uint8_t)(h >> (24 - 8 * i));
    }
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 6;   // RGBA
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering
    ihdr[12] = 0;  // not interlaced
    chunk(png, "IHDR", ihdr, 13);
    for (size_t o = 0; o < zlen; o += 0x7FFFFFFFu) chunk(png, "IDAT", z.data() + o, std::min<size_t>(zlen - o, 0x7FFFFFFFu));
    chunk(png, "IEND", nullptr, 0);
    *out_len = png.size();
    if (png.size() > out_cap) return TX_ERR_OUTPUT_SHORT;
    memcpy(out, png.data(), png.size());
    return TX_OK;
}

}  // extern "C"
