// tx_dds.hip -- DDS container for compressed textures (host code only).
//
// texpresso's CLI wraps the encoder output in a DDS file through the ddsfile
// crate (cli/src/main.rs:143-164): a DX10 extended header, Texture2D,
// DXGI BC1/BC2/BC3 *_UNORM_SRGB or BC4/BC5 *_UNORM (format_to_dxgiformat,
// cli/src/main.rs:220-228), alpha mode premultiplied for BC1 and straight
// otherwise.  Its reader (cli/src/main.rs:167-196) accepts those five DXGI
// formats (dxgiformat_to_format, 230-239) with a Texture2D resource
// dimension, or a legacy DXT1/DXT3/DXT5 FourCC header (d3dformat_to_format,
// 241-248).  The byte layout below is the documented
// DDS one (DDS_HEADER, DDS_PIXELFORMAT, DDS_HEADER_DXT10), written
// little-endian; mip levels follow level 0 back to back.
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "texpresso_hip.h"

namespace {

constexpr uint32_t kMagic = 0x20534444u;  // "DDS "
constexpr uint32_t kFourCCDX10 = 0x30315844u, kFourCCDXT1 = 0x31545844u, kFourCCDXT3 = 0x33545844u,
                   kFourCCDXT5 = 0x35545844u;
constexpr uint32_t kFlagCaps = 0x1, kFlagHeight = 0x2, kFlagWidth = 0x4, kFlagPixelFormat = 0x1000,
                   kFlagMipCount = 0x20000, kFlagLinearSize = 0x80000;
constexpr uint32_t kCapsComplex = 0x8, kCapsTexture = 0x1000, kCapsMipmap = 0x400000;
constexpr uint32_t kPfFourCC = 0x4;
constexpr uint32_t kDimTexture2D = 3;
constexpr uint32_t kAlphaStraight = 1, kAlphaPremultiplied = 2;
constexpr size_t kHeaderBytes = 4 + 124 + 20;

constexpr uint32_t kDxgi[5] = {72, 75, 78, 80, 83};  // BC1..BC3 UNORM_SRGB, BC4/BC5 UNORM

void put32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

uint32_t get32(const uint8_t* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

}  // namespace

extern "C" {

size_t tx_dds_header_size(void) { return kHeaderBytes; }

int tx_dds_write_header(int format, size_t width, size_t height, uint32_t levels, uint8_t* out, size_t out_len)
{
    if (format < TX_BC1 || format > TX_BC5) return TX_ERR_FORMAT;
    if (width == 0) return TX_ERR_ZERO_WIDTH;
    if (height == 0 || width > 0xFFFFFFFFu || height > 0xFFFFFFFFu) return TX_ERR_INVALID;
    if (levels == 0 || levels > tx_mip_levels(width, height)) return TX_ERR_INVALID;
    if (!out) return TX_ERR_INVALID;
    if (out_len < kHeaderBytes) return TX_ERR_OUTPUT_SHORT;
    memset(out, 0, kHeaderBytes);
    uint8_t* h = out + 4;
    put32(out, kMagic);
    uint32_t flags = kFlagCaps | kFlagHeight | kFlagWidth | kFlagPixelFormat | kFlagLinearSize;
    uint32_t caps = kCapsTexture;
    if (levels > 1) {
        flags |= kFlagMipCount;
        caps |= kCapsComplex | kCapsMipmap;
    }
    const size_t top = tx_compressed_size(format, width, height);
    put32(h + 0, 124);                                                      // dwSize
    put32(h + 4, flags);                                                    // dwFlags
    put32(h + 8, (uint32_t)height);                                         // dwHeight
    put32(h + 12, (uint32_t)width);                                         // dwWidth
    put32(h + 16, (uint32_t)std::min<size_t>(top, 0xFFFFFFFFu));            // dwPitchOrLinearSize
    put32(h + 24, levels > 1 ? levels : 0);                                 // dwMipMapCount
    put32(h + 72, 32);                                                      // ddspf.dwSize
    put32(h + 76, kPfFourCC);                                               // ddspf.dwFlags
    put32(h + 80, kFourCCDX10);                                             // ddspf.dwFourCC
    put32(h + 104, caps);                                                   // dwCaps
    uint8_t* x = out + 128;                                                 // DDS_HEADER_DXT10
    put32(x + 0, kDxgi[format]);                                            // dxgiFormat
    put32(x + 4, kDimTexture2D);                                            // resourceDimension
    put32(x + 12, 1);                                                       // arraySize
    put32(x + 16, format == TX_BC1 ? kAlphaPremultiplied : kAlphaStraight);  // miscFlags2
    return TX_OK;
}

int tx_dds_parse_header(const uint8_t* data, size_t len, int* format, size_t* width, size_t* height,
                        uint32_t* levels, size_t* data_offset)
{
    if (!data || !format || !width || !height || !levels || !data_offset) return TX_ERR_INVALID;
    if (len < 128 || get32(data) != kMagic || get32(data + 4) != 124) return TX_ERR_INVALID;
    const uint8_t* h = data + 4;
    const uint32_t flags = get32(h + 4);
    const uint32_t fourcc = get32(h + 80);
    if (!(get32(h + 76) & kPfFourCC)) return TX_ERR_FORMAT;
    int fmt = -1;
    size_t off = 128;
    if (fourcc == kFourCCDX10) {
        if (len < 148) return TX_ERR_INPUT_SHORT;
        const uint32_t dxgi = get32(data + 128);
        if (get32(data + 132) != kDimTexture2D) return TX_ERR_FORMAT;
        for (int f = 0; f < 5; ++f)
            if (kDxgi[f] == dxgi) fmt = f;
        off = 148;
    } else if (fourcc == kFourCCDXT1) {
        fmt = TX_BC1;
    } else if (fourcc == kFourCCDXT3) {
        fmt = TX_BC2;
    } else if (fourcc == kFourCCDXT5) {
        fmt = TX_BC3;
    }
    if (fmt < 0) return TX_ERR_FORMAT;
    *format = fmt;
    *height = get32(h + 8);
    *width = get32(h + 12);
    const uint32_t mips = get32(h + 24);
    *levels = (flags & kFlagMipCount) && mips > 0 ? mips : 1;
    *data_offset = off;
    return TX_OK;
}

}  // extern "C"
