"""A PNG writer for the tests, written from the PNG specification apart
from the library's decoder: any colour type and bit depth, a chosen filter
per row (all five), optional Adam7 interlacing, PLTE and tRNS chunks.  It
also gives the RGBA8 pixels texpresso's CLI sees for the samples it wrote
(png crate EXPAND | STRIP_16, then the CLI's conversion to RGBA,
cli/src/image/png.rs:59-99), computed straight from the samples."""
import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def chunk(kind, body):
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def pack_row(samples, depth):
    """Samples of one row (flat, channel-interleaved) to bytes at `depth`."""
    s = np.asarray(samples, np.uint32).reshape(-1)
    if depth == 8:
        return s.astype(np.uint8).tobytes()
    if depth == 16:
        return s.astype(">u2").tobytes()
    per = 8 // depth
    pad = (-s.size) % per
    s = np.concatenate([s, np.zeros(pad, np.uint32)]).reshape(-1, per)
    shifts = (8 - depth * (np.arange(per) + 1)).astype(np.uint32)
    return (s << shifts).sum(axis=1).astype(np.uint8).tobytes()


def filter_row(ftype, row, prev, bpp):
    """Filtered bytes of one row (PNG section 9.2), from the unfiltered bytes."""
    x = np.frombuffer(row, np.uint8).astype(np.int32)
    b = np.frombuffer(prev, np.uint8).astype(np.int32) if prev is not None else np.zeros_like(x)
    a = np.zeros_like(x)  # byte bpp to the left, 0 before the row start
    c = np.zeros_like(x)  # the same above
    if x.size > bpp:
        a[bpp:] = x[:-bpp]
        c[bpp:] = b[:-bpp]
    if ftype == 0:
        pred = np.zeros_like(x)
    elif ftype == 1:
        pred = a
    elif ftype == 2:
        pred = b
    elif ftype == 3:
        pred = (a + b) >> 1
    else:
        p = a + b - c
        pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
        pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
    return bytes([ftype]) + ((x - pred) & 255).astype(np.uint8).tobytes()


def encode(samples, ctype, depth, interlace=False, filters=(0, 1, 2, 3, 4), plte=None, trns=None, split_idat=0):
    """PNG bytes of `samples` (h x w x channels, integers at `depth`).  Row r
    of each pass uses filters[r % len(filters)]."""
    h, w = samples.shape[:2]
    nch = CHANNELS[ctype]
    samples = samples.reshape(h, w, nch)
    bpp = max(1, nch * depth // 8)
    raw = bytearray()
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = None
        for r in range(sub.shape[0]):
            row = pack_row(sub[r], depth)
            raw += filter_row(filters[r % len(filters)], row, prev, bpp)
            prev = row
    z = zlib.compress(bytes(raw), 9)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    out += chunk(b"tEXt", b"Comment\x00test image")  # an ancillary chunk the reader skips
    if plte is not None:
        out += chunk(b"PLTE", np.asarray(plte, np.uint8).tobytes())
    if trns is not None:
        out += chunk(b"tRNS", trns)
    if split_idat:
        for o in range(0, len(z), split_idat):
            out += chunk(b"IDAT", z[o:o + split_idat])
    else:
        out += chunk(b"IDAT", z)
    return out + chunk(b"IEND", b"")


def expected_rgba(samples, ctype, depth, plte=None, trns=None):
    """RGBA8 (h x w x 4) texpresso's CLI reads from such a file."""
    h, w = samples.shape[:2]
    s = samples.reshape(h, w, CHANNELS[ctype]).astype(np.uint32)

    def to8(v):
        if depth == 16:
            return (v >> 8).astype(np.uint8)
        if depth < 8:
            return (v * (255 // ((1 << depth) - 1))).astype(np.uint8)
        return v.astype(np.uint8)

    out = np.zeros((h, w, 4), np.uint8)
    if ctype == 3:
        pal = np.zeros((256, 3), np.uint8)
        p = np.asarray(plte, np.uint8).reshape(-1, 3)
        pal[:len(p)] = p
        alpha = np.full(256, 255, np.uint8)
        if trns is not None:
            alpha[:len(trns)] = np.frombuffer(trns, np.uint8)
        idx = s[..., 0]
        out[..., :3] = pal[idx]
        out[..., 3] = alpha[idx]
        return out
    if ctype in (0, 4):
        g = to8(s[..., 0])
        out[..., 0] = out[..., 1] = out[..., 2] = g
        out[..., 3] = to8(s[..., 1]) if ctype == 4 else 255
        if ctype == 0 and trns is not None:
            key = struct.unpack(">H", trns)[0]
            out[..., 3] = np.where(s[..., 0] == key, 0, 255)
        return out
    out[..., :3] = to8(s[..., :3])
    out[..., 3] = to8(s[..., 3]) if ctype == 6 else 255
    if ctype == 2 and trns is not None:
        key = np.array(struct.unpack(">HHH", trns), np.uint32)
        out[..., 3] = np.where(np.all(s == key, axis=-1), 0, 255)
    return out
