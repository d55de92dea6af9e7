"""PNG reading/writing and the command line of the `texpresso` tool, on the
CPU (the tool's failures before any GPU work, and the image I/O it uses).

PNG files come from an independent writer (tests/pngenc.py: every colour
type and depth, all five row filters, Adam7) with the expected RGBA computed
from the samples by the rules of the reference CLI's reader
(cli/src/image/png.rs:59-99); 8-bit files are also read with Pillow as a
second decoder.  Parity with the png crate's own decoder is pinned by those
rules, not by running it (no Rust toolchain here)."""
import io
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import png

from . import pngenc

CASES = [  # (colour type, depth)
    (0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8), (4, 8), (4, 16),
    (6, 8), (6, 16),
]


def samples_for(ctype, depth, w, h, seed):
    rng = np.random.default_rng(seed)
    top = (1 << depth) - 1 if ctype != 3 else min((1 << depth) - 1, 40)
    return rng.integers(0, top + 1, size=(h, w, pngenc.CHANNELS[ctype]), dtype=np.uint32)


def palette(n=41):
    return (np.arange(3 * n, dtype=np.uint32) * 37 % 256).astype(np.uint8)


@pytest.mark.parametrize("ctype,depth", CASES, ids=lambda v: str(v))
@pytest.mark.parametrize("interlace", [False, True], ids=["plain", "adam7"])
@pytest.mark.parametrize("w,h", [(1, 1), (3, 5), (13, 9), (33, 17)], ids=lambda v: str(v))
def test_read_matches_the_sample_rules(ctype, depth, interlace, w, h):
    s = samples_for(ctype, depth, w, h, 1000 * ctype + depth + w + h)
    plte = palette() if ctype == 3 else None
    data = pngenc.encode(s, ctype, depth, interlace=interlace, plte=plte, split_idat=7 if w > 8 else 0)
    rgba, pw, ph = png.read(data)
    assert (pw, ph) == (w, h)
    assert np.array_equal(rgba.reshape(h, w, 4), pngenc.expected_rgba(s, ctype, depth, plte=plte))


@pytest.mark.parametrize("ftype", range(5))
def test_each_filter_alone(ftype):
    s = samples_for(6, 8, 21, 11, ftype)
    data = pngenc.encode(s, 6, 8, filters=(ftype,))
    rgba, _, _ = png.read(data)
    assert np.array_equal(rgba.reshape(11, 21, 4), s.astype(np.uint8))


@pytest.mark.parametrize("ctype,depth,trns", [
    (0, 8, struct.pack(">H", 7)), (0, 16, struct.pack(">H", 300)), (0, 2, struct.pack(">H", 1)),
    (2, 8, struct.pack(">HHH", 1, 2, 3)), (2, 16, struct.pack(">HHH", 256, 512, 768)),
    (3, 4, bytes([0, 128, 255, 7])), (3, 8, bytes(range(0, 200, 5))),
], ids=lambda v: v.hex() if isinstance(v, bytes) else str(v))
def test_transparency_chunk(ctype, depth, trns):
    w, h = 17, 6
    rng = np.random.default_rng(depth + ctype)
    if ctype == 3:
        s = rng.integers(0, 41 if depth == 8 else 16, size=(h, w, 1), dtype=np.uint32)
    else:
        key = struct.unpack(">H" if ctype == 0 else ">HHH", trns)
        s = rng.integers(0, (1 << depth), size=(h, w, pngenc.CHANNELS[ctype]), dtype=np.uint32)
        s[::3, ::2] = key  # plant the colour key
    plte = palette() if ctype == 3 else None
    data = pngenc.encode(s, ctype, depth, plte=plte, trns=trns)
    rgba, _, _ = png.read(data)
    want = pngenc.expected_rgba(s, ctype, depth, plte=plte, trns=trns)
    assert np.array_equal(rgba.reshape(h, w, 4), want)
    assert (want[..., 3] == 0).any() or ctype == 3


@pytest.mark.parametrize("mode", ["L", "LA", "RGB", "RGBA", "P", "1"])
def test_read_agrees_with_pillow(mode):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(len(mode))
    base = rng.integers(0, 256, size=(23, 37, 4), dtype=np.uint8)
    im = Image.fromarray(base, "RGBA")
    im = im.convert(mode) if mode != "P" else im.convert("RGB").quantize(50)
    bio = io.BytesIO()
    im.save(bio, format="PNG")
    rgba, w, h = png.read(bio.getvalue())
    assert (w, h) == (37, 23)
    assert np.array_equal(rgba.reshape(h, w, 4), np.asarray(im.convert("RGBA")))


def test_write_round_trip_and_pillow():
    rng = np.random.default_rng(5)
    for w, h in [(1, 1), (7, 3), (64, 33)]:
        px = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
        data = png.write(px, w, h)
        back, bw, bh = png.read(data)
        assert (bw, bh) == (w, h) and np.array_equal(back.reshape(h, w, 4), px)
        Image = pytest.importorskip("PIL.Image")
        im = Image.open(io.BytesIO(data))
        assert im.mode == "RGBA" and np.array_equal(np.asarray(im), px)


def test_malformed_files_are_refused():
    good = pngenc.encode(samples_for(2, 8, 9, 9, 1), 2, 8)
    assert png.read(good)[1:] == (9, 9)
    bad = [
        b"", good[:7], b"\x89PNH" + good[4:],          # signature
        good[:-12],                                      # no IEND
        good[:40] + bytes([good[40] ^ 1]) + good[41:],   # CRC of a chunk body
    ]
    for b in bad:
        with pytest.raises(tx.TexpressoError):
            png.read(b)
    # a row with filter type 5
    raw = bytes([5]) + bytes(9 * 3)
    z = zlib.compress(raw)
    f = b"\x89PNG\r\n\x1a\n" + pngenc.chunk(b"IHDR", struct.pack(">IIBBBBB", 9, 1, 8, 2, 0, 0, 0))
    f += pngenc.chunk(b"IDAT", z) + pngenc.chunk(b"IEND", b"")
    with pytest.raises(tx.TexpressoError):
        png.read(f)
    # truncated image data, a depth the colour type does not allow, an
    # unknown critical chunk
    f = b"\x89PNG\r\n\x1a\n" + pngenc.chunk(b"IHDR", struct.pack(">IIBBBBB", 9, 4, 8, 2, 0, 0, 0))
    with pytest.raises(tx.TexpressoError):
        png.read(f + pngenc.chunk(b"IDAT", zlib.compress(raw[:10])) + pngenc.chunk(b"IEND", b""))
    f = b"\x89PNG\r\n\x1a\n" + pngenc.chunk(b"IHDR", struct.pack(">IIBBBBB", 9, 1, 4, 2, 0, 0, 0))
    with pytest.raises(tx.TexpressoError):
        png.read(f + pngenc.chunk(b"IDAT", z) + pngenc.chunk(b"IEND", b""))
    f = good[:33] + pngenc.chunk(b"ZZZZ", b"x") + good[33:]
    with pytest.raises(tx.TexpressoError):
        png.read(f)


# ---- command line: argument handling and failures before any GPU work


def run_cli(args, cwd):
    return subprocess.run([tx.cli_path()] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=60)


def test_cli_is_built():
    assert os.access(tx.cli_path(), os.X_OK)


@pytest.mark.parametrize("args,code,text", [
    (["--version"], 0, "2.0.2"),
    (["help", "compress"], 0, "--weigh-colour-by-alpha"),
    (["compress", "--help"], 0, "[default: balanced]"),
    ([], 2, "Usage"),
    (["frobnicate"], 2, "unrecognized subcommand"),
    (["compress", "a.png"], 2, "--format <FORMAT>"),
    (["compress", "-f", "bc6", "a.png"], 2, "possible values"),
    (["compress", "-f", "bc1", "-p", "fast", "a.png"], 2, "possible values"),
    (["compress", "-f", "bc1", "-f", "bc2", "a.png"], 2, "cannot be used multiple times"),
    (["compress", "-f", "bc1", "-w", "abc", "a.png"], 2, "invalid float literal"),
    (["compress", "-f", "bc1", "-w", "0x10", "a.png"], 2, "invalid float literal"),
    (["compress", "-f", "bc1"], 2, "<INFILE>"),
    (["compress", "-f", "bc1", "a.png", "b.png"], 2, "unexpected argument"),
    (["compress", "-f", "bc1", "--bogus", "a.png"], 2, "unexpected argument"),
    (["decompress", "-f", "bc1", "a.dds"], 2, "unexpected argument"),
    (["compress", "-f", "bc1", "-w", "1", "-w", "2", "a.png"], 101, "Weights must have 3 values"),
    (["compress", "-f", "bc1", "noext"], 101, "no extension"),
    (["compress", "-f", "bc1", "img.gif"], 101, "Unrecognized image format"),
    (["compress", "-f", "bc1", "img.JPG"], 101, "JPEG"),
    (["compress", "-f", "bc1", "missing.png"], 101, "Failed to open file"),
    (["compress", "-f", "bc1", "notpng.png"], 101, "Is this really a PNG file?"),
    (["decompress", "missing.dds"], 101, "Failed to open file"),
    (["decompress", "notdds.dds"], 101, "DDS"),
    (["decompress", "huge.dds"], 101, "out of range"),
])
def test_cli_arguments(tmp_path, args, code, text):
    (tmp_path / "notpng.png").write_bytes(b"not a png at all")
    (tmp_path / "notdds.dds").write_bytes(b"DDS " + bytes(40))
    # a legacy DXT1 header claiming 2^32-1 x 2^32-1 pixels (refused before
    # anything is allocated or sent to the GPU)
    huge = bytearray(128)
    struct.pack_into("<4s7I", huge, 0, b"DDS ", 124, 0x1007, 0xFFFFFFFF, 0xFFFFFFFF, 0, 0, 0)
    struct.pack_into("<2I4s", huge, 4 + 72, 32, 0x4, b"DXT1")
    (tmp_path / "huge.dds").write_bytes(bytes(huge) + bytes(64))
    r = run_cli(args, tmp_path)
    assert r.returncode == code, r.stderr
    assert text in r.stdout + r.stderr
    assert not (tmp_path / "a.dds").exists()
