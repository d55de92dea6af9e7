"""DDS container and mip-chain bookkeeping (host code of the C ABI; no GPU).

The header is checked field by field with a struct-based reader written
from the documented DDS layout (DDS_HEADER, DDS_PIXELFORMAT,
DDS_HEADER_DXT10), not with the library's own parser; the formats and alpha
modes are those of texpresso's CLI (cli/src/main.rs:143-164, 220-248).  The
exact bytes the ddsfile crate writes are not available here (parity
unpinned for the header's optional fields)."""
import struct

import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import dds

DXGI = {0: 72, 1: 75, 2: 78, 3: 80, 4: 83}


def read_header(b):
    magic, = struct.unpack_from("<4s", b, 0)
    (size, flags, height, width, linear, depth, mips) = struct.unpack_from("<7I", b, 4)
    pf = struct.unpack_from("<8I", b, 4 + 72)
    caps = struct.unpack_from("<5I", b, 4 + 104)
    dx10 = struct.unpack_from("<5I", b, 128)
    return dict(magic=magic, size=size, flags=flags, height=height, width=width, linear=linear, depth=depth,
                mips=mips, pf=pf, caps=caps, dx10=dx10)


def full_levels(w, h):
    n = 1
    while w > 1 or h > 1:
        w, h, n = max(1, w >> 1), max(1, h >> 1), n + 1
    return n


@pytest.mark.parametrize("fmt", range(5))
@pytest.mark.parametrize("w,h,levels", [(1, 1, 1), (5, 3, 1), (256, 64, 9), (257, 129, 3), (4096, 4096, 13)])
def test_header_fields(fmt, w, h, levels):
    b = dds.header(fmt, w, h, levels).tobytes()
    assert len(b) == 148 == tx.lib().tx_dds_header_size()
    f = read_header(b)
    assert f["magic"] == b"DDS " and f["size"] == 124
    assert f["width"] == w and f["height"] == h
    required = 0x1 | 0x2 | 0x4 | 0x1000  # CAPS | HEIGHT | WIDTH | PIXELFORMAT
    assert f["flags"] & required == required
    assert f["flags"] & 0x80000 and f["linear"] == tx.Format(fmt).compressed_size(w, h)  # LINEARSIZE
    assert f["pf"][0] == 32 and f["pf"][1] & 0x4 and f["pf"][2] == struct.unpack("<I", b"DX10")[0]
    assert f["caps"][0] & 0x1000  # TEXTURE
    if levels > 1:
        assert f["flags"] & 0x20000 and f["mips"] == levels
        assert f["caps"][0] & 0x8 and f["caps"][0] & 0x400000  # COMPLEX | MIPMAP
    else:
        assert not f["flags"] & 0x20000
    dxgi, dim, misc, array, misc2 = f["dx10"]
    assert dxgi == DXGI[fmt] and dim == 3 and misc == 0 and array == 1
    assert misc2 == (2 if fmt == 0 else 1)  # premultiplied for BC1, straight otherwise


def test_parse_round_trip_and_levels():
    w, h = 100, 37
    n = full_levels(w, h)
    assert tx.mip_levels(w, h) == n
    data = np.arange(tx.Format.Bc3.mip_chain_size(w, h), dtype=np.uint32).astype(np.uint8)
    blob = dds.wrap(tx.Format.Bc3, w, h, data, n)
    fmt, pw, ph, pn, views = dds.parse(blob)
    assert (fmt, pw, ph, pn) == (tx.Format.Bc3, w, h, n)
    assert np.array_equal(np.concatenate(views), data)
    for k, v in enumerate(views):
        lw, lh = max(1, w >> k), max(1, h >> k)
        assert tx.mip_dims(w, h, k) == (lw, lh)
        assert v.size == tx.Format.Bc3.compressed_size(lw, lh)


@pytest.mark.parametrize("fourcc,fmt", [(b"DXT1", 0), (b"DXT3", 1), (b"DXT5", 2)])
def test_parse_legacy_fourcc(fourcc, fmt):
    b = bytearray(128)
    struct.pack_into("<4s7I", b, 0, b"DDS ", 124, 0x1007, 8, 12, 0, 0, 0)
    struct.pack_into("<3I", b, 4 + 72, 32, 0x4, struct.unpack("<I", fourcc)[0])
    f, w, h, n, views = dds.parse(bytes(b) + bytes(tx.Format(fmt).compressed_size(12, 8)))
    assert (int(f), w, h, n) == (fmt, 12, 8, 1)


@pytest.mark.parametrize("dxgi,dim", [(71, 3), (77, 3), (84, 3), (98, 3), (72, 4), (78, 2)])
def test_parse_rejects_what_the_cli_rejects(dxgi, dim):
    # BC1_UNORM (71), BC3_UNORM (77), BC5_SNORM (84), BC7 (98), or a resource
    # dimension other than Texture2D: the reference CLI panics on these
    b = bytearray(dds.header(0, 8, 8).tobytes())
    struct.pack_into("<2I", b, 128, dxgi, dim)
    with pytest.raises(tx.TexpressoError):
        dds.parse(bytes(b))


def test_header_rejects_bad_arguments():
    out = np.zeros(148, np.uint8)
    L = tx.lib()
    assert L.tx_dds_write_header(7, 4, 4, 1, out.ctypes.data, 148) == -1
    assert L.tx_dds_write_header(0, 0, 4, 1, out.ctypes.data, 148) == -2
    assert L.tx_dds_write_header(0, 4, 4, 4, out.ctypes.data, 148) < 0  # 4x4 has 3 levels
    assert L.tx_dds_write_header(0, 4, 4, 1, out.ctypes.data, 147) < 0
    with pytest.raises(tx.TexpressoError):
        dds.parse(b"DDS " + bytes(60))


@pytest.mark.parametrize("w,h", [(1, 1), (1, 9), (7, 3), (64, 64), (4096, 4096), (3000, 17)])
def test_chain_sizes(w, h):
    n = full_levels(w, h)
    dims = [(max(1, w >> k), max(1, h >> k)) for k in range(n)]
    assert tx.lib().tx_mip_chain_rgba_bytes(w, h, n) == sum(4 * a * b for a, b in dims[1:])
    for fmt in range(5):
        f = tx.Format(fmt)
        assert f.mip_chain_size(w, h) == sum(f.compressed_size(a, b) for a, b in dims)


def test_mip_entry_points_reject_bad_sizes():
    # argument checks run before any device work, so these need no GPU
    import ctypes
    L = tx.lib()
    p = tx.Params()._c()
    big = 0x7FFFFFFF
    assert L.tx_compress_mips(0, None, 0, big, big, ctypes.byref(p), 1, None, 0) == -10  # 4*w*h overflows
    assert L.tx_compress_mips(0, None, 0, 0, 4, ctypes.byref(p), 1, None, 0) == -2       # zero width
    assert L.tx_compress_mips(0, None, 0, 4, 4, ctypes.byref(p), 4, None, 0) == -10      # 4x4 has 3 levels
    assert L.tx_compress_mips(9, None, 0, 4, 4, ctypes.byref(p), 1, None, 0) == -1       # bad format
    assert L.tx_mip_chain_rgba_bytes(2 ** 40, 2 ** 40, 3) == 2 ** 64 - 1                  # saturates
    assert L.tx_mip_levels(0, 5) == 0 and L.tx_mip_levels(1, 1) == 1 and L.tx_mip_levels(4096, 3) == 13
