"""DDS files around the encoder's output (the container texpresso's CLI
writes and reads, cli/src/main.rs:122-196), through the library's
tx_dds_write_header / tx_dds_parse_header.

    data = dds.encode(Format.Bc3, rgba, w, h, Params(), mips=True)
    fmt, w, h, levels = dds.parse(data)[:4]
    rgba0 = dds.decode_level0(data)          # what `texpresso decompress` does
"""
import ctypes

import numpy as np

from . import Format, _check, _ptr, lib, mip_dims, mip_levels


def header(fmt, width, height, levels=1):
    out = np.zeros(lib().tx_dds_header_size(), np.uint8)
    _check(lib().tx_dds_write_header(int(fmt), width, height, levels, out.ctypes.data, out.nbytes), "dds header")
    return out


def wrap(fmt, width, height, data, levels=1):
    """A DDS file (bytes) holding `data`: the compressed levels back to back."""
    h = header(fmt, width, height, levels)
    return h.tobytes() + bytes(np.ascontiguousarray(data, np.uint8).reshape(-1).data)


def encode(fmt, rgba, width, height, params, mips=False):
    """Compress `rgba` (level 0, plus the GPU-generated chain with mips=True) into a DDS file."""
    fmt = Format(fmt)
    if mips:
        levels = mip_levels(width, height)
        data = fmt.compress_mips(rgba, width, height, params, levels)
    else:
        levels = 1
        data = np.zeros(fmt.compressed_size(width, height), np.uint8)
        fmt.compress(rgba, width, height, params, data)
    return wrap(fmt, width, height, data, levels)


def parse(data):
    """(format, width, height, levels, [level data views]) of a DDS file."""
    buf = np.frombuffer(bytes(data), np.uint8)
    f, w, h, n, off = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint32(), ctypes.c_size_t()
    p, l, _keep = _ptr(buf)
    _check(lib().tx_dds_parse_header(p, l, ctypes.byref(f), ctypes.byref(w), ctypes.byref(h), ctypes.byref(n),
                                     ctypes.byref(off)), "dds parse")
    fmt = Format(f.value)
    views, pos = [], off.value
    for k in range(n.value):
        lw, lh = mip_dims(w.value, h.value, k)
        size = fmt.compressed_size(lw, lh)
        views.append(buf[pos:pos + size])
        pos += size
    return fmt, w.value, h.value, n.value, views


def decode_level0(data):
    """RGBA8 pixels of level 0 (cli/src/main.rs:167-196)."""
    fmt, w, h, _, views = parse(data)
    out = np.zeros(4 * w * h, np.uint8)
    fmt.decompress(views[0], w, h, out)
    return out
