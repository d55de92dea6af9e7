"""PNG files of the command-line tool (cli/src/image/png.rs) through the
library's tx_png_* functions: read gives RGBA8 as texpresso's CLI sees it
(png crate EXPAND | STRIP_16, then grey / grey-alpha / RGB to RGBA), write
stores RGBA8.

    rgba, w, h = png.read(open("in.png", "rb").read())
    data = png.write(rgba, w, h)
"""
import ctypes

import numpy as np

from . import TexpressoError, _check, _ptr, lib


def read(data):
    """(RGBA8 array of 4*w*h bytes, w, h) of a PNG file's bytes."""
    buf = np.frombuffer(bytes(data), np.uint8)
    w, h = ctypes.c_size_t(), ctypes.c_size_t()
    p, n, _keep = _ptr(buf)
    _check(lib().tx_png_info(p, n, ctypes.byref(w), ctypes.byref(h)), "png info")
    out = np.zeros(4 * w.value * h.value, np.uint8)
    _check(lib().tx_png_decode(p, n, out.ctypes.data, out.nbytes), "png decode")
    return out, w.value, h.value


def write(rgba, width, height):
    """PNG file bytes holding RGBA8 pixels (4*width*height bytes)."""
    src = np.ascontiguousarray(rgba, np.uint8).reshape(-1)
    if src.size < 4 * width * height:
        raise TexpressoError(-3, "png encode")
    cap = lib().tx_png_encode_bound(width, height)
    if cap == 0:
        raise TexpressoError(-10, "png encode")
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t()
    _check(lib().tx_png_encode(src.ctypes.data, width, height, out.ctypes.data, cap, ctypes.byref(n)), "png encode")
    return out[:n.value].tobytes()
