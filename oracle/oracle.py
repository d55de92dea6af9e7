"""ctypes wrapper around liboracle.so, the CPU restatement of texpresso 2.0.2.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the texpresso_amd package.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC1, BC2, BC3, BC4, BC5 = range(5)
RANGE_FIT, CLUSTER_FIT, ITERATIVE_CLUSTER_FIT = range(3)
PERCEPTUAL = (0.2126, 0.7152, 0.0722)
UNIFORM = (1.0, 1.0, 1.0)


class OraParams(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("weights", ctypes.c_float * 3),
        ("weigh_colour_by_alpha", ctypes.c_uint8),
    ]


class OraBlockStats(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int32),
        ("transparent", ctypes.c_int32),
        ("fit", ctypes.c_int32),
        ("searched3", ctypes.c_int32),
        ("searched4", ctypes.c_int32),
    ]


STATS_DTYPE = np.dtype(
    [("count", "<i4"), ("transparent", "<i4"), ("fit", "<i4"), ("searched3", "<i4"), ("searched4", "<i4")]
)

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def load(path):
    """Loads a build of the restatement from `path` with its argument types set."""
    L = ctypes.CDLL(path)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    L.ora_compress_block_masked.argtypes = [ctypes.c_int, u8p, ctypes.c_uint32, ctypes.POINTER(OraParams), u8p,
                                            ctypes.POINTER(OraBlockStats)]
    L.ora_compress.argtypes = [ctypes.c_int, u8p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                               ctypes.POINTER(OraParams), u8p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    L.ora_compress_blocks.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.POINTER(OraParams), ctypes.c_void_p, ctypes.c_int]
    L.ora_decompress_block.argtypes = [ctypes.c_int, u8p, u8p]
    L.ora_decompress.argtypes = [ctypes.c_int, u8p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, u8p,
                                 ctypes.c_size_t]
    L.ora_single_lut.argtypes = [u8p]
    L.ora_block_size.restype = ctypes.c_size_t
    L.ora_num_blocks.restype = ctypes.c_size_t
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def swapped(path):
    """Runs the enclosed code against another build of the restatement (the
    planted-mistake checks build copies of bcn_oracle.c with one change)."""
    global _lib
    old = lib()
    _lib = load(path)
    try:
        yield
    finally:
        _lib = old


def params(algorithm=CLUSTER_FIT, weights=PERCEPTUAL, weigh_colour_by_alpha=False):
    p = OraParams()
    p.algorithm = algorithm
    for i in range(3):
        p.weights[i] = weights[i]
    p.weigh_colour_by_alpha = 1 if weigh_colour_by_alpha else 0
    return p


def _u8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def block_size(fmt):
    return 8 if fmt in (BC1, BC4) else 16


def num_blocks(n):
    return (n + 3) // 4


def compressed_size(fmt, w, h):
    return num_blocks(w) * num_blocks(h) * block_size(fmt)


def compress_block_masked(fmt, rgba16, mask, p):
    rgba = np.ascontiguousarray(np.asarray(rgba16, dtype=np.uint8).reshape(64))
    out = np.zeros(block_size(fmt), np.uint8)
    st = OraBlockStats()
    rc = lib().ora_compress_block_masked(fmt, _u8(rgba), mask, ctypes.byref(p), _u8(out), ctypes.byref(st))
    if rc != 0:
        raise ValueError("oracle rejected block (rc=%d)" % rc)
    return out, st


def compress(fmt, rgba, w, h, p, out_len=None, threads=1, stats=False):
    rgba = np.ascontiguousarray(np.asarray(rgba, dtype=np.uint8).reshape(-1))
    if out_len is None:
        out_len = compressed_size(fmt, w, h)
    out = np.zeros(out_len, np.uint8)
    st = np.zeros(out_len // block_size(fmt), STATS_DTYPE) if stats else None
    rc = lib().ora_compress(fmt, _u8(rgba), rgba.size, w, h, ctypes.byref(p), _u8(out), out_len, threads,
                            st.ctypes.data if stats else None)
    if rc != 0:
        raise ValueError("oracle compress failed (rc=%d)" % rc)
    return (out, st) if stats else out


def compress_blocks(fmt, blocks, masks, p, threads=None):
    """Many compress_block_masked calls; blocks (n, 64) uint8, masks (n,) uint32."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
    masks = np.ascontiguousarray(masks, dtype=np.uint32).reshape(-1)
    out = np.zeros((blocks.shape[0], block_size(fmt)), np.uint8)
    rc = lib().ora_compress_blocks(fmt, blocks.ctypes.data, masks.ctypes.data, blocks.shape[0], ctypes.byref(p),
                               out.ctypes.data, threads or default_threads())
    if rc != 0:
        raise ValueError("oracle compress_blocks failed (rc=%d)" % rc)
    return out


def default_threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def decompress_block(fmt, block):
    b = np.ascontiguousarray(np.asarray(block, dtype=np.uint8))
    out = np.zeros(64, np.uint8)
    lib().ora_decompress_block(fmt, _u8(b), _u8(out))
    return out


def decompress(fmt, data, w, h):
    d = np.ascontiguousarray(np.asarray(data, dtype=np.uint8).reshape(-1))
    out = np.zeros(4 * w * h, np.uint8)
    rc = lib().ora_decompress(fmt, _u8(d), d.size, w, h, _u8(out), out.size)
    if rc != 0:
        raise ValueError("oracle decompress failed (rc=%d)" % rc)
    return out


def single_lut():
    out = np.zeros(4 * 256 * 6, np.uint8)
    lib().ora_single_lut(_u8(out))
    return out.tobytes()
