"""texpresso_amd -- MI355X (gfx950) BC1-BC5 texture encoder.

Python mirror of texpresso 2.0.2's public API (lib/src/lib.rs:38-336) over the
C ABI in include/texpresso_hip.h.  All encoding and decoding runs in the HIP
kernels of libtexpresso_hip.so; there is no CPU fallback.  Where the reference
panics (short buffers, zero width) these functions raise TexpressoError.

    from texpresso_amd import Format, Params, Algorithm
    out = bytearray(Format.Bc1.compressed_size(w, h))
    Format.Bc1.compress(rgba, w, h, Params(), out)
"""
import ctypes
import enum
import os

__all__ = [
    "Format", "Algorithm", "Params", "TexpressoError", "COLOUR_WEIGHTS_UNIFORM", "COLOUR_WEIGHTS_PERCEPTUAL",
    "num_blocks", "lib", "library_path",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtexpresso_hip.so")

# lib/src/lib.rs:71-74
COLOUR_WEIGHTS_UNIFORM = (1.0, 1.0, 1.0)
COLOUR_WEIGHTS_PERCEPTUAL = (0.2126, 0.7152, 0.0722)


class TexpressoError(ValueError):
    """A call the reference would panic on, or a device failure."""

    def __init__(self, code, what=""):
        self.code = code
        msg = lib().tx_error_string(code).decode() if _lib is not None else str(code)
        super().__init__("%s%s (code %d)" % (what + ": " if what else "", msg, code))


class Algorithm(enum.IntEnum):
    """lib/src/lib.rs:49-59."""
    RangeFit = 0
    ClusterFit = 1
    IterativeClusterFit = 2


class _TxParams(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("weights", ctypes.c_float * 3),
        ("weigh_colour_by_alpha", ctypes.c_uint8),
    ]


class Params:
    """lib/src/lib.rs:76-100; defaults ClusterFit, perceptual weights, no alpha weighting."""

    __slots__ = ("algorithm", "weights", "weigh_colour_by_alpha")

    def __init__(self, algorithm=Algorithm.ClusterFit, weights=COLOUR_WEIGHTS_PERCEPTUAL, weigh_colour_by_alpha=False):
        self.algorithm = Algorithm(algorithm)
        self.weights = tuple(float(w) for w in weights)
        if len(self.weights) != 3:
            raise ValueError("weights must have 3 entries")
        self.weigh_colour_by_alpha = bool(weigh_colour_by_alpha)

    def _c(self):
        p = _TxParams()
        p.algorithm = int(self.algorithm)
        for i in range(3):
            p.weights[i] = self.weights[i]
        p.weigh_colour_by_alpha = 1 if self.weigh_colour_by_alpha else 0
        return p

    def __repr__(self):
        return "Params(algorithm=%s, weights=%r, weigh_colour_by_alpha=%r)" % (
            self.algorithm.name, self.weights, self.weigh_colour_by_alpha)


_lib = None


def library_path():
    return LIB_PATH


def lib():
    """Load libtexpresso_hip.so (built by texpresso_amd.build); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtexpresso_hip.so is not built; run `python -m texpresso_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        c_u8p, c_sz, c_int, c_vp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p
        pp = ctypes.POINTER(_TxParams)
        sigs = {
            "tx_params_default": (_TxParams, []),
            "tx_num_blocks": (c_sz, [c_sz]),
            "tx_block_size": (c_sz, [c_int]),
            "tx_compressed_size": (c_sz, [c_int, c_sz, c_sz]),
            "tx_compress": (c_int, [c_int, c_u8p, c_sz, c_sz, c_sz, pp, c_u8p, c_sz]),
            "tx_decompress": (c_int, [c_int, c_u8p, c_sz, c_sz, c_sz, c_u8p, c_sz]),
            "tx_compress_block_masked": (c_int, [c_int, c_u8p, ctypes.c_uint32, pp, c_u8p, c_sz]),
            "tx_compress_blocks": (c_int, [c_int, c_u8p, c_vp, c_sz, pp, c_u8p]),
            "tx_decompress_block": (c_int, [c_int, c_u8p, c_sz, c_u8p]),
            "tx_compress_device": (c_int, [c_int, c_vp, c_sz, c_sz, pp, c_vp, c_sz, c_vp]),
            "tx_decompress_device": (c_int, [c_int, c_vp, c_sz, c_sz, c_vp, c_vp]),
            "tx_single_lut": (None, [c_u8p]),
            "tx_util_malloc": (c_int, [ctypes.POINTER(ctypes.c_void_p), c_sz]),
            "tx_util_free": (c_int, [c_vp]),
            "tx_util_upload": (c_int, [c_vp, c_vp, c_sz]),
            "tx_util_download": (c_int, [c_vp, c_vp, c_sz]),
            "tx_util_copy": (c_int, [c_vp, c_vp, c_sz]),
            "tx_util_valu_probe": (c_int, [ctypes.POINTER(ctypes.c_double)]),
            "tx_util_time_compress": (c_int, [c_int, c_vp, c_sz, c_sz, pp, c_vp, c_sz, c_int,
                                              ctypes.POINTER(ctypes.c_float)]),
            "tx_error_string": (ctypes.c_char_p, [c_int]),
            "tx_device_arch": (ctypes.c_char_p, []),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(buf, writable=False):
    """Address and length of a bytes-like object or numpy array."""
    try:
        import numpy as np
        if isinstance(buf, np.ndarray):
            if not buf.flags["C_CONTIGUOUS"]:
                raise ValueError("array must be C-contiguous")
            if writable and not buf.flags["WRITEABLE"]:
                raise ValueError("output array is read-only")
            return buf.ctypes.data, buf.nbytes, buf
    except ImportError:
        pass
    if isinstance(buf, bytes):
        if writable:
            raise ValueError("output must be writable")
        keep = ctypes.create_string_buffer(buf, len(buf))
        return ctypes.addressof(keep), len(buf), keep
    mv = memoryview(buf).cast("B")
    if writable and mv.readonly:
        raise ValueError("output must be writable")
    arr = (ctypes.c_uint8 * len(mv)).from_buffer(mv) if not mv.readonly else \
        ctypes.create_string_buffer(mv.tobytes(), len(mv))
    return ctypes.addressof(arr), len(mv), arr


def _check(rc, what):
    if rc != 0:
        raise TexpressoError(rc, what)


def num_blocks(size):
    """lib/src/lib.rs:103-105."""
    return (int(size) + 3) // 4


class Format(enum.IntEnum):
    """lib/src/lib.rs:39-46 with the methods of lib/src/lib.rs:117-336."""
    Bc1 = 0
    Bc2 = 1
    Bc3 = 2
    Bc4 = 3
    Bc5 = 4

    def block_size(self):
        """lib/src/lib.rs:159-168."""
        return 8 if self in (Format.Bc1, Format.Bc4) else 16

    def compressed_size(self, width, height):
        """lib/src/lib.rs:175-179."""
        return num_blocks(width) * num_blocks(height) * self.block_size()

    def compress(self, rgba, width, height, params, output):
        """lib/src/lib.rs:287-335: encode `rgba` (RGBA8, row-major) into `output`."""
        p = params._c()
        ip, il, _k1 = _ptr(rgba)
        op, ol, _k2 = _ptr(output, writable=True)
        _check(lib().tx_compress(int(self), ip, il, width, height, ctypes.byref(p), op, ol), "compress")

    def decompress(self, data, width, height, output):
        """lib/src/lib.rs:124-156."""
        ip, il, _k1 = _ptr(data)
        op, ol, _k2 = _ptr(output, writable=True)
        _check(lib().tx_decompress(int(self), ip, il, width, height, op, ol), "decompress")

    def compress_block_masked(self, rgba, mask, params, output):
        """lib/src/lib.rs:188-234: `rgba` is 16 RGBA8 pixels (64 bytes or 16x4 values)."""
        flat = bytes(bytearray(v for px in rgba for v in px)) if _is_nested(rgba) else bytes(rgba)
        if len(flat) != 64:
            raise ValueError("a block is 16 RGBA8 pixels")
        p = params._c()
        ip, _, _k1 = _ptr(flat)
        op, ol, _k2 = _ptr(output, writable=True)
        _check(lib().tx_compress_block_masked(int(self), ip, mask & 0xFFFFFFFF, ctypes.byref(p), op, ol),
               "compress_block_masked")

    def decompress_block(self, block):
        """lib/src/lib.rs:240-277: returns 16 [r, g, b, a] lists."""
        ip, il, _k = _ptr(bytes(block))
        out = (ctypes.c_uint8 * 64)()
        _check(lib().tx_decompress_block(int(self), ip, il, out), "decompress_block")
        return [list(out[4 * i:4 * i + 4]) for i in range(16)]

    # -- device-resident entry points (pointers are device addresses) --------

    def compress_device(self, d_rgba, width, height, params, d_output, out_len, stream=None):
        p = params._c()
        _check(lib().tx_compress_device(int(self), d_rgba, width, height, ctypes.byref(p), d_output, out_len,
                                        stream), "compress_device")

    def decompress_device(self, d_data, width, height, d_output, stream=None):
        _check(lib().tx_decompress_device(int(self), d_data, width, height, d_output, stream), "decompress_device")


def _is_nested(x):
    try:
        return len(x) == 16 and hasattr(x[0], "__len__")
    except TypeError:
        return False


class DeviceBuffer:
    """Device memory on the current GPU (tx_util_malloc); `ptr` is the address."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().tx_util_malloc(ctypes.byref(p), nbytes), "device malloc")
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, arr):
        import numpy as np
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr):
        import numpy as np
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _check(lib().tx_util_upload(self.ptr, arr.ctypes.data, arr.nbytes), "upload")

    def download(self, nbytes=None):
        import numpy as np
        out = np.empty(self.nbytes if nbytes is None else nbytes, np.uint8)
        _check(lib().tx_util_download(out.ctypes.data, self.ptr, out.nbytes), "download")
        return out

    def free(self):
        if self.ptr:
            lib().tx_util_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def time_compress(fmt, d_rgba, width, height, params, d_out, out_len, iters):
    """Milliseconds for `iters` back-to-back device launches (HIP events)."""
    ms = ctypes.c_float()
    p = params._c()
    _check(lib().tx_util_time_compress(int(fmt), d_rgba, width, height, ctypes.byref(p), d_out, out_len, iters,
                                       ctypes.byref(ms)), "time_compress")
    return ms.value


def compress_blocks(fmt, blocks, masks, params):
    """Encode many masked 4x4 blocks in one launch (tx_compress_blocks).

    `blocks` is a uint8 array of shape (n, 64), `masks` a uint32 array (n,).
    Returns a uint8 array (n, block_size)."""
    import numpy as np
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
    masks = np.ascontiguousarray(masks, dtype=np.uint32).reshape(-1)
    if masks.size != blocks.shape[0]:
        raise ValueError("one mask per block")
    fmt = Format(fmt)
    out = np.zeros((blocks.shape[0], fmt.block_size()), np.uint8)
    p = params._c()
    _check(lib().tx_compress_blocks(int(fmt), blocks.ctypes.data, masks.ctypes.data, blocks.shape[0],
                                    ctypes.byref(p), out.ctypes.data), "compress_blocks")
    return out
