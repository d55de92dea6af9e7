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
    "num_blocks", "lib", "library_path", "mip_levels", "mip_dims",
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


class _TxTexture(ctypes.Structure):
    _fields_ = [
        ("format", ctypes.c_int32),
        ("params", _TxParams),
        ("d_rgba", ctypes.c_void_p),
        ("d_output", ctypes.c_void_p),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
    ]


class _TxTuning(ctypes.Structure):
    _fields_ = [
        ("chunk_blocks", ctypes.c_uint32),
        ("waves_per_group", ctypes.c_uint32),
        ("groups_per_cu", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
        ("pipeline", ctypes.c_uint32),
        ("lanes_per_block", ctypes.c_uint32),
    ]


class _TxServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_size_t),
        ("max_inflight", ctypes.c_uint32),
        ("max_jobs_per_launch", ctypes.c_uint32),
        ("max_pending", ctypes.c_uint32),
        ("tuning", _TxTuning),
    ]


class _TxJob(ctypes.Structure):
    _fields_ = [
        ("format", ctypes.c_int32),
        ("params", _TxParams),
        ("rgba", ctypes.c_void_p),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
    ]


class _TxJobInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("t_submit_ns", ctypes.c_uint64),
        ("t_admit_ns", ctypes.c_uint64),
        ("t_launch_ns", ctypes.c_uint64),
        ("t_done_ns", ctypes.c_uint64),
        ("d_output", ctypes.c_void_p),
        ("out_len", ctypes.c_size_t),
        ("launch_id", ctypes.c_uint32),
        ("launch_jobs", ctypes.c_uint32),
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


def cli_path():
    """The `texpresso` command-line tool built beside the library."""
    return os.path.join(HERE, "texpresso")


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
            "tx_compress_device_noalloc": (c_int, [c_int, c_vp, c_sz, c_sz, pp, c_vp, c_sz, c_vp]),
            "tx_single_lut": (None, [c_u8p]),
            "tx_util_malloc": (c_int, [ctypes.POINTER(ctypes.c_void_p), c_sz]),
            "tx_util_free": (c_int, [c_vp]),
            "tx_util_upload": (c_int, [c_vp, c_vp, c_sz]),
            "tx_util_download": (c_int, [c_vp, c_vp, c_sz]),
            "tx_util_copy": (c_int, [c_vp, c_vp, c_sz]),
            "tx_util_elementary": (c_int, [c_vp, c_sz, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
            "tx_set_pipeline_threshold": (ctypes.c_uint64, [ctypes.c_uint64]),
            "tx_util_time_compress": (c_int, [c_int, c_vp, c_sz, c_sz, pp, c_vp, c_sz, c_int,
                                              ctypes.POINTER(ctypes.c_float)]),
            "tx_util_time_phases": (c_int, [c_int, c_vp, c_sz, c_sz, pp, c_vp, c_sz, c_int,
                                            ctypes.POINTER(ctypes.c_float), ctypes.POINTER(c_int)]),
            "tx_batch_tuning_default": (_TxTuning, []),
            "tx_batch_create": (c_int, [ctypes.POINTER(c_vp), ctypes.POINTER(_TxTexture), c_sz,
                                        ctypes.POINTER(_TxTuning)]),
            "tx_batch_run": (c_int, [c_vp, c_vp]),
            "tx_batch_run_timed": (c_int, [c_vp, c_int, ctypes.POINTER(ctypes.c_float)]),
            "tx_batch_destroy": (c_int, [c_vp]),
            "tx_batch_info": (c_int, [c_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
            "tx_batch_autotune": (c_int, [ctypes.POINTER(_TxTexture), c_sz, ctypes.POINTER(_TxTuning), c_sz, c_int,
                                          ctypes.POINTER(_TxTuning), ctypes.POINTER(ctypes.c_float)]),
            "tx_util_time_per_texture": (c_int, [ctypes.POINTER(_TxTexture), c_sz, c_int,
                                                 ctypes.POINTER(ctypes.c_float)]),
            "tx_server_config_default": (_TxServerConfig, []),
            "tx_server_create": (c_int, [ctypes.POINTER(c_vp), ctypes.POINTER(_TxServerConfig)]),
            "tx_server_submit": (c_int, [c_vp, ctypes.POINTER(_TxJob), ctypes.POINTER(ctypes.c_uint64)]),
            "tx_server_wait": (c_int, [c_vp, ctypes.c_uint64, c_int, ctypes.POINTER(_TxJobInfo)]),
            "tx_server_query": (c_int, [c_vp, ctypes.c_uint64, ctypes.POINTER(_TxJobInfo)]),
            "tx_server_fetch": (c_int, [c_vp, ctypes.c_uint64, c_vp, c_sz]),
            "tx_server_release": (c_int, [c_vp, ctypes.c_uint64]),
            "tx_server_destroy": (c_int, [c_vp]),
            "tx_mip_levels": (ctypes.c_uint32, [c_sz, c_sz]),
            "tx_mip_dims": (None, [c_sz, c_sz, ctypes.c_uint32, ctypes.POINTER(c_sz), ctypes.POINTER(c_sz)]),
            "tx_mip_chain_rgba_bytes": (c_sz, [c_sz, c_sz, ctypes.c_uint32]),
            "tx_mip_chain_compressed_size": (c_sz, [c_int, c_sz, c_sz, ctypes.c_uint32]),
            "tx_generate_mips_device": (c_int, [c_vp, c_sz, c_sz, ctypes.c_uint32, c_vp, c_sz, c_vp]),
            "tx_compress_mips_device": (c_int, [c_int, c_vp, c_sz, c_sz, pp, ctypes.c_uint32, c_vp, c_sz, c_vp, c_sz,
                                                c_vp]),
            "tx_compress_mips": (c_int, [c_int, c_u8p, c_sz, c_sz, c_sz, pp, ctypes.c_uint32, c_u8p, c_sz]),
            "tx_dds_header_size": (c_sz, []),
            "tx_dds_write_header": (c_int, [c_int, c_sz, c_sz, ctypes.c_uint32, c_u8p, c_sz]),
            "tx_dds_parse_header": (c_int, [c_u8p, c_sz, ctypes.POINTER(c_int), ctypes.POINTER(c_sz),
                                            ctypes.POINTER(c_sz), ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(c_sz)]),
            "tx_png_info": (c_int, [c_u8p, c_sz, ctypes.POINTER(c_sz), ctypes.POINTER(c_sz)]),
            "tx_png_decode": (c_int, [c_u8p, c_sz, c_u8p, c_sz]),
            "tx_png_encode_bound": (c_sz, [c_sz, c_sz]),
            "tx_png_encode": (c_int, [c_u8p, c_sz, c_sz, c_u8p, c_sz, ctypes.POINTER(c_sz)]),
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

    # -- mip chains (no counterpart in texpresso: it encodes one level) ------

    def mip_chain_size(self, width, height, levels=None):
        """Compressed bytes of levels 0..levels-1 back to back (default: full chain)."""
        levels = mip_levels(width, height) if levels is None else levels
        return lib().tx_mip_chain_compressed_size(int(self), width, height, levels)

    def compress_mips(self, rgba, width, height, params, levels=None):
        """Generate the mip chain of `rgba` on the GPU and compress every level;
        returns the levels' blocks back to back (level 0 first) as a uint8 array."""
        import numpy as np
        levels = mip_levels(width, height) if levels is None else levels
        out = np.zeros(max(1, self.mip_chain_size(width, height, levels)), np.uint8)
        p = params._c()
        ip, il, _k1 = _ptr(rgba)
        _check(lib().tx_compress_mips(int(self), ip, il, width, height, ctypes.byref(p), levels, out.ctypes.data,
                                      out.nbytes), "compress_mips")
        return out[:self.mip_chain_size(width, height, levels)]

    def compress_mips_device(self, d_rgba, width, height, params, levels, d_chain, chain_len, d_output, out_len,
                             stream=None):
        p = params._c()
        _check(lib().tx_compress_mips_device(int(self), d_rgba, width, height, ctypes.byref(p), levels, d_chain,
                                             chain_len, d_output, out_len, stream), "compress_mips_device")


def mip_levels(width, height):
    """Levels of a full mip chain down to 1x1."""
    return lib().tx_mip_levels(width, height)


def mip_dims(width, height, level):
    w, h = ctypes.c_size_t(), ctypes.c_size_t()
    lib().tx_mip_dims(width, height, level, ctypes.byref(w), ctypes.byref(h))
    return w.value, h.value


def mip_chain_rgba_bytes(width, height, levels):
    return lib().tx_mip_chain_rgba_bytes(width, height, levels)


def generate_mips_device(d_rgba, width, height, levels, d_chain, chain_len, stream=None):
    """Levels 1..levels-1 of the device image d_rgba into d_chain."""
    _check(lib().tx_generate_mips_device(d_rgba, width, height, levels, d_chain, chain_len, stream),
           "generate_mips_device")


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


def set_pipeline_threshold(min_blocks):
    """ClusterFit textures with at least `min_blocks` blocks use the binned
    pipeline (0: always, 2**64-1: never).  Returns the previous value."""
    return lib().tx_set_pipeline_threshold(int(min_blocks))


def elementary(values):
    """The device's evaluation of the encoder's exactly specified operations
    (tx_util_elementary): a dict with the u8/255 table, sqrtf of the counts
    0..16 and of k/256 for k = 1..4096, and for each float32 in `values` the
    5-/6-bit grid snaps and roundf clamped to [0, 255]."""
    import numpy as np
    v = np.ascontiguousarray(values, dtype=np.float32)
    n = int(v.size)
    u8f = np.zeros(256, np.float32)
    sqc = np.zeros(17, np.float32)
    sqa = np.zeros(4096, np.float32)
    s31 = np.zeros(max(n, 1), np.float32)
    s63 = np.zeros(max(n, 1), np.float32)
    r255 = np.zeros(max(n, 1), np.int32)
    _check(lib().tx_util_elementary(v.ctypes.data if n else None, n, u8f.ctypes.data, sqc.ctypes.data,
                                    sqa.ctypes.data, s31.ctypes.data, s63.ctypes.data, r255.ctypes.data),
           "elementary")
    return {"u8f": u8f, "sqrt_counts": sqc, "sqrt_alpha": sqa, "snap31": s31[:n], "snap63": s63[:n],
            "round255": r255[:n]}


def time_compress(fmt, d_rgba, width, height, params, d_out, out_len, iters):
    """Milliseconds for `iters` back-to-back device launches (HIP events)."""
    ms = ctypes.c_float()
    p = params._c()
    _check(lib().tx_util_time_compress(int(fmt), d_rgba, width, height, ctypes.byref(p), d_out, out_len, iters,
                                       ctypes.byref(ms)), "time_compress")
    return ms.value


def time_phases(fmt, d_rgba, width, height, params, d_out, out_len, iters):
    """Per-phase milliseconds of the binned ClusterFit pipeline summed over
    `iters` launches: {"classify", "bin", "search3", "search4"}, or None when
    the texture goes through the direct kernel."""
    ms = (ctypes.c_float * 4)()
    used = ctypes.c_int()
    p = params._c()
    _check(lib().tx_util_time_phases(int(fmt), d_rgba, width, height, ctypes.byref(p), d_out, out_len, iters, ms,
                                     ctypes.byref(used)), "time_phases")
    if not used.value:
        return None
    return dict(zip(("classify", "bin", "search3", "search4"), (ms[0], ms[1], ms[2], ms[3])))


class Tuning:
    """Batch kernel tuning: blocks per queue chunk, waves (x4 blocks) per
    workgroup, persistent workgroups per CU, hipGraph replay, whether
    ClusterFit blocks go through the binned pipeline, and the lanes that
    encode one BC4/BC5 block (1 or 16)."""

    __slots__ = ("chunk_blocks", "waves_per_group", "groups_per_cu", "use_graph", "pipeline", "lanes_per_block")

    def __init__(self, chunk_blocks=64, waves_per_group=1, groups_per_cu=12, use_graph=True, pipeline=False,
                 lanes_per_block=1):
        self.chunk_blocks = int(chunk_blocks)
        self.waves_per_group = int(waves_per_group)
        self.groups_per_cu = int(groups_per_cu)
        self.use_graph = bool(use_graph)
        self.pipeline = bool(pipeline)
        self.lanes_per_block = int(lanes_per_block)

    def _c(self):
        return _TxTuning(self.chunk_blocks, self.waves_per_group, self.groups_per_cu, 1 if self.use_graph else 0,
                         1 if self.pipeline else 0, self.lanes_per_block)

    @classmethod
    def _from_c(cls, t):
        return cls(t.chunk_blocks, t.waves_per_group, t.groups_per_cu, bool(t.use_graph), bool(t.pipeline),
                   t.lanes_per_block or 1)

    def as_dict(self):
        return {k: getattr(self, k) for k in self.__slots__}

    def __repr__(self):
        return "Tuning(%s)" % ", ".join("%s=%r" % kv for kv in self.as_dict().items())


def _textures(items):
    """items: iterable of (format, params, d_rgba_ptr, d_out_ptr, width, height)."""
    items = list(items)
    arr = (_TxTexture * max(1, len(items)))()
    for i, (fmt, params, d_in, d_out, w, h) in enumerate(items):
        arr[i].format = int(fmt)
        arr[i].params = params._c()
        arr[i].d_rgba = d_in
        arr[i].d_output = d_out
        arr[i].width = w
        arr[i].height = h
    return arr, len(items)


class Batch:
    """Many textures / mip levels encoded by one persistent launch
    (tx_batch_*).  `textures` holds (format, params, d_rgba, d_output, w, h)
    with device addresses (e.g. DeviceBuffer.ptr)."""

    def __init__(self, textures, tuning=None):
        self._arr, self._n = _textures(textures)
        h = ctypes.c_void_p()
        t = tuning._c() if tuning is not None else None
        _check(lib().tx_batch_create(ctypes.byref(h), self._arr, self._n, ctypes.byref(t) if t else None),
               "batch create")
        self._h = h.value
        self.tuning = tuning or Tuning._from_c(lib().tx_batch_tuning_default())

    def run(self, stream=None):
        _check(lib().tx_batch_run(self._h, stream), "batch run")

    def run_timed(self, iters=1):
        ms = ctypes.c_float()
        _check(lib().tx_batch_run_timed(self._h, iters, ctypes.byref(ms)), "batch run")
        return ms.value

    def info(self):
        c, b = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().tx_batch_info(self._h, ctypes.byref(c), ctypes.byref(b)), "batch info")
        return {"chunks": c.value, "blocks": b.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().tx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def time_per_texture(textures, iters=1):
    """Milliseconds for `iters` passes of one plain launch per texture."""
    arr, n = _textures(textures)
    ms = ctypes.c_float()
    _check(lib().tx_util_time_per_texture(arr, n, iters, ctypes.byref(ms)), "per-texture timing")
    return ms.value


def autotune(textures, candidates=None, iters=3):
    """Times every candidate tuning (default: the built-in grid of chunk size
    x waves per workgroup x workgroups per CU) and returns (best Tuning, ms)."""
    arr, n = _textures(textures)
    best, ms = _TxTuning(), ctypes.c_float()
    if candidates:
        cand = (_TxTuning * len(candidates))(*[c._c() for c in candidates])
        nc = len(candidates)
    else:
        cand, nc = None, 0
    _check(lib().tx_batch_autotune(arr, n, cand, nc, iters, ctypes.byref(best), ctypes.byref(ms)), "autotune")
    return Tuning._from_c(best), ms.value


class JobInfo:
    """Result record of a served job (timestamps in CLOCK_MONOTONIC ns, the
    clock of time.monotonic_ns())."""

    def __init__(self, c):
        self.status = c.status
        self.t_submit_ns, self.t_admit_ns = c.t_submit_ns, c.t_admit_ns
        self.t_launch_ns, self.t_done_ns = c.t_launch_ns, c.t_done_ns
        self.d_output = c.d_output
        self.out_len = c.out_len
        self.launch_id, self.launch_jobs = c.launch_id, c.launch_jobs


class Server:
    """Continuous-batching compression server (tx_server_*): submit host
    textures, jobs are admitted by free device memory and folded into
    persistent-kernel launches that overlap on the GPU."""

    def __init__(self, hbm_budget=None, max_inflight=None, max_jobs_per_launch=None, max_pending=None,
                 tuning=None):
        cfg = lib().tx_server_config_default()
        if hbm_budget is not None:
            cfg.hbm_budget = int(hbm_budget)
        if max_inflight is not None:
            cfg.max_inflight = int(max_inflight)
        if max_jobs_per_launch is not None:
            cfg.max_jobs_per_launch = int(max_jobs_per_launch)
        if max_pending is not None:
            cfg.max_pending = int(max_pending)
        if tuning is not None:
            t = tuning._c()
            t.use_graph = 0
            cfg.tuning = t
        h = ctypes.c_void_p()
        _check(lib().tx_server_create(ctypes.byref(h), ctypes.byref(cfg)), "server create")
        self._h = h.value
        self._keep = {}

    def submit(self, fmt, rgba, width, height, params):
        """Queue one texture; `rgba` (numpy uint8, width*height*4) is kept alive
        until the job is released.  Returns the job id."""
        import numpy as np
        arr = np.ascontiguousarray(rgba, dtype=np.uint8).reshape(-1)
        job = _TxJob(int(fmt), params._c(), arr.ctypes.data, width, height)
        jid = ctypes.c_uint64()
        _check(lib().tx_server_submit(self._h, ctypes.byref(job), ctypes.byref(jid)), "submit")
        self._keep[jid.value] = arr
        return jid.value

    def wait(self, jid, timeout_ms=-1):
        info = _TxJobInfo()
        _check(lib().tx_server_wait(self._h, jid, timeout_ms, ctypes.byref(info)), "wait")
        return JobInfo(info)

    def query(self, jid):
        info = _TxJobInfo()
        _check(lib().tx_server_query(self._h, jid, ctypes.byref(info)), "query")
        return JobInfo(info)

    def fetch(self, jid, out_len):
        import numpy as np
        out = np.zeros(out_len, np.uint8)
        _check(lib().tx_server_fetch(self._h, jid, out.ctypes.data, out_len), "fetch")
        return out

    def release(self, jid):
        _check(lib().tx_server_release(self._h, jid), "release")
        self._keep.pop(jid, None)

    def close(self):
        if getattr(self, "_h", None):
            lib().tx_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TUNING_TABLE = os.path.join(HERE, "tuning.json")


def tuned(signature, textures=None, table_path=TUNING_TABLE, retune=False, iters=3):
    """Tuning for a workload `signature` from the autotuner table; runs the
    autotuner on `textures` (and records the result) when absent."""
    import json
    table = {}
    if os.path.exists(table_path):
        with open(table_path) as f:
            table = json.load(f)
    if signature in table and not retune:
        return Tuning(**table[signature]["tuning"]), table[signature].get("ms")
    if textures is None:
        return Tuning(), None
    best, ms = autotune(textures, iters=iters)
    table[signature] = {"tuning": best.as_dict(), "ms": ms}
    try:
        with open(table_path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return best, ms


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
