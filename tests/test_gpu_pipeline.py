"""The binned ClusterFit pipeline (classify -> sort by colour count ->
independent per-block searches) gives the oracle's bytes, and the same bytes
as the direct lockstep kernel, for every format, algorithm and size."""
import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import texgen
from tests.blockgen import make_blocks

pytestmark = pytest.mark.gpu

ALGS = [tx.Algorithm.ClusterFit, tx.Algorithm.IterativeClusterFit]
NEVER = 2 ** 64 - 1


@pytest.fixture
def pipeline_always(gpu):
    old = tx.set_pipeline_threshold(0)
    yield
    tx.set_pipeline_threshold(old)


def encode(fmt, img, w, h, p, out_len=None):
    f = tx.Format(fmt)
    out = np.zeros(out_len if out_len is not None else f.compressed_size(w, h), np.uint8)
    f.compress(np.ascontiguousarray(img).reshape(-1), w, h, p, out)
    return out


@pytest.mark.parametrize("fmt", [0, 1, 2], ids=["bc1", "bc2", "bc3"])
@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.name)
@pytest.mark.parametrize("size", [(64, 64), (131, 77), (5, 3), (1, 1), (37, 128)], ids=lambda s: "%dx%d" % s)
def test_pipeline_matches_oracle(pipeline_always, oracle, fmt, alg, size):
    w, h = size
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 31 + w * 3 + h), 7 + w)
    p = tx.Params(alg)
    got = encode(fmt, img, w, h, p)
    want = oracle.compress(fmt, img, w, h, oracle.params(int(alg), p.weights, False), threads=oracle.default_threads())
    assert np.array_equal(got, want)


@pytest.mark.parametrize("fmt", [0, 2], ids=["bc1", "bc3"])
@pytest.mark.parametrize("alpha", [False, True], ids=["plain", "alphaw"])
def test_pipeline_adversarial_blocks_as_image(pipeline_always, oracle, fmt, alpha):
    # the adversarial block set laid out as one image (masks become full)
    blocks, _ = make_blocks(77 + fmt, per_class=64)
    nb = len(blocks) - len(blocks) % 16
    img = blocks[:nb].reshape(nb // 16, 16, 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(nb // 16 * 4, 64, 4)
    h, w = img.shape[:2]
    p = tx.Params(tx.Algorithm.IterativeClusterFit, (0.9, 0.05, 0.37), alpha)
    got = encode(fmt, img, w, h, p)
    want = oracle.compress(fmt, img, w, h, oracle.params(2, p.weights, alpha), threads=oracle.default_threads())
    assert np.array_equal(got, want)


def test_pipeline_classify_every_fit_on_adversarial_blocks(gpu, oracle):
    # every format x algorithm x alpha weighting in one pipeline batch, so
    # the one-block-per-lane classify takes each fit path (single colour,
    # range fit, ClusterFit records, transparent BC1, NaN axes, ties) and
    # RangeFit textures (which single-texture calls never send here) as well
    items = []
    for fmt in (0, 1, 2):
        blocks, _ = make_blocks(301 + fmt, per_class=24)
        nb = len(blocks) - len(blocks) % 16
        img = np.ascontiguousarray(blocks[:nb].reshape(nb // 16, 16, 4, 4, 4).transpose(0, 2, 1, 3, 4)
                                   .reshape(nb // 16 * 4, 64, 4))
        for alg in (0, 1, 2):
            for alpha in (False, True):
                items.append((fmt, tx.Params(tx.Algorithm(alg), (0.9, 0.05, 0.37), alpha), img))
    bufs, descs = [], []
    for fmt, p, img in items:
        h, w = img.shape[:2]
        d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
        n = tx.Format(fmt).compressed_size(w, h)
        d_out = tx.DeviceBuffer(n)
        bufs.append((d_in, d_out, n))
        descs.append((fmt, p, d_in.ptr, d_out.ptr, w, h))
    batch = tx.Batch(descs, tx.Tuning(chunk_blocks=64, waves_per_group=1, groups_per_cu=12, use_graph=False,
                                      pipeline=True))
    batch.run()
    for (fmt, p, img), (_, d_out, n) in zip(items, bufs):
        h, w = img.shape[:2]
        op = oracle.params(int(p.algorithm), p.weights, p.weigh_colour_by_alpha)
        want = oracle.compress(fmt, img, w, h, op, threads=oracle.default_threads())
        assert np.array_equal(d_out.download(n), want), (fmt, p.algorithm, p.weigh_colour_by_alpha)
    batch.close()


def test_pipeline_oversized_output(pipeline_always, oracle):
    img = texgen.rgb_texture(9, 6, 4)
    n = tx.Format.Bc1.compressed_size(9, 6) + 5 * 8
    got = encode(0, img, 9, 6, tx.Params(), out_len=n)
    want = oracle.compress(0, img, 9, 6, oracle.params(1), out_len=n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("fmt,alg", [(2, tx.Algorithm.IterativeClusterFit), (0, tx.Algorithm.ClusterFit),
                                     (0, tx.Algorithm.IterativeClusterFit)], ids=["bc3-iter", "bc1", "bc1-iter"])
def test_pipeline_equals_direct_kernel_at_full_size(gpu, fmt, alg):
    # a size-independent property at the benchmark size: both device paths agree byte for byte
    w = h = 1024
    seed = texgen.SEED_BASE + 1
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, seed), seed)
    p = tx.Params(alg)
    old = tx.set_pipeline_threshold(0)
    try:
        a = encode(fmt, img, w, h, p)
        tx.set_pipeline_threshold(NEVER)
        b = encode(fmt, img, w, h, p)
    finally:
        tx.set_pipeline_threshold(old)
    assert np.array_equal(a, b)


def test_pipeline_device_entry_repeats(pipeline_always, oracle):
    w, h = 200, 120
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 12), 12)
    d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
    n = tx.Format.Bc3.compressed_size(w, h)
    d_out = tx.DeviceBuffer(n)
    p = tx.Params(tx.Algorithm.IterativeClusterFit)
    want = oracle.compress(2, img, w, h, oracle.params(2))
    for _ in range(3):
        d_out.upload(np.full(n, 0x5A, np.uint8))
        tx.Format.Bc3.compress_device(d_in.ptr, w, h, p, d_out.ptr, n)
        assert np.array_equal(d_out.download(), want)


@pytest.mark.parametrize("fmt", [0, 2], ids=["bc1", "bc3"])
@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.name)
@pytest.mark.parametrize("alpha", [False, True], ids=["plain", "alphaw"])
def test_pipeline_random_weights(pipeline_always, oracle, fmt, alg, alpha):
    # the records carry Params::weights to the searches: seeded non-uniform,
    # tiny, large and one-zero-channel triples through the binned pipeline
    from tests.test_gpu_parity import _random_weights
    w = h = 256
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 600 + fmt + int(alg)), 601 + fmt)
    for weights in _random_weights(70 + fmt + 2 * int(alg) + (1 if alpha else 0)):
        p = tx.Params(alg, weights, alpha)
        print("pipeline weights:", fmt, alg.name, alpha, weights, flush=True)
        got = encode(fmt, img, w, h, p)
        want = oracle.compress(fmt, img, w, h, oracle.params(int(alg), p.weights, alpha),
                               threads=oracle.default_threads())
        assert np.array_equal(got, want), weights


def test_pipeline_scratch_shared_across_streams(pipeline_always, oracle):
    # The pipeline's record buffers are shared by every stream of the device:
    # calls on callers' streams wait for the previous use and record theirs,
    # calls on the library stream are ordered by that stream and recorded only
    # when another stream comes next.  Alternate between two caller streams
    # and the library stream with nothing synchronised in between, and check
    # every texture against the oracle.
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    streams = []
    for _ in range(2):
        s = ctypes.c_void_p()
        assert hip.hipStreamCreateWithFlags(ctypes.byref(s), 1) == 0  # non-blocking, default priority
        streams.append(s)
    try:
        order = [streams[0], None, streams[1], None, None, streams[0], streams[1], None]
        jobs = []
        for k, s in enumerate(order):
            w, h = 128, 96 + 4 * k
            fmt = tx.Format.Bc3 if k % 2 == 0 else tx.Format.Bc1
            img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 900 + k), 40 + k)
            p = tx.Params(tx.Algorithm.IterativeClusterFit)
            d_in = tx.DeviceBuffer.from_array(np.ascontiguousarray(img).reshape(-1))
            n = fmt.compressed_size(w, h)
            d_out = tx.DeviceBuffer(n)
            jobs.append((fmt, img, w, h, p, d_in, d_out, n))
        for (fmt, img, w, h, p, d_in, d_out, n), s in zip(jobs, order):
            fmt.compress_device(d_in.ptr, w, h, p, d_out.ptr, n, s)
        assert hip.hipDeviceSynchronize() == 0
        for fmt, img, w, h, p, d_in, d_out, n in jobs:
            want = oracle.compress(int(fmt), img, w, h, oracle.params(2, p.weights, False),
                                   threads=oracle.default_threads())
            assert np.array_equal(d_out.download(), want)
    finally:
        assert hip.hipDeviceSynchronize() == 0
        for s in streams:
            hip.hipStreamDestroy(s)
