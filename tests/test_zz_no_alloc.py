"""tx_compress_device_noalloc, the entry point that allocates nothing: its
output equals tx_compress_device's (and the oracle's) on two calls in a row
that both go through the binned pipeline's fixed scratch, the second with a
smaller texture; textures the scratch cannot hold are refused before
anything runs, and paths that use no scratch have no size limit."""
import ctypes

import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import texgen

pytestmark = pytest.mark.gpu

MAX_BLOCKS = 262144  # TX_NOALLOC_MAX_PIPELINE_BLOCKS
ERR_TOO_LARGE = -12


def noalloc(fmt, d_in, w, h, p, d_out, n):
    pc = p._c()
    return tx.lib().tx_compress_device_noalloc(int(fmt), d_in.ptr, w, h, ctypes.byref(pc), d_out.ptr, n, None)


def case(fmt, alg, w, h, seed):
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, seed), seed)
    n = tx.Format(fmt).compressed_size(w, h)
    return img, tx.DeviceBuffer.from_array(img.reshape(-1)), n


def device_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_noalloc_matches_compress_device_two_calls(gpu, oracle):
    # 1: the 2048 x 2048 BC3 IterativeClusterFit texture (262,144 blocks, the
    #    scratch's full capacity) through the pipeline at the default threshold;
    # 2: a smaller BC1 ClusterFit texture, pipeline forced by a zero threshold
    #    (the scratch still holds the first call's records)
    calls = [(tx.Format.Bc3, tx.Algorithm.IterativeClusterFit, 2048, 2048, texgen.SEED_BASE + 1, None),
             (tx.Format.Bc1, tx.Algorithm.ClusterFit, 301, 187, 77, 0)]
    for fmt, alg, w, h, seed, threshold in calls:
        print("noalloc %s %s %dx%d" % (fmt.name, alg.name, w, h), flush=True)
        img, d_in, n = case(fmt, alg, w, h, seed)
        p = tx.Params(alg)
        d_ref, d_got = tx.DeviceBuffer(n), tx.DeviceBuffer(n)
        d_got.upload(np.full(n, 0xA5, np.uint8))
        old = tx.set_pipeline_threshold(threshold) if threshold is not None else None
        try:
            fmt.compress_device(d_in.ptr, w, h, p, d_ref.ptr, n)
            ref = d_ref.download(n)
            assert noalloc(fmt, d_in, w, h, p, d_got, n) == 0
            got = d_got.download(n)
        finally:
            if old is not None:
                tx.set_pipeline_threshold(old)
        bs = fmt.block_size()
        bad = int(np.count_nonzero(np.any(got.reshape(-1, bs) != ref.reshape(-1, bs), axis=1)))
        assert bad == 0, "%d/%d blocks differ from tx_compress_device" % (bad, n // bs)
        if w * h <= 1 << 16:
            want = oracle.compress(int(fmt), img, w, h, oracle.params(int(alg), p.weights, False))
            assert np.array_equal(got, want)


def test_noalloc_limits_and_no_device_allocation(gpu):
    # a texture with more pipeline blocks than the scratch holds is refused
    w, h = 4096, 2048
    assert (w // 4) * (h // 4) > MAX_BLOCKS
    d_in = tx.DeviceBuffer(4 * w * h)
    n = tx.Format.Bc3.compressed_size(w, h)
    d_out = tx.DeviceBuffer(n)
    p = tx.Params(tx.Algorithm.IterativeClusterFit)
    old = tx.set_pipeline_threshold(0)
    try:
        assert noalloc(tx.Format.Bc3, d_in, w, h, p, d_out, n) == ERR_TOO_LARGE
    finally:
        tx.set_pipeline_threshold(old)
    # RangeFit uses no scratch: any size runs (and matches tx_compress_device)
    img, d_img, n1 = case(tx.Format.Bc1, tx.Algorithm.RangeFit, 1200, 1100, 91)
    pr = tx.Params(tx.Algorithm.RangeFit)
    a, b = tx.DeviceBuffer(n1), tx.DeviceBuffer(n1)
    tx.Format.Bc1.compress_device(d_img.ptr, 1200, 1100, pr, a.ptr, n1)
    assert noalloc(tx.Format.Bc1, d_img, 1200, 1100, pr, b, n1) == 0
    assert np.array_equal(a.download(n1), b.download(n1))
    # repeated calls of both kinds leave the device's free memory unchanged
    small, d_small, ns = case(tx.Format.Bc3, tx.Algorithm.ClusterFit, 640, 480, 92)
    pc = tx.Params(tx.Algorithm.ClusterFit)
    d_out2 = tx.DeviceBuffer(ns)
    old = tx.set_pipeline_threshold(0)
    try:
        assert noalloc(tx.Format.Bc3, d_small, 640, 480, pc, d_out2, ns) == 0
        assert ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize() == 0
        before = device_free_bytes()
        for _ in range(3):
            assert noalloc(tx.Format.Bc3, d_small, 640, 480, pc, d_out2, ns) == 0
            assert noalloc(tx.Format.Bc1, d_img, 1200, 1100, pr, b, n1) == 0
        assert ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize() == 0
        after = device_free_bytes()
    finally:
        tx.set_pipeline_threshold(old)
    print("device free bytes before/after repeated calls: %d / %d" % (before, after), flush=True)
    assert after >= before
