"""Alpha codebook fits where distances tie: every BC3/BC4/BC5 alpha path
(the 16-lane group kernel, the one-block-per-lane BC4/BC5 kernel and the
pipeline's classify pass) against the oracle on blocks whose values sit on
narrow ranges, so that many pixels lie exactly halfway between two codes or
on a code shared by two slots (alpha.rs:79-118 keeps the first strict
minimum), plus ranges that touch 0 and 255 (the 5-point book's extra codes)
and ranges narrower than the books' minimum spans (alpha.rs:223-236)."""
import numpy as np
import pytest

import texpresso_amd as tx
from tests.blockgen import RECT_MASKS

pytestmark = pytest.mark.gpu


def tie_blocks(seed, count=4096):
    """Blocks whose R, G and A channels each hold 16 values drawn from a
    narrow range [lo, lo + span]; RGB random elsewhere."""
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 256, size=(count, 16, 4), dtype=np.int64)
    for ch in (0, 1, 3):
        span = rng.integers(0, 24, size=count)
        span[: count // 8] = rng.integers(0, 8, size=count // 8)  # below the books' minimum spans
        lo = rng.integers(0, 256, size=count)
        lo = np.minimum(lo, 255 - span)
        edge = rng.random(count)
        lo = np.where(edge < 0.1, 0, lo)  # ranges that start at 0
        lo = np.where(edge > 0.9, 255 - span, lo)  # ranges that end at 255
        vals = lo[:, None] + rng.integers(0, 1 << 30, size=(count, 16)) % (span[:, None] + 1)
        px[:, :, ch] = vals
    blocks = px.astype(np.uint8).reshape(count, 64)
    masks = np.full(count, 0xFFFF, np.uint32)
    masks[::5] = rng.choice(np.array(RECT_MASKS, np.uint32), size=len(masks[::5]))
    return blocks, masks


@pytest.mark.parametrize("fmt", [2, 3, 4], ids=["bc3", "bc4", "bc5"])
def test_alpha_ties_blocks(gpu, oracle, fmt):
    blocks, masks = tie_blocks(900 + fmt)
    p = tx.Params(tx.Algorithm.RangeFit)
    print("alpha ties:", fmt, len(blocks), flush=True)
    got = tx.compress_blocks(fmt, blocks, masks, p)
    want = oracle.compress_blocks(fmt, blocks, masks, oracle.params(int(p.algorithm), p.weights, False))
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert bad.size == 0, "%d blocks differ, first %d" % (bad.size, bad[0])


@pytest.mark.parametrize("fmt", [2, 3, 4], ids=["bc3", "bc4", "bc5"])
def test_alpha_ties_texture(gpu, oracle, fmt):
    # the same blocks tiled into a 256 x 256 texture: the texture kernels,
    # and for BC3 the pipeline's classify pass (forced on by a zero threshold)
    blocks, _ = tie_blocks(950 + fmt)
    w = h = 256
    img = blocks.reshape(h // 4, w // 4, 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(h, w, 4)
    img = np.ascontiguousarray(img)
    p = tx.Params(tx.Algorithm.ClusterFit)
    f = tx.Format(fmt)
    old = tx.set_pipeline_threshold(0)
    try:
        got = np.zeros(f.compressed_size(w, h), np.uint8)
        f.compress(img.reshape(-1), w, h, p, got)
    finally:
        tx.set_pipeline_threshold(old)
    want = oracle.compress(fmt, img, w, h, oracle.params(int(p.algorithm), p.weights, False),
                           threads=oracle.default_threads())
    assert np.array_equal(got, want)


def two_value_texture(seed):
    """Every pair (a, b), a <= b, as one block holding a and b in a seeded
    layout (R: a/b, G: b/a, A: a/b; B random), tiled row-major into a
    texture of 184 x 179 blocks (the last blocks repeat pairs)."""
    rng = np.random.default_rng(seed)
    pairs = np.array([(a, b) for a in range(256) for b in range(a, 256)], np.int64)
    bw, bh = 184, 179
    n = bw * bh
    pairs = np.concatenate([pairs, pairs[rng.integers(0, len(pairs), n - len(pairs))]])
    pat = rng.integers(0, 2, size=(n, 16)).astype(bool)
    pat[::7] = np.arange(16) < rng.integers(1, 16, size=(len(pat[::7]), 1))  # k copies of a, 16 - k of b
    px = np.zeros((n, 16, 4), np.int64)
    px[:, :, 0] = np.where(pat, pairs[:, :1], pairs[:, 1:])
    px[:, :, 1] = np.where(pat, pairs[:, 1:], pairs[:, :1])
    px[:, :, 2] = rng.integers(0, 256, size=(n, 16))
    px[:, :, 3] = px[:, :, 0]
    img = px.astype(np.uint8).reshape(bh, bw, 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(4 * bh, 4 * bw, 4)
    return np.ascontiguousarray(img)


@pytest.mark.parametrize("fmt", [2, 3, 4], ids=["bc3", "bc4", "bc5"])
def test_alpha_two_value_blocks_texture(gpu, oracle, fmt):
    # every two-value block through the texture kernels (BC4/BC5: one block
    # per lane; BC3: the pipeline's classify pass, forced on), all pixels valid
    img = two_value_texture(960 + fmt)
    h, w = img.shape[:2]
    # ClusterFit: BC3 colour goes through the pipeline (RangeFit would not)
    p = tx.Params(tx.Algorithm.ClusterFit)
    f = tx.Format(fmt)
    old = tx.set_pipeline_threshold(0)
    try:
        got = np.zeros(f.compressed_size(w, h), np.uint8)
        f.compress(img.reshape(-1), w, h, p, got)
    finally:
        tx.set_pipeline_threshold(old)
    want = oracle.compress(fmt, img, w, h, oracle.params(int(p.algorithm), p.weights, False),
                           threads=oracle.default_threads())
    bs = f.block_size()
    bad = np.nonzero(np.any(got.reshape(-1, bs) != want.reshape(-1, bs), axis=1))[0]
    assert bad.size == 0, "%d blocks differ, first %d" % (bad.size, bad[0])
