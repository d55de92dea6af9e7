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
