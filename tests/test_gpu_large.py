"""Large textures: an 8190 x 8189 image (4.2M blocks; records, list
positions and output offsets far beyond the smaller tests) through the
binned pipeline, checked against the oracle on bands of block rows at the
top, middle and bottom.  Blocks are independent, so a band encoded alone by
the oracle gives exactly the bytes of those block rows; the bottom band
ends in partially masked blocks (8189 is not a multiple of 4)."""
import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import texgen

pytestmark = pytest.mark.gpu

W, H = 8190, 8189


@pytest.fixture(scope="module")
def large_image():
    base = texgen.with_alpha_field(texgen.rgb_texture(1024, 1024, 4242), 4243)
    img = np.tile(base, (8, 8, 1))[:H, :W].copy()
    # make the tiles differ: a per-tile offset in the green channel
    ty, tx_ = np.meshgrid(np.arange(H) // 1024, np.arange(W) // 1024, indexing="ij")
    img[..., 1] = (img[..., 1].astype(np.uint32) + (ty * 8 + tx_) * 7).astype(np.uint8)
    return img


@pytest.mark.parametrize("fmt,alg", [(0, tx.Algorithm.ClusterFit), (2, tx.Algorithm.IterativeClusterFit)],
                         ids=["bc1-cluster", "bc3-iterative"])
def test_large_texture_bands_match_oracle(gpu, oracle, large_image, fmt, alg):
    f = tx.Format(fmt)
    p = tx.Params(alg)
    print("large:", fmt, alg.name, W, H, flush=True)
    out = np.zeros(f.compressed_size(W, H), np.uint8)
    f.compress(large_image.reshape(-1), W, H, p, out)
    row_bytes = tx.num_blocks(W) * f.block_size()
    op = oracle.params(int(alg), p.weights, False)
    for r0 in (0, 4096, 8128):
        rows = min(64, H - r0)
        band = np.ascontiguousarray(large_image[r0:r0 + rows])
        want = oracle.compress(fmt, band, W, rows, op, threads=oracle.default_threads())
        got = out[(r0 // 4) * row_bytes:(r0 // 4) * row_bytes + want.size]
        assert np.array_equal(got, want), "rows %d..%d" % (r0, r0 + rows)
