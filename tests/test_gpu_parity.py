"""Bit-exact parity of the HIP encoder/decoder with the CPU oracle, through
the C ABI.  Bar: every output byte identical (integer/byte work and the
FP32 fits are all required to be bit-exact); a failing case reports the
number of differing blocks and the decoded RGB/alpha MSE delta."""
import json
import os

import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import texgen
from tests.blockgen import make_blocks

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KAT = json.load(open(os.path.join(GOLDEN, "kat.json")))
ALGS = [tx.Algorithm.RangeFit, tx.Algorithm.ClusterFit, tx.Algorithm.IterativeClusterFit]


def _params(alg, weights=tx.COLOUR_WEIGHTS_PERCEPTUAL, alpha=False):
    return tx.Params(alg, weights, alpha)


def _oparams(O, p):
    return O.params(int(p.algorithm), p.weights, p.weigh_colour_by_alpha)


def gpu_compress(fmt, img, w, h, p, out_len=None):
    fmt = tx.Format(fmt)
    out = np.zeros(out_len if out_len is not None else fmt.compressed_size(w, h), np.uint8)
    fmt.compress(np.ascontiguousarray(img).reshape(-1), w, h, p, out)
    return out


def report(fmt, got, want, O):
    """Number of differing blocks and the MSE delta of their decodings."""
    bs = tx.Format(fmt).block_size()
    g = got.reshape(-1, bs)
    r = want.reshape(-1, bs)
    bad = np.nonzero(np.any(g != r, axis=1))[0]
    if len(bad) == 0:
        return "0 mismatches"
    dg = np.stack([O.decompress_block(fmt, g[b]) for b in bad[:256]]).astype(np.float64)
    dr = np.stack([O.decompress_block(fmt, r[b]) for b in bad[:256]]).astype(np.float64)
    d = (dg - dr).reshape(-1, 16, 4)
    return "%d/%d blocks differ (first %s); RGB MSE delta %.3f, alpha MSE delta %.3f" % (
        len(bad), len(g), list(bad[:8]), float((d[..., :3] ** 2).mean()), float((d[..., 3] ** 2).mean()))


@pytest.mark.parametrize("blk", KAT["blocks"], ids=[b["name"] for b in KAT["blocks"]])
@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.name)
def test_kat(gpu, blk, alg):
    out = gpu_compress(blk["format"], np.array(blk["decoded"], np.uint8), 4, 4, _params(alg, tx.COLOUR_WEIGHTS_UNIFORM))
    assert bytes(out) == bytes(blk["encoded"])


@pytest.mark.parametrize("blk", KAT["blocks"], ids=[b["name"] for b in KAT["blocks"]])
def test_decode_kat(gpu, blk):
    fmt = tx.Format(blk["format"])
    out = bytearray(64)
    fmt.decompress(bytes(blk["encoded"]), 4, 4, out)
    assert bytes(out) == bytes(blk["decoded"])
    assert [v for px in fmt.decompress_block(bytes(blk["encoded"])) for v in px] == blk["decoded"]


def test_decode_height_not_multiple_of_4(gpu):
    d = KAT["decode_height6"]
    out = bytearray(4 * d["width"] * d["height"])
    tx.Format.Bc1.decompress(bytes(d["encoded"]), d["width"], d["height"], out)
    assert bytes(out) == bytes(d["pixel"]) * (d["width"] * d["height"])


SIZES = [(64, 64), (131, 77), (5, 3), (1, 1), (2, 7), (37, 128)]


@pytest.mark.parametrize("fmt", [0, 1, 2], ids=["bc1", "bc2", "bc3"])
@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.name)
@pytest.mark.parametrize("size", SIZES, ids=lambda s: "%dx%d" % s)
def test_texture_parity_colour(gpu, oracle, fmt, alg, size):
    w, h = size
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 11 + w * 7 + h), 5 + w)
    p = _params(alg)
    got = gpu_compress(fmt, img, w, h, p)
    want = oracle.compress(fmt, img, w, h, _oparams(oracle, p), threads=oracle.default_threads())
    assert np.array_equal(got, want), report(fmt, got, want, oracle)


@pytest.mark.parametrize("fmt", [3, 4], ids=["bc4", "bc5"])
@pytest.mark.parametrize("size", SIZES, ids=lambda s: "%dx%d" % s)
def test_texture_parity_alpha_formats(gpu, oracle, fmt, size):
    w, h = size
    img = texgen.normal_map(w, h, 3 + w) if fmt == 4 else texgen.rgb_texture(w, h, 9 + h)
    p = _params(tx.Algorithm.ClusterFit)
    got = gpu_compress(fmt, img, w, h, p)
    want = oracle.compress(fmt, img, w, h, _oparams(oracle, p))
    assert np.array_equal(got, want), report(fmt, got, want, oracle)


WEIGHTS = [tx.COLOUR_WEIGHTS_PERCEPTUAL, tx.COLOUR_WEIGHTS_UNIFORM, (0.9, 0.05, 0.37)]


@pytest.mark.parametrize("fmt", [0, 1, 2, 3, 4], ids=["bc1", "bc2", "bc3", "bc4", "bc5"])
@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.name)
@pytest.mark.parametrize("wi", range(len(WEIGHTS)), ids=["perceptual", "uniform", "skewed"])
@pytest.mark.parametrize("alpha", [False, True], ids=["plain", "alphaw"])
def test_adversarial_blocks(gpu, oracle, fmt, alg, wi, alpha):
    if fmt in (3, 4) and (alg != tx.Algorithm.ClusterFit or wi or alpha):
        pytest.skip("BC4/BC5 ignore the colour parameters")
    blocks, masks = make_blocks(1000 + 17 * fmt + 5 * int(alg) + wi, per_class=128)
    p = _params(alg, WEIGHTS[wi], alpha)
    got = tx.compress_blocks(fmt, blocks, masks, p)
    want = oracle.compress_blocks(fmt, blocks, masks, _oparams(oracle, p))
    assert np.array_equal(got, want), report(fmt, got.reshape(-1), want.reshape(-1), oracle)


def _random_weights(seed):
    """Seeded colour-weight triples: unit-sum, tiny, large, one channel zero."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        w = rng.random(3) + 0.05
        out.append(tuple(float(x) for x in w / w.sum()))
    out.append(tuple(float(x) for x in rng.random(3) * 1e-3 + 1e-4))
    out.append(tuple(float(x) for x in rng.random(3) * 40.0 + 1.0))
    w = rng.random(3)
    w[int(rng.integers(0, 3))] = 0.0
    out.append(tuple(float(x) for x in w))
    return out


@pytest.mark.parametrize("fmt", [0, 2], ids=["bc1", "bc3"])
@pytest.mark.parametrize("alg", [tx.Algorithm.ClusterFit, tx.Algorithm.IterativeClusterFit], ids=lambda a: a.name)
@pytest.mark.parametrize("alpha", [False, True], ids=["plain", "alphaw"])
def test_random_weight_sets(gpu, oracle, fmt, alg, alpha):
    # Params::weights is a free float triple: the fits multiply by it in
    # their error sums, so non-uniform, very small, very large and zero
    # weights each take other rounding paths through the same expressions
    blocks, masks = make_blocks(2000 + 7 * fmt + int(alg) + (50 if alpha else 0), per_class=48)
    for wi, weights in enumerate(_random_weights(31 + fmt + 3 * int(alg) + (1 if alpha else 0))):
        p = _params(alg, weights, alpha)
        print("weights:", fmt, alg.name, alpha, wi, weights, flush=True)
        got = tx.compress_blocks(fmt, blocks, masks, p)
        want = oracle.compress_blocks(fmt, blocks, masks, _oparams(oracle, p))
        assert np.array_equal(got, want), (weights, report(fmt, got.reshape(-1), want.reshape(-1), oracle))


def test_bc1_clusterfit_512_texture(gpu, oracle):
    # BC1 ClusterFit on the seeded 512x512 texture, compared byte for byte
    img = texgen.rgb_texture(512, 512, texgen.SEED_BASE)
    p = tx.Params()
    got = gpu_compress(0, img, 512, 512, p)
    want = oracle.compress(0, img, 512, 512, _oparams(oracle, p), threads=oracle.default_threads())
    assert np.array_equal(got, want), report(0, got, want, oracle)


@pytest.mark.parametrize("fmt", [0, 2, 3, 4], ids=["bc1", "bc3", "bc4", "bc5"])
def test_oversized_output_gets_masked_blocks(gpu, oracle, fmt):
    # lib.rs:300-305: whole extra rows of the output are encoded as fully masked blocks
    img = texgen.rgb_texture(9, 6, 4)
    f = tx.Format(fmt)
    n = f.compressed_size(9, 6) + 7 * f.block_size()
    p = tx.Params()
    got = gpu_compress(fmt, img, 9, 6, p, out_len=n)
    want = oracle.compress(fmt, img, 9, 6, _oparams(oracle, p), out_len=n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("fmt", [0, 1, 2, 3, 4], ids=["bc1", "bc2", "bc3", "bc4", "bc5"])
def test_decode_parity_random_blocks(gpu, oracle, fmt):
    rng = np.random.default_rng(77 + fmt)
    f = tx.Format(fmt)
    w, h = 67, 45
    data = rng.integers(0, 256, f.compressed_size(w, h)).astype(np.uint8)
    got = np.zeros(4 * w * h, np.uint8)
    f.decompress(data, w, h, got)
    want = oracle.decompress(fmt, data, w, h)
    assert np.array_equal(got, want)


def test_device_entry_points_match_host_path(gpu, oracle):
    w, h = 100, 68
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 8), 8)
    d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
    for fmt in range(5):
        f = tx.Format(fmt)
        p = _params(tx.Algorithm.IterativeClusterFit)
        n = f.compressed_size(w, h)
        d_out = tx.DeviceBuffer(n)
        d_dec = tx.DeviceBuffer(4 * w * h)
        f.compress_device(d_in.ptr, w, h, p, d_out.ptr, n)
        f.decompress_device(d_out.ptr, w, h, d_dec.ptr)
        want = oracle.compress(fmt, img, w, h, _oparams(oracle, p))
        assert np.array_equal(d_out.download(), want)
        assert np.array_equal(d_dec.download(), oracle.decompress(fmt, want, w, h))
        ms = tx.time_compress(f, d_in.ptr, w, h, p, d_out.ptr, n, 3)
        assert ms > 0 and np.array_equal(d_out.download(), want)


def test_repeat_runs_identical_and_tiling_independent(gpu):
    # the same texture through the image path twice and through the block-list path
    w, h = 96, 52
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 21), 21)
    p = _params(tx.Algorithm.IterativeClusterFit)
    a = gpu_compress(2, img, w, h, p)
    b = gpu_compress(2, img, w, h, p)
    assert np.array_equal(a, b)
    blocks = img.reshape(h // 4, 4, w // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 64)
    c = tx.compress_blocks(2, blocks, np.full(len(blocks), 0xFFFF, np.uint32), p)
    assert np.array_equal(a, c.reshape(-1))


def adversarial_texture(seed, bw=131, bh=101, cut=(3, 2)):
    """The adversarial block set tiled row-major into a texture of bw x bh
    blocks (13,231 blocks, repeating the set), cut by `cut` pixels on the
    right and bottom so the last column and row are partial blocks."""
    blocks, _ = make_blocks(seed, per_class=128)
    reps = -(-(bw * bh) // len(blocks))
    tiled = np.tile(blocks, (reps, 1))[: bw * bh]
    img = tiled.reshape(bh, bw, 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(4 * bh, 4 * bw, 4)
    return np.ascontiguousarray(img[: 4 * bh - cut[1], : 4 * bw - cut[0]])


@pytest.mark.parametrize("fmt", [0, 1, 2], ids=["bc1", "bc2", "bc3"])
@pytest.mark.parametrize("wi", range(len(WEIGHTS)), ids=["perceptual", "uniform", "skewed"])
@pytest.mark.parametrize("alpha", [False, True], ids=["plain", "alphaw"])
def test_range_fit_lane_kernel_adversarial_texture(gpu, oracle, fmt, wi, alpha):
    # RangeFit textures of at least 12,288 blocks take the one-block-per-lane
    # kernel: every adversarial class (single colours, NaN axes, ties,
    # transparency, alpha extremes, small palettes, noise) in one texture,
    # with partial blocks in its last row and column
    img = adversarial_texture(3000 + 11 * fmt + wi)
    h, w = img.shape[:2]
    assert tx.num_blocks(w) * tx.num_blocks(h) >= 12288
    p = _params(tx.Algorithm.RangeFit, WEIGHTS[wi], alpha)
    got = gpu_compress(fmt, img, w, h, p)
    want = oracle.compress(fmt, img, w, h, _oparams(oracle, p), threads=oracle.default_threads())
    assert np.array_equal(got, want), report(fmt, got, want, oracle)
