"""Device mip-chain generation and chain compression.

The chain is compared with an independent numpy filter (edge replication by
padding, then a reshape-sum of 2x2 boxes, (sum + 2) >> 2); every compressed
level must equal what Format.compress gives for that level alone, and the
oracle's bytes on sampled levels; the DDS file built from it decodes level 0
like texpresso's `decompress` command."""
import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import dds, texgen

pytestmark = pytest.mark.gpu


def np_down(img):
    h, w = img.shape[:2]
    nw, nh = max(1, w >> 1), max(1, h >> 1)
    # pad to 2*nw x 2*nh by repeating the last column/row (edge clamp)
    pad_w, pad_h = max(0, 2 * nw - w), max(0, 2 * nh - h)
    p = np.pad(img.astype(np.uint32), ((0, pad_h), (0, pad_w), (0, 0)), mode="edge")[:2 * nh, :2 * nw]
    s = p.reshape(nh, 2, nw, 2, 4).sum(axis=(1, 3))
    return ((s + 2) >> 2).astype(np.uint8)


def np_chain(img, levels):
    out = [img]
    for _ in range(levels - 1):
        out.append(np_down(out[-1]))
    return out


SIZES = [(1, 1), (1, 7), (5, 3), (33, 17), (257, 129), (1000, 3), (512, 512)]


@pytest.mark.parametrize("w,h", SIZES, ids=lambda v: str(v))
def test_generate_mips_matches_numpy(gpu, w, h):
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 90 + w + h), 91 + w)
    n = tx.mip_levels(w, h)
    want = np_chain(img, n)
    d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
    chain_len = tx.mip_chain_rgba_bytes(w, h, n)
    d_chain = tx.DeviceBuffer(max(4, chain_len))
    tx.generate_mips_device(d_in.ptr, w, h, n, d_chain.ptr, chain_len)
    got = d_chain.download(chain_len) if chain_len else np.zeros(0, np.uint8)
    off = 0
    for k in range(1, n):
        lw, lh = tx.mip_dims(w, h, k)
        lvl = got[off:off + 4 * lw * lh].reshape(lh, lw, 4)
        assert np.array_equal(lvl, want[k]), "level %d" % k
        off += 4 * lw * lh


@pytest.mark.parametrize("fmt,alg", [(0, tx.Algorithm.ClusterFit), (1, tx.Algorithm.RangeFit),
                                     (2, tx.Algorithm.IterativeClusterFit), (3, tx.Algorithm.ClusterFit),
                                     (4, tx.Algorithm.ClusterFit)], ids=["bc1", "bc2", "bc3", "bc4", "bc5"])
@pytest.mark.parametrize("w,h", [(300, 75), (64, 64), (2, 9)], ids=lambda v: str(v))
def test_compress_mips_equals_per_level(gpu, oracle, fmt, alg, w, h):
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 7 * w + h), 3 + h)
    p = tx.Params(alg)
    f = tx.Format(fmt)
    n = tx.mip_levels(w, h)
    got = f.compress_mips(img.reshape(-1), w, h, p)
    assert got.size == f.mip_chain_size(w, h)
    off = 0
    for k, lvl in enumerate(np_chain(img, n)):
        lh, lw = lvl.shape[:2]
        size = f.compressed_size(lw, lh)
        alone = np.zeros(size, np.uint8)
        f.compress(np.ascontiguousarray(lvl).reshape(-1), lw, lh, p, alone)
        assert np.array_equal(got[off:off + size], alone), "level %d" % k
        if k in (0, n - 1, n // 2):
            want = oracle.compress(fmt, lvl, lw, lh, oracle.params(int(alg), p.weights, False))
            assert np.array_equal(alone, want), "level %d vs oracle" % k
        off += size


def test_compress_mips_device_and_partial_chain(gpu):
    w, h = 200, 120
    img = texgen.rgb_texture(w, h, 5)
    p = tx.Params(tx.Algorithm.ClusterFit)
    f = tx.Format.Bc1
    levels = 4
    chain_len = tx.mip_chain_rgba_bytes(w, h, levels)
    out_len = f.mip_chain_size(w, h, levels)
    d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
    d_chain = tx.DeviceBuffer(chain_len)
    d_out = tx.DeviceBuffer(out_len)
    f.compress_mips_device(d_in.ptr, w, h, p, levels, d_chain.ptr, chain_len, d_out.ptr, out_len)
    assert np.array_equal(d_out.download(), f.compress_mips(img.reshape(-1), w, h, p, levels))
    with pytest.raises(tx.TexpressoError):
        f.compress_mips_device(d_in.ptr, w, h, p, levels, d_chain.ptr, chain_len - 4, d_out.ptr, out_len)
    with pytest.raises(tx.TexpressoError):
        f.compress_mips_device(d_in.ptr, w, h, p, 99, d_chain.ptr, chain_len, d_out.ptr, out_len)


@pytest.mark.parametrize("fmt", range(5))
def test_dds_file_decodes_like_the_cli(gpu, fmt):
    w, h = 96, 40
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, 40 + fmt), 41)
    blob = dds.encode(fmt, img.reshape(-1), w, h, tx.Params(), mips=True)
    f, pw, ph, n, views = dds.parse(blob)
    assert (int(f), pw, ph, n) == (fmt, w, h, tx.mip_levels(w, h))
    level0 = np.zeros(tx.Format(fmt).compressed_size(w, h), np.uint8)
    tx.Format(fmt).compress(img.reshape(-1), w, h, tx.Params(), level0)
    assert np.array_equal(views[0], level0)
    pixels = np.zeros(4 * w * h, np.uint8)
    tx.Format(fmt).decompress(level0, w, h, pixels)
    assert np.array_equal(dds.decode_level0(blob), pixels)
