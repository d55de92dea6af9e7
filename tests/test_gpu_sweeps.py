"""Parity sweeps of the weight options at scale: uniform and perceptual
weights, with and without alpha-weighted colour sets (lib/src/lib.rs:85-89,
lib/src/colourset.rs:70-97), for BC1 and BC3 ClusterFit and
IterativeClusterFit, through both device paths (the lockstep kernel and the
binned pipeline), byte-compared with the oracle on 768x512 textures; and one
2048x2048 alpha-weighted BC1 case where both device paths must agree and
match the oracle."""
import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import texgen

pytestmark = pytest.mark.gpu

NEVER = 2 ** 64 - 1
WEIGHTS = {"uniform": tx.COLOUR_WEIGHTS_UNIFORM, "perceptual": tx.COLOUR_WEIGHTS_PERCEPTUAL}


@pytest.fixture
def threshold(gpu):
    old = tx.set_pipeline_threshold(NEVER)
    yield lambda v: tx.set_pipeline_threshold(v)
    tx.set_pipeline_threshold(old)


def encode(fmt, img, w, h, p):
    out = np.zeros(tx.Format(fmt).compressed_size(w, h), np.uint8)
    tx.Format(fmt).compress(np.ascontiguousarray(img).reshape(-1), w, h, p, out)
    return out


@pytest.mark.parametrize("path", ["lockstep", "pipeline"])
@pytest.mark.parametrize("alpha", [False, True], ids=["plain", "alphaw"])
@pytest.mark.parametrize("weights", ["uniform", "perceptual"])
@pytest.mark.parametrize("alg", [tx.Algorithm.ClusterFit, tx.Algorithm.IterativeClusterFit], ids=lambda a: a.name)
@pytest.mark.parametrize("fmt", [0, 2], ids=["bc1", "bc3"])
def test_weight_sweep(threshold, oracle, fmt, alg, weights, alpha, path):
    w, h = 768, 512
    seed = texgen.SEED_BASE + 50 + fmt
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, seed), seed)
    p = tx.Params(alg, WEIGHTS[weights], alpha)
    threshold(0 if path == "pipeline" else NEVER)
    got = encode(fmt, img, w, h, p)
    want = oracle.compress(fmt, img, w, h, oracle.params(int(alg), p.weights, alpha), threads=oracle.default_threads())
    bs = tx.Format(fmt).block_size()
    bad = np.nonzero(np.any(got.reshape(-1, bs) != want.reshape(-1, bs), axis=1))[0]
    assert bad.size == 0, "first mismatching blocks %s" % bad[:8].tolist()


def test_alpha_weighted_bc1_full_size(threshold, oracle):
    w = h = 2048
    seed = texgen.SEED_BASE + 60
    img = texgen.with_alpha_field(texgen.rgb_texture(w, h, seed), seed)
    p = tx.Params(tx.Algorithm.IterativeClusterFit, tx.COLOUR_WEIGHTS_PERCEPTUAL, True)
    threshold(0)
    a = encode(0, img, w, h, p)
    threshold(NEVER)
    b = encode(0, img, w, h, p)
    assert np.array_equal(a, b)
    want = oracle.compress(0, img, w, h, oracle.params(2, p.weights, True), threads=oracle.default_threads())
    assert np.array_equal(a, want)
