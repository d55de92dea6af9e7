"""Persistent batch kernel: byte parity with the oracle for mixed formats,
algorithms, odd sizes and mip chains; hipGraph replays (checked one by one,
with eager work between them); identical output for every tuning; autotuner."""
import numpy as np
import pytest

import texpresso_amd as tx
from texpresso_amd import texgen

pytestmark = pytest.mark.gpu


def make_set(seed, n, max_side):
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n):
        w = int(rng.integers(1, max_side + 1))
        h = int(rng.integers(1, max_side + 1))
        fmt = int(rng.integers(0, 5))
        alg = tx.Algorithm(int(rng.integers(0, 3)))
        weights = tx.COLOUR_WEIGHTS_PERCEPTUAL if rng.random() < 0.7 else tx.COLOUR_WEIGHTS_UNIFORM
        p = tx.Params(alg, weights, bool(rng.random() < 0.2))
        base = texgen.with_alpha_field(texgen.rgb_texture(w, h, seed * 100 + i), seed + i)
        levels = texgen.mip_chain(base) if rng.random() < 0.5 else [base]
        for lv in levels:
            items.append((fmt, p, np.ascontiguousarray(lv)))
    return items


def upload(items):
    bufs = []
    for fmt, p, img in items:
        h, w = img.shape[:2]
        d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
        n = tx.Format(fmt).compressed_size(w, h)
        d_out = tx.DeviceBuffer(max(n, 16))
        bufs.append((d_in, d_out, n))
    descs = [(fmt, p, b[0].ptr, b[1].ptr, img.shape[1], img.shape[0]) for (fmt, p, img), b in zip(items, bufs)]
    return bufs, descs


def expected(oracle, items):
    out = []
    for fmt, p, img in items:
        h, w = img.shape[:2]
        op = oracle.params(int(p.algorithm), p.weights, p.weigh_colour_by_alpha)
        out.append(oracle.compress(fmt, img, w, h, op, threads=oracle.default_threads()))
    return out


def clear(bufs):
    for _, d_out, n in bufs:
        d_out.upload(np.full(d_out.nbytes, 0xA5, np.uint8))


def check(bufs, want):
    bad = []
    for i, ((_, d_out, n), w) in enumerate(zip(bufs, want)):
        got = d_out.download(n)
        if not np.array_equal(got, w):
            bad.append(i)
    return bad


def test_batch_parity_and_graph_replays(gpu, oracle):
    items = make_set(3, 24, 90)
    bufs, descs = upload(items)
    want = expected(oracle, items)
    batch = tx.Batch(descs, tx.Tuning(chunk_blocks=16, waves_per_group=1, groups_per_cu=4, use_graph=True,
                                      pipeline=True))
    assert batch.info()["blocks"] == sum(tx.num_blocks(d[4]) * tx.num_blocks(d[5]) for d in descs)
    scratch_img = texgen.rgb_texture(33, 17, 5)
    for rep in range(3):
        clear(bufs)
        batch.run()
        assert check(bufs, want) == [], "replay %d" % rep
        # eager work between replays
        out = np.zeros(tx.Format.Bc1.compressed_size(33, 17), np.uint8)
        tx.Format.Bc1.compress(scratch_img.reshape(-1), 33, 17, tx.Params(), out)
    batch.close()


@pytest.mark.parametrize("tuning", [
    tx.Tuning(4, 1, 2, False, False), tx.Tuning(32, 2, 6, True, False), tx.Tuning(256, 4, 3, True, False),
    tx.Tuning(64, 1, 12, True, False), tx.Tuning(4, 1, 12, False, True), tx.Tuning(64, 1, 12, True, True),
    tx.Tuning(256, 1, 12, True, True), tx.Tuning(32, 2, 6, True, False, 16), tx.Tuning(16, 4, 3, False, False, 16),
    tx.Tuning(64, 1, 12, True, True, 16),
], ids=["c4w1", "c32w2", "c256w4", "c64w1g12", "pipe-c4-nograph", "pipe-c64", "pipe-c256", "c32w2-alpha16",
        "c16w4-alpha16-nograph", "pipe-c64-alpha16"])
def test_batch_identical_across_tunings(gpu, oracle, tuning):
    items = make_set(9, 16, 70)
    bufs, descs = upload(items)
    want = expected(oracle, items)
    clear(bufs)
    batch = tx.Batch(descs, tuning)
    batch.run()
    assert check(bufs, want) == []
    ms = batch.run_timed(2)
    assert ms > 0 and check(bufs, want) == []
    batch.close()


def test_autotune_picks_a_candidate(gpu):
    items = make_set(5, 8, 64)
    bufs, descs = upload(items)
    cands = [tx.Tuning(16, 1, 4), tx.Tuning(64, 2, 6)]
    best, ms = tx.autotune(descs, cands, iters=2)
    assert ms > 0 and best.as_dict() in [c.as_dict() for c in cands]
    assert tx.time_per_texture(descs, 1) > 0


def test_batch_rejects_bad_texture(gpu):
    d_in = tx.DeviceBuffer(64)
    d_out = tx.DeviceBuffer(16)
    with pytest.raises(tx.TexpressoError):
        tx.Batch([(7, tx.Params(), d_in.ptr, d_out.ptr, 4, 4)])
    with pytest.raises(tx.TexpressoError):
        tx.Batch([(0, tx.Params(), d_in.ptr + 1, d_out.ptr, 4, 4)])
