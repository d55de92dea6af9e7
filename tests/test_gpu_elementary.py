"""The encoder's exactly specified elementary operations, as the device
evaluates them (tx_util_elementary), against numpy's IEEE binary32 evaluation
(correctly rounded division, square root, multiply and add), exhaustively or
on every rounding boundary's neighbourhood:

- u / 255 for every byte u: the point coordinates of the colour set
  (lib/src/colourset.rs:60-66);
- sqrtf(c) for the duplicate counts c = 0..16 and sqrtf(k / 256) for every
  summed alpha weight k = 1..4096 (lib/src/colourset.rs:103-110);
- the 5- and 6-bit grid snap, clamp to [0, 1] then truncf(g * v + 0.5) *
  (1/g) (lib/src/colourfit/range.rs:84-95, cluster.rs:207-216), on a dense
  sweep of [-0.25, 1.25], +-64 ulps around every decision point (k - 0.5)/g,
  and the non-finite values;
- roundf(a) clamped to [0, 255] with NaN to 0 (lib/src/math.rs:100-102), +-64
  ulps around every integer and half-integer of [-3, 258], on every
  x * 255 the single-colour fit rounds, and on the non-finite values.

These do not go through the oracle: they pin the device arithmetic (the
build's division and square root, the clamp lowering, trunc) on its own.
"""
import numpy as np
import pytest

import texpresso_amd as tx

f32 = np.float32


def around(centres, ulps=64):
    """Every float32 within `ulps` steps of each (float32) centre."""
    c = np.asarray(centres, dtype=np.float32)
    steps = np.arange(-ulps, ulps + 1, dtype=np.int64)
    out = []
    for x in c:
        if x == 0:
            pos = (np.arange(0, ulps + 1, dtype=np.int64)).astype(np.int32).view(np.float32)
            out.append(pos)
            out.append(-pos)
            continue
        bits = np.int64(np.array([x], np.float32).view(np.int32)[0])
        out.append((bits + steps).astype(np.int32).view(np.float32))
    return np.concatenate(out)


SPECIAL = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-45, -1e-45, 1e-38, 3e38, -3e38, 1e10, 0.5, 1.0,
                    254.5, 255.5, 255.49998, -0.5, 2.5], dtype=np.float32)


def host_snap(v, g):
    c = np.fmin(f32(1), np.fmax(f32(0), v))  # fminf/fmaxf: NaN -> the other operand
    t = (f32(g) * c).astype(np.float32)
    t = (t + f32(0.5)).astype(np.float32)
    q = np.trunc(t).astype(np.float32)
    rcp = (f32(1) / f32(g)).astype(np.float32)
    return (q * rcp).astype(np.float32)


def host_round_clamped(a, limit):
    a64 = a.astype(np.float64)
    r = np.sign(a64) * np.floor(np.abs(a64) + 0.5)  # roundf: half away from zero, exact in float64
    r = np.fmin(np.fmax(r, 0.0), float(limit))  # f32::max/min ignore NaN
    r = np.where(np.isnan(a64), 0.0, r)
    return r.astype(np.int32)


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


@pytest.fixture(scope="module")
def device_tables(gpu):
    return tx.elementary(np.zeros(0, np.float32))


@pytest.mark.gpu
def test_u8_over_255_every_byte(device_tables):
    want = (np.arange(256, dtype=np.float32) / f32(255)).astype(np.float32)
    assert same_bits(device_tables["u8f"], want)


@pytest.mark.gpu
def test_sqrt_of_counts_and_alpha_weights(device_tables):
    want_c = np.sqrt(np.arange(17, dtype=np.float32)).astype(np.float32)
    assert same_bits(device_tables["sqrt_counts"], want_c)
    k = np.arange(1, 4097, dtype=np.float32)
    want_a = np.sqrt((k / f32(256)).astype(np.float32)).astype(np.float32)
    assert same_bits(device_tables["sqrt_alpha"], want_a)


@pytest.mark.gpu
def test_grid_snap_sweep_and_decision_points(gpu):
    sweep = np.linspace(-0.25, 1.25, 300001, dtype=np.float64).astype(np.float32)
    edges = [f32((k - 0.5) / g) for g in (31, 63) for k in range(0, g + 2)]
    v = np.concatenate([sweep, around(edges), around([0.0, 1.0]), SPECIAL])
    got = tx.elementary(v)
    for g, key in ((31, "snap31"), (63, "snap63")):
        want = host_snap(v, g)
        bad = np.flatnonzero(got[key].view(np.int32) != want.view(np.int32))
        assert bad.size == 0, "snap %d differs at %s: device %s host %s" % (g, v[bad[:5]], got[key][bad[:5]],
                                                                           want[bad[:5]])


@pytest.mark.gpu
def test_round_clamped_every_boundary(gpu):
    ints = [f32(k) for k in range(-3, 259)]
    halves = [f32(k + 0.5) for k in range(-3, 259)]
    x255 = (np.arange(256, dtype=np.float32) / f32(255)).astype(np.float32) * f32(255)
    v = np.concatenate([around(ints), around(halves), x255.astype(np.float32), SPECIAL])
    got = tx.elementary(v)["round255"]
    want = host_round_clamped(v, 255)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, "roundf clamp differs at %s: device %s host %s" % (v[bad[:5]], got[bad[:5]],
                                                                            want[bad[:5]])


def test_planted_host_mistakes_would_fail():
    """The host references are not vacuous: a round-half-to-even, an
    unclamped snap and a reciprocal-times u/255 each differ from them on the
    inputs used above."""
    halves = np.array([0.5, 1.5, 2.5, 254.5], dtype=np.float32)
    assert not np.array_equal(np.round(halves).astype(np.int32), host_round_clamped(halves, 255))
    v = np.array([1.25, -0.25], dtype=np.float32)
    unclamped = (np.trunc(f32(31) * v + f32(0.5)) * (f32(1) / f32(31))).astype(np.float32)
    assert not same_bits(unclamped, host_snap(v, 31))
    u = np.arange(256, dtype=np.float32)
    assert not same_bits((u * (f32(1) / f32(255))).astype(np.float32), (u / f32(255)).astype(np.float32))
