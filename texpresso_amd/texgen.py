"""Seeded synthetic textures for the parity tests and benchmarks.

The workloads (deterministic given the seed; numpy PCG64, float32 arithmetic
with only +, -, *, / and rounding):
  * rgb_texture: per channel, 4-octave value noise (lattice periods
    64/32/16/8 px, amplitudes 96/48/24/12, smoothstep-bilinear) around 128,
    plus a linear gradient (random direction, +-100) and 24 random
    axis-aligned hard-edged rectangles (8-256 px, per-channel offset +-80),
    clamped and rounded to u8; alpha 255.
  * with_alpha_field: alpha = 2-octave smooth noise (periods 128/64); a
    seeded 10% of the 4x4 blocks are binarised (>= 128 -> 255, else 0).
  * normal_map: 5-octave fBm height field, central-difference tangent-space
    normal normalize(-s*dh/dx, -s*dh/dy, 1) encoded to u8 RGB, alpha 255.
  * mip_chain: w' = max(1, w >> 1), 2x2 box filter (sum + 2) >> 2 with edge
    clamp, down to 1x1.
Images are numpy uint8 arrays of shape (height, width, 4).
"""
import hashlib

import numpy as np

SEED_BASE = 0x7E570000


def _smooth(t):
    return t * t * (3.0 - 2.0 * t)


def value_noise(rng, w, h, period, amp):
    """Smoothstep-bilinear value noise with lattice spacing `period` px."""
    gw = w // period + 2
    gh = h // period + 2
    lat = rng.uniform(-1.0, 1.0, size=(gh, gw)).astype(np.float32)
    xs = np.arange(w, dtype=np.float32) / np.float32(period)
    ys = np.arange(h, dtype=np.float32) / np.float32(period)
    ix = np.floor(xs).astype(np.int64)
    iy = np.floor(ys).astype(np.int64)
    sx = _smooth(xs - ix).astype(np.float32)
    sy = _smooth(ys - iy).astype(np.float32)
    rows = lat[:, ix] * (np.float32(1) - sx) + lat[:, ix + 1] * sx  # (gh, w)
    out = rows[iy] * (np.float32(1) - sy)[:, None] + rows[iy + 1] * sy[:, None]
    return out * np.float32(amp)


def _to_u8(x):
    return np.clip(np.rint(x), 0, 255).astype(np.uint8)


def rgb_texture(w, h, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    img = np.zeros((h, w, 4), np.uint8)
    theta = rng.uniform(0.0, 2.0 * np.pi)
    dx, dy = np.float32(np.cos(theta)), np.float32(np.sin(theta))
    gx = (np.arange(w, dtype=np.float32) / np.float32(max(w, 1)) - np.float32(0.5)) * dx
    gy = (np.arange(h, dtype=np.float32) / np.float32(max(h, 1)) - np.float32(0.5)) * dy
    grad = (gy[:, None] + gx[None, :]) * np.float32(200.0)
    rects = []
    for _ in range(24):
        rw = int(rng.integers(8, 257))
        rh = int(rng.integers(8, 257))
        x0 = int(rng.integers(0, max(w, 1)))
        y0 = int(rng.integers(0, max(h, 1)))
        off = rng.uniform(-80.0, 80.0, size=3).astype(np.float32)
        rects.append((x0, y0, rw, rh, off))
    for c in range(3):
        acc = np.full((h, w), 128.0, np.float32)
        for period, amp in ((64, 96.0), (32, 48.0), (16, 24.0), (8, 12.0)):
            acc += value_noise(rng, w, h, period, amp)
        acc += grad
        for x0, y0, rw, rh, off in rects:
            acc[y0:y0 + rh, x0:x0 + rw] += off[c]
        img[:, :, c] = _to_u8(acc)
    img[:, :, 3] = 255
    return img


def with_alpha_field(img, seed, binary_fraction=0.1):
    h, w = img.shape[:2]
    rng = np.random.Generator(np.random.PCG64(seed ^ 0xA1FA))
    a = np.full((h, w), 128.0, np.float32)
    a += value_noise(rng, w, h, 128, 100.0)
    a += value_noise(rng, w, h, 64, 50.0)
    a8 = _to_u8(a)
    bh, bw = (h + 3) // 4, (w + 3) // 4
    pick = rng.random((bh, bw)) < binary_fraction
    full = np.repeat(np.repeat(pick, 4, axis=0), 4, axis=1)[:h, :w]
    a8 = np.where(full, np.where(a8 >= 128, 255, 0), a8).astype(np.uint8)
    out = img.copy()
    out[:, :, 3] = a8
    return out


def normal_map(w, h, seed, strength=64.0):
    rng = np.random.Generator(np.random.PCG64(seed ^ 0x4E4D))
    hf = np.zeros((h, w), np.float32)
    amp = 1.0
    for period in (256, 128, 64, 32, 16):
        hf += value_noise(rng, w, h, period, amp)
        amp *= 0.5
    xp = np.concatenate([hf[:, 1:], hf[:, -1:]], axis=1)
    xm = np.concatenate([hf[:, :1], hf[:, :-1]], axis=1)
    yp = np.concatenate([hf[1:, :], hf[-1:, :]], axis=0)
    ym = np.concatenate([hf[:1, :], hf[:-1, :]], axis=0)
    s = np.float32(strength)
    nx = -s * (xp - xm) * np.float32(0.5)
    ny = -s * (yp - ym) * np.float32(0.5)
    nz = np.ones_like(nx)
    inv = np.float32(1.0) / np.sqrt(nx * nx + ny * ny + nz * nz)
    img = np.zeros((h, w, 4), np.uint8)
    for c, comp in enumerate((nx, ny, nz)):
        img[:, :, c] = _to_u8((comp * inv * np.float32(0.5) + np.float32(0.5)) * np.float32(255.0))
    img[:, :, 3] = 255
    return img


def downsample(img):
    """One mip step: (sum of the 2x2 box + 2) >> 2, edge-clamped."""
    h, w = img.shape[:2]
    nw, nh = max(1, w >> 1), max(1, h >> 1)
    xs0 = np.minimum(2 * np.arange(nw), w - 1)
    xs1 = np.minimum(2 * np.arange(nw) + 1, w - 1)
    ys0 = np.minimum(2 * np.arange(nh), h - 1)
    ys1 = np.minimum(2 * np.arange(nh) + 1, h - 1)
    src = img.astype(np.uint16)
    acc = src[ys0][:, xs0] + src[ys0][:, xs1] + src[ys1][:, xs0] + src[ys1][:, xs1]
    return ((acc + 2) >> 2).astype(np.uint8)


def mip_chain(img):
    levels = [img]
    while levels[-1].shape[0] > 1 or levels[-1].shape[1] > 1:
        levels.append(downsample(levels[-1]))
    return levels


def sha256(img):
    return hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest()


def log_uniform_sides(rng, n, lo, hi):
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(np.int64)
