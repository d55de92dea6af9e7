"""Seeded adversarial 4x4 blocks for parity tests (the block classes that
stress distinct code paths of the encoder):

  single     one colour (every grey, random RGB): the lookup-table path
  nan_axis   two colours with equal channel sums (red/green): NaN principal axis
  ties       mirrored colours and repeated projections: stable-sort ties
  transp     BC1 blocks with some or all alpha < 128
  masks      partial blocks (rectangles and random masks)
  alpha      alpha in {0,255}, constant, narrow ranges, random
  palette    2..16 colours drawn from a small palette (ClusterFit, n < 16)
  noise      16 distinct colours (the bulk of natural textures)
Returns (blocks uint8 (n, 64), masks uint32 (n,)).
"""
import numpy as np


def _rect_masks():
    out = []
    for w in range(1, 5):
        for h in range(1, 5):
            m = 0
            for y in range(h):
                for x in range(w):
                    m |= 1 << (4 * y + x)
            out.append(m)
    return out


RECT_MASKS = _rect_masks()


def make_blocks(seed, per_class=256):
    rng = np.random.default_rng(seed)
    blocks, masks = [], []

    def add(px, mask=0xFFFF):
        blocks.append(np.asarray(px, np.uint8).reshape(64))
        masks.append(mask)

    # single colour
    for v in range(256):
        px = np.zeros((16, 4), np.uint8)
        px[:, :3] = v
        px[:, 3] = 255
        add(px)
    for _ in range(per_class):
        px = np.zeros((16, 4), np.uint8)
        px[:, :3] = rng.integers(0, 256, 3)
        px[:, 3] = rng.integers(0, 256)
        add(px, int(rng.choice(RECT_MASKS + [0xFFFF])))
    # NaN axis: equal channel sums
    for _ in range(per_class):
        a, b = rng.integers(0, 256, 2)
        c = rng.integers(0, 256)
        px = np.zeros((16, 4), np.uint8)
        pat = rng.integers(0, 2, 16)
        px[:, 0] = np.where(pat, a, b)
        px[:, 1] = np.where(pat, b, a)
        px[:, 2] = c
        px[:, 3] = 255
        add(px)
    # ties: mirrored colours / repeated projections
    for _ in range(per_class):
        base = rng.integers(0, 256, (3, 3))
        cols = np.concatenate([base, base[:, [0, 2, 1]], base[:, [1, 0, 2]]])
        px = np.zeros((16, 4), np.uint8)
        px[:, :3] = cols[rng.integers(0, len(cols), 16)]
        px[:, 3] = 255
        add(px)
    for _ in range(per_class // 2):
        g = np.sort(rng.integers(0, 256, 4))
        px = np.zeros((16, 4), np.uint8)
        px[:, :3] = g[rng.integers(0, 4, 16)][:, None]
        px[:, 3] = 255
        add(px)
    # transparency
    for _ in range(per_class):
        px = rng.integers(0, 256, (16, 4)).astype(np.uint8)
        frac = rng.choice([0.0, 0.2, 0.5, 1.0])
        px[:, 3] = np.where(rng.random(16) < frac, rng.integers(0, 128, 16), rng.integers(128, 256, 16))
        add(px, int(rng.choice([0xFFFF, int(rng.integers(0, 1 << 16))])))
    # partial blocks
    for _ in range(per_class):
        px = rng.integers(0, 256, (16, 4)).astype(np.uint8)
        add(px, int(rng.choice(RECT_MASKS)) if rng.random() < 0.5 else int(rng.integers(0, 1 << 16)))
    # alpha edge cases
    for _ in range(per_class):
        px = np.zeros((16, 4), np.uint8)
        pal = rng.integers(0, 256, (rng.integers(1, 5), 3))
        px[:, :3] = pal[rng.integers(0, len(pal), 16)]
        kind = rng.integers(0, 4)
        if kind == 0:
            px[:, 3] = rng.choice([0, 255], 16)
        elif kind == 1:
            px[:, 3] = rng.integers(0, 256)
        elif kind == 2:
            lo = rng.integers(0, 250)
            px[:, 3] = rng.integers(lo, lo + rng.integers(1, 8), 16).clip(0, 255)
        else:
            px[:, 3] = rng.integers(0, 256, 16)
        add(px, 0xFFFF if rng.random() < 0.7 else int(rng.integers(0, 1 << 16)))
    # small palettes
    for _ in range(per_class * 2):
        ncol = int(rng.integers(2, 17))
        pal = rng.integers(0, 256, (ncol, 3))
        px = np.zeros((16, 4), np.uint8)
        px[:, :3] = pal[rng.integers(0, ncol, 16)]
        px[:, 3] = rng.integers(0, 256, 16) if rng.random() < 0.3 else 255
        add(px)
    # 16 distinct colours, smooth and noisy
    for _ in range(per_class * 2):
        if rng.random() < 0.5:
            px = rng.integers(0, 256, (16, 4)).astype(np.uint8)
        else:
            c0 = rng.integers(0, 200, 3)
            d = rng.integers(-6, 7, (16, 3))
            px = np.zeros((16, 4), np.uint8)
            px[:, :3] = np.clip(c0 + np.cumsum(d, 0), 0, 255)
        px[:, 3] = 255
        add(px)
    return np.stack(blocks), np.array(masks, np.uint32)
