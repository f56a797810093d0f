"""Seeded synthetic video for the ME workloads.

Reference pictures are smooth fields (a sum of 8 random-phase 2-D sinusoids
with periods 16-256 px, envelope amplitude 40) plus Gaussian noise around 128;
current pictures displace blocks of a reference by seeded integer or
quarter-pel motion vectors and add fresh noise.
"""
import numpy as np


def smooth_field(w, h, seed, noise=4.0, amp=40.0, n_waves=8):
    rng = np.random.default_rng(seed)
    yy = np.arange(h, dtype=np.float64)[:, None]
    xx = np.arange(w, dtype=np.float64)[None, :]
    f = np.zeros((h, w), dtype=np.float64)
    for _ in range(n_waves):
        period = np.exp(rng.uniform(np.log(16.0), np.log(256.0)))
        theta = rng.uniform(0, 2 * np.pi)
        phase = rng.uniform(0, 2 * np.pi)
        fx, fy = np.cos(theta) / period, np.sin(theta) / period
        f += np.sin(2 * np.pi * (fx * xx + fy * yy) + phase)
    f = 128.0 + (amp / n_waves) * f + rng.normal(0.0, noise, (h, w))
    return np.clip(np.rint(f), 0, 255).astype(np.uint8)


def displaced(ref, block, mv_range, seed, noise=3.0):
    """Current picture: every block x block tile is ref shifted by a seeded
    integer MV in [-mv_range, mv_range]^2 (pred = ref[p + mv]), plus noise."""
    h, w = ref.shape
    rng = np.random.default_rng(seed)
    pad = mv_range + 1
    big = np.pad(ref, pad, mode="edge")
    cur = np.empty_like(ref)
    mvs = rng.integers(-mv_range, mv_range + 1, size=((h + block - 1) // block, (w + block - 1) // block, 2))
    for by in range(0, h, block):
        for bx in range(0, w, block):
            dx, dy = mvs[by // block, bx // block]
            y0, x0 = by + dy + pad, bx + dx + pad
            hh, ww = min(block, h - by), min(block, w - bx)
            cur[by:by + hh, bx:bx + ww] = big[y0:y0 + hh, x0:x0 + ww]
    cur = np.clip(cur.astype(np.float64) + rng.normal(0.0, noise, cur.shape), 0, 255)
    return np.rint(cur).astype(np.uint8), mvs


def multi_ref_current(refs, region, mv_range, seed, noise=3.0):
    """Current picture built from per-region random mixtures: each region x region
    tile copies one of the references displaced by a seeded MV."""
    h, w = refs[0].shape
    rng = np.random.default_rng(seed)
    pad = mv_range + 1
    bigs = [np.pad(r, pad, mode="edge") for r in refs]
    cur = np.empty_like(refs[0])
    for by in range(0, h, region):
        for bx in range(0, w, region):
            k = int(rng.integers(0, len(refs)))
            dx, dy = rng.integers(-mv_range, mv_range + 1, size=2)
            hh, ww = min(region, h - by), min(region, w - bx)
            cur[by:by + hh, bx:bx + ww] = bigs[k][by + dy + pad:by + dy + pad + hh, bx + dx + pad:bx + dx + pad + ww]
    cur = np.clip(cur.astype(np.float64) + rng.normal(0.0, noise, cur.shape), 0, 255)
    return np.rint(cur).astype(np.uint8)


# PU partitions of a 64x64 CTU used by the mixed-shape workloads, with the
# shapes of the sub-pel configuration: 8x8, 16x16, 32x32, 64x64, 64x16, 16x64,
# 32x8, 8x32.
def _split(x, y, s, rng, out, small_bias):
    # split probabilities chosen so that the expected area is ~10 % 64x64,
    # ~30 % 32x32 and asymmetric (64x16, 16x64, 32x8, 8x32), ~60 % 8x8/16x16
    r = rng.random()
    if s == 64:
        if r < 0.10:
            out.append((x, y, 64, 64)); return
        if r < 0.125:
            out += [(x, y, 64, 16), (x, y + 16, 64, 16), (x, y + 32, 64, 16), (x, y + 48, 64, 16)]; return
        if r < 0.15:
            out += [(x, y, 16, 64), (x + 16, y, 16, 64), (x + 32, y, 16, 64), (x + 48, y, 16, 64)]; return
    if s == 32:
        if r < 0.15:
            out.append((x, y, 32, 32)); return
        if r < 0.222:
            out += [(x, y + k, 32, 8) for k in range(0, 32, 8)]; return
        if r < 0.294:
            out += [(x + k, y, 8, 32) for k in range(0, 32, 8)]; return
    if s == 16:
        if r < 0.5:
            out.append((x, y, 16, 16)); return
        out += [(x, y, 8, 8), (x + 8, y, 8, 8), (x, y + 8, 8, 8), (x + 8, y + 8, 8, 8)]; return
    h = s // 2
    for dy in (0, h):
        for dx in (0, h):
            _split(x + dx, y + dy, h, rng, out, small_bias)


def mixed_pus(w, h, seed):
    """Seeded PU tiling of a w x h picture (64x64 CTUs, partial CTUs at the
    right/bottom edge are tiled with 16x16 and 8x8)."""
    rng = np.random.default_rng(seed)
    out = []
    for cy in range(0, h, 64):
        for cx in range(0, w, 64):
            if cx + 64 <= w and cy + 64 <= h:
                _split(cx, cy, 64, rng, out, 0)
            else:
                for y in range(cy, min(cy + 64, h), 16):
                    for x in range(cx, min(cx + 64, w), 16):
                        if x + 16 <= w and y + 16 <= h:
                            out.append((x, y, 16, 16))
                        else:
                            for yy in range(y, min(y + 16, h), 8):
                                for xx in range(x, min(x + 16, w), 8):
                                    if xx + 8 <= w and yy + 8 <= h:
                                        out.append((xx, yy, 8, 8))
    return out


def shape_tiling(w, h, pw, ph):
    return [(x, y, pw, ph) for y in range(0, h - ph + 1, ph) for x in range(0, w - pw + 1, pw)]


def moving_sequence(w, h, n_frames, seed, pan=2, local=32, block=32, noise=3.0):
    """A moving smooth+noise sequence: frame t is a crop of one larger smooth
    field panned by a seeded global motion of at most `pan` px per frame and
    axis, with every block x block tile displaced by a seeded local MV in
    [-local, local]^2 (drawn afresh per frame) and fresh noise."""
    rng = np.random.default_rng(seed)
    gx, gy = (int(v) for v in rng.integers(-pan, pan + 1, size=2))
    margin = local + pan * n_frames + 1
    big = smooth_field(w + 2 * margin, h + 2 * margin, seed=seed + 1)
    nby, nbx = (h + block - 1) // block, (w + block - 1) // block
    yy = np.arange(h)[:, None]
    xx = np.arange(w)[None, :]
    frames = []
    for t in range(n_frames):
        mv = rng.integers(-local, local + 1, size=(nby, nbx, 2))
        dx = np.repeat(np.repeat(mv[:, :, 0], block, 0), block, 1)[:h, :w]
        dy = np.repeat(np.repeat(mv[:, :, 1], block, 0), block, 1)[:h, :w]
        f = big[yy + margin + gy * t + dy, xx + margin + gx * t + dx].astype(np.float64)
        f += rng.normal(0.0, noise, f.shape)
        frames.append(np.clip(np.rint(f), 0, 255).astype(np.uint8))
    return frames
