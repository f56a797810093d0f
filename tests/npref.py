"""Independent numpy formulations of the ME cost path used to pin the oracle.

Each function is written from the closed form of the computation (matrix
products, vectorised filters, brute-force minima), not from the loop structure
of the C restatement, so that a dropped term, a wrong sign or index, or a
transposed operand in either one makes the two disagree.
"""
import numpy as np

LUMA_TAPS = np.array([[0, 0, 0, 64, 0, 0, 0, 0],
                      [-1, 4, -10, 57, 19, -7, 3, -1],
                      [-1, 4, -11, 40, 40, -11, 4, -1],
                      [-1, 3, -7, 19, 57, -10, 4, -1]], dtype=np.int64)
CHROMA_TAPS = np.array([[0, 64, 0, 0], [-4, 62, 6, 0], [-6, 56, 15, -1], [-5, 47, 25, -3],
                        [-4, 36, 36, -4], [-3, 25, 47, -5], [-1, 15, 56, -6], [0, 6, 62, -4]],
                       dtype=np.int64)

H4 = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]], dtype=np.int64)


def sad(a, b):
    return int(np.abs(a.astype(np.int64) - b.astype(np.int64)).sum())


def satd(a, b):
    """Half the sum over 4x4 tiles of |H4 D H4^T| (row order of H4 is irrelevant)."""
    d = a.astype(np.int64) - b.astype(np.int64)
    h, w = d.shape
    tiles = d.reshape(h // 4, 4, w // 4, 4).transpose(0, 2, 1, 3)
    t = np.einsum("ij,abjk,lk->abil", H4, tiles, H4)
    return int(np.abs(t).sum()) // 2


def satd_tiles_even(a, b):
    d = a.astype(np.int64) - b.astype(np.int64)
    h, w = d.shape
    tiles = d.reshape(h // 4, 4, w // 4, 4).transpose(0, 2, 1, 3)
    t = np.einsum("ij,abjk,lk->abil", H4, tiles, H4)
    return bool(np.all(np.abs(t).sum(axis=(2, 3)) % 2 == 0))


def se_len(v):
    """Length of the signed Exp-Golomb code se(v): codeNum k = 2|v| - (v > 0)."""
    k = 2 * abs(v) - (1 if v > 0 else 0)
    return 2 * int(np.floor(np.log2(k + 1))) + 1


def clamp_picture(img, pad):
    """Edge-replicated padding of a picture by pad pixels on every side."""
    return np.pad(img, pad, mode="edge")


def interp_plane(img, dx, dy):
    """Interpolated plane (dx, dy) of picture img on x in [-4, W+4), y in [-4, H+4).

    Returns an array of shape (H+8, W+8); element [y+4, x+4] is the value at (x, y).
    Horizontal-only / vertical-only: clip((sum C*P + 32) >> 6).  Both: the int16
    horizontal sums filtered vertically with (sum + 2048) >> 12.
    """
    h, w = img.shape
    P = clamp_picture(img.astype(np.int64), 16)          # P[y+16, x+16] = img(clamp)
    ys = np.arange(-4, h + 4)
    xs = np.arange(-4, w + 4)

    def hsum(yy):  # horizontal 8-tap sums at rows yy (any), columns xs
        acc = np.zeros((len(yy), len(xs)), dtype=np.int64)
        for k in range(8):
            acc += LUMA_TAPS[dx][k] * P[np.ix_(yy + 16, xs + k - 3 + 16)]
        return acc

    if dy == 0:
        return np.clip((hsum(ys) + 32) >> 6, 0, 255).astype(np.uint8)
    if dx == 0:
        acc = np.zeros((len(ys), len(xs)), dtype=np.int64)
        for k in range(8):
            acc += LUMA_TAPS[dy][k] * P[np.ix_(ys + k - 3 + 16, xs + 16)]
        return np.clip((acc + 32) >> 6, 0, 255).astype(np.uint8)
    hs = hsum(np.arange(-7, h + 8)).astype(np.int16).astype(np.int64)   # rows -7 .. h+7
    acc = np.zeros((len(ys), len(xs)), dtype=np.int64)
    for k in range(8):
        acc += LUMA_TAPS[dy][k] * hs[ys + k - 3 + 7]
    return np.clip((acc + 2048) >> 12, 0, 255).astype(np.uint8)


def interp_plane_padded(img, k, pad=80):
    """Plane k = 4*dy + dx over the full padded area [-pad, W+pad) x [-pad, H+pad),
    using the border rule plane(x, y) = interp(clamp(x, -4, W+3), clamp(y, -4, H+3))."""
    h, w = img.shape
    core = interp_plane(img, k & 3, k >> 2)
    yy = np.clip(np.arange(-pad, h + pad), -4, h + 3) + 4
    xx = np.clip(np.arange(-pad, w + pad), -4, w + 3) + 4
    return core[np.ix_(yy, xx)]


def block_filter_luma(src_pad, x0, y0, w, h, dx, dy):
    """Block MC from a padded array src_pad whose picture origin is (x0, y0)
    relative to the block: the block's top-left full-pel sample is src_pad[y0, x0]."""
    S = src_pad.astype(np.int64)
    if dx == 0 and dy == 0:
        return S[y0:y0 + h, x0:x0 + w].astype(np.uint8)
    if dy == 0:
        acc = sum(LUMA_TAPS[dx][k] * S[y0:y0 + h, x0 + k - 3:x0 + k - 3 + w] for k in range(8))
        return np.clip((acc + 32) >> 6, 0, 255).astype(np.uint8)
    if dx == 0:
        acc = sum(LUMA_TAPS[dy][k] * S[y0 + k - 3:y0 + k - 3 + h, x0:x0 + w] for k in range(8))
        return np.clip((acc + 32) >> 6, 0, 255).astype(np.uint8)
    hs = sum(LUMA_TAPS[dx][k] * S[y0 - 3:y0 + h + 4, x0 + k - 3:x0 + k - 3 + w] for k in range(8))
    hs = hs.astype(np.int16).astype(np.int64)
    acc = sum(LUMA_TAPS[dy][k] * hs[k:k + h] for k in range(8))
    return np.clip((acc + 2048) >> 12, 0, 255).astype(np.uint8)


def block_filter_chroma(src_pad, x0, y0, w, h, dx, dy):
    S = src_pad.astype(np.int64)
    if dx == 0 and dy == 0:
        return S[y0:y0 + h, x0:x0 + w].astype(np.uint8)
    if dy == 0:
        acc = sum(CHROMA_TAPS[dx][k] * S[y0:y0 + h, x0 + k - 1:x0 + k - 1 + w] for k in range(4))
        return np.clip((acc + 32) >> 6, 0, 255).astype(np.uint8)
    if dx == 0:
        acc = sum(CHROMA_TAPS[dy][k] * S[y0 + k - 1:y0 + k - 1 + h, x0:x0 + w] for k in range(4))
        return np.clip((acc + 32) >> 6, 0, 255).astype(np.uint8)
    hs = sum(CHROMA_TAPS[dx][k] * S[y0 - 1:y0 + h + 2, x0 + k - 1:x0 + k - 1 + w] for k in range(4))
    acc = sum(CHROMA_TAPS[dy][k] * hs[k:k + h] for k in range(4))
    return np.clip((acc + 2048) >> 12, 0, 255).astype(np.uint8)


def dct_matrix(n):
    """AVS2 integer DCT basis: 32 for the DC row, else round(32*sqrt(2)*cos((2j+1)k*pi/(2n)))."""
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    m = np.rint(32 * np.sqrt(2) * np.cos((2 * j + 1) * k * np.pi / (2 * n))).astype(np.int64)
    m[0, :] = 32
    return m


def _wrap16(x):
    return ((x + 32768) % 65536) - 32768


def dct2(block, n):
    """Two-stage integer DCT: Y = T X T^T with the stage roundings and int16 wraps."""
    T = dct_matrix(n)
    l = int(np.log2(n))
    s1, s2 = l - 2, l + 5
    X = block.astype(np.int64)
    A = X @ T.T                       # A[j, k] = sum_m X[j, m] T[k, m]  (rows transformed)
    A = _wrap16((A + ((1 << s1) >> 1)) >> s1)
    B = T @ A                         # B[k2, k] = sum_j T[k2, j] A[j, k] (columns)
    return _wrap16((B + ((1 << s2) >> 1)) >> s2).astype(np.int16)


def idct2(coef, n):
    T = dct_matrix(n)
    C = coef.astype(np.int64)
    A = T.T @ C                       # columns first: A[m, j] = sum_k T[k, m] C[k, j]
    A = np.clip((A + 16) >> 5, -32768, 32767)
    B = A @ T                         # rows: B[m, m2] = sum_j A[m, j] T[j, m2]
    return np.clip((B + 2048) >> 12, -256, 255).astype(np.int16)


def quant(coef, scale, shift, add):
    c = coef.astype(np.int64)
    q = np.where(c < 0, -1, 1) * ((np.abs(c) * scale + add) >> shift)
    return _wrap16(q).astype(np.int16)


def dequant(coef, scale, shift):
    c = coef.astype(np.int64)
    d = np.clip((c * scale + (1 << (shift - 1))) >> shift, -32768, 32767)
    return np.where(c != 0, d, 0).astype(np.int16)


def mv_window(frame_w, frame_h, x, y, w, h, mvr_x=(-8192, 8191), mvr_y=(-2048, 2047)):
    """Legal full-pel MV box of a PU: [min, max] per axis (md_inter.c:910-926)."""
    def axis(pos, size, frame, lo_hi):
        lo = min(max((-64 - pos) * 4, lo_hi[0]), lo_hi[1])
        hi = min(max((frame + 64 - pos - size) * 4, lo_hi[0]), lo_hi[1])
        return (lo >> 2) + 6, (hi >> 2) - 6
    return axis(x, w, frame_w, mvr_x), axis(y, h, frame_h, mvr_y)


def fs_brute(cur, ref_pad, pad, pu, R, lam_f, mvp=(0, 0), frame_wh=None):
    """Integer full search by brute force: the lexicographic minimum of
    (cost, evaluation order) over the MV-candidate list and the raster.
    ref_pad: padded reference (picture origin at [pad, pad])."""
    x, y, w, h = pu
    fw, fh = frame_wh if frame_wh is not None else (cur.shape[1], cur.shape[0])
    (xmin, xmax), (ymin, ymax) = mv_window(fw, fh, x, y, w, h)
    blk = cur[y:y + h, x:x + w].astype(np.int64)

    def bits(v):
        return 1 if v == 0 else se_len(v)

    def cost(mx, my):
        r = ref_pad[pad + y + my:pad + y + my + h, pad + x + mx:pad + x + mx + w].astype(np.int64)
        return int(np.abs(blk - r).sum()) + ((lam_f * (bits(4 * mx - mvp[0]) + bits(4 * my - mvp[1]))) >> 16)

    def legal(mx, my):
        return xmin <= mx <= xmax and ymin <= my <= ymax

    cands = []
    for q in (mvp, (0, 0)):
        c = (min(max((q[0] + 2) >> 2, xmin), xmax), min(max((q[1] + 2) >> 2, ymin), ymax))
        if c not in cands:
            cands.append(c)
    best = None
    for order, (mx, my) in enumerate(cands):
        if legal(mx, my):
            key = (cost(mx, my), order)
            best = key + ((mx, my),) if best is None or key < best[:2] else best
    if best is None:
        return None
    omx, omy = best[2]
    order = 2
    for j in range(-R, R):
        for i in range(-R, R):
            mx, my = omx + i, omy + j
            if legal(mx, my):
                key = (cost(mx, my), order)
                if key < best[:2]:
                    best = key + ((mx, my),)
            order += 1
    return best[0], best[2]
