"""Pins the oracle's restatement of the per-call g_funcs entries that have no
batched counterpart: the whole dct_t table (transform.c:1599-1643: square,
asymmetric and 64-wide sizes, dct_half, the wavelet flag in the stride) and
the interpolation members with their own taps (mc.c:275-547).

* Every table size: a constant block transforms to DC = 128 c (the two
  passes' shifts sum to log2(W) + log2(H) + 3, which the DC fixes) and back.
* Non-64 sizes: numpy integer matrix products T_W B^T, T_H (.)^T with the
  shifts written from the reference's named constants (B*_IN_BIT,
  FACTO_BIT, LIMIT_BIT) per function, int16 casts and clips, on random and
  int16-extreme inputs, with and without the wavelet flag.
* The lifting wavelet on hand-worked ramps (the mirror at the far edge
  gives the one odd value), and its inverse.
* dct_half: the full transform's low 16x16, zeros elsewhere.
* Interpolation with random taps against numpy separable FIRs, including taps
  large enough that the luma int16 intermediate wraps."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref

B_IN_BIT = {4: 2, 8: 3, 16: 4, 32: 5}
FACTO_BIT, LIMIT_BIT, BIT_DEPTH = 5, 16, 8
SIZES = [(4, 4), (8, 8), (16, 16), (32, 32), (16, 4), (4, 16), (32, 8), (8, 32)]
SIZES64 = [(64, 64), (64, 16), (16, 64)]


def i16(a):
    return ((np.asarray(a, np.int64) + 32768) % 65536 - 32768).astype(np.int64)


def ref_shifts(w, h, flag):
    """(shift1, shift2) of dct_WxH_c as the reference writes them (transform.c:1123-1490)."""
    if w == h:
        s1 = B_IN_BIT[w] + FACTO_BIT + BIT_DEPTH + 1 - LIMIT_BIT + (flag if w == 32 else 0)
        return s1, B_IN_BIT[w] + FACTO_BIT
    if (w, h) == (16, 4):
        return B_IN_BIT[16] + FACTO_BIT + BIT_DEPTH + 1 - LIMIT_BIT, B_IN_BIT[16] + FACTO_BIT - 2
    if (w, h) == (4, 16):
        return B_IN_BIT[16] + FACTO_BIT + BIT_DEPTH + 1 - LIMIT_BIT - 2, B_IN_BIT[16] + FACTO_BIT
    if (w, h) == (32, 8):
        return B_IN_BIT[32] + FACTO_BIT + BIT_DEPTH + 1 - LIMIT_BIT, B_IN_BIT[32] + FACTO_BIT - 2 - flag
    return B_IN_BIT[32] + FACTO_BIT + BIT_DEPTH + 1 - LIMIT_BIT - 2 + flag, B_IN_BIT[32] + FACTO_BIT   # 8x32


def np_dct(w, h, blk, flag=0):
    s1, s2 = ref_shifts(w, h, flag)
    Tw, Th = npref.dct_matrix(w), npref.dct_matrix(h)
    a = i16((blk.astype(np.int64) @ Tw.T + ((1 << s1) >> 1)) >> s1)          # rows: h x w
    return i16((Th @ a + ((1 << s2) >> 1)) >> s2)                              # columns


def np_idct(w, h, coef, flag=0):
    Tw, Th = npref.dct_matrix(w), npref.dct_matrix(h)
    a = np.clip((Th.T @ coef.astype(np.int64) + 16) >> 5, -32768, 32767)        # columns, 16 bits
    s2 = 20 - BIT_DEPTH - flag
    hi = (1 << (BIT_DEPTH + flag)) - 1
    return np.clip((a @ Tw + (1 << (s2 - 1))) >> s2, -hi - 1, hi)


def at(a, off):
    """Address of element `off` (in elements) of array a."""
    return O.C.c_void_p(a.ctypes.data + off * a.itemsize)


def call(fn, w, h, src, stride, out_n):
    dst = np.zeros(out_n, np.int16)
    rc = fn(w, h, O.ptr(np.ascontiguousarray(src, np.int16)), O.ptr(dst), stride)
    return rc, dst


@pytest.mark.parametrize("w,h", SIZES + SIZES64)
@pytest.mark.parametrize("c", [1, -3, 7, 40, -255])
def test_constant_block_dc_gain_and_back(w, h, c):
    """DC = 128 c and back to c.  64x16 is the exception: its 32x8 stage takes the
    wavelet flag off shift2 (transform.c:1416) where the other flagged stages add it
    to a shift, so the low band (2c) comes out at 512 c, and the inverse (shift 11 on
    the 32x8 stage, then the wavelet) returns 4 c.  At c = -255 its first pass
    (256 c = -65280) wraps to 256 in int16, and the DC is 512."""
    L = O.lib()
    blk = np.full(w * h, c, np.int16)
    rc, coef = call(L.xo_dct_part, w, h, blk, w, w * h)
    assert rc == 0
    cw, ch = (w // 2, h // 2) if 64 in (w, h) else (w, h)
    coef = coef[: cw * ch].reshape(ch, cw)
    if (w, h) == (64, 16):
        assert coef[0, 0] == (512 if c == -255 else 512 * c)
    else:
        assert coef[0, 0] == 128 * c
    assert np.count_nonzero(coef) == 1
    rc, back = call(L.xo_idct_part, w, h, coef.ravel(), w, w * h)
    assert rc == 0
    if (w, h) == (64, 16):
        if c != -255:
            assert np.all(back == 4 * c)
    else:
        assert np.all(back == c)


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("kind", ["random", "extreme"])
def test_rectangular_transforms_match_numpy(w, h, kind):
    L = O.lib()
    rng = np.random.default_rng(w * 100 + h)
    flags = [0, 1] if (w, h) in ((32, 32), (32, 8), (8, 32)) else [0]
    for t in range(6):
        if kind == "random":
            blk = rng.integers(-255, 256, (h, w))
        else:
            blk = rng.choice([-32768, 32767, 0, 255, -255], (h, w))
        for flag in flags:
            stride = w + 6 if not flags[-1] else w          # a padded stride where none is masked
            src = np.zeros((h, stride), np.int16)
            src[:, :w] = blk
            rc, got = call(L.xo_dct_part, w, h, src.ravel(), stride | flag, w * h)
            assert rc == 0
            exp = np_dct(w, h, blk, flag)
            assert np.array_equal(got.reshape(h, w), exp), (t, flag)
            coef = rng.integers(-4000, 4001, (h, w)) if kind == "random" else rng.choice([-32768, 32767, 0], (h, w))
            rc, back = call(L.xo_idct_part, w, h, coef.ravel(), stride | flag, stride * h)
            assert np.array_equal(back.reshape(h, stride)[:, :w], np_idct(w, h, coef, flag)), (t, flag)


def test_wavelet_rows_ramp_known_answer():
    """16x64, every row 0, 4, ..., 60.  Rows: the odd samples are 0 except the last,
    60 - (56 + 56) / 2 = 4 (x[16] mirrors x[14]); the even ones keep their value except
    x[14] + (0 + 4 + 2) >> 2 = 57.  Columns are constant: 2x.  Low band: every row
    0, 16, 32, 48, 64, 80, 96, 114."""
    blk = np.tile(np.arange(0, 64, 4, dtype=np.int16), (64, 1))
    low = np.zeros(8 * 32, np.int16)
    O.lib().xo_wavelet_fwd(16, 64, O.ptr(np.ascontiguousarray(blk)), O.ptr(low))
    assert np.array_equal(low.reshape(32, 8), np.tile([0, 16, 32, 48, 64, 80, 96, 114], (32, 1)))


def test_wavelet_columns_ramp_known_answer_and_inverse():
    """64x16, every column 0, 4, ..., 60: rows pass unchanged (constant rows); the
    column pass doubles the even samples and adds (left + right + 1) >> 1 of the odd
    ones, which are 0 but the last (4): 2 * 56 + (0 + 4 + 1) >> 1 = 114.
    Inverse of that low band: halve (0, 8, ..., 48, 57), odd samples the floor mean,
    the mirrored last one (57 + 57) >> 1."""
    blk = np.repeat(np.arange(0, 64, 4, dtype=np.int16)[:, None], 64, axis=1)
    low = np.zeros(32 * 8, np.int16)
    O.lib().xo_wavelet_fwd(64, 16, O.ptr(np.ascontiguousarray(blk)), O.ptr(low))
    col = np.array([0, 16, 32, 48, 64, 80, 96, 114])
    assert np.array_equal(low.reshape(8, 32), np.repeat(col[:, None], 32, axis=1))
    out = np.zeros(64 * 16, np.int16)
    O.lib().xo_wavelet_inv(64, 16, O.ptr(low), O.ptr(out))
    exp = np.array([0, 4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 52, 57, 57])
    assert np.array_equal(out.reshape(16, 64), np.repeat(exp[:, None], 64, axis=1))


def test_wavelet_64x64_row_edges_known_answer():
    """64x64 with every row x = 0..63 (values x): odd samples 0 but the last,
    63 - (62 + 62) >> 1 = 1; the first odd (mirror, x[-1] = 1, x[-2] = 2) is
    1 - (2 + 0) >> 1 = 0; the last even 62 + (0 + 1 + 2) >> 2 = 62.  Columns are
    constant: 2x.  Low band row: 0, 4, 8, ..., 124."""
    blk = np.tile(np.arange(64, dtype=np.int16), (64, 1))
    low = np.zeros(32 * 32, np.int16)
    O.lib().xo_wavelet_fwd(64, 64, O.ptr(np.ascontiguousarray(blk)), O.ptr(low))
    assert np.array_equal(low.reshape(32, 32), np.tile(np.arange(0, 128, 4), (32, 1)))


@pytest.mark.parametrize("w,h", SIZES64)
def test_64_wide_sizes_are_the_half_size_transform_of_the_low_band(w, h):
    L = O.lib()
    rng = np.random.default_rng(w + 3 * h)
    blk = rng.integers(-255, 256, w * h).astype(np.int16)
    low = np.zeros(w * h // 4, np.int16)
    L.xo_wavelet_fwd(w, h, O.ptr(blk), O.ptr(low))
    rc, got = call(L.xo_dct_part, w, h, blk, 12345, w * h)         # the stride is not read
    assert rc == 0
    exp = np_dct(w // 2, h // 2, low.reshape(h // 2, w // 2).astype(np.int64), flag=1)
    assert np.array_equal(got[: w * h // 4].reshape(h // 2, w // 2), exp)
    coef = rng.integers(-3000, 3001, (h // 2, w // 2))
    rc, back = call(L.xo_idct_part, w, h, coef.ravel(), 777, w * h)
    low2 = np_idct(w // 2, h // 2, coef, flag=1).astype(np.int16)
    out = np.zeros(w * h, np.int16)
    L.xo_wavelet_inv(w, h, O.ptr(np.ascontiguousarray(low2.ravel())), O.ptr(out))
    assert np.array_equal(back, out)


@pytest.mark.parametrize("w,h", [(32, 32), (64, 64)])
def test_dct_half_keeps_the_low_16x16(w, h):
    L = O.lib()
    blk = np.random.default_rng(9).integers(-255, 256, w * h).astype(np.int16)
    _, full = call(L.xo_dct_part, w, h, blk, w, 1024)
    rc, half = call(L.xo_dct_half_part, w, h, blk, w, 1024)
    assert rc == 0
    full, half = full.reshape(32, 32), half.reshape(32, 32)
    assert np.array_equal(half[:16, :16], full[:16, :16]) and np.count_nonzero(full[16:]) > 0
    assert not half[16:].any() and not half[:, 16:].any()


def test_sizes_without_a_table_entry_are_refused():
    L = O.lib()
    for w, h in ((8, 4), (4, 8), (16, 8), (64, 32), (12, 16)):
        assert call(L.xo_dct_part, w, h, np.zeros(w * h, np.int16), w, w * h)[0] == -1
        assert call(L.xo_idct_part, w, h, np.zeros(w * h, np.int16), w, w * h)[0] == -1
    assert call(L.xo_dct_half_part, 16, 16, np.zeros(256, np.int16), 16, 256)[0] == -1


# ---------------------------------------------------------------------------
# interpolation members with their own taps

def fir(src, taps, axis, first):
    """sum_k taps[k] * src shifted by (k + first) along axis, for the valid core."""
    n = len(taps)
    S = src.astype(np.int64)
    L = S.shape[axis] - (n - 1)
    acc = 0
    for k in range(n):
        sl = [slice(None), slice(None)]
        sl[axis] = slice(k, k + L)
        acc = acc + taps[k] * S[tuple(sl)]
    return acc


def taps8(rng, big=False):
    return rng.integers(-128, 128, 8) if big else rng.integers(-12, 60, 8)


@pytest.mark.parametrize("big", [False, True])
def test_luma_hor_ver_ext_with_own_taps(big):
    L = O.lib()
    rng = np.random.default_rng(5 + big)
    w, h = 40, 12
    src = rng.integers(0, 256, (h + 7, w + 7)).astype(np.uint8)       # 3 before, 4 after on each axis
    c = taps8(rng, big).astype(np.int8)
    dst = np.zeros((h, w), np.uint8)
    tmp = np.zeros((h, w + 3), np.int16)
    L.xo_intpl_luma_hor(O.ptr(dst), w, O.ptr(tmp), w + 3, at(src, 3 * (w + 7) + 3), w + 7, w, h, O.ptr(c))
    acc = fir(src[3:3 + h], c, 1, -3)
    assert np.array_equal(tmp[:, :w], i16(acc))
    assert np.array_equal(dst, np.clip((acc + 32) >> 6, 0, 255))
    if big:
        assert np.any(i16(acc) != acc)              # the int16 intermediate wraps
    L.xo_intpl_luma_ver(O.ptr(dst), w, at(src, 3 * (w + 7) + 3), w + 7, w, h, O.ptr(c))
    acc = fir(src[:, 3:3 + w], c, 0, -3)
    assert np.array_equal(dst, np.clip((acc + 32) >> 6, 0, 255))
    t16 = rng.integers(-6000, 20000, (h + 7, w)).astype(np.int16)
    L.xo_intpl_luma_ext(O.ptr(dst), w, at(t16, 3 * w), w, w, h, O.ptr(c))
    acc = fir(t16, c, 0, -3)
    assert np.array_equal(dst, np.clip((acc + 2048) >> 12, 0, 255))


def test_luma_block_ext_and_chroma_blocks_with_own_taps():
    L = O.lib()
    rng = np.random.default_rng(17)
    w, h = 24, 16
    src = rng.integers(0, 256, (h + 7, w + 7)).astype(np.uint8)
    ch, cv = taps8(rng).astype(np.int8), taps8(rng).astype(np.int8)
    dst = np.zeros((h, w), np.uint8)
    L.xo_intpl_luma_block_ext(O.ptr(dst), w, at(src, 3 * (w + 7) + 3), w + 7, w, h, O.ptr(ch), O.ptr(cv))
    mid = i16(fir(src[:, :], ch, 1, -3))                     # rows -3 .. h+3
    assert np.array_equal(dst, np.clip((fir(mid, cv, 0, -3) + 2048) >> 12, 0, 255))
    w, h = 16, 8
    s4 = rng.integers(0, 256, (h + 3, w + 3)).astype(np.uint8)   # 1 before, 2 after
    c4h, c4v = rng.integers(-20, 70, 4).astype(np.int8), rng.integers(-20, 70, 4).astype(np.int8)
    d = np.zeros((h, w), np.uint8)
    base = at(s4, (w + 3) + 1)
    L.xo_intpl_chroma_block_hor(O.ptr(d), w, base, w + 3, w, h, O.ptr(c4h))
    assert np.array_equal(d, np.clip((fir(s4[1:1 + h], c4h, 1, -1) + 32) >> 6, 0, 255))
    L.xo_intpl_chroma_block_ver(O.ptr(d), w, base, w + 3, w, h, O.ptr(c4v))
    assert np.array_equal(d, np.clip((fir(s4[:, 1:1 + w], c4v, 0, -1) + 32) >> 6, 0, 255))
    L.xo_intpl_chroma_block_ext(O.ptr(d), w, base, w + 3, w, h, O.ptr(c4h), O.ptr(c4v))
    mid = fir(s4, c4h, 1, -1)                                # int32: no wrap
    assert np.array_equal(d, np.clip((fir(mid, c4v, 0, -1) + 2048) >> 12, 0, 255))
