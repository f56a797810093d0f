"""Pins the oracle's pixel cost primitives (SAD, SAD_x3/x4, SATD, AVG) against
independent numpy formulations and closed-form special cases."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref


@pytest.mark.parametrize("part", range(25))
def test_sad_satd_random_and_extreme(part):
    w, h = O.PU_W[part], O.PU_H[part]
    rng = np.random.default_rng(1000 + part)
    for trial in range(6):
        if trial == 0:
            a = np.zeros((h, w), np.uint8); b = np.full((h, w), 255, np.uint8)
        elif trial == 1:   # checkerboard extremes: maximal Hadamard energy
            yy, xx = np.mgrid[0:h, 0:w]
            a = np.where((yy + xx) % 2 == 0, 255, 0).astype(np.uint8)
            b = 255 - a
        else:
            a = rng.integers(0, 256, (h, w), dtype=np.uint8)
            b = rng.integers(0, 256, (h, w), dtype=np.uint8)
        assert O.sad(w, h, a, b) == npref.sad(a, b)
        assert O.satd(w, h, a, b) == npref.satd(a, b)
        assert npref.satd_tiles_even(a, b)


def test_satd_constant_difference_closed_form():
    # D = d everywhere: only the DC of each 4x4 tile is nonzero (16|d|), SATD = 8|d| per tile.
    for d in (-255, -17, -1, 0, 1, 5, 255):
        a = np.full((16, 16), 128 if abs(d) < 128 else (255 if d > 0 else 0), np.uint8)
        b = (a.astype(int) - d).clip(0, 255).astype(np.uint8)
        dd = int(a[0, 0]) - int(b[0, 0])
        assert O.satd(16, 16, a, b) == 16 * 8 * abs(dd)


def test_satd_single_impulse():
    # a single nonzero difference e spreads to |e| in all 16 coefficients: 16|e|/2 = 8|e|
    a = np.zeros((8, 8), np.uint8); b = np.zeros((8, 8), np.uint8)
    a[5, 2] = 200
    assert O.satd(8, 8, a, b) == 8 * 200


def test_sad_x3_x4_use_fenc_stride():
    rng = np.random.default_rng(7)
    fenc = rng.integers(0, 256, (64, O.FENC_STRIDE), dtype=np.uint8)
    ref = rng.integers(0, 256, (96, 200), dtype=np.uint8)
    L = O.lib()
    for part in (2, 19, 24, 14):
        w, h = O.PU_W[part], O.PU_H[part]
        offs = [(3, 5), (0, 0), (17, 9), (30, 1)]
        ptrs = [O.C.c_void_p(ref.ctypes.data + y * ref.strides[0] + x) for x, y in offs]
        r3 = np.zeros(3, np.int32); r4 = np.zeros(4, np.int32)
        L.xo_sad_x3(w, h, O.ptr(fenc), ptrs[0], ptrs[1], ptrs[2], ref.strides[0], O.ptr(r3))
        L.xo_sad_x4(w, h, O.ptr(fenc), *ptrs, ref.strides[0], O.ptr(r4))
        exp = [npref.sad(fenc[:h, :w], ref[y:y + h, x:x + w]) for x, y in offs]
        assert list(r3) == exp[:3]
        assert list(r4) == exp


def test_avg_rounds_up():
    a = np.array([[0, 1, 254, 255]], np.uint8); b = np.array([[1, 1, 255, 255]], np.uint8)
    out = np.zeros((1, 4), np.uint8)
    O.lib().xo_avg(4, 1, O.ptr(out), 4, O.ptr(a), 4, O.ptr(b), 4)
    assert list(out[0]) == [1, 1, 255, 255]


def test_mvbits_is_exp_golomb_length():
    for R in (32, 64, 256, 300):
        maxmvd = O.lib().xo_mvbits_max_mvd(R)
        assert maxmvd >= 4 * (2 * max(R, 256) + 3)
        for v in list(range(-300, 301)) + [maxmvd, -maxmvd, 8191, -8192, 4096, -4095]:
            if abs(v) <= maxmvd:
                assert O.mvbits(R, v) == (1 if v == 0 else npref.se_len(v)), (R, v)
    assert O.lib().xo_mvbits_max_mvd(64) == 16383


def test_lambda_factor():
    assert O.lambda_factor(16.0) == 4 * 65536   # BASELINE lambda 4 means sqrt(lambda_mode) = 4
    assert O.lambda_factor(2.0) == int(65536 * 2 ** 0.5 + 0.5)
