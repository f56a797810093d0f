"""GPU parity of the per-call faces with the exact g_funcs signatures, called
through ctypes function pointers as an encoder would after installing them:
dctf.dct / idct / dct_half for every size of the reference's table (with and
without the stride's wavelet flag, in place and out of place), dctf.quant /
dequant, pixf.avg, and every interpolation member (intpl_luma_hor/ver/ext and
their _x3 forms, the luma and chroma block filters) with the reference's taps
and with arbitrary taps, against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests import npref
from xavs2_amd import lib as X

pytestmark = pytest.mark.gpu

DCT_FT = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_int)
TABLE_SIZES = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (16, 4), (4, 16), (32, 8), (8, 32), (64, 16), (16, 64)]


def at(a, off):
    return C.c_void_p(a.ctypes.data + off * a.itemsize)


def tables():
    d, i, h = (C.c_void_p * 25)(), (C.c_void_p * 25)(), (C.c_void_p * 25)()
    assert X.lib().xme_get_dct_table(d, i, h) == 0
    return d, i, h


def test_dct_table_entries_are_the_references():
    d, i, h = tables()
    for part in range(25):
        wh = (X.PU_W[part], X.PU_H[part])
        assert bool(d[part]) == (wh in TABLE_SIZES) and bool(i[part]) == (wh in TABLE_SIZES), wh
        assert bool(h[part]) == (wh in ((32, 32), (64, 64))), wh


@pytest.mark.parametrize("w,h", TABLE_SIZES)
def test_dct_idct_half_match_oracle(w, h):
    d, i, hf = tables()
    part = X.part_index(w, h)
    L = O.lib()
    rng = np.random.default_rng(w * 7 + h)
    flagged = (w, h) in ((32, 32), (32, 8), (8, 32))
    wave = 64 in (w, h)
    n_out = (w // 2) * (h // 2) if wave else w * h
    for trial in range(3):
        blk = rng.integers(-255, 256, (h, w)).astype(np.int16)
        if trial == 2:
            blk = rng.choice(np.array([-32768, 32767, 0], np.int16), (h, w))
        for flag in ([0, 1] if flagged else [0]):
            stride = w if (wave or flagged) else w + 5
            src = np.zeros((h, stride), np.int16)
            src[:, :w] = blk
            got = np.zeros(w * h, np.int16)
            DCT_FT(d[part])(O.ptr(src), O.ptr(got), stride | flag)
            exp = np.zeros(w * h, np.int16)
            assert L.xo_dct_part(w, h, O.ptr(src), O.ptr(exp), stride | flag) == 0
            assert np.array_equal(got[:n_out], exp[:n_out]), (trial, flag)
            # in place (rdo.c:769 passes one buffer as src and dst)
            inplace = src.copy()
            DCT_FT(d[part])(O.ptr(inplace), O.ptr(inplace), stride | flag)
            assert np.array_equal(inplace.ravel()[:n_out], exp[:n_out])
            coef = rng.integers(-3000, 3001, n_out).astype(np.int16)
            if trial == 2:
                coef = rng.choice(np.array([-32768, 32767, 0], np.int16), n_out)
            gi = np.zeros(h * stride, np.int16)
            ei = np.zeros(h * stride, np.int16)
            DCT_FT(i[part])(O.ptr(coef), O.ptr(gi), stride | flag)
            L.xo_idct_part(w, h, O.ptr(coef), O.ptr(ei), stride | flag)
            assert np.array_equal(gi, ei), (trial, flag)
            if hf[part]:
                DCT_FT(hf[part])(O.ptr(src), O.ptr(got), stride | flag)
                L.xo_dct_half_part(w, h, O.ptr(src), O.ptr(exp), stride | flag)
                assert np.array_equal(got[:n_out], exp[:n_out])


def test_quant_dequant_match_oracle():
    q, dq = C.c_void_p(), C.c_void_p()
    assert X.lib().xme_get_quant_funcs(C.byref(q), C.byref(dq)) == 0
    quant = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int)(q.value)
    dequant = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int)(dq.value)
    rng = np.random.default_rng(3)
    for n, qp in ((16, 22), (64, 32), (256, 37), (1024, 42), (1024, 0), (4096, 63)):
        log2n = int(np.log2(np.sqrt(n)))
        scale, shift, add = (C.c_int(), C.c_int(), C.c_int())
        O.lib().xo_quant_params(min(log2n, 5), qp, 0, C.byref(scale), C.byref(shift), C.byref(add))
        c = rng.integers(-32768, 32768, n).astype(np.int16)
        c[rng.random(n) < 0.3] = 0
        g, e = c.copy(), c.copy()
        assert quant(O.ptr(g), n, scale.value, shift.value, add.value) == \
            O.lib().xo_quant(O.ptr(e), n, scale.value, shift.value, add.value)
        assert np.array_equal(g, e)
        ds, dsh = C.c_int(), C.c_int()
        O.lib().xo_dequant_params(min(log2n, 5), qp, C.byref(ds), C.byref(dsh))
        dequant(O.ptr(g), n, ds.value, dsh.value)
        O.lib().xo_dequant(O.ptr(e), n, ds.value, dsh.value)
        assert np.array_equal(g, e)


def test_avg_table_matches_oracle():
    t = (C.c_void_p * 25)()
    assert X.lib().xme_get_avg_table(t) == 0
    ft = C.CFUNCTYPE(None, C.c_void_p, C.c_ssize_t, C.c_void_p, C.c_ssize_t, C.c_void_p, C.c_ssize_t, C.c_int)
    rng = np.random.default_rng(4)
    a = rng.integers(0, 256, (80, 100), dtype=np.uint8)
    b = rng.integers(0, 256, (70, 90), dtype=np.uint8)
    for part in range(25):
        w, h = X.PU_W[part], X.PU_H[part]
        got = np.zeros((64, 72), np.uint8)
        ft(t[part])(O.ptr(got), 72, at(a, 3 * 100 + 5), 100, at(b, 2 * 90 + 7), 90, 32)
        exp = np.zeros((64, 72), np.uint8)
        O.lib().xo_avg(w, h, O.ptr(exp), 72, at(a, 3 * 100 + 5), 100, at(b, 2 * 90 + 7), 90)
        assert np.array_equal(got, exp), part


def intpl_funcs():
    f = X.IntplFuncs()
    assert X.lib().xme_get_intpl_funcs(C.byref(f)) == 0
    return f


P, I = C.c_void_p, C.c_int
HOR_T = C.CFUNCTYPE(None, P, I, P, I, P, I, I, I, P)
VER_T = C.CFUNCTYPE(None, P, I, P, I, I, I, P)
EXT_T = C.CFUNCTYPE(None, P, I, P, I, I, I, P)
X3_VER_T = C.CFUNCTYPE(None, P, I, P, I, I, I, P)
X3_HOR_T = C.CFUNCTYPE(None, P, I, P, I, P, I, I, I, P)
X3_EXT_T = C.CFUNCTYPE(None, P, I, P, I, I, I, P)
BLK_T = C.CFUNCTYPE(None, P, I, P, I, I, I, P)
BLKX_T = C.CFUNCTYPE(None, P, I, P, I, I, I, P, P)


@pytest.mark.parametrize("taps", ["reference", "arbitrary"])
def test_luma_interpolation_members_match_oracle(taps):
    f = intpl_funcs()
    L = O.lib()
    rng = np.random.default_rng(11 if taps == "reference" else 12)
    sets = [np.ascontiguousarray(npref.LUMA_TAPS[k], np.int8) for k in (1, 2, 3)] if taps == "reference" else \
        [rng.integers(-128, 128, 8).astype(np.int8) for _ in range(3)]
    W, H, S = 300, 40, 320
    # rows 5 - 3 .. 5 + h + 4 and columns 6 - 3 .. 6 + w + 4 of every case inside src: a
    # read past it would take whatever memory follows, which may change between the GPU
    # call and the oracle's
    src = rng.integers(0, 256, (H + 32, S), dtype=np.uint8)
    s0 = at(src, 5 * S + 6)                   # the block's first sample, margins inside src
    for w, h in ((300, 40), (64, 64 - 16), (8, 4), (17, 3)):
        assert 5 + h + 4 <= src.shape[0] and 6 + w + 4 <= src.shape[1], (w, h)
        c = sets[0]
        gd, ed = np.zeros((h, w + 9), np.uint8), np.zeros((h, w + 9), np.uint8)
        gt, et = np.zeros((h, w + 3), np.int16), np.zeros((h, w + 3), np.int16)
        HOR_T(f.intpl_luma_hor)(O.ptr(gd), w + 9, O.ptr(gt), w + 3, s0, S, w, h, O.ptr(c))
        L.xo_intpl_luma_hor(O.ptr(ed), w + 9, O.ptr(et), w + 3, s0, S, w, h, O.ptr(c))
        assert np.array_equal(gd, ed) and np.array_equal(gt, et), (w, h)
        VER_T(f.intpl_luma_ver)(O.ptr(gd), w + 9, s0, S, w, h, O.ptr(c))
        L.xo_intpl_luma_ver(O.ptr(ed), w + 9, s0, S, w, h, O.ptr(c))
        assert np.array_equal(gd, ed)
        BLK_T(f.intpl_luma_block_hor)(O.ptr(gd), w + 9, s0, S, w, h, O.ptr(c))
        L.xo_intpl_luma_hor(O.ptr(ed), w + 9, None, 0, s0, S, w, h, O.ptr(c))
        assert np.array_equal(gd, ed)
        BLK_T(f.intpl_luma_block_ver)(O.ptr(gd), w + 9, s0, S, w, h, O.ptr(c))
        L.xo_intpl_luma_ver(O.ptr(ed), w + 9, s0, S, w, h, O.ptr(c))
        assert np.array_equal(gd, ed)
        if w <= 64 and h <= 64:
            BLKX_T(f.intpl_luma_block_ext)(O.ptr(gd), w + 9, s0, S, w, h, O.ptr(sets[1]), O.ptr(sets[2]))
            L.xo_intpl_luma_block_ext(O.ptr(ed), w + 9, s0, S, w, h, O.ptr(sets[1]), O.ptr(sets[2]))
            assert np.array_equal(gd, ed)
        tmp = rng.integers(-5000, 22000, (h + 8, w + 2)).astype(np.int16)
        t0 = at(tmp, 3 * (w + 2) + 1)
        EXT_T(f.intpl_luma_ext)(O.ptr(gd), w + 9, t0, w + 2, w, h, O.ptr(c))
        L.xo_intpl_luma_ext(O.ptr(ed), w + 9, t0, w + 2, w, h, O.ptr(c))
        assert np.array_equal(gd, ed)
        # the x3 forms: three planes at once, the same as three single calls
        cptr = (C.c_void_p * 3)(*[s.ctypes.data for s in sets])
        outs = [np.zeros((h, w + 9), np.uint8) for _ in range(3)]
        tmps = [np.zeros((h, w + 3), np.int16) for _ in range(3)]
        dptr = (C.c_void_p * 3)(*[o.ctypes.data for o in outs])
        tptr = (C.c_void_p * 3)(*[t.ctypes.data for t in tmps])
        X3_HOR_T(f.intpl_luma_hor_x3)(dptr, w + 9, tptr, w + 3, s0, S, w, h, cptr)
        for k in range(3):
            L.xo_intpl_luma_hor(O.ptr(ed), w + 9, O.ptr(et), w + 3, s0, S, w, h, O.ptr(sets[k]))
            assert np.array_equal(outs[k], ed) and np.array_equal(tmps[k], et), k
        X3_VER_T(f.intpl_luma_ver_x3)(dptr, w + 9, s0, S, w, h, cptr)
        for k in range(3):
            L.xo_intpl_luma_ver(O.ptr(ed), w + 9, s0, S, w, h, O.ptr(sets[k]))
            assert np.array_equal(outs[k], ed), k
        X3_EXT_T(f.intpl_luma_ext_x3)(dptr, w + 9, t0, w + 2, w, h, cptr)
        for k in range(3):
            L.xo_intpl_luma_ext(O.ptr(ed), w + 9, t0, w + 2, w, h, O.ptr(sets[k]))
            assert np.array_equal(outs[k], ed), k


def test_chroma_block_members_match_oracle():
    f = intpl_funcs()
    L = O.lib()
    rng = np.random.default_rng(13)
    src = rng.integers(0, 256, (48, 60), dtype=np.uint8)
    s0 = at(src, 2 * 60 + 3)
    for dx, dy in ((1, 0), (0, 5), (3, 7), (6, 2)):
        ch = np.ascontiguousarray(npref.CHROMA_TAPS[dx], np.int8)
        cv = np.ascontiguousarray(npref.CHROMA_TAPS[dy], np.int8)
        for w, h in ((2, 2), (4, 8), (16, 16), (32, 32), (6, 12)):
            g, e = np.zeros((h, 40), np.uint8), np.zeros((h, 40), np.uint8)
            BLK_T(f.intpl_chroma_block_hor)(O.ptr(g), 40, s0, 60, w, h, O.ptr(ch))
            L.xo_intpl_chroma_block_hor(O.ptr(e), 40, s0, 60, w, h, O.ptr(ch))
            assert np.array_equal(g, e)
            BLK_T(f.intpl_chroma_block_ver)(O.ptr(g), 40, s0, 60, w, h, O.ptr(cv))
            L.xo_intpl_chroma_block_ver(O.ptr(e), 40, s0, 60, w, h, O.ptr(cv))
            assert np.array_equal(g, e)
            BLKX_T(f.intpl_chroma_block_ext)(O.ptr(g), 40, s0, 60, w, h, O.ptr(ch), O.ptr(cv))
            L.xo_intpl_chroma_block_ext(O.ptr(e), 40, s0, 60, w, h, O.ptr(ch), O.ptr(cv))
            assert np.array_equal(g, e)
