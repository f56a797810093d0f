"""Pins the oracle's integer DCT/IDCT and flat quant/dequant."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tests import npref

REF = "/root/reference/source/common"


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_matrix_closed_form_and_butterfly_symmetries(n):
    T = O.tmat(n).astype(np.int64)
    assert np.array_equal(T, npref.dct_matrix(n))
    # T_N[k] = T_32[k*32/N] restricted to the first N columns
    assert np.array_equal(T, O.tmat(32)[:: 32 // n, :n])
    # even/odd symmetry T[k][N-1-j] = (-1)^k T[k][j]: the partial butterflies of
    # transform.c rely on it, and with it they compute exactly this matrix product
    for k in range(n):
        assert np.array_equal(T[k, ::-1], (-1) ** k * T[k])
    # and recursively for the even half: T[2k][:N/2] behaves as an N/2-point basis
    if n > 4:
        assert np.array_equal(T[::2, : n // 2], O.tmat(n // 2))
    # near-orthogonality: every row has squared norm ~ 32^2 * N (the DC row exactly)
    G = T @ T.T
    d = np.diag(G)
    assert d[0] == 1024 * n
    assert np.all(np.abs(d - 1024 * n) <= 0.02 * 1024 * n)
    off = G - np.diag(d)
    assert np.abs(off).max() <= 0.02 * 1024 * n


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not mounted")
def test_tables_equal_reference_text():
    """Read the reference's tables as text (no build) and compare entry by entry."""
    src = open(os.path.join(REF, "transform.c"), encoding="latin-1").read()
    for n in (4, 8, 16, 32):
        m = re.search(r"g_T%d\[%d\]\[%d\]\s*=\s*\{(.*?)\};" % (n, n, n), src, re.S)
        vals = [int(v) for v in re.findall(r"-?\d+", re.sub(r"//[^\n]*", "", m.group(1)))]
        assert np.array_equal(np.array(vals).reshape(n, n), O.tmat(n)), n
    q = open(os.path.join(REF, "quant.c"), encoding="latin-1").read()
    L = O.lib()
    for name, arr, typ in (("tab_Q_TAB", "xo_Q_TAB", O.C.c_uint16), ("tab_IQ_TAB", "xo_IQ_TAB", O.C.c_uint16),
                           ("tab_IQ_SHIFT", "xo_IQ_SHIFT", O.C.c_uint8)):
        m = re.search(name + r"\[80\]\s*=\s*\{(.*?)\};", q, re.S)
        vals = [int(v) for v in re.findall(r"\d+", m.group(1))]
        ours = list((typ * 80).in_dll(L, arr))
        assert vals == ours, name
    mc = open(os.path.join(REF, "mc.c"), encoding="latin-1").read()
    m = re.search(r"INTPL_FILTERS\[4\]\[8\]\)\s*=\s*\{(.*?)\};", mc, re.S)
    assert [int(v) for v in re.findall(r"-?\d+", m.group(1))] == list(npref.LUMA_TAPS.ravel())
    m = re.search(r"INTPL_FILTERS_C\[8\]\[4\]\)\s*=\s*\{(.*?)\};", mc, re.S)
    vals = [int(v) for v in re.findall(r"-?\d+", re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S))]
    assert vals == list(npref.CHROMA_TAPS.ravel())


def test_quant_table_relations():
    L = O.lib()
    Q = np.array(list((O.C.c_uint16 * 80).in_dll(L, "xo_Q_TAB")), dtype=np.float64)
    IQ = np.array(list((O.C.c_uint16 * 80).in_dll(L, "xo_IQ_TAB")), dtype=np.float64)
    S = np.array(list((O.C.c_uint8 * 80).in_dll(L, "xo_IQ_SHIFT")), dtype=np.float64)
    # forward and inverse scales are reciprocal up to 2^(15 + shift), within the
    # rounding of the (small, for high QP) integer forward scale
    rel = Q * IQ / 2.0 ** (15 + S)
    assert np.all(np.abs(rel - 1) <= 0.6 / Q + 1e-3)
    # the step doubles every 8 QP
    assert np.all(np.abs(Q[8:] / Q[:-8] - 0.5) < 0.01)
    assert np.all(np.diff(Q) < 0)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_dct_idct_match_matrix_form(n):
    rng = np.random.default_rng(n)
    for t in range(20):
        if t == 0:
            blk = np.full((n, n), 255, np.int16)
        elif t == 1:   # adversarial: maximal alternating residual
            blk = (np.where((np.add.outer(np.arange(n), np.arange(n)) % 2) == 0, 255, -255)).astype(np.int16)
        elif t == 2:   # beyond the residual range to exercise the int16 wraps of the first stage
            blk = rng.integers(-32768, 32767, (n, n)).astype(np.int16)
        else:
            blk = rng.integers(-255, 256, (n, n)).astype(np.int16)
        c = O.dct(n, blk)
        assert np.array_equal(c, npref.dct2(blk, n)), (n, t)
        coef = rng.integers(-2000, 2000, (n, n)).astype(np.int16) if t % 2 else c
        assert np.array_equal(O.idct(n, coef), npref.idct2(coef, n)), (n, t)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_dct_orientation_dc_and_round_trip(n):
    # a flat block: only the DC is nonzero and it is exact
    for v in (-255, -3, 1, 77, 255):
        c = O.dct(n, np.full((n, n), v, np.int16))
        l = int(np.log2(n))
        s1, s2 = l - 2, l + 5
        a = (32 * v * n + ((1 << s1) >> 1)) >> s1
        dc = (32 * a * n + ((1 << s2) >> 1)) >> s2
        assert c[0, 0] == dc and np.count_nonzero(c) == (1 if dc else 0)
    # variation along x only: every row identical -> only the first coefficient row is nonzero
    row = np.random.default_rng(3).integers(-200, 200, n)
    c = O.dct(n, np.tile(row, (n, 1)).astype(np.int16))
    assert np.count_nonzero(c[1:]) == 0 and np.count_nonzero(c[0]) > 0
    c = O.dct(n, np.tile(row[:, None], (1, n)).astype(np.int16))
    assert np.count_nonzero(c[:, 1:]) == 0 and np.count_nonzero(c[:, 0]) > 0
    # forward then inverse reproduces the residual up to the integer basis' small
    # non-orthogonality (a transposed stage or a wrong shift is off by hundreds)
    blk = np.random.default_rng(n + 5).integers(-255, 256, (n, n)).astype(np.int16)
    err = np.abs(O.idct(n, O.dct(n, blk)).astype(int) - blk)
    assert err.max() <= 8 and err.mean() <= 1.5


def test_dct32_stride_flag():
    n = 32
    blk = np.random.default_rng(1).integers(-255, 256, (n, n)).astype(np.int16)
    out0 = np.zeros((n, n), np.int16); out1 = np.zeros((n, n), np.int16)
    O.lib().xo_dct(32, O.ptr(blk), O.ptr(out0), 32)
    O.lib().xo_dct(32, O.ptr(blk), O.ptr(out1), 33)          # low bit: one more bit of first shift
    T = npref.dct_matrix(32)
    A = blk.astype(np.int64) @ T.T
    A = ((A + 8) >> 4)
    A = ((A + 32768) % 65536) - 32768
    B = T @ A
    assert np.array_equal(out1, ((((B + 512) >> 10) + 32768) % 65536 - 32768).astype(np.int16))
    assert not np.array_equal(out0, out1)


@pytest.mark.parametrize("n,qp,b_intra", [(4, 22, 0), (8, 27, 1), (16, 32, 0), (32, 37, 0), (32, 42, 1),
                                          (4, 0, 0), (32, 79, 0)])
def test_quant_dequant_formula(n, qp, b_intra):
    l = int(np.log2(n))
    scale, shift, add = O.quant_params(l, qp, b_intra)
    assert shift == 22 - l and add == ((1 << (shift + b_intra)) * 5) // 31
    rng = np.random.default_rng(qp)
    c = rng.integers(-32768, 32768, n * n).astype(np.int16)
    c[:5] = [0, -1, 1, -32768, 32767]
    q = c.copy()
    nz = O.lib().xo_quant(O.ptr(q), n * n, scale, shift, add)
    assert np.array_equal(q, npref.quant(c, scale, shift, add))
    assert nz == np.count_nonzero(q)
    dscale, dshift = O.dequant_params(l, qp)
    d = q.copy()
    O.lib().xo_dequant(O.ptr(d), n * n, dscale, dshift)
    assert np.array_equal(d, npref.dequant(q, dscale, dshift))


def test_residual_batch_composes_the_stages():
    rng = np.random.default_rng(4)
    for n in (4, 8, 16, 32):
        blocks = rng.integers(-255, 256, (5, n, n)).astype(np.int16)
        blocks[1] = 0
        level, recon, nnz = O.residual_batch(n, 32, 0, blocks)
        l = int(np.log2(n))
        s, sh, a = O.quant_params(l, 32, 0)
        ds, dsh = O.dequant_params(l, 32)
        for b in range(5):
            lv = npref.quant(npref.dct2(blocks[b], n), s, sh, a)
            assert np.array_equal(level[b], lv)
            assert nnz[b] == np.count_nonzero(lv)
            assert np.array_equal(recon[b], npref.idct2(npref.dequant(lv, ds, dsh), n))
        assert nnz[1] == 0 and not recon[1].any()
