"""Pins the oracle's luma inter reconstruction (xo_inter_recon_batch: the TU
path of encoder/rdo.c:1333-1366, tu_recon_inter_luma, on the prediction error of
rdo.c:1751-1753 with the prediction of rdo.c:1616-1633):

* a known answer: the reference holds the picture moved by d and the MV is
  4d, so the prediction is exact: zero levels, nnz 0, the reconstruction is
  the picture and the SSD is 0;
* an independent numpy prediction (separable FIR planes) composed with the
  residual stages (pinned on their own in test_oracle_transform.py), numpy
  add / clip and SSD, for every TU size and several QPs, including MVs that
  get_mv_for_mc clamps with the PU's size;
* the full-pel plane alone (block filters) gives the same results as the
  interpolated planes.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref
from tests.test_oracle_sym import cur_plane, planes_of, texture


def tiles(w, h, n, rng, mv_span=60, far=False):
    """Non-overlapping TUs: PUs of 2n x 2n (n x n for n = 32) split into n x n TUs, one MV per PU."""
    pu = n if n == 32 else 2 * n
    out = []
    for py in range(0, h - pu + 1, pu):
        for px in range(0, w - pu + 1, pu):
            mv = (int(rng.integers(-mv_span, mv_span + 1)), int(rng.integers(-mv_span, mv_span + 1)))
            if far and rng.random() < 0.3:
                mv = (4 * int(rng.integers(-300, 300)) + int(rng.integers(0, 4)),
                      4 * int(rng.integers(-300, 300)) + int(rng.integers(0, 4)))
            for ty in range(py, py + pu, n):
                for tx in range(px, px + pu, n):
                    out.append((tx, ty, px, py, pu, pu, mv[0], mv[1]))
    return np.array(out, np.int32)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_exact_prediction_known_answer(n):
    w, h = 128, 96
    cur = texture(w, h, 1)
    d = (3, -2)
    planes = planes_of(np.roll(cur, (d[1], d[0]), axis=(0, 1)))
    tus = tiles(w, h, n, np.random.default_rng(0))
    tus[:, 6], tus[:, 7] = 4 * d[0], 4 * d[1]
    # keep TUs whose reference pixels do not wrap around the rolled picture
    keep = (tus[:, 0] >= 4) & (tus[:, 1] >= 4) & (tus[:, 0] + n <= w - 4) & (tus[:, 1] + n <= h - 4)
    tus = tus[keep]
    level, rec, nnz, ssd = O.inter_recon_batch(n, 32, cur_plane(cur), planes, w, h, tus)
    assert not level.any() and not nnz.any() and not ssd.any()
    for t in tus:
        assert np.array_equal(rec[t[1]:t[1] + n, t[0]:t[0] + n], cur[t[1]:t[1] + n, t[0]:t[0] + n])


def numpy_recon(n, qp, cur, ref, tus):
    h, w = cur.shape
    pad = O.PAD
    pl = [npref.interp_plane_padded(ref, k, pad) for k in range(16)]
    res = np.zeros((len(tus), n, n), np.int16)
    preds = []
    for i, (tx, ty, px, py, pw, ph, mx, my) in enumerate(tus):
        qx = O.lib().xo_get_mc_pos(w, int(pw), int(px), int(mx))
        qy = O.lib().xo_get_mc_pos(h, int(ph), int(py), int(my))
        x0, y0 = (qx >> 2) + (tx - px), (qy >> 2) + (ty - py)
        p = pl[(qy & 3) * 4 + (qx & 3)][pad + y0:pad + y0 + n, pad + x0:pad + x0 + n].astype(np.int32)
        preds.append(p)
        res[i] = cur[ty:ty + n, tx:tx + n].astype(np.int32) - p
    level, rr, nnz = O.residual_batch(n, qp, 0, res)
    rec = np.zeros_like(cur)
    ssd = np.zeros(len(tus), np.int32)
    for i, (tx, ty, *_rest) in enumerate(tus):
        v = np.clip(preds[i] + rr[i].astype(np.int32), 0, 255)
        rec[ty:ty + n, tx:tx + n] = v
        ssd[i] = int(((cur[ty:ty + n, tx:tx + n].astype(np.int64) - v) ** 2).sum())
    return level, rec, nnz, ssd


@pytest.mark.parametrize("n", [4, 8, 16, 32])
@pytest.mark.parametrize("qp", [22, 37, 51])
def test_matches_numpy_composition(n, qp):
    w, h = 128, 96
    cur, ref = texture(w, h, 2), texture(w, h, 3)
    tus = tiles(w, h, n, np.random.default_rng(n + qp), far=True)
    got = O.inter_recon_batch(n, qp, cur_plane(cur), planes_of(ref), w, h, tus)
    exp = numpy_recon(n, qp, cur, ref, tus)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    assert got[3].any() and got[0].any()


def test_full_pel_plane_alone_gives_the_same_results():
    w, h = 128, 96
    cur, ref = texture(w, h, 4), texture(w, h, 5)
    full = planes_of(ref)
    only0 = [full[0]] + [None] * 15
    tus = tiles(w, h, 8, np.random.default_rng(7), far=True)
    a = O.inter_recon_batch(8, 27, cur_plane(cur), full, w, h, tus)
    b = O.inter_recon_batch(8, 27, cur_plane(cur), only0, w, h, tus)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_bi_prediction_known_answer_and_reductions():
    """Two references moved by d and e with MVs 4d and 4e average to the
    picture (zero residual); the same reference and MV twice reduce to the
    uni-predicted result; and the average is numpy's (a + b + 1) >> 1."""
    w, h = 128, 96
    cur = texture(w, h, 6)
    d, e = (2, -3), (-1, 2)
    p1 = planes_of(np.roll(cur, (d[1], d[0]), axis=(0, 1)))
    p2 = planes_of(np.roll(cur, (e[1], e[0]), axis=(0, 1)))
    tus = tiles(w, h, 8, np.random.default_rng(1))
    keep = (tus[:, 2] >= 8) & (tus[:, 3] >= 8) & (tus[:, 2] + tus[:, 4] <= w - 8) & (tus[:, 3] + tus[:, 5] <= h - 8)
    tus = tus[keep]
    tus[:, 6], tus[:, 7] = 4 * d[0], 4 * d[1]
    mv2 = np.tile(np.array([4 * e[0], 4 * e[1]], np.int32), (len(tus), 1))
    level, rec, nnz, ssd = O.inter_recon_batch(8, 27, cur_plane(cur), p1, w, h, tus, p2, mv2)
    assert not level.any() and not nnz.any() and not ssd.any()
    # the same reference and MV twice
    r1, r2 = texture(w, h, 7), texture(w, h, 8)
    q1, q2 = planes_of(r1), planes_of(r2)
    tus = tiles(w, h, 16, np.random.default_rng(2), far=True)
    same = O.inter_recon_batch(16, 30, cur_plane(cur), q1, w, h, tus, q1, tus[:, 6:8])
    uni = O.inter_recon_batch(16, 30, cur_plane(cur), q1, w, h, tus)
    for a, b in zip(same, uni):
        assert np.array_equal(a, b)
    # numpy: the average of the two numpy predictions, then the pinned residual stages
    mv2 = np.random.default_rng(3).integers(-50, 51, (len(tus), 2)).astype(np.int32)
    got = O.inter_recon_batch(16, 30, cur_plane(cur), q1, w, h, tus, q2, mv2)
    pad = O.PAD
    pl1 = [npref.interp_plane_padded(r1, k, pad) for k in range(16)]
    pl2 = [npref.interp_plane_padded(r2, k, pad) for k in range(16)]

    def pred(pl, t, mx, my):
        tx, ty, px, py, pw, ph = (int(v) for v in t[:6])
        qx = O.lib().xo_get_mc_pos(w, pw, px, int(mx))
        qy = O.lib().xo_get_mc_pos(h, ph, py, int(my))
        x0, y0 = (qx >> 2) + (tx - px), (qy >> 2) + (ty - py)
        return pl[(qy & 3) * 4 + (qx & 3)][pad + y0:pad + y0 + 16, pad + x0:pad + x0 + 16].astype(np.int32)

    res = np.zeros((len(tus), 16, 16), np.int16)
    preds = []
    for i, t in enumerate(tus):
        p = (pred(pl1, t, t[6], t[7]) + pred(pl2, t, mv2[i, 0], mv2[i, 1]) + 1) >> 1
        preds.append(p)
        res[i] = cur[t[1]:t[1] + 16, t[0]:t[0] + 16].astype(np.int32) - p
    lv, rr, nz = O.residual_batch(16, 30, 0, res)
    assert np.array_equal(got[0], lv) and np.array_equal(got[2], nz)
    for i, t in enumerate(tus):
        v = np.clip(preds[i] + rr[i].astype(np.int32), 0, 255)
        assert np.array_equal(got[1][t[1]:t[1] + 16, t[0]:t[0] + 16], v)
        assert got[3][i] == int(((cur[t[1]:t[1] + 16, t[0]:t[0] + 16].astype(np.int64) - v) ** 2).sum())


@pytest.mark.parametrize("qp", [37, 51])
def test_reconstruction_clips_to_8_bits(qp):
    """A picture of 0 / 255 noise predicted from a flat 128 reference: the
    quantised residual overshoots, and add_ps clips pred + residual to 0..255."""
    w, h = 64, 64
    rng = np.random.default_rng(qp)
    cur = (rng.integers(0, 2, (h, w)) * 255).astype(np.uint8)
    ref = np.full((h, w), 128, np.uint8)
    tus = tiles(w, h, 8, np.random.default_rng(1), mv_span=8)
    got = O.inter_recon_batch(8, qp, cur_plane(cur), planes_of(ref), w, h, tus)
    exp = numpy_recon(8, qp, cur, ref, tus)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    # the unclipped sums leave the 8-bit range somewhere
    res = np.array([cur[t[1]:t[1] + 8, t[0]:t[0] + 8].astype(np.int16) - 128 for t in tus])
    _, rr, _ = O.residual_batch(8, qp, 0, res)
    assert (128 + rr.astype(np.int32) > 255).any() or (128 + rr.astype(np.int32) < 0).any()


@pytest.mark.parametrize("d", [3, 20])
def test_dct_half_drops_a_checkerboard_residual(d):
    """dct_half (rdo.c:764-771, transform.c:1276-1286) keeps only the low 16x16 of a 32x32
    TU's coefficients.  A residual that is a +-d checkerboard has all its energy at the
    highest frequency, so with dct_half every level is 0, nnz is 0, the reconstruction is
    the prediction and the SSD is 1024 d^2; the full DCT keeps it (nonzero levels)."""
    w, h = 96, 64
    ref = texture(w, h, 5)
    cur = ref.copy().astype(np.int32)
    chk = np.where((np.add.outer(np.arange(32), np.arange(32)) & 1) == 0, d, -d)
    cur[16:48, 32:64] = np.clip(cur[16:48, 32:64] + chk, 0, 255)
    cur = cur.astype(np.uint8)
    planes = planes_of(ref)
    tus = np.array([(32, 16, 32, 16, 32, 32, 0, 0)], np.int32)
    level, rec, nnz, ssd = O.inter_recon_batch(32 | O.TU_DCT_HALF, 22, cur_plane(cur), planes, w, h, tus)
    assert not level.any() and nnz[0] == 0
    assert np.array_equal(rec[16:48, 32:64], ref[16:48, 32:64])
    exp = int(((cur[16:48, 32:64].astype(np.int64) - ref[16:48, 32:64]) ** 2).sum())
    assert ssd[0] == exp and (exp == 1024 * d * d or d == 20)      # +-20 may clip at 0 / 255
    full = O.inter_recon_batch(32, 22, cur_plane(cur), planes, w, h, tus)
    assert full[2][0] > 0


def test_dct_half_keeps_a_smooth_residual_like_the_full_transform():
    """A residual whose coefficients all lie in the low 16x16 (a constant offset: DC only)
    gives the same levels, reconstruction and SSD with dct_half as with the full DCT."""
    w, h = 96, 64
    ref = texture(w, h, 6)
    cur = np.clip(ref.astype(np.int32) + 9, 0, 255).astype(np.uint8)
    planes = planes_of(ref)
    tus = np.array([(0, 0, 0, 0, 32, 32, 0, 0), (32, 32, 32, 32, 32, 32, 0, 0)], np.int32)
    a = O.inter_recon_batch(32 | O.TU_DCT_HALF, 27, cur_plane(cur), planes, w, h, tus)
    b = O.inter_recon_batch(32, 27, cur_plane(cur), planes, w, h, tus)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def pred_blocks(w, h, n, rng, keep=0.7):
    """Non-overlapping n x n blocks of a w x h plane, a random subset, in random order."""
    out = [(x, y) for y in range(0, h - n + 1, n) for x in range(0, w - n + 1, n) if rng.random() < keep]
    rng.shuffle(out)
    return np.array(out, np.int32).reshape(-1, 2)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_recon_pred_known_answer(n):
    """A prediction equal to the picture (cu_recon_chroma, rdo.c:807-838, on an exact
    prediction): zero levels, nnz 0, the reconstruction is the picture, SSD 0."""
    w, h = 96, 64
    cur = texture(w, h, 11)
    blocks = pred_blocks(w, h, n, np.random.default_rng(n))
    level, rec, nnz, ssd = O.recon_pred_batch(n, 30, cur, cur, blocks)
    assert not level.any() and not nnz.any() and not ssd.any()
    for x, y in blocks:
        assert np.array_equal(rec[y:y + n, x:x + n], cur[y:y + n, x:x + n])


@pytest.mark.parametrize("n", [4, 8, 16, 32])
@pytest.mark.parametrize("qp", [22, 39, 51])
def test_recon_pred_matches_numpy_composition(n, qp):
    """cur - pred, the residual stages (pinned on their own), numpy add / clip and SSD."""
    w, h = 96, 64
    cur, pred = texture(w, h, 12), texture(w, h, 13)
    blocks = pred_blocks(w, h, n, np.random.default_rng(qp + n))
    level, rec, nnz, ssd = O.recon_pred_batch(n, qp, cur, pred, blocks)
    res = np.array([cur[y:y + n, x:x + n].astype(np.int16) - pred[y:y + n, x:x + n] for x, y in blocks])
    lv, rr, nz = O.residual_batch(n, qp, 0, res)
    assert np.array_equal(level, lv) and np.array_equal(nnz, nz)
    exp = np.zeros_like(cur)
    for i, (x, y) in enumerate(blocks):
        v = np.clip(pred[y:y + n, x:x + n].astype(np.int32) + rr[i], 0, 255)
        exp[y:y + n, x:x + n] = v
        assert ssd[i] == int(((cur[y:y + n, x:x + n].astype(np.int64) - v) ** 2).sum())
    assert np.array_equal(rec, exp) and nnz.any()


@pytest.mark.parametrize("n_arg", [8, 32, 32 | O.TU_DCT_HALF])
def test_recon_pred_equals_the_luma_path_at_mv_zero(n_arg):
    """At MV (0, 0) the luma TU's prediction is the reference picture at the TU, so the
    luma path (pinned above, dct_half included) and the prediction-picture path agree."""
    n = n_arg & 0xff
    w, h = 128, 96
    cur, ref = texture(w, h, 14), texture(w, h, 15)
    blocks = pred_blocks(w, h, n, np.random.default_rng(n_arg), keep=1.0)
    tus = np.array([(x, y, x, y, n, n, 0, 0) for x, y in blocks], np.int32)
    a = O.inter_recon_batch(n_arg, 33, cur_plane(cur), planes_of(ref), w, h, tus)
    b = O.recon_pred_batch(n_arg, 33, cur, ref, blocks)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_chroma_qp():
    """cu_get_chroma_qp (cudata.h:118-125): the identity to 41, then the runs of
    tab_qp_scale_chroma (cudata.c:98-106), read off the table at these points; the luma QP
    is clipped to 0..63 first."""
    hand = {0: 0, 17: 17, 41: 41, 42: 42, 43: 42, 44: 43, 47: 44, 48: 45, 51: 46, 53: 47, 54: 48, 56: 48,
            57: 49, 59: 49, 60: 50, 62: 50, 63: 51, -3: 0, 64: 51, 79: 51}
    for q, c in hand.items():
        assert O.chroma_qp(q) == c, q
    seq = [O.chroma_qp(q) for q in range(64)]
    assert all(0 <= b - a <= 1 for a, b in zip(seq, seq[1:]))


def np_recon64(qp, half, cur_blk, pred_blk):
    """A 64x64 TU from pinned pieces: the lifting wavelet (xo_wavelet_fwd / _inv, pinned on
    hand-worked ramps), numpy's 32x32 transforms with the wavelet flag, and numpy quant /
    dequant at level 6, whose shifts are level 5's (pinned) moved by one, as
    tu_quant_forward / tu_quant_inverse (rdo.c:658-718) write them."""
    from tests.test_oracle_percall import np_dct, np_idct
    res = cur_blk.astype(np.int16) - pred_blk.astype(np.int16)
    low = np.zeros(32 * 32, np.int16)
    O.lib().xo_wavelet_fwd(64, 64, O.ptr(np.ascontiguousarray(res)), O.ptr(low))
    coef = np_dct(32, 32, low.reshape(32, 32).astype(np.int64), flag=1)
    if half:
        coef[16:, :] = 0
        coef[:, 16:] = 0
    qs, qsh5, _ = O.quant_params(5, qp, 0)
    ds, dsh5 = O.dequant_params(5, qp)
    qsh = qsh5 - 1
    lv = npref.quant(coef, qs, qsh, ((1 << qsh) * 5) // 31)
    dq = npref.dequant(lv, ds, dsh5 + 1)
    rr_low = np_idct(32, 32, dq.astype(np.int64), flag=1).astype(np.int16)
    rr = np.zeros(64 * 64, np.int16)
    O.lib().xo_wavelet_inv(64, 64, O.ptr(np.ascontiguousarray(rr_low.ravel())), O.ptr(rr))
    v = np.clip(pred_blk.astype(np.int32) + rr.reshape(64, 64), 0, 255).astype(np.uint8)
    return lv, v, int(np.count_nonzero(lv)), int(((cur_blk.astype(np.int64) - v) ** 2).sum())


@pytest.mark.parametrize("half", [0, 1])
@pytest.mark.parametrize("qp", [22, 37, 51])
def test_recon_64_matches_the_pinned_pieces(qp, half):
    """64x64 TUs (dct_64x64 / dct_64x64_half and idct_64x64, transform.c:1514-1541, at quant
    level 6): 32x32 levels, the reconstruction and the SSD."""
    w, h = 192, 128
    cur, pred = texture(w, h, 21), texture(w, h, 22)
    blocks = np.array([(0, 0), (128, 64), (64, 0), (0, 64)], np.int32)
    n_arg = 64 | (O.TU_DCT_HALF if half else 0)
    level, rec, nnz, ssd = O.recon_pred_batch(n_arg, qp, cur, pred, blocks)
    assert level.shape == (4, 32, 32)
    for i, (x, y) in enumerate(blocks):
        lv, v, nz, e = np_recon64(qp, half, cur[y:y + 64, x:x + 64], pred[y:y + 64, x:x + 64])
        assert np.array_equal(level[i], lv) and nnz[i] == nz and ssd[i] == e
        assert np.array_equal(rec[y:y + 64, x:x + 64], v)
    assert nnz.any()


def test_recon_64_known_answer_and_the_luma_path():
    """An exact prediction gives zero levels and the picture back; at MV (0, 0) a 64x64 PU's
    TU through the luma path equals the prediction-picture path (dct_half too)."""
    w, h = 128, 128
    cur, ref = texture(w, h, 23), texture(w, h, 24)
    blocks = np.array([(0, 0), (64, 64)], np.int32)
    level, rec, nnz, ssd = O.recon_pred_batch(64, 30, cur, cur, blocks)
    assert not level.any() and not nnz.any() and not ssd.any()
    assert np.array_equal(rec[:64, :64], cur[:64, :64]) and np.array_equal(rec[64:, 64:], cur[64:, 64:])
    tus = np.array([(x, y, x, y, 64, 64, 0, 0) for x, y in blocks], np.int32)
    for n_arg in (64, 64 | O.TU_DCT_HALF):
        a = O.inter_recon_batch(n_arg, 33, cur_plane(cur), planes_of(ref), w, h, tus)
        b = O.recon_pred_batch(n_arg, 33, cur, ref, blocks)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
