"""GPU parity of the luma inter reconstruction (xme_inter_recon_batch: MC at
the PU's clamped position -> sub_ps -> DCT / quant / dequant / IDCT ->
add_ps -> SSD, encoder/rdo.c:1333-1366, :1751-1753 and :1616-1633) against the oracle, for
every TU size at several QPs, with MVs that get_mv_for_mc clamps, from the
interpolated planes and from the full-pel plane alone (block filters), and a
1080p frame of 8x8 TUs."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_me import device_cur, oracle_cur, oracle_ref
from tests.test_oracle_recon import tiles
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu


def gpu_recon(n_arg, qp, cur, ref_img, tus, frac, intermediates=True, want_intermediates=False):
    """xme_inter_recon_batch; n_arg may carry XME_TU_DCT_HALF.  With frac 2 every plane is
    resident (the one-launch path), with frac 0 the block filters predict (staged path)."""
    n = n_arg & 0xff
    nl = min(n, 32)
    h, w = cur.shape
    dref = X.DeviceRef(w, h, frac=frac)
    dref.load(ref_img)
    dcur = device_cur(cur)
    ntu = len(tus)
    d_tus = X.upload(np.ascontiguousarray(tus, np.int32))
    resid, level, rres = (X.DeviceBuffer(2 * ntu * n * n) for _ in range(3))
    rec = X.DevicePlane(w, h, pad=0, stride=dcur.stride)
    rec.buf.zero()
    nnz, ssd = X.DeviceBuffer(4 * ntu), X.DeviceBuffer(4 * ntu)
    rd = dref.desc()
    X.check(X.lib().xme_inter_recon_batch(n_arg, qp, C.c_void_p(dcur.origin), dcur.stride, C.byref(rd), w, h,
                                          d_tus.at(), ntu, resid.at() if intermediates else None, level.at(),
                                          rres.at() if intermediates else None, C.c_void_p(rec.origin),
                                          rec.stride, nnz.at(), ssd.at(), None), "xme_inter_recon_batch")
    X.check(X.lib().xme_stream_sync(None), "sync")
    full = rec.buf.download(rec.rows * rec.stride, np.uint8).reshape(rec.rows, rec.stride)
    out = (level.download(ntu * nl * nl, np.int16).reshape(ntu, nl, nl), full[:h, :w],
           nnz.download(ntu, np.int32), ssd.download(ntu, np.int32))
    if want_intermediates:
        out += (resid.download(ntu * n * n, np.int16), rres.download(ntu * n * n, np.int16))
    return out


@pytest.mark.parametrize("frac", [2, 0])
@pytest.mark.parametrize("qp", [22, 37])
@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_matches_oracle(n, qp, frac):
    w, h = 192, 128
    ref = synth.smooth_field(w, h, seed=11)
    cur = synth.multi_ref_current([ref], 8, 8, seed=12)
    tus = tiles(w, h, n, np.random.default_rng(100 * n + qp), far=True)
    got = gpu_recon(n, qp, cur, ref, tus, frac)
    exp = O.inter_recon_batch(n, qp, oracle_cur(cur), oracle_ref(ref, 2), w, h, tus)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name
    assert got[2].any() and got[3].any()


def test_arguments_and_empty_batch():
    ref = X.DeviceRef(64, 64, frac=2)
    rd = ref.desc()
    for n, qp in ((12, 30), (8, -1), (8, 80)):
        rc = X.lib().xme_inter_recon_batch(n, qp, None, 64, C.byref(rd), 64, 64, None, 1, None, None, None, None,
                                           64, None, None, None)
        assert rc == X.XME_ERR_ARG
    assert X.lib().xme_inter_recon_batch(8, 30, None, 64, C.byref(rd), 64, 64, None, 0, None, None, None, None,
                                         64, None, None, None) == X.XME_OK


def test_1080p_8x8():
    W, H = 1920, 1080
    ref = synth.smooth_field(W, H, seed=21)
    cur = synth.multi_ref_current([ref], 16, 16, seed=22)
    tus = tiles(W, H, 8, np.random.default_rng(23), mv_span=120, far=True)
    got = gpu_recon(8, 32, cur, ref, tus, 2)
    exp = O.inter_recon_batch(8, 32, oracle_cur(cur), oracle_ref(ref, 2), W, H, tus)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name


def gpu_recon_bi(n_arg, qp, cur, r1_img, r2_img, tus, mv2, frac):
    n = n_arg & 0xff
    nl = min(n, 32)
    h, w = cur.shape
    d1, d2 = X.DeviceRef(w, h, frac=frac), X.DeviceRef(w, h, frac=frac)
    d1.load(r1_img)
    d2.load(r2_img)
    dcur = device_cur(cur)
    ntu = len(tus)
    d_tus, d_mv2 = X.upload(np.ascontiguousarray(tus, np.int32)), X.upload(np.ascontiguousarray(mv2, np.int32))
    resid, level, rres = (X.DeviceBuffer(2 * ntu * n * n) for _ in range(3))
    rec = X.DevicePlane(w, h, pad=0, stride=dcur.stride)
    rec.buf.zero()
    nnz, ssd = X.DeviceBuffer(4 * ntu), X.DeviceBuffer(4 * ntu)
    a, b = d1.desc(), d2.desc()
    X.check(X.lib().xme_inter_recon_bi_batch(n_arg, qp, C.c_void_p(dcur.origin), dcur.stride, C.byref(a), C.byref(b),
                                             w, h, d_tus.at(), d_mv2.at(), ntu, resid.at(), level.at(), rres.at(),
                                             C.c_void_p(rec.origin), rec.stride, nnz.at(), ssd.at(), None),
            "xme_inter_recon_bi_batch")
    X.check(X.lib().xme_stream_sync(None), "sync")
    full = rec.buf.download(rec.rows * rec.stride, np.uint8).reshape(rec.rows, rec.stride)
    return (level.download(ntu * nl * nl, np.int16).reshape(ntu, nl, nl), full[:h, :w],
            nnz.download(ntu, np.int32), ssd.download(ntu, np.int32))


@pytest.mark.parametrize("frac", [2, 0])
@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_bi_prediction_matches_oracle(n, frac):
    w, h = 192, 128
    r1 = synth.smooth_field(w, h, seed=31)
    r2 = synth.smooth_field(w, h, seed=32)
    cur = synth.multi_ref_current([r1, r2], 8, 8, seed=33)
    rng = np.random.default_rng(200 + n)
    tus = tiles(w, h, n, rng, far=True)
    # one second MV per PU, shared by its TUs; some far outside the picture
    mv2 = np.zeros((len(tus), 2), np.int32)
    per_pu = {}
    for i, t in enumerate(tus):
        key = (int(t[2]), int(t[3]))
        if key not in per_pu:
            per_pu[key] = ((int(rng.integers(-60, 61)), int(rng.integers(-60, 61))) if rng.random() < 0.8 else
                           (4 * int(rng.integers(-300, 300)) + 1, 4 * int(rng.integers(-300, 300)) + 2))
        mv2[i] = per_pu[key]
    got = gpu_recon_bi(n, 32, cur, r1, r2, tus, mv2, frac)
    exp = O.inter_recon_batch(n, 32, oracle_cur(cur), oracle_ref(r1, 2), w, h, tus, oracle_ref(r2, 2), mv2)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name
    assert got[3].any()


def test_tus_outside_their_limits_are_skipped():
    w, h = 64, 64
    ref = synth.smooth_field(w, h, seed=41)
    cur = synth.multi_ref_current([ref], 4, 4, seed=42)
    good = (16, 16, 16, 16, 16, 16, 4, -4)
    bad = [(18, 16, 16, 16, 16, 16, 0, 0),      # column not a multiple of 4
           (56, 16, 16, 16, 8, 8, 0, 0),        # TU not inside its PU
           (60, 56, 56, 56, 16, 16, 0, 0),      # beyond the picture's right edge
           (0, 0, 0, 0, 128, 8, 0, 0)]          # PU wider than 64
    tus = np.array([good] + bad, np.int32)
    level, rec, nnz, ssd = gpu_recon(8, 30, cur, ref, tus, 2)
    assert ssd[0] >= 0 and list(ssd[1:]) == [-1, -1, -1, -1]
    assert not level[1:].any() and not nnz[1:].any()
    exp = O.inter_recon_batch(8, 30, oracle_cur(cur), oracle_ref(ref, 2), w, h, tus[:1])
    assert np.array_equal(level[0], exp[0][0]) and ssd[0] == exp[3][0]
    # only the good TU's pixels are written
    assert np.array_equal(rec, exp[1])


def test_reconstruction_clips_to_8_bits():
    w, h = 64, 64
    rng = np.random.default_rng(51)
    cur = (rng.integers(0, 2, (h, w)) * 255).astype(np.uint8)
    ref = np.full((h, w), 128, np.uint8)
    tus = tiles(w, h, 8, np.random.default_rng(1), mv_span=8)
    got = gpu_recon(8, 51, cur, ref, tus, 2)
    exp = O.inter_recon_batch(8, 51, oracle_cur(cur), oracle_ref(ref, 2), w, h, tus)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name
    assert (got[1] == 0).any() and (got[1] == 255).any()


@pytest.mark.parametrize("frac", [2, 0])
def test_dct_half_matches_oracle(frac):
    """32x32 TUs with XME_TU_DCT_HALF (dct_half, rdo.c:764-771), uni- and bi-predicted."""
    w, h = 192, 128
    ref = synth.smooth_field(w, h, seed=61)
    r2 = synth.smooth_field(w, h, seed=62)
    cur = synth.multi_ref_current([ref, r2], 8, 8, seed=63)
    tus = tiles(w, h, 32, np.random.default_rng(64), far=True)
    n_arg = 32 | X.XME_TU_DCT_HALF
    got = gpu_recon(n_arg, 27, cur, ref, tus, frac)
    exp = O.inter_recon_batch(32 | O.TU_DCT_HALF, 27, oracle_cur(cur), oracle_ref(ref, 2), w, h, tus)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name
    assert not got[0][:, 16:, :].any() and not got[0][:, :, 16:].any()
    full = gpu_recon(32, 27, cur, ref, tus, frac)
    assert not np.array_equal(full[0], got[0])
    mv2 = np.array([(int(t[6]) // 2 + 3, -int(t[7]) // 2 + 1) for t in tus], np.int32)
    got = gpu_recon_bi(n_arg, 27, cur, ref, r2, tus, mv2, frac)
    exp = O.inter_recon_batch(32 | O.TU_DCT_HALF, 27, oracle_cur(cur), oracle_ref(ref, 2), w, h, tus,
                              oracle_ref(r2, 2), mv2)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_one_launch_path_without_intermediates(n):
    """With every plane resident the path needs no prediction-error or residual buffers;
    when given, they hold what the staged path (block filters) writes there."""
    w, h = 192, 128
    ref = synth.smooth_field(w, h, seed=71)
    cur = synth.multi_ref_current([ref], 8, 8, seed=72)
    tus = tiles(w, h, n, np.random.default_rng(73 + n), far=True)
    exp = O.inter_recon_batch(n, 32, oracle_cur(cur), oracle_ref(ref, 2), w, h, tus)
    got = gpu_recon(n, 32, cur, ref, tus, 2, intermediates=False)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name
    fused = gpu_recon(n, 32, cur, ref, tus, 2, want_intermediates=True)
    staged = gpu_recon(n, 32, cur, ref, tus, 0, want_intermediates=True)
    assert np.array_equal(fused[4], staged[4]) and np.array_equal(fused[5], staged[5])
    # the staged path needs them
    dref = X.DeviceRef(w, h, frac=0)
    rd = dref.desc()
    dcur = device_cur(cur)
    d_tus = X.upload(np.ascontiguousarray(tus, np.int32))
    bufs = [X.DeviceBuffer(2 * len(tus) * n * n), X.DeviceBuffer(4 * len(tus)), X.DeviceBuffer(4 * len(tus))]
    rc = X.lib().xme_inter_recon_batch(n, 32, C.c_void_p(dcur.origin), dcur.stride, C.byref(rd), w, h, d_tus.at(),
                                       len(tus), None, bufs[0].at(), None, C.c_void_p(dcur.origin), dcur.stride,
                                       bufs[1].at(), bufs[2].at(), None)
    assert rc == X.XME_ERR_ARG


def gpu_recon_pred(n_arg, qp, cur, pred, blocks, pred_stride=None, d_pred=None):
    """xme_recon_pred_batch on one plane; pred is uploaded at pred_stride (any), unless a
    device plane d_pred already holds it."""
    n = n_arg & 0xff
    nl = min(n, 32)
    h, w = cur.shape
    dcur = device_cur(cur)
    if d_pred is None:
        d_pred = X.DevicePlane(w, h, pad=0, stride=pred_stride or w)
        d_pred.upload_picture(pred)
    nblk = len(blocks)
    d_blocks = X.upload(np.ascontiguousarray(blocks, np.int32).reshape(-1, 2))
    level = X.DeviceBuffer(2 * max(nblk, 1) * n * n)
    rec = X.DevicePlane(w, h, pad=0, stride=dcur.stride)
    rec.buf.zero()
    nnz, ssd = X.DeviceBuffer(4 * max(nblk, 1)), X.DeviceBuffer(4 * max(nblk, 1))
    X.check(X.lib().xme_recon_pred_batch(n_arg, qp, C.c_void_p(dcur.origin), dcur.stride, C.c_void_p(d_pred.origin),
                                         d_pred.stride, w, h, d_blocks.at(), nblk, level.at(), C.c_void_p(rec.origin),
                                         rec.stride, nnz.at(), ssd.at(), None), "xme_recon_pred_batch")
    X.check(X.lib().xme_stream_sync(None), "sync")
    full = rec.buf.download(rec.rows * rec.stride, np.uint8).reshape(rec.rows, rec.stride)
    return (level.download(nblk * nl * nl, np.int16).reshape(nblk, nl, nl), full[:h, :w],
            nnz.download(nblk, np.int32), ssd.download(nblk, np.int32))


def gpu_recon_chroma(n_arg, qp, curs, d_preds, blocks):
    """xme_recon_chroma_batch: U and V of the same blocks in one launch; returns the
    (level, rec, nnz, ssd) of each plane."""
    n = n_arg & 0xff
    h, w = curs[0].shape
    dcur = [device_cur(c) for c in curs]
    nblk = len(blocks)
    d_blocks = X.upload(np.ascontiguousarray(blocks, np.int32).reshape(-1, 2))
    level = X.DeviceBuffer(2 * 2 * max(nblk, 1) * n * n)
    rec = [X.DevicePlane(w, h, pad=0, stride=dcur[0].stride) for _ in range(2)]
    for r in rec:
        r.buf.zero()
    nnz, ssd = X.DeviceBuffer(8 * max(nblk, 1)), X.DeviceBuffer(8 * max(nblk, 1))
    X.check(X.lib().xme_recon_chroma_batch(n_arg, qp, C.c_void_p(dcur[0].origin), C.c_void_p(dcur[1].origin),
                                           dcur[0].stride, C.c_void_p(d_preds[0].origin),
                                           C.c_void_p(d_preds[1].origin), d_preds[0].stride, w, h, d_blocks.at(),
                                           nblk, level.at(), C.c_void_p(rec[0].origin), C.c_void_p(rec[1].origin),
                                           rec[0].stride, nnz.at(), ssd.at(), None), "xme_recon_chroma_batch")
    X.check(X.lib().xme_stream_sync(None), "sync")
    lv = level.download(2 * nblk * n * n, np.int16).reshape(2, nblk, n, n)
    nz, sd = nnz.download(2 * nblk, np.int32).reshape(2, nblk), ssd.download(2 * nblk, np.int32).reshape(2, nblk)
    out = []
    for p in range(2):
        full = rec[p].buf.download(rec[p].rows * rec[p].stride, np.uint8).reshape(rec[p].rows, rec[p].stride)
        out.append((lv[p], full[:h, :w], nz[p], sd[p]))
    return out


def pred_grid(w, h, n, rng, keep=0.8):
    out = [(x, y) for y in range(0, h - n + 1, n) for x in range(0, w - n + 1, n) if rng.random() < keep]
    rng.shuffle(out)
    return np.array(out, np.int32).reshape(-1, 2)


@pytest.mark.parametrize("qp", [22, 43])
@pytest.mark.parametrize("n_arg", [4, 8, 16, 32, 32 | X.XME_TU_DCT_HALF, 64, 64 | X.XME_TU_DCT_HALF])
def test_recon_pred_matches_oracle(n_arg, qp):
    """xme_recon_pred_batch (cu_recon_chroma, rdo.c:807-838; 64x64 luma TUs through the
    wavelet, transform.c:1514-1541) on a prediction picture with its own stride, blocks in
    random order."""
    n = n_arg & 0xff
    w, h = 192, 128
    pred = synth.smooth_field(w, h, seed=81)
    cur = synth.multi_ref_current([pred], 8, 8, seed=82)
    blocks = pred_grid(w, h, n, np.random.default_rng(83 + n_arg), keep=0.8 if n < 64 else 1.0)
    got = gpu_recon_pred(n_arg, qp, cur, pred, blocks, pred_stride=w + 12)
    exp = O.recon_pred_batch(n_arg, qp, cur, pred, blocks)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name
    assert got[2].any()


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_chroma_of_inter_cus(n):
    """The chroma of 2n x 2n inter CUs end to end: xme_mc_chroma_batch writes each CU's
    n x n U and V predictions into prediction pictures at the CU's chroma position, then
    xme_recon_pred_batch reconstructs each plane at the chroma QP of luma QP 45
    (xme_chroma_qp), 32x32 blocks with dct_half as a 64x64 CU's first RD pass does."""
    W, H = 256, 128
    cw, ch = W // 2, H // 2
    rng = np.random.default_rng(90 + n)
    src_u, src_v = synth.smooth_field(cw, ch, seed=91), synth.smooth_field(cw, ch, seed=92)
    pu_o, pv_o = O.Plane(cw, ch, pad=40), O.Plane(cw, ch, pad=40)
    for p, img in ((pu_o, src_u), (pv_o, src_v)):
        p.set_picture(img)
        p.expand_border()
    du, dv = X.DevicePlane(cw, ch, pad=40), X.DevicePlane(cw, ch, pad=40)
    du.upload_padded(pu_o.view()[: ch + 80, : cw + 80])
    dv.upload_padded(pv_o.view()[: ch + 80, : cw + 80])
    cus = []
    for y in range(0, H, 2 * n):
        for x in range(0, W, 2 * n):
            mv = (int(rng.integers(-40, 41)), int(rng.integers(-40, 41)))
            if rng.random() < 0.2:
                mv = (4 * int(rng.integers(-200, 200)) + 3, 4 * int(rng.integers(-200, 200)) + 1)
            cus.append((x, y, 2 * n, 2 * n) + mv)
    cus = np.array(cus, np.int32)
    ncu = len(cus)
    pred_u, pred_v = X.DevicePlane(cw, ch, pad=0, stride=cw + 8), X.DevicePlane(cw, ch, pad=0, stride=cw + 8)
    off = np.array([(int(c[1]) // 2) * pred_u.stride + int(c[0]) // 2 for c in cus], np.int64)
    d_cus, d_off = X.upload(cus.ravel()), X.upload(off)
    X.check(X.lib().xme_mc_chroma_batch(C.c_void_p(du.origin), C.c_void_p(dv.origin), du.stride, d_cus.at(), ncu, W,
                                        H, C.c_void_p(pred_u.origin), C.c_void_p(pred_v.origin), pred_u.stride,
                                        d_off.at(), None), "xme_mc_chroma_batch")
    # the oracle's prediction pictures
    eu, ev = np.zeros((ch, cw), np.uint8), np.zeros((ch, cw), np.uint8)
    for x, y, w, h, mx, my in cus:
        qx = O.lib().xo_get_mc_pos(W, int(w), int(x), int(mx))
        qy = O.lib().xo_get_mc_pos(H, int(h), int(y), int(my))
        bu, bv = O.mc_chroma(qx, qy, int(w) // 2, int(h) // 2, pu_o, pv_o)
        eu[y // 2:y // 2 + n, x // 2:x // 2 + n], ev[y // 2:y // 2 + n, x // 2:x // 2 + n] = bu, bv
    qp_c = X.lib().xme_chroma_qp(45)
    assert qp_c == O.chroma_qp(45) == 43
    n_arg = 32 | X.XME_TU_DCT_HALF if n == 32 else n
    blocks = cus[:, :2] // 2
    curs, exps = [], []
    for seed, exp_pred, d_pred in ((93, eu, pred_u), (94, ev, pred_v)):
        cur = synth.multi_ref_current([exp_pred], 4, 4, seed=seed)
        got = gpu_recon_pred(n_arg, qp_c, cur, None, blocks, d_pred=d_pred)
        exp = O.recon_pred_batch(n_arg, qp_c, cur, exp_pred, blocks)
        for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
            assert np.array_equal(g, e), name
        curs.append(cur)
        exps.append(exp)
    # both planes in one launch (xme_recon_chroma_batch)
    both = gpu_recon_chroma(n_arg, qp_c, curs, (pred_u, pred_v), blocks)
    for p in range(2):
        for name, g, e in zip(("level", "rec", "nnz", "ssd"), both[p], exps[p]):
            assert np.array_equal(g, e), (p, name)


def test_recon_pred_skips_blocks_outside_the_plane():
    w, h = 64, 48
    pred = synth.smooth_field(w, h, seed=95)
    cur = synth.multi_ref_current([pred], 4, 4, seed=96)
    blocks = np.array([(16, 8), (18, 8), (60, 0), (0, 44), (-8, 0), (0, -8)], np.int32)
    level, rec, nnz, ssd = gpu_recon_pred(8, 30, cur, pred, blocks)
    assert ssd[0] >= 0 and list(ssd[1:]) == [-1] * 5
    assert not level[1:].any() and not nnz[1:].any()
    exp = O.recon_pred_batch(8, 30, cur, pred, blocks[:1])
    assert np.array_equal(rec, exp[1]) and np.array_equal(level[0], exp[0][0]) and ssd[0] == exp[3][0]


def test_recon_pred_arguments_and_empty_batch():
    buf = X.DeviceBuffer(4096)
    p = buf.at()
    for n, qp in ((12, 30), (8, -1), (8, 80), (16 | X.XME_TU_DCT_HALF, 30)):
        rc = X.lib().xme_recon_pred_batch(n, qp, p, 64, p, 64, 64, 16, p, 1, p, p, 64, p, p, None)
        assert rc == X.XME_ERR_ARG
    # a NULL buffer, a misaligned current picture or stride
    assert X.lib().xme_recon_pred_batch(8, 30, None, 64, p, 64, 64, 16, p, 1, p, p, 64, p, p, None) == X.XME_ERR_ARG
    assert X.lib().xme_recon_pred_batch(8, 30, buf.at(2), 64, p, 64, 64, 16, p, 1, p, p, 64, p, p,
                                        None) == X.XME_ERR_ARG
    assert X.lib().xme_recon_pred_batch(8, 30, p, 66, p, 64, 64, 16, p, 1, p, p, 64, p, p, None) == X.XME_ERR_ARG
    # the prediction picture is read as aligned dwords too
    assert X.lib().xme_recon_pred_batch(8, 30, p, 64, buf.at(1), 64, 64, 16, p, 1, p, p, 64, p, p,
                                        None) == X.XME_ERR_ARG
    assert X.lib().xme_recon_pred_batch(8, 30, p, 64, p, 65, 64, 16, p, 1, p, p, 64, p, p, None) == X.XME_ERR_ARG
    assert X.lib().xme_recon_pred_batch(8, 30, None, 64, None, 64, 64, 16, None, 0, None, None, 64, None, None,
                                        None) == X.XME_OK
    # the two-plane call: a missing V plane, a misaligned V reconstruction, bad sizes
    L = X.lib()
    assert L.xme_recon_chroma_batch(8, 30, p, None, 64, p, p, 64, 64, 16, p, 1, p, p, p, 64, p, p,
                                    None) == X.XME_ERR_ARG
    assert L.xme_recon_chroma_batch(8, 30, p, p, 64, p, None, 64, 64, 16, p, 1, p, p, p, 64, p, p,
                                    None) == X.XME_ERR_ARG
    assert L.xme_recon_chroma_batch(8, 30, p, p, 64, p, p, 64, 64, 16, p, 1, p, p, buf.at(1), 64, p, p,
                                    None) == X.XME_ERR_ARG
    assert L.xme_recon_chroma_batch(12, 30, p, p, 64, p, p, 64, 64, 16, p, 1, p, p, p, 64, p, p,
                                    None) == X.XME_ERR_ARG
    assert L.xme_recon_chroma_batch(8, 30, None, None, 64, None, None, 64, 64, 16, None, 0, None, None, None, 64,
                                    None, None, None) == X.XME_OK


@pytest.mark.parametrize("n_arg", [4, 16, 32 | X.XME_TU_DCT_HALF])
def test_recon_chroma_matches_two_single_plane_calls(n_arg):
    """U and V in one launch equal two xme_recon_pred_batch calls, with a ragged block
    count (not a multiple of a tile's blocks) so that the planes meet inside a tile."""
    n = n_arg & 0xff
    w, h = 160, 96
    rng = np.random.default_rng(97 + n_arg)
    blocks = pred_grid(w, h, n, rng, keep=0.9)
    per_tile = (32 // n) ** 2
    if per_tile > 1 and len(blocks) % per_tile == 0:
        blocks = blocks[:-1]
    curs, preds, d_preds, exps = [], [], [], []
    for seed in (98, 99):
        pred = synth.smooth_field(w, h, seed=seed)
        cur = synth.multi_ref_current([pred], 8, 8, seed=seed + 10)
        d = X.DevicePlane(w, h, pad=0, stride=w + 4)
        d.upload_picture(pred)
        curs.append(cur)
        d_preds.append(d)
        exps.append(gpu_recon_pred(n_arg, 37, cur, None, blocks, d_pred=d))
    both = gpu_recon_chroma(n_arg, 37, curs, d_preds, blocks)
    for p in range(2):
        for name, g, e in zip(("level", "rec", "nnz", "ssd"), both[p], exps[p]):
            assert np.array_equal(g, e), (p, name)


def tus64(w, h, rng, far=True):
    """64x64 PUs tiling the picture, each one TU (TU_SPLIT_NON of a 64x64 CU), one MV per PU,
    some far outside the picture."""
    out = []
    for y in range(0, h - 63, 64):
        for x in range(0, w - 63, 64):
            mv = (int(rng.integers(-60, 61)), int(rng.integers(-60, 61)))
            if far and rng.random() < 0.3:
                mv = (4 * int(rng.integers(-300, 300)) + int(rng.integers(0, 4)),
                      4 * int(rng.integers(-300, 300)) + int(rng.integers(0, 4)))
            out.append((x, y, x, y, 64, 64) + mv)
    return np.array(out, np.int32)


@pytest.mark.parametrize("qp", [22, 37])
@pytest.mark.parametrize("half", [0, 1])
def test_64x64_tus_uni_and_bi(half, qp):
    """64x64 TUs of 64x64 PUs through xme_inter_recon_batch and _bi_batch (one launch with
    the planes resident; dct_64x64 and its dct_half at quant level 6)."""
    w, h = 256, 192
    r1 = synth.smooth_field(w, h, seed=111)
    r2 = synth.smooth_field(w, h, seed=112)
    cur = synth.multi_ref_current([r1, r2], 8, 8, seed=113)
    rng = np.random.default_rng(114 + qp + half)
    tus = tus64(w, h, rng)
    n_arg = 64 | (X.XME_TU_DCT_HALF if half else 0)
    got = gpu_recon(n_arg, qp, cur, r1, tus, 2, intermediates=False)
    exp = O.inter_recon_batch(n_arg, qp, oracle_cur(cur), oracle_ref(r1, 2), w, h, tus)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), name
    assert got[2].any() and (got[3] > 0).all()
    mv2 = np.array([(int(rng.integers(-40, 41)), int(rng.integers(-40, 41))) for _ in tus], np.int32)
    got = gpu_recon_bi(n_arg, qp, cur, r1, r2, tus, mv2, 2)
    exp = O.inter_recon_batch(n_arg, qp, oracle_cur(cur), oracle_ref(r1, 2), w, h, tus, oracle_ref(r2, 2), mv2)
    for name, g, e in zip(("level", "rec", "nnz", "ssd"), got, exp):
        assert np.array_equal(g, e), ("bi", name)


def test_64x64_tus_limits():
    """A 64x64 TU needs a 64x64 PU inside the picture and the interpolated planes."""
    w, h = 192, 128
    ref = synth.smooth_field(w, h, seed=121)
    cur = synth.multi_ref_current([ref], 8, 8, seed=122)
    tus = np.array([(64, 64, 64, 64, 64, 64, 5, -7),      # good
                    (64, 0, 64, 0, 64, 32, 0, 0),         # PU 64x32: the TU is not inside it
                    (160, 0, 160, 0, 64, 64, 0, 0),       # past the right edge
                    (66, 0, 64, 0, 64, 64, 0, 0)],        # column not a multiple of 4 / not inside
                   np.int32)
    level, rec, nnz, ssd = gpu_recon(64, 30, cur, ref, tus, 2, intermediates=False)
    assert ssd[0] > 0 and list(ssd[1:]) == [-1, -1, -1]
    assert not level[1:].any() and not nnz[1:].any()
    exp = O.inter_recon_batch(64, 30, oracle_cur(cur), oracle_ref(ref, 2), w, h, tus[:1])
    assert np.array_equal(rec, exp[1]) and np.array_equal(level[0], exp[0][0])
    rd_ref = X.DeviceRef(w, h, frac=0)
    rd = rd_ref.desc()
    buf = X.DeviceBuffer(1 << 16)
    p = buf.at()
    assert X.lib().xme_inter_recon_batch(64, 30, p, 192, C.byref(rd), w, h, p, 1, p, p, p, p, 192, p, p,
                                         None) == X.XME_ERR_ARG


@pytest.mark.parametrize("bi", [False, True])
def test_2160p_records_and_prediction_picture_agree(bi):
    """At 3840x2160 with the mixed PU tiling, every PU split into its largest square TUs
    (64x64 PUs one wavelet TU), two independent paths give the same levels, nnz, SSD and
    reconstruction: the record form (xme_inter_recon_batch / _bi_batch, the prediction
    formed inside the fused kernel) and the prediction picture (xme_mc_luma_batch /
    _bi_batch over every PU, then xme_recon_pred_batch)."""
    W, H = 3840, 2160
    rng = np.random.default_rng(131 + bi)
    r1 = synth.smooth_field(W, H, seed=132)
    r2 = synth.smooth_field(W, H, seed=133)
    cur = synth.multi_ref_current([r1, r2], 32, 16, seed=134)
    pus = synth.mixed_pus(W, H, seed=135)
    mv = rng.integers(-64, 65, (len(pus), 4)).astype(np.int32)
    far = rng.random(len(pus)) < 0.05
    mv[far, 0] = 4 * rng.integers(-3000, 3000, far.sum()) + 3
    d1, d2 = X.DeviceRef(W, H, frac=2), X.DeviceRef(W, H, frac=2)
    d1.load(r1)
    d2.load(r2)
    dcur = device_cur(cur)
    a, b = d1.desc(), d2.desc()
    pred = X.DevicePlane(W, H, pad=0, stride=dcur.stride)
    blocks = np.array([(x, y, pw, ph) + tuple(int(v) for v in mv[i]) for i, (x, y, pw, ph) in enumerate(pus)],
                      np.int32)
    off = np.array([int(bb[1]) * pred.stride + int(bb[0]) for bb in blocks], np.int64)
    d_off = X.upload(off)
    if bi:
        d_pb = X.upload(blocks.ravel())
        X.check(X.lib().xme_mc_luma_bi_batch(C.byref(a), C.byref(b), d_pb.at(), len(blocks), W, H,
                                             C.c_void_p(pred.origin), pred.stride, d_off.at(), None), "mc bi")
    else:
        d_pb = X.upload(np.ascontiguousarray(blocks[:, :6]).ravel())
        X.check(X.lib().xme_mc_luma_batch(C.byref(a), d_pb.at(), len(blocks), W, H, C.c_void_p(pred.origin),
                                          pred.stride, d_off.at(), None), "mc")
    by_n = {}
    for i, (x, y, pw, ph) in enumerate(pus):
        n = 64
        while pw % n or ph % n:
            n >>= 1
        for ty in range(y, y + ph, n):
            for tx in range(x, x + pw, n):
                by_n.setdefault(n, []).append((tx, ty, x, y, pw, ph, int(mv[i, 0]), int(mv[i, 1]), int(mv[i, 2]),
                                               int(mv[i, 3])))
    rec_a = X.DevicePlane(W, H, pad=0, stride=dcur.stride)
    rec_b = X.DevicePlane(W, H, pad=0, stride=dcur.stride)
    rec_a.buf.zero()
    rec_b.buf.zero()
    for n, rows in sorted(by_n.items()):
        t = np.array(rows, np.int32)
        ntu, nl = len(t), min(n, 32)
        d_t = X.upload(np.ascontiguousarray(t[:, :8]))
        d_blk = X.upload(np.ascontiguousarray(t[:, :2]))
        outs = []
        for _ in range(2):
            outs.append((X.DeviceBuffer(2 * ntu * nl * nl), X.DeviceBuffer(4 * ntu), X.DeviceBuffer(4 * ntu)))
        (la, na, sa), (lb, nb, sb) = outs
        if bi:
            d_mv2 = X.upload(np.ascontiguousarray(t[:, 8:10]))
            X.check(X.lib().xme_inter_recon_bi_batch(n, 32, C.c_void_p(dcur.origin), dcur.stride, C.byref(a),
                                                     C.byref(b), W, H, d_t.at(), d_mv2.at(), ntu, None, la.at(), None,
                                                     C.c_void_p(rec_a.origin), rec_a.stride, na.at(), sa.at(), None),
                    "records bi")
        else:
            X.check(X.lib().xme_inter_recon_batch(n, 32, C.c_void_p(dcur.origin), dcur.stride, C.byref(a), W, H,
                                                  d_t.at(), ntu, None, la.at(), None, C.c_void_p(rec_a.origin),
                                                  rec_a.stride, na.at(), sa.at(), None), "records")
        X.check(X.lib().xme_recon_pred_batch(n, 32, C.c_void_p(dcur.origin), dcur.stride, C.c_void_p(pred.origin),
                                             pred.stride, W, H, d_blk.at(), ntu, lb.at(), C.c_void_p(rec_b.origin),
                                             rec_b.stride, nb.at(), sb.at(), None), "prediction picture")
        X.check(X.lib().xme_stream_sync(None), "sync")
        for name, x, y, cnt, dt in (("level", la, lb, ntu * nl * nl, np.int16), ("nnz", na, nb, ntu, np.int32),
                                    ("ssd", sa, sb, ntu, np.int32)):
            assert np.array_equal(x.download(cnt, dt), y.download(cnt, dt)), (n, name)
        assert (sa.download(ntu, np.int32) >= 0).all(), n
    ga = rec_a.buf.download(rec_a.rows * rec_a.stride, np.uint8)
    gb = rec_b.buf.download(rec_b.rows * rec_b.stride, np.uint8)
    assert np.array_equal(ga, gb)
