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
    out = (level.download(ntu * n * n, np.int16).reshape(ntu, n, n), full[:h, :w],
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
    return (level.download(ntu * n * n, np.int16).reshape(ntu, n, n), full[:h, :w],
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
