"""GPU parity of the bi-prediction refinements against the oracle: the
symmetric one (xme_me_search_sym_batch, xavs2_me_search_sym,
encoder/me.c:1190-1262) and the bi-directional one (xme_me_search_bid_batch,
xavs2_me_search_bid, me.c:1278-1392).  For each:

* random jobs over every LumaPU shape, with SATD and SAD, use_fractional_me
  0 / 1 / 2 and PMVR (sym: B and P slices, distances 1 to 8; bid: backward
  MVs on planes that are not built, which report status 2);
* start MVs at the edges of the qpel window and a narrowed MV range, so that
  the start, its mirror (sym), the backward MV (bid) or a Spiral point leaves
  the window;
* a known answer (sym: the current picture is the average of two references
  moved by +d and -d; bid: both references reproduce the PU at the two MVs;
  the start scores its MV rates alone);
* jobs the entry point does not evaluate (status 2) and argument errors;
* a 1080p batch of mixed shapes against the oracle job by job.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_me import LAM, device_cur, oracle_cur, oracle_ref
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu


def oracle_planes(img, frac):
    pl = oracle_ref(img, frac)
    if frac == 1:
        # use_fractional_me 1 builds planes 2, 8 and 10 only (frame.c:343-350)
        pl = [p if k in (0, 2, 8, 10) else None for k, p in enumerate(pl)]
    return pl


def device_ref(img, frac):
    h, w = img.shape
    d = X.DeviceRef(w, h, frac=frac)
    d.load(img)
    return d


def jobs_array(rows):
    a = np.zeros(len(rows), dtype=X.SYM_JOB_DTYPE)
    for i, r in enumerate(rows):
        for k, v in r.items():
            a[k if k not in ("pix_x", "pix_y") else k[-1]][i] = v
    return a


def job(x, y, w, h, mv, mvp=(0, 0), mvp1=(0, 0), dist=(1, 1), b_slice=1):
    return dict(pix_x=x, pix_y=y, w=w, h=h, mv_x=mv[0], mv_y=mv[1], mvp_x=mvp[0], mvp_y=mvp[1],
                mvp1_x=mvp1[0], mvp1_y=mvp1[1], dist_fwd=dist[0], dist_bwd=dist[1], b_slice=b_slice)


def run_sym(w, h, cur, r1, r2, rows, R=16, **kw):
    """Runs the jobs on the GPU and through the oracle; returns both."""
    frac = kw.get("use_fractional_me", 2)
    prm_o = O.me_param(w, h, R, LAM, **kw)
    prm_d = X.me_params(w, h, R, LAM, **kw)
    got = X.me_search_sym_batch(prm_d, jobs_array(rows), device_cur(cur), device_ref(r1, frac),
                                device_ref(r2, frac))
    p1, p2, oc = oracle_planes(r1, frac), oracle_planes(r2, frac), oracle_cur(cur)
    exp = np.array([O.me_search_sym(prm_o, j, oc, p1, p2) for j in rows], dtype=np.int64)
    return got, exp


def assert_same(got, exp, rows):
    assert np.all(got["status"] == 0)
    g = np.stack([got["cost"], got["mv_x"], got["mv_y"], got["mvcost"]], axis=1).astype(np.int64)
    bad = np.nonzero(np.any(g != exp, axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} of {len(rows)} differ; first {rows[bad[0]]}: {g[bad[0]]} vs {exp[bad[0]]}"


def random_rows(w, h, n, rng, mv_span=48):
    rows = []
    for i in range(n):
        part = i % 25
        pw, ph = X.PU_W[part], X.PU_H[part]
        x = int(rng.integers(0, (w - pw) // 4 + 1)) * 4
        y = int(rng.integers(0, h - ph + 1))
        mv = (int(rng.integers(-mv_span, mv_span + 1)), int(rng.integers(-mv_span, mv_span + 1)))
        rows.append(job(x, y, pw, ph, mv, mvp=(int(rng.integers(-40, 41)), int(rng.integers(-40, 41))),
                        mvp1=(int(rng.integers(-12, 13)) + mv[0], int(rng.integers(-12, 13)) + mv[1]),
                        dist=(int(rng.integers(1, 9)), int(rng.integers(1, 9))), b_slice=int(rng.integers(0, 2))))
    return rows


def pictures(w, h, seed):
    r1 = synth.smooth_field(w, h, seed=seed)
    r2 = synth.smooth_field(w, h, seed=seed + 1)
    cur = synth.multi_ref_current([r1, r2], 8, 8, seed=seed + 2)
    return cur, r1, r2


@pytest.mark.parametrize("kw", [dict(), dict(use_hadamard=0), dict(use_fractional_me=1),
                                dict(use_fractional_me=0), dict(enable_pmvr=1),
                                dict(use_hadamard=0, enable_pmvr=1)],
                         ids=["satd", "sad", "frac1", "frac0", "pmvr", "sad-pmvr"])
def test_random_jobs_every_shape(kw):
    w, h = 192, 160
    cur, r1, r2 = pictures(w, h, 11)
    rows = random_rows(w, h, 200, np.random.default_rng(5))
    got, exp = run_sym(w, h, cur, r1, r2, rows, **kw)
    assert_same(got, exp, rows)


def test_window_edges_and_narrow_range():
    """Starts whose points straddle the qpel window: 64 pixels beyond each
    picture edge, and a +-40 qpel MV range."""
    w, h = 128, 96
    cur, r1, r2 = pictures(w, h, 21)
    rows = []
    for (pw, ph) in ((16, 16), (8, 8), (64, 64), (32, 8)):
        for (x, y) in ((0, 0), (w - pw, h - ph), (0, h - ph), (w - pw, 0)):
            lo_x, hi_x = (-64 - x) * 4, (w + 64 - x - pw) * 4
            lo_y, hi_y = (-64 - y) * 4, (h + 64 - y - ph) * 4
            for mv in ((lo_x, lo_y), (hi_x, hi_y), (lo_x + 1, hi_y - 1), (hi_x - 2, lo_y + 2), (lo_x - 1, 0)):
                for b in (0, 1):
                    rows.append(job(x, y, pw, ph, mv, mvp=mv, mvp1=mv, dist=(2, 1), b_slice=b))
    got, exp = run_sym(w, h, cur, r1, r2, rows)
    assert_same(got, exp, rows)
    assert np.any(got["cost"] == X.MAX_DISTORTION) and np.any(got["cost"] < X.MAX_DISTORTION)
    rows2 = random_rows(w, h, 100, np.random.default_rng(8), mv_span=44)
    got, exp = run_sym(w, h, cur, r1, r2, rows2, mv_range_x=(-40, 40), mv_range_y=(-40, 40))
    assert_same(got, exp, rows2)
    assert np.any(got["cost"] == X.MAX_DISTORTION)


@pytest.mark.parametrize("d", [(3, -2), (-4, 1)])
def test_mirrored_shift_known_answer(d):
    w, h = 96, 80
    cur = synth.smooth_field(w, h, seed=31)
    r1 = np.roll(cur, (d[1], d[0]), axis=(0, 1))       # r1[p + d] = cur[p]
    r2 = np.roll(cur, (-d[1], -d[0]), axis=(0, 1))     # r2[p - d] = cur[p]
    mv = (4 * d[0], 4 * d[1])
    rows = [job(32, 24, 16, 16, mv), job(32, 24, 16, 16, mv, b_slice=0)]
    got, exp = run_sym(w, h, cur, r1, r2, rows)
    assert_same(got, exp, rows)

    def bits(v):
        return 1 if v == 0 else 2 * int(abs(v)).bit_length() + 1
    b = (LAM * (bits(mv[0]) + bits(mv[1]))) >> 16
    assert (got["cost"][0], got["mv_x"][0], got["mv_y"][0]) == (b, mv[0], mv[1])
    assert got["cost"][1] > b


def test_jobs_not_evaluated_and_bad_arguments():
    w, h = 96, 80
    cur, r1, r2 = pictures(w, h, 41)
    rows = [job(2, 8, 8, 8, (0, 0)),                        # x % 4 != 0
            job(92, 8, 8, 8, (0, 0)),                       # beyond the right edge
            job(8, 8, 8, 8, (0, 0), dist=(0, 1)),           # dist_fwd 0
            job(8, 8, 8, 8, (9000, 0)),                     # |MV| > 8191
            job(8, 8, 8, 8, (4000, 0), dist=(1, 3)),        # mirrored MV beyond 8191
            job(8, 8, 8, 8, (8000, 0), mvp=(-9000, 0)),     # MV difference beyond the mvbits table
            job(8, 8, 8, 8, (4, -4))]                       # a good one
    prm = X.me_params(w, h, 16, LAM)
    d1, d2 = device_ref(r1, 2), device_ref(r2, 2)
    got = X.me_search_sym_batch(prm, jobs_array(rows), device_cur(cur), d1, d2)
    assert list(got["status"]) == [2, 2, 2, 2, 2, 2, 0]
    assert np.all(got["cost"][:6] == X.MAX_DISTORTION)
    assert got["cost"][6] < X.MAX_DISTORTION
    # an empty batch is a no-op; mismatched strides and a misaligned current plane are errors
    assert len(X.me_search_sym_batch(prm, jobs_array([]), device_cur(cur), d1, d2)) == 0
    odd = X.DeviceRef(w + 512, h, frac=2)
    odd.load(np.zeros((h, w + 512), np.uint8))
    with pytest.raises(RuntimeError):
        X.me_search_sym_batch(prm, jobs_array(rows), device_cur(cur), d1, odd)
    dc = device_cur(cur)
    jb = X.upload(jobs_array(rows))
    res = X.DeviceBuffer(16 * len(rows))
    r1d, r2d = d1.desc(), d2.desc()
    rc = X.lib().xme_me_search_sym_batch(C.byref(prm), C.c_void_p(jb.ptr), len(rows), C.c_void_p(dc.origin + 1),
                                         dc.stride, C.byref(r1d), C.byref(r2d), C.c_void_p(res.ptr), None)
    assert rc == X.XME_ERR_ARG


def test_1080p_mixed_shapes():
    W, H = 1920, 1080
    cur, r1, r2 = pictures(W, H, 51)
    rng = np.random.default_rng(52)
    rows = []
    for (x, y, pw, ph) in synth.mixed_pus(W, H, 53):
        mv = (int(rng.integers(-96, 97)), int(rng.integers(-96, 97)))
        rows.append(job(x, y, pw, ph, mv, mvp=(int(rng.integers(-64, 65)), int(rng.integers(-64, 65))),
                        mvp1=mv, dist=(int(rng.integers(1, 5)), int(rng.integers(1, 5))),
                        b_slice=int(rng.integers(0, 2))))
    got, exp = run_sym(W, H, cur, r1, r2, rows, R=64, enable_pmvr=1)
    assert_same(got, exp, rows)


# ---- bi-directional refinement (xme_me_search_bid_batch, me.c:1278-1392) ----

def bid_job(x, y, w, h, fwd, bwd, mvp1=(0, 0), mvp2=(0, 0)):
    return dict(pix_x=x, pix_y=y, w=w, h=h, fwd_x=fwd[0], fwd_y=fwd[1], bwd_x=bwd[0], bwd_y=bwd[1],
                mvp1_x=mvp1[0], mvp1_y=mvp1[1], mvp2_x=mvp2[0], mvp2_y=mvp2[1])


def bid_array(rows):
    a = np.zeros(len(rows), dtype=X.BID_JOB_DTYPE)
    for i, r in enumerate(rows):
        for k, v in r.items():
            a[k if k not in ("pix_x", "pix_y") else k[-1]][i] = v
    return a


def built(frac, mx, my):
    k = (my & 3) * 4 + (mx & 3)
    return frac == 2 or k == 0 or (frac == 1 and k in (2, 8, 10))


def run_bid(w, h, cur, r1, r2, rows, R=16, **kw):
    """GPU and oracle results, and the status the GPU must report: 2 where the
    backward MV is inside the window and its plane is not built (the
    reference's mc_luma fallback), else 0."""
    frac = kw.get("use_fractional_me", 2)
    prm_o = O.me_param(w, h, R, LAM, **kw)
    prm_d = X.me_params(w, h, R, LAM, **kw)
    got = X.me_search_bid_batch(prm_d, bid_array(rows), device_cur(cur), device_ref(r1, frac), device_ref(r2, frac))
    p1, p2, oc = oracle_planes(r1, frac), oracle_planes(r2, frac), oracle_cur(cur)
    exp = np.array([O.me_search_bid(prm_o, j, oc, p1, p2) for j in rows], dtype=np.int64)
    mx_lo, mx_hi = kw.get("mv_range_x", (-8192, 8191))
    my_lo, my_hi = kw.get("mv_range_y", (-2048, 2047))
    status = []
    for j in rows:
        x, y, bw, bh = j["pix_x"], j["pix_y"], j["w"], j["h"]
        legal = (max(mx_lo, (-64 - x) * 4) <= j["bwd_x"] <= min(mx_hi, (w + 64 - x - bw) * 4) and
                 max(my_lo, (-64 - y) * 4) <= j["bwd_y"] <= min(my_hi, (h + 64 - y - bh) * 4))
        status.append(2 if legal and not built(frac, j["bwd_x"], j["bwd_y"]) else 0)
    return got, exp, np.array(status)


def assert_bid_same(got, exp, status, rows):
    assert np.array_equal(got["status"], status)
    ok = status == 0
    assert np.any(ok)
    g = np.stack([got["cost"], got["mv_x"], got["mv_y"], got["mvcost"]], axis=1).astype(np.int64)
    bad = np.nonzero(ok & np.any(g != exp, axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} of {len(rows)} differ; first {rows[bad[0]]}: {g[bad[0]]} vs {exp[bad[0]]}"
    assert np.all(got["cost"][~ok] == X.MAX_DISTORTION)


def random_bid_rows(w, h, n, rng, span=48):
    rows = []
    for i in range(n):
        part = i % 25
        pw, ph = X.PU_W[part], X.PU_H[part]
        fwd = (int(rng.integers(-span, span + 1)), int(rng.integers(-span, span + 1)))
        rows.append(bid_job(int(rng.integers(0, (w - pw) // 4 + 1)) * 4, int(rng.integers(0, h - ph + 1)), pw, ph,
                            fwd, (int(rng.integers(-span, span + 1)), int(rng.integers(-span, span + 1))),
                            mvp1=(fwd[0] + int(rng.integers(-8, 9)), fwd[1] + int(rng.integers(-8, 9))),
                            mvp2=(int(rng.integers(-40, 41)), int(rng.integers(-40, 41)))))
    return rows


@pytest.mark.parametrize("kw", [dict(), dict(use_hadamard=0), dict(use_fractional_me=1),
                                dict(use_fractional_me=0), dict(enable_pmvr=1),
                                dict(use_hadamard=0, enable_pmvr=1)],
                         ids=["satd", "sad", "frac1", "frac0", "pmvr", "sad-pmvr"])
def test_bid_random_jobs_every_shape(kw):
    w, h = 192, 160
    cur, r1, r2 = pictures(w, h, 61)
    rows = random_bid_rows(w, h, 200, np.random.default_rng(6))
    got, exp, status = run_bid(w, h, cur, r1, r2, rows, **kw)
    assert_bid_same(got, exp, status, rows)


def test_bid_window_edges_and_narrow_range():
    w, h = 128, 96
    cur, r1, r2 = pictures(w, h, 71)
    rows = []
    for (pw, ph) in ((16, 16), (8, 8), (64, 64), (32, 8)):
        for (x, y) in ((0, 0), (w - pw, h - ph), (0, h - ph), (w - pw, 0)):
            lo_x, hi_x = (-64 - x) * 4, (w + 64 - x - pw) * 4
            lo_y, hi_y = (-64 - y) * 4, (h + 64 - y - ph) * 4
            for fwd in ((lo_x, lo_y), (hi_x, hi_y), (lo_x + 1, hi_y - 1), (hi_x - 2, lo_y + 2)):
                for bwd in ((lo_x, hi_y), (0, 0), (lo_x - 4, 0)):
                    rows.append(bid_job(x, y, pw, ph, fwd, bwd, mvp1=fwd, mvp2=bwd))
    got, exp, status = run_bid(w, h, cur, r1, r2, rows)
    assert_bid_same(got, exp, status, rows)
    assert np.any(got["cost"] == X.MAX_DISTORTION) and np.any(got["cost"] < X.MAX_DISTORTION)
    rows2 = random_bid_rows(w, h, 100, np.random.default_rng(9), span=44)
    got, exp, status = run_bid(w, h, cur, r1, r2, rows2, mv_range_x=(-40, 40), mv_range_y=(-40, 40))
    assert_bid_same(got, exp, status, rows2)
    assert np.any(got["cost"] == X.MAX_DISTORTION)


@pytest.mark.parametrize("d,e", [((3, -2), (-1, 4)), ((-4, 1), (2, 2))])
def test_bid_known_answer(d, e):
    """Both references reproduce the PU at (fwd, bwd): the target is the PU
    itself and the start scores the two MV rates alone."""
    w, h = 96, 80
    cur = synth.smooth_field(w, h, seed=81)
    r1 = np.roll(cur, (d[1], d[0]), axis=(0, 1))
    r2 = np.roll(cur, (e[1], e[0]), axis=(0, 1))
    fwd, bwd = (4 * d[0], 4 * d[1]), (4 * e[0], 4 * e[1])
    rows = [bid_job(32, 24, 16, 16, fwd, bwd, (3, -5), (-7, 2))]
    got, exp, status = run_bid(w, h, cur, r1, r2, rows)
    assert_bid_same(got, exp, status, rows)

    def bits(v):
        return 1 if v == 0 else 2 * int(abs(v)).bit_length() + 1
    rate = ((LAM * (bits(fwd[0] - 3) + bits(fwd[1] + 5))) >> 16) + ((LAM * (bits(bwd[0] + 7) + bits(bwd[1] - 2))) >> 16)
    assert (got["cost"][0], got["mv_x"][0], got["mv_y"][0], got["mvcost"][0]) == (rate, fwd[0], fwd[1], rate)


def test_bid_jobs_not_evaluated():
    w, h = 96, 80
    cur, r1, r2 = pictures(w, h, 91)
    rows = [bid_job(2, 8, 8, 8, (0, 0), (0, 0)),                   # x % 4 != 0
            bid_job(8, 76, 8, 8, (0, 0), (0, 0)),                  # beyond the bottom edge
            bid_job(8, 8, 8, 8, (9000, 0), (0, 0)),                # |fwd| > 8191
            bid_job(8, 8, 8, 8, (0, 0), (0, -8200)),               # |bwd| > 8191
            bid_job(8, 8, 8, 8, (0, 0), (8000, 0), mvp2=(-9000, 0)),   # backward MV difference beyond the table
            bid_job(8, 8, 8, 8, (4, -4), (8, 2))]                  # a good one
    prm = X.me_params(w, h, 16, LAM)
    got = X.me_search_bid_batch(prm, bid_array(rows), device_cur(cur), device_ref(r1, 2), device_ref(r2, 2))
    assert list(got["status"]) == [2, 2, 2, 2, 2, 0]
    assert got["cost"][5] < X.MAX_DISTORTION


def test_bid_1080p_mixed_shapes():
    W, H = 1920, 1080
    cur, r1, r2 = pictures(W, H, 101)
    rng = np.random.default_rng(102)
    rows = []
    for (x, y, pw, ph) in synth.mixed_pus(W, H, 103):
        fwd = (int(rng.integers(-96, 97)), int(rng.integers(-96, 97)))
        rows.append(bid_job(x, y, pw, ph, fwd, (int(rng.integers(-96, 97)), int(rng.integers(-96, 97))),
                            mvp1=fwd, mvp2=(int(rng.integers(-64, 65)), int(rng.integers(-64, 65)))))
    got, exp, status = run_bid(W, H, cur, r1, r2, rows, R=64, enable_pmvr=1)
    assert_bid_same(got, exp, status, rows)
