"""GPU parity of the callers of the bi-prediction refinements against the
oracle: xme_pred_inter_search_bi_batch (pred_inter_search_bi,
encoder/md_inter.c:1036-1128, B slices) and xme_pred_inter_search_dual_batch
(pred_inter_search_dual, md_inter.c:1133-1219, F slices), frame sequences:

* random PUs over every LumaPU shape with SATD and SAD, use_fractional_me 1
  and 2, PMVR, one and several frame threads, LCU levels 4 to 6, start MVs
  and predictors drawn so that every gate (check_mvd, check_mv_range's box,
  check_mv_range_sym's mirror) both passes and fails;
* each gate on its hand-worked boundary, as the oracle's own tests pin them;
* the symmetric known answer (the picture is the average of two references
  moved by +d and -d) reached through the caller;
* dual over 2, 3 and 4 references: REF_COST per reference, the first strict
  minimum, the second MVs;
* the PUs the entry points do not evaluate (status 2) and argument errors.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_me import LAM, device_cur, oracle_cur, oracle_ref
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu

W, H = 192, 128
MAX = X.MAX_DISTORTION
BI_FIELDS = ("sym_cost", "sym_mv_x", "sym_mv_y", "sym_mvcost", "bid_cost", "fwd_x", "fwd_y", "bwd_x", "bwd_y",
             "bid_mvcost")


def frames(n, seed):
    return [synth.smooth_field(W, H, seed=seed + k, noise=6.0) for k in range(n)]


def oracle_planes(img, frac):
    pl = oracle_ref(img, frac)
    if frac == 1:
        # use_fractional_me 1 builds planes 2, 8 and 10 only (frame.c:343-350)
        pl = [p if k in (0, 2, 8, 10) else None for k, p in enumerate(pl)]
    return pl


def device_ref(img, frac):
    d = X.DeviceRef(W, H, frac=frac)
    d.load(img)
    return d


def bi_jobs(rows):
    a = np.zeros(len(rows), dtype=X.BI_JOB_DTYPE)
    for i, r in enumerate(rows):
        for k, v in r.items():
            a[{"pix_x": "x", "pix_y": "y"}.get(k, k)][i] = v
    return a


def box_fpel(x, y, w, h, lcu_level):
    """check_mv_range's box in full pels for a PU (predict.h:70-74)."""
    b = 1 << lcu_level
    return -(x + b), -(y + b), (W - (x + w)) + b, (H - (y + h)) + b


def near(rng, lo, hi, span):
    """A value around one of the box edges or inside it."""
    k = rng.integers(0, 4)
    if k == 0:
        return int(lo + rng.integers(-2, 3))
    if k == 1:
        return int(hi + rng.integers(-2, 3))
    return int(rng.integers(-span, span + 1))


def random_bi_rows(n, rng, lcu_level):
    rows = []
    for i in range(n):
        part = i % 25
        w, h = X.PU_W[part], X.PU_H[part]
        x = int(rng.integers(0, (W - w) // 4 + 1)) * 4
        y = int(rng.integers(0, H - h + 1))
        x0, y0, x1, y1 = box_fpel(x, y, w, h, lcu_level)
        imv = (near(rng, x0, x1, 12), near(rng, y0, y1, 12))
        bwd = (4 * near(rng, x0, x1, 12) + int(rng.integers(-3, 4)),
               4 * near(rng, y0, y1, 12) + int(rng.integers(-3, 4)))
        mvp1 = (4 * imv[0] + int(rng.integers(-24, 25)), 4 * imv[1] + int(rng.integers(-24, 25)))
        if rng.integers(0, 6) == 0:      # the mirror or check_mvd gates
            mvp1 = (int(rng.choice([-4097, -4096, 4095, 4096, 4 * x0 - 4, 4 * x1 + 4])), mvp1[1])
        rows.append(dict(pix_x=x, pix_y=y, w=w, h=h, fwd_imv_x=imv[0], fwd_imv_y=imv[1], bwd_x=bwd[0], bwd_y=bwd[1],
                         mvp_x=int(rng.integers(-60, 61)), mvp_y=int(rng.integers(-60, 61)), mvp1_x=mvp1[0],
                         mvp1_y=mvp1[1], mvp2_x=bwd[0] + int(rng.integers(-40, 41)),
                         mvp2_y=bwd[1] + int(rng.integers(-40, 41))))
    return rows


def run_bi(rows, imgs, dist, frac, lcu_level=6, frame_threads=1, R=32, **kw):
    cur, r1, r2 = imgs
    prm_o = O.me_param(W, H, R, LAM, use_fractional_me=frac, **kw)
    prm_d = X.me_params(W, H, R, LAM, use_fractional_me=frac, **kw)
    got = X.pred_inter_search_bi_batch(prm_d, bi_jobs(rows), dist[0], dist[1], device_cur(cur),
                                       device_ref(r1, frac), device_ref(r2, frac), lcu_level=lcu_level,
                                       frame_threads=frame_threads)
    oc, p1, p2 = oracle_cur(cur), oracle_planes(r1, frac), oracle_planes(r2, frac)
    exp = [O.pred_inter_search_bi(prm_o, j, dist[0], dist[1], oc, p1, p2, lcu_level=lcu_level,
                                  frame_threads=frame_threads) for j in rows]
    return got, exp


def compare_bi(got, exp, rows, allow_status2=None):
    """Every evaluated PU equals the oracle field by field; status 2 only
    where allow_status2(row) says the entry point may decline."""
    n2 = 0
    for i, (g, e) in enumerate(zip(got, exp)):
        if g["status"] != 0:
            assert g["status"] == 2 and allow_status2 is not None and allow_status2(rows[i]), (i, rows[i], g)
            n2 += 1
            continue
        gv = tuple(int(g[f]) for f in BI_FIELDS)
        ev = tuple(int(e[f]) for f in BI_FIELDS)
        assert gv == ev, f"PU {i} {rows[i]}:\n got {dict(zip(BI_FIELDS, gv))}\n exp {dict(zip(BI_FIELDS, ev))}"
    return n2


@pytest.fixture(scope="module")
def imgs():
    return frames(3, 11)


@pytest.mark.parametrize("hadamard,pmvr,lcu_level,frame_threads,dist", [
    (1, 1, 6, 1, (2, 2)),
    (0, 1, 6, 1, (1, 3)),
    (1, 0, 5, 2, (3, 1)),
    (1, 1, 4, 4, (4, 4)),
])
def test_bi_random_matches_oracle(imgs, hadamard, pmvr, lcu_level, frame_threads, dist):
    rng = np.random.default_rng(100 * lcu_level + frame_threads + 7 * hadamard)
    rows = random_bi_rows(150, rng, lcu_level)
    got, exp = run_bi(rows, imgs, dist, 2, lcu_level=lcu_level, frame_threads=frame_threads,
                      use_hadamard=hadamard, enable_pmvr=pmvr)
    assert compare_bi(got, exp, rows) == 0
    # the draw exercises both outcomes of both gates
    sym_ran = np.sum(got["sym_mvcost"] >= 0)
    bid_ran = np.sum(got["bid_mvcost"] >= 0)
    assert 0 < sym_ran < len(rows) and 0 < bid_ran < len(rows), (sym_ran, bid_ran)
    assert np.any(got["sym_cost"] == MAX) and np.any(got["bid_cost"] < MAX)


def test_bi_half_pel_only(imgs):
    """use_fractional_me 1: planes 2, 8 and 10 only; a backward MV on a plane
    that is not built is predicted with mc_luma in the reference (not
    provided here): those PUs report status 2 when the bid search runs."""
    rng = np.random.default_rng(5)
    rows = random_bi_rows(100, rng, 6)
    got, exp = run_bi(rows, imgs, (2, 2), 1)
    built = {(0, 0), (2, 0), (0, 2), (2, 2)}
    n2 = compare_bi(got, exp, rows, allow_status2=lambda r: (r["bwd_x"] & 3, r["bwd_y"] & 3) not in built)
    assert n2 < len(rows)


@pytest.mark.parametrize("field,value", [
    ("mvp1_x", 448), ("mvp1_x", 452), ("mvp1_x", 4095), ("mvp1_x", 4096), ("mvp1_x", -4096), ("mvp1_x", -4097),
    ("fwd_imv_x", 192), ("fwd_imv_x", 193), ("fwd_imv_y", -96), ("fwd_imv_y", -97),
    ("bwd_x", 768), ("bwd_x", 769), ("mvp2_y", 4095), ("mvp2_y", 4096),
])
@pytest.mark.parametrize("frame_threads", [1, 2])
def test_bi_gate_boundaries(imgs, field, value, frame_threads):
    """The PU (48, 32) 16x16 at LCU level 6 in the 192x128 picture: the qpel
    box is x in [-448, 768], y in [-384, 576] (check_mv_range,
    predict.h:70-79); the mirror of mvp1 in a B slice at equal distances is
    -mvp1."""
    row = dict(pix_x=48, pix_y=32, w=16, h=16, fwd_imv_x=3, fwd_imv_y=-2, bwd_x=-12, bwd_y=8, mvp_x=0, mvp_y=0,
               mvp1_x=0, mvp1_y=0, mvp2_x=0, mvp2_y=0)
    row[field] = value
    got, exp = run_bi([row], imgs, (2, 2), 2, frame_threads=frame_threads)
    compare_bi(got, exp, [row])


def test_bi_known_answer():
    """ref1 moved by +d and ref2 by -d average to the picture: the symmetric MV
    4d predicts it exactly and scores its MV rate alone."""
    d = (3, -2)
    cur = synth.smooth_field(W, H, seed=3, noise=6.0)
    r1 = np.roll(cur, (d[1], d[0]), axis=(0, 1))
    r2 = np.roll(cur, (-d[1], -d[0]), axis=(0, 1))
    row = dict(pix_x=48, pix_y=32, w=16, h=16, fwd_imv_x=d[0], fwd_imv_y=d[1], bwd_x=-4 * d[0], bwd_y=-4 * d[1],
               mvp_x=0, mvp_y=0, mvp1_x=0, mvp1_y=0, mvp2_x=0, mvp2_y=0)
    got, exp = run_bi([row], (cur, r1, r2), (2, 2), 2)
    compare_bi(got, exp, [row])
    rate = (LAM * (O.mvbits(32, 4 * d[0]) + O.mvbits(32, 4 * d[1]))) >> 16
    assert (int(got["sym_mv_x"][0]), int(got["sym_mv_y"][0])) == (4 * d[0], 4 * d[1])
    assert int(got["sym_cost"][0]) == rate == int(got["sym_mvcost"][0])


def dual_jobs(rows):
    a = np.zeros(len(rows), dtype=X.DUAL_JOB_DTYPE)
    for i, r in enumerate(rows):
        a["x"][i], a["y"][i], a["w"][i], a["h"][i] = r["pix_x"], r["pix_y"], r["w"], r["h"]
        a["mvp_x"][i], a["mvp_y"][i] = r["mvp_x"], r["mvp_y"]
        for k, (mv1, imv) in enumerate(zip(r["mvp1"], r["imv"])):
            a["mvp1"][i, k] = mv1
            a["imv"][i, k] = imv
    return a


def random_dual_rows(n, n_ref, rng, lcu_level):
    rows = []
    for i in range(n):
        part = i % 25
        w, h = X.PU_W[part], X.PU_H[part]
        x = int(rng.integers(0, (W - w) // 4 + 1)) * 4
        y = int(rng.integers(0, H - h + 1))
        x0, y0, x1, y1 = box_fpel(x, y, w, h, lcu_level)
        imv = [(near(rng, x0, x1, 10), near(rng, y0, y1, 10)) for _ in range(n_ref)]
        mvp1 = [(4 * a + int(rng.integers(-20, 21)), 4 * b + int(rng.integers(-20, 21))) for a, b in imv]
        if rng.integers(0, 8) == 0:
            k = int(rng.integers(0, n_ref))
            mvp1[k] = (int(rng.choice([-4097, 4095, 4096])), mvp1[k][1])
        mvp = (int(rng.integers(-60, 61)), int(rng.integers(-60, 61)))
        if rng.integers(0, 12) == 0:     # p_me->mvp fails check_mvd: no reference passes
            mvp = (4096, mvp[1])
        rows.append(dict(pix_x=x, pix_y=y, w=w, h=h, mvp_x=mvp[0], mvp_y=mvp[1], mvp1=mvp1, imv=imv))
    return rows


def run_dual(rows, imgs, dist, lcu_level=6, frame_threads=1, R=32, **kw):
    cur, refs = imgs[0], imgs[1:1 + len(dist)]
    prm_o = O.me_param(W, H, R, LAM, **kw)
    prm_d = X.me_params(W, H, R, LAM, **kw)
    got = X.pred_inter_search_dual_batch(prm_d, dual_jobs(rows), dist, device_cur(cur),
                                         [device_ref(r, 2) for r in refs], lcu_level=lcu_level,
                                         frame_threads=frame_threads)
    oc, pls = oracle_cur(cur), [oracle_planes(r, 2) for r in refs]
    exp = [O.pred_inter_search_dual(prm_o, j, dist, oc, pls, lcu_level=lcu_level, frame_threads=frame_threads)
           for j in rows]
    return got, exp


def compare_dual(got, exp, rows, n_ref):
    for i, (g, e) in enumerate(zip(got, exp)):
        assert g["status"] == 0, (i, rows[i])
        gv = (int(g["cost"]), int(g["fst_ref"]), int(g["snd_ref"]), int(g["mvcost"]),
              [tuple(int(v) for v in g["fst"][r]) for r in range(n_ref)],
              [tuple(int(v) for v in g["snd"][r]) for r in range(n_ref)])
        assert gv == tuple(e), f"PU {i} {rows[i]}:\n got {gv}\n exp {e}"
        assert not np.any(g["fst"][n_ref:]) and not np.any(g["snd"][n_ref:])


@pytest.fixture(scope="module")
def imgs5():
    return frames(5, 21)


@pytest.mark.parametrize("dist,lcu_level,frame_threads,hadamard", [
    ((1, 2), 6, 1, 1),
    ((2, 4, 6), 5, 2, 1),
    ((1, 3, 5, 7), 6, 1, 0),
    ((3, 1, 2, 4), 4, 3, 1),
])
def test_dual_random_matches_oracle(imgs5, dist, lcu_level, frame_threads, hadamard):
    n_ref = len(dist)
    rng = np.random.default_rng(10 * n_ref + lcu_level)
    rows = random_dual_rows(120, n_ref, rng, lcu_level)
    got, exp = run_dual(rows, imgs5, dist, lcu_level=lcu_level, frame_threads=frame_threads, use_hadamard=hadamard)
    compare_dual(got, exp, rows, n_ref)
    # every reference wins somewhere, and some PUs have none
    assert set(range(n_ref)) <= set(int(v) for v in got["fst_ref"])
    assert np.any(got["fst_ref"] == -1) and np.all((got["fst_ref"] == -1) == (got["cost"] == MAX))


def test_dual_ref_cost_breaks_ties(imgs5):
    """Identical references: the searches score alike, REF_COST(0) = lambda *
    1 >> 16 is the smallest, so reference 0 wins with its cost plus 4."""
    cur, r = imgs5[0], imgs5[1]
    row = dict(pix_x=64, pix_y=32, w=16, h=16, mvp_x=0, mvp_y=0, mvp1=[(0, 0)] * 3, imv=[(1, -1)] * 3)
    got, exp = run_dual([row], (cur, r, r, r), (2, 2, 2))
    compare_dual(got, exp, [row], 3)
    assert int(got["fst_ref"][0]) == 0 and int(got["snd_ref"][0]) == 1


def test_status_2_and_argument_errors(imgs):
    cur, r1, r2 = imgs
    prm = X.me_params(W, H, 32, LAM)
    dc, d1, d2 = device_cur(cur), device_ref(r1, 2), device_ref(r2, 2)
    # a PU the refinements do not take (x not a multiple of 4, a 6-pixel
    # width) and a mirror whose scale_mv_skip product leaves int
    rows = [dict(pix_x=2, pix_y=0, w=8, h=8), dict(pix_x=0, pix_y=0, w=6, h=8),
            dict(pix_x=16, pix_y=16, w=8, h=8, fwd_imv_x=140)]
    got = X.pred_inter_search_bi_batch(prm, bi_jobs(rows[:2]), 2, 2, dc, d1, d2)
    assert list(got["status"]) == [2, 2]
    # 140 * 4 * 255 * 16384 > 2^31: undefined in the reference's int arithmetic
    got = X.pred_inter_search_bi_batch(prm, bi_jobs(rows[2:]), 1, 255, dc, d1, d2)
    assert list(got["status"]) == [2]
    got = X.pred_inter_search_dual_batch(prm, dual_jobs([dict(pix_x=2, pix_y=0, w=8, h=8, mvp_x=0, mvp_y=0,
                                                              mvp1=[(0, 0)] * 2, imv=[(0, 0)] * 2)]),
                                         (1, 2), dc, [d1, d2])
    assert list(got["status"]) == [2]
    # empty batches are fine
    assert len(X.pred_inter_search_bi_batch(prm, bi_jobs([]), 2, 2, dc, d1, d2)) == 0
    # argument errors
    L = X.lib()
    jobs = X.upload(bi_jobs(rows[:1]))
    res = X.DeviceBuffer(64)
    rd1, rd2 = d1.desc(), d2.desc()

    def bi(lcu=6, ft=1, df=2, db=2, ref2=rd2):
        return L.xme_pred_inter_search_bi_batch(C.byref(prm), lcu, ft, C.c_void_p(jobs.ptr), 1, df, db,
                                                C.c_void_p(dc.origin), dc.stride, C.byref(rd1), C.byref(ref2),
                                                C.c_void_p(res.ptr), None)
    assert bi() == X.XME_OK
    for kw in (dict(lcu=2), dict(lcu=7), dict(ft=0), dict(df=0), dict(db=256), dict(df=-256)):
        assert bi(**kw) == X.XME_ERR_ARG, kw
    other = X.RefDesc()
    other.planes[0] = rd2.planes[0]
    other.stride = rd2.stride + 64
    assert bi(ref2=other) == X.XME_ERR_ARG
    descs = (X.RefDesc * 4)(rd1, rd2, rd1, rd2)
    djobs = X.upload(dual_jobs([dict(pix_x=0, pix_y=0, w=8, h=8, mvp_x=0, mvp_y=0, mvp1=[(0, 0)] * 2,
                                     imv=[(0, 0)] * 2)]))

    def dual(n_ref=2, dist=(1, 2, 0, 0)):
        return L.xme_pred_inter_search_dual_batch(C.byref(prm), 6, 1, C.c_void_p(djobs.ptr), 1, n_ref,
                                                  (C.c_int * 4)(*dist), C.c_void_p(dc.origin), dc.stride, descs,
                                                  C.c_void_p(res.ptr), None)
    assert dual() == X.XME_OK
    for kw in (dict(n_ref=1), dict(n_ref=5), dict(n_ref=3, dist=(1, 2, 0, 0)), dict(dist=(0, 2, 0, 0))):
        assert dual(**kw) == X.XME_ERR_ARG, kw
    assert L.xme_stream_sync(None) == X.XME_OK
