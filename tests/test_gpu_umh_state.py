"""GPU parity of the UMH search's encoder-state batch (xme_umh_prepare_batch:
get_mvp_default_sad, encoder/md_inter.c:527-608, and fast_me_prepare_info,
md_inter.c:640-718) against the oracle: random jobs over every mode, the P,
B and F slice rules, co-located MVs, int32-wrapping predicted-SAD squares,
refused indices and modes, zero distances; h->umh_bsize
(xavs2_me_init_umh_threshold, me.c:647-665) for every QP and slice type;
and the states fed to a UMH plan on the device (xme_plan_set_umh_states)
give the oracle's search."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_mvp import random_jobs
from tests.test_gpu_pattern import run_umh_case
from xavs2_amd import lib as X

pytestmark = pytest.mark.gpu

P_DPOC = [2, 4, 6, 8]
P_MULTI = [16384 // d for d in P_DPOC]
B_DPOC = [2, 6, 0, 0]
B_MULTI = [8192, 2731, 0, 0]
# oracle status -> xme_umh_state status
STATUS = {0: 0, -1: 3, -2: 2, -3: 4}


def random_umh_jobs(n, rng, max_ref, big_sad_frac=0.05):
    jobs = np.zeros(n, X.UMH_JOB_DTYPE)
    jobs["mvp"] = random_jobs(n, rng, max_ref)
    jobs["mode"] = rng.integers(1, 8, n)
    jobs["nb_sad"] = rng.integers(0, 8000, (n, 3))
    jobs["mincost_up"] = rng.integers(0, 16000, (n, 3))
    jobs["best_up"] = rng.integers(-200, 201, (n, 3, 2))
    jobs["mincost_prev"] = rng.integers(0, 8000, n)
    big = rng.random(n) < big_sad_frac
    jobs["mincost_prev"][big] = rng.integers(46341, 200000, int(big.sum()))
    jobs["best_prev"] = rng.integers(-300, 301, (n, 2))
    jobs["col_valid"] = rng.random(n) < 0.7
    jobs["col_ref"] = rng.integers(-1, max_ref + 1, n)
    jobs["col_mv"] = rng.integers(-300, 301, (n, 2))
    jobs["col_dpoc"] = rng.choice([-6, -4, -2, 2, 4, 6, 8, 10], n)
    return jobs


def oracle_job(j):
    m = j["mvp"]
    mj = dict(ref_idx=int(m["ref_idx"]), bwd_2nd=int(m["bwd_2nd"]),
              cb=(int(m["cb_x"]), int(m["cb_y"]), int(m["cb_w"]), int(m["cb_h"])),
              ur_available=int(m["ur_available"]), ref=[int(v) for v in m["ref"]],
              mv=[[int(v) for v in mv] for mv in m["mv"]])
    col = (int(j["col_ref"]), tuple(int(v) for v in j["col_mv"]), int(j["col_dpoc"])) if j["col_valid"] else None
    return O.umh_job(mj, int(j["mode"]), nb_sad=[int(v) for v in j["nb_sad"]],
                     mincost_up=[int(v) for v in j["mincost_up"]],
                     best_up=[tuple(int(v) for v in b) for b in j["best_up"]], mincost_prev=int(j["mincost_prev"]),
                     best_prev=tuple(int(v) for v in j["best_prev"]), col=col)


def run_gpu(jobs, slice_type, dpoc, multi, bsize):
    d_jobs = X.upload(jobs)
    d_out = X.DeviceBuffer(max(1, len(jobs)) * X.UMH_STATE_DTYPE.itemsize)
    dp = (C.c_int32 * X.XME_MVP_REFS)(*dpoc)
    mu = (C.c_int32 * X.XME_MVP_REFS)(*multi)
    bs = (C.c_double * 8)(*bsize)
    X.check(X.lib().xme_umh_prepare_batch(d_jobs.at(), len(jobs), slice_type, dp, mu, bs, d_out.at(), None),
            "xme_umh_prepare_batch")
    X.check(X.lib().xme_stream_sync(None), "sync")
    return d_out.download(len(jobs), X.UMH_STATE_DTYPE), d_out


def gpu_bsize(qp, slice_type):
    b = (C.c_double * 8)()
    X.check(X.lib().xme_umh_bsize(qp, slice_type, b), "xme_umh_bsize")
    return list(b)


def expected(jobs, slice_type, dpoc, multi, bsize):
    exp = np.zeros(len(jobs), X.UMH_STATE_DTYPE)
    for i, j in enumerate(jobs):
        o = O.umh_prepare(oracle_job(j), slice_type, dpoc, multi, bsize)
        e = exp[i]
        e["status"] = STATUS[o.status]
        if o.status == -2:
            continue
        e["mvp_x"], e["mvp_y"], e["pred_sad_space"] = o.pmv_x, o.pmv_y, o.pred_sad_space
        for f in X.UMH_CTX_DTYPE.names:
            if f != "reserved":
                e["ctx"][f] = getattr(o.ctx, f)
    return exp


def compare(got, exp):
    for f in ("status", "mvp_x", "mvp_y", "pred_sad_space"):
        bad = np.nonzero(got[f] != exp[f])[0]
        assert len(bad) == 0, (f, bad[:5], got[f][bad[:5]], exp[f][bad[:5]])
    for f in X.UMH_CTX_DTYPE.names:
        g, e = got["ctx"][f], exp["ctx"][f]
        # the betas bit for bit: the same IEEE operations in the same order
        same = (g.view(np.int64) == e.view(np.int64)) if g.dtype == np.float64 else (g == e)
        bad = np.nonzero(~same)[0]
        assert len(bad) == 0, (f, bad[:5], g[bad[:5]], e[bad[:5]])


@pytest.mark.parametrize("slice_type", [1, 2, 3], ids=["P", "B", "F"])
@pytest.mark.parametrize("qp", [22, 45])
def test_random_jobs(slice_type, qp):
    dpoc, multi = (B_DPOC, B_MULTI) if slice_type == 2 else (P_DPOC, P_MULTI)
    rng = np.random.default_rng(100 * slice_type + qp)
    jobs = random_umh_jobs(3000, rng, 1 if slice_type == 2 else 3)
    bsize = O.umh_bsize(qp, slice_type)
    got, _ = run_gpu(jobs, slice_type, dpoc, multi, bsize)
    exp = expected(jobs, slice_type, dpoc, multi, bsize)
    compare(got, exp)
    # the draw reaches every route: wrapped squares (negative betas), co-located MVs, B_FWD
    assert (got["ctx"]["beta2"] < 0).any() and (got["ctx"]["mvp3_x"] != 0).any()
    assert len(np.unique(got["pred_sad_space"])) > 500
    assert (got["status"] == 0).mean() > 0.5


def test_refused_and_zero_distances():
    """Indices outside the lists and modes outside 1..7 give status 2 and an
    all-zero record; a zero distance to the nearest reference or of the
    co-located block gives status 4 with the fields derived before it."""
    rng = np.random.default_rng(5)
    jobs = random_umh_jobs(64, rng, 3)
    jobs["mvp"]["ref"][:] = 0
    jobs["mvp"]["ref_idx"][:8] = 4
    jobs["mode"][8:16] = [0, 8, 9, -1, 0, 8, 100, 0]
    jobs["mvp"]["ref_idx"][16:24] = 1
    jobs["col_valid"][24:32] = 1
    jobs["col_ref"][24:32] = 0
    jobs["col_dpoc"][24:32] = 0
    dpoc = [0, 4, 6, 8]
    bsize = O.umh_bsize(33, 1)
    got, _ = run_gpu(jobs, 1, dpoc, P_MULTI, bsize)
    exp = expected(jobs, 1, dpoc, P_MULTI, bsize)
    compare(got, exp)
    assert (got["status"][:16] == 2).all()
    assert (got["status"][16:24] == 4).all() and (got["status"][24:32] == 4).all()
    zeroed = got[:16].copy()
    zeroed["status"] = 0
    assert (zeroed.view(np.uint8) == 0).all()


def test_umh_bsize_every_qp_and_slice_type():
    for st in range(4):
        for qp in range(80):
            assert gpu_bsize(qp, st) == list(O.umh_bsize(qp, st)), (qp, st)


def test_empty_batch_and_bad_arguments():
    dp = (C.c_int32 * X.XME_MVP_REFS)(*P_DPOC)
    bs = (C.c_double * 8)()
    assert X.lib().xme_umh_prepare_batch(None, 0, 1, dp, dp, bs, None, None) == X.XME_OK
    assert X.lib().xme_umh_prepare_batch(None, 4, 1, dp, dp, bs, None, None) != X.XME_OK
    assert X.lib().xme_umh_prepare_batch(None, 0, 4, dp, dp, bs, None, None) != X.XME_OK
    assert X.lib().xme_umh_bsize(80, 1, bs) != X.XME_OK


def test_states_drive_the_umh_plan():
    """16x16 and 8x8 PUs over two references: the states the batch writes on the
    device set the plan's context (xme_plan_set_umh_states) and its search
    equals the oracle's with the oracle's states."""
    w, h, n_ref = 128, 96, 2
    pus = [(x, y, 16, 16, (x % 48) - 20, (y % 40) - 16) for y in range(0, h, 16) for x in range(0, w, 16)]
    pus += [(x, y, 8, 8, 4, -8) for y in range(0, 32, 8) for x in range(0, 64, 8)]
    rng = np.random.default_rng(17)
    jobs = random_umh_jobs(len(pus) * n_ref, rng, 1, big_sad_frac=0.0)
    jobs["mvp"]["ref_idx"] = np.arange(len(jobs)) % n_ref
    bsize = O.umh_bsize(30, 1)
    got, d_states = run_gpu(jobs, 1, P_DPOC, P_MULTI, bsize)
    exp = expected(jobs, 1, P_DPOC, P_MULTI, bsize)
    compare(got, exp)
    assert (got["status"] == 0).all()
    ctx = np.ascontiguousarray(exp["ctx"])
    for level in (0, 2):
        run_umh_case(w, h, pus, 32, n_ref, 41 + level, level, ctx=ctx, d_states=d_states.at())
