"""GPU parity of the MV predictor batch (xme_get_mvp_batch, get_mvp_default,
encoder/md_inter.c:470-524) against the oracle: random neighbourhoods in P
and B slices (every type rule, invalid and unavailable neighbours, scaling
between references), scaling products that leave int (status 3, flagged by
the oracle too), and reference indices outside the lists (status 2)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from xavs2_amd import lib as X

pytestmark = pytest.mark.gpu

P_DPOC = [2, 4, 6, 8]
P_MULTI = [16384 // d for d in P_DPOC]
B_DPOC = [2, 6, 0, 0]
B_MULTI = [8192, 2731, 0, 0]


def run_gpu(jobs, b_slice, dpoc, multi):
    d_jobs = X.upload(jobs)
    d_out = X.DeviceBuffer(max(1, len(jobs)) * X.MVP_RESULT_DTYPE.itemsize)
    dp = (C.c_int32 * X.XME_MVP_REFS)(*dpoc)
    mu = (C.c_int32 * X.XME_MVP_REFS)(*multi)
    X.check(X.lib().xme_get_mvp_batch(d_jobs.at(), len(jobs), b_slice, dp, mu, d_out.at(), None), "xme_get_mvp_batch")
    X.check(X.lib().xme_stream_sync(None), "sync")
    return d_out.download(len(jobs), X.MVP_RESULT_DTYPE)


def random_jobs(n, rng, max_ref, mv_span=200):
    jobs = np.zeros(n, X.MVP_JOB_DTYPE)
    shapes = [(16, 16), (8, 16), (16, 8), (32, 8), (8, 32), (32, 32)]
    for i in range(n):
        w, h = shapes[int(rng.integers(0, len(shapes)))]
        jobs["cb_w"][i], jobs["cb_h"][i] = w, h
        jobs["cb_x"][i] = int(rng.integers(0, 2)) * w
        jobs["cb_y"][i] = int(rng.integers(0, 2)) * h
        jobs["ref_idx"][i] = int(rng.integers(0, max_ref + 1))
        jobs["bwd_2nd"][i] = int(rng.integers(0, 2))
        jobs["ur_available"][i] = int(rng.integers(0, 2))
        jobs["ref"][i] = rng.integers(-1, max_ref + 1, 4)
        jobs["mv"][i] = rng.integers(-mv_span, mv_span + 1, (4, 2))
    return jobs


def oracle_of(j, b_slice, dpoc, multi):
    return O.get_mvp_default(dict(ref_idx=int(j["ref_idx"]), bwd_2nd=int(j["bwd_2nd"]),
                                  cb=(int(j["cb_x"]), int(j["cb_y"]), int(j["cb_w"]), int(j["cb_h"])),
                                  ur_available=int(j["ur_available"]), ref=[int(v) for v in j["ref"]],
                                  mv=[[int(v) for v in m] for m in j["mv"]]), b_slice, dpoc, multi)


@pytest.mark.parametrize("b_slice", [0, 1])
def test_random_neighbourhoods(b_slice):
    dpoc, multi = (B_DPOC, B_MULTI) if b_slice else (P_DPOC, P_MULTI)
    rng = np.random.default_rng(7 + b_slice)
    jobs = random_jobs(4000, rng, 1 if b_slice else 3)
    got = run_gpu(jobs, b_slice, dpoc, multi)
    for i, j in enumerate(jobs):
        rc, x, y = oracle_of(j, b_slice, dpoc, multi)
        assert rc == 0 and got["status"][i] == 0, i
        assert (got["mvp_x"][i], got["mvp_y"][i]) == (x, y), (i, j)
    assert len(set(zip(got["mvp_x"].tolist(), got["mvp_y"].tolist()))) > 1000


def test_overflow_and_bad_reference_indices():
    rng = np.random.default_rng(3)
    jobs = random_jobs(64, rng, 3, mv_span=30000)
    dpoc, multi = [512] * 4, [16384] * 4
    got = run_gpu(jobs, 0, dpoc, multi)
    flagged = 0
    for i, j in enumerate(jobs):
        rc, x, y = oracle_of(j, 0, dpoc, multi)
        assert (got["status"][i] == 3) == (rc == -1), i
        assert (got["mvp_x"][i], got["mvp_y"][i]) == (x, y), i
        flagged += rc == -1
    assert flagged > 0
    bad = random_jobs(6, rng, 3)
    bad["ref_idx"][0] = 9
    bad["ref"][1][2] = 8
    bad["ref_idx"][2] = 4                  # MAX_REFS = 4 entries in the lists
    bad["ref"][3][0] = 4
    bad["ref"][4][3] = 5
    bad["ur_available"][4] = 1
    got = run_gpu(bad, 0, P_DPOC, P_MULTI)
    for i in range(6):
        rc, _, _ = oracle_of(bad[i], 0, P_DPOC, P_MULTI)
        assert (got["status"][i] == 2) == (rc == -2), i
    assert list(got["status"][[0, 2, 3, 4]]) == [2, 2, 2, 2]
    assert got["status"][1] == (2 if bad["ur_available"][1] == 0 else got["status"][1])
