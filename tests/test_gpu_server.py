"""GPU tests of the continuous-batching ME server: results against the oracle,
the same results alone and in a mix of concurrent streams, memory-budget
admission and rejection, and the latency record."""
import threading

import numpy as np
import pytest

from oracle import oracle as O
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu

LAM = 4 * 65536
GIB = 1 << 30


def oracle_frame_results(frames, f, pus, n_ref, R, me_method=O.ME_FS):
    """What the server must return for frame f: PU-major, reference r = frame f-1-r."""
    h, w = frames[0].shape
    prm = O.me_param(w, h, R, LAM, me_method=me_method)
    arr = np.array([(p[0], p[1], p[2], p[3], p[4], p[5]) for p in pus], dtype=O.PU_DTYPE)
    cur = O.Plane(w, h)
    cur.set_picture(frames[f])
    k = min(f, n_ref)
    out = np.zeros(len(pus) * k, dtype=X.RESULT_DTYPE)
    for r in range(k):
        p = O.Plane(w, h)
        p.set_picture(frames[f - 1 - r])
        p.expand_border()
        res = O.me_search_batch(prm, arr, cur, O.interp_planes(p, 2), nthreads=8)
        for name in X.RESULT_DTYPE.names:
            out[name][r::k] = res[name]
    return out


def pu_tuples(w, h, seed):
    rng = np.random.default_rng(seed)
    return [(x, y, pw, ph, int(rng.integers(-12, 13)), int(rng.integers(-12, 13)))
            for (x, y, pw, ph) in synth.mixed_pus(w, h, seed)]


def test_stream_results_match_oracle():
    w, h, R, n_ref = 192, 128, 16, 2
    frames = synth.moving_sequence(w, h, 5, seed=11, local=8)
    pus = pu_tuples(w, h, 4)
    srv = X.Server(4 * GIB, max_inflight=2, search_range=R, n_ref=n_ref)
    sid = srv.open_stream(w, h, X.make_pus(pus))
    jobs = [srv.submit(sid, fr) for fr in frames]
    for f, j in enumerate(jobs):
        got, lat = srv.wait(j)
        assert lat > 0
        exp = oracle_frame_results(frames, f, pus, n_ref, R)
        assert len(got) == len(pus) * min(f, n_ref)
        bad = np.nonzero(got != exp)[0]
        assert len(bad) == 0, f"frame {f}: {len(bad)} mismatches, first {bad[0]}: {got[bad[0]]} vs {exp[bad[0]]}"
    st = srv.stats()
    assert st["submitted"] == 5 and st["completed"] == 5 and st["rejected"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    srv.close_stream(sid)
    assert srv.stats()["resident_bytes"] == 0
    srv.close()


@pytest.mark.parametrize("method", [X.ME_DIA, X.ME_HEX])
def test_pattern_search_server_matches_oracle(method):
    """A server configured for the diamond or hexagon search: every frame's
    results equal the oracle's pattern search."""
    w, h, R, n_ref = 192, 128, 32, 2
    frames = synth.moving_sequence(w, h, 4, seed=13, local=8)
    pus = pu_tuples(w, h, 6)
    srv = X.Server(4 * GIB, max_inflight=2, search_range=R, n_ref=n_ref, me_method=method)
    sid = srv.open_stream(w, h, X.make_pus(pus))
    jobs = [srv.submit(sid, fr) for fr in frames]
    for f, j in enumerate(jobs):
        got, _ = srv.wait(j)
        exp = oracle_frame_results(frames, f, pus, n_ref, R, method)
        bad = np.nonzero(got != exp)[0]
        assert len(bad) == 0, f"frame {f}: {len(bad)} mismatches, first {bad[0]}: {got[bad[0]]} vs {exp[bad[0]]}"
    srv.close_stream(sid)
    srv.close()
    with pytest.raises(RuntimeError):
        X.Server(GIB, max_inflight=2, search_range=16, n_ref=2, me_method=X.ME_UMH)


def run_streams(specs, n_frames, interleave, max_inflight, budget=8 * GIB, R=32, **server_kw):
    """specs: list of (w, h, seed).  Returns {stream index: [results per frame]}."""
    srv = X.Server(budget, max_inflight=max_inflight, search_range=R, n_ref=2, **server_kw)
    seqs = [synth.moving_sequence(w, h, n_frames, seed=s) for (w, h, s) in specs]
    sids = [srv.open_stream(w, h, X.make_pus(pu_tuples(w, h, s))) for (w, h, s) in specs]
    out = {i: [None] * n_frames for i in range(len(specs))}
    if interleave:
        order = [(f, i) for f in range(n_frames) for i in range(len(specs))]
    else:
        order = [(f, i) for i in range(len(specs)) for f in range(n_frames)]
    jobs = [(i, f, srv.submit(sids[i], seqs[i][f])) for (f, i) in order]
    for i, f, j in jobs:
        out[i][f] = srv.wait(j)[0]
    srv.close()
    return out


def test_same_results_alone_and_in_the_mix():
    a = (320, 192, 21)
    others = [(256, 128, 22), (640, 360 - 360 % 8, 23), (128, 96, 24)]
    alone = run_streams([a], 4, interleave=False, max_inflight=1)[0]
    mixed = run_streams([a] + others, 4, interleave=True, max_inflight=4)[0]
    for f in range(4):
        assert np.array_equal(alone[f], mixed[f]), f"frame {f} differs between alone and mix"


def test_capped_and_uncapped_streams_keep_results():
    """A 50 % workgroup cap on the streams of large frames (search work per
    reference of at least 3e8 absdiffs at R = 32: 640x352 has 640 * 352 * 64^2
    = 9.2e8, 320x192 has 2.5e8), split launches, and the small streams on the
    whole device: every stream's results equal its results alone."""
    a = (320, 192, 21)
    others = [(256, 128, 22), (640, 360 - 360 % 8, 23), (128, 96, 24)]
    alone = run_streams([a, others[1]], 3, interleave=False, max_inflight=1)
    mixed = run_streams([a] + others, 3, interleave=True, max_inflight=4, workgroup_share_pct=50,
                        launch_split_absdiff=int(2e8), full_share_below_absdiff=int(3e8))
    for f in range(3):
        assert np.array_equal(alone[0][f], mixed[0][f]), f"small stream, frame {f}"
        assert np.array_equal(alone[1][f], mixed[2][f]), f"capped stream, frame {f}"
    with pytest.raises(X.XmeError):
        X.Server(GIB, max_inflight=2, search_range=16, n_ref=2, full_share_below_absdiff=-1)


def test_concurrent_submitters_and_waiters():
    specs = [(256, 192, 31), (384, 256, 32), (192, 128, 33)]
    srv = X.Server(8 * GIB, max_inflight=3, search_range=32, n_ref=2)
    seqs = [synth.moving_sequence(w, h, 4, seed=s) for (w, h, s) in specs]
    sids = [srv.open_stream(w, h, X.make_pus(pu_tuples(w, h, s))) for (w, h, s) in specs]
    got = {}
    errors = []

    def enc(i):
        try:
            js = [srv.submit(sids[i], fr) for fr in seqs[i]]
            got[i] = [srv.wait(j)[0] for j in js]
        except Exception as e:          # reported below
            errors.append(e)

    th = [threading.Thread(target=enc, args=(i,)) for i in range(len(specs))]
    for t in th:
        t.start()
    for t in th:
        t.join(120)
    assert not errors
    srv.close()
    ref = run_streams(specs, 4, interleave=False, max_inflight=1)
    for i in range(len(specs)):
        for f in range(4):
            assert np.array_equal(got[i][f], ref[i][f])


def test_memory_budget_admission_and_rejection():
    w, h = 256, 128
    pus = X.make_pus(pu_tuples(w, h, 41))
    frames = synth.moving_sequence(w, h, 6, seed=41)
    # find the stream's resident size with a roomy server
    srv = X.Server(4 * GIB, max_inflight=4, search_range=16, n_ref=2)
    sid = srv.open_stream(w, h, pus)
    resident = srv.stats()["resident_bytes"]
    srv.close()
    assert resident > 0
    # a budget below the resident size: the stream is refused
    srv = X.Server(resident - 1, max_inflight=4, search_range=16, n_ref=2)
    with pytest.raises(X.XmeError) as ei:
        srv.open_stream(w, h, pus)
    assert "(-4)" in str(ei.value)
    srv.close()
    # room for the stream but not for a frame: frames are rejected
    srv = X.Server(resident + 1024, max_inflight=4, search_range=16, n_ref=2)
    sid = srv.open_stream(w, h, pus)
    assert srv.submit(sid, frames[0]) is None
    assert srv.stats()["rejected"] == 1
    srv.close()
    # room for one frame at a time: everything runs, one frame in flight at most
    job_bytes = (((w + 63) // 64) * 64 * h + 255) // 256 * 256 + ((len(pus) * 2 * 24 + 255) // 256) * 256
    srv = X.Server(resident + job_bytes + 4096, max_inflight=4, search_range=16, n_ref=2)
    sid = srv.open_stream(w, h, pus)
    js = [srv.submit(sid, fr) for fr in frames]
    assert all(j is not None for j in js)
    outs = [srv.wait(j)[0] for j in js]
    st = srv.stats()
    assert st["completed"] == 6 and st["peak_bytes"] <= resident + job_bytes + 4096
    srv.close()
    ref = run_streams([(w, h, 41)], 6, interleave=False, max_inflight=1, R=16)[0]
    for f in range(6):
        assert np.array_equal(outs[f], ref[f])


def test_bad_arguments():
    srv = X.Server(GIB, max_inflight=2, search_range=16, n_ref=2)
    pus = X.make_pus([(0, 0, 16, 16, 0, 0)])
    with pytest.raises(X.XmeError):
        srv.open_stream(100, 64, pus)            # width not a multiple of 8
    with pytest.raises(X.XmeError):
        srv.submit(12345, np.zeros((64, 64), np.uint8))
    with pytest.raises(X.XmeError):
        srv.wait(999)
    sid = srv.open_stream(64, 64, pus)
    srv.close_stream(sid)
    with pytest.raises(X.XmeError):
        srv.close_stream(sid)
    srv.close()
