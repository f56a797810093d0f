"""GPU parity at the two largest workloads, against the oracle:

* batched multi-reference ME at 3840x2160: 4 references, the mixed PU tiling
  (8x8 to 64x64 with the asymmetric shapes), full search [-64, 64)^2 and
  quarter-pel SATD refinement, run through the captured HIP graph and
  replayed, every PU of every reference compared;
* the serving loop with 16 concurrent encodes (6 x 1280x720, 6 x 1920x1080,
  4 x 3840x2160), 2 references, R = 64, 3 frames each, submitted interleaved:
  one stream of each resolution gives the same results alone as in the mix,
  and the mix's results equal the oracle's (whole frames at 720p and 1080p,
  the first CTU row of a 2160p frame)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_me import LAM, device_cur
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def oracle_refs(refs):
    out = []
    for r in refs:
        h, w = r.shape
        p = O.Plane(w, h)
        p.set_picture(r)
        p.expand_border()
        out.append(O.interp_planes(p, 2))
    return out


def oracle_search(w, h, R, pus, cur, planes_by_ref):
    """Results PU-major, one entry per reference in the given order."""
    prm = O.me_param(w, h, R, LAM)
    arr = np.array([(p[0], p[1], p[2], p[3], p[4], p[5]) for p in pus], dtype=O.PU_DTYPE)
    cp = O.Plane(w, h)
    cp.set_picture(cur)
    n_ref = len(planes_by_ref)
    out = np.zeros(len(pus) * n_ref, dtype=X.RESULT_DTYPE)
    for r, planes in enumerate(planes_by_ref):
        res = O.me_search_batch(prm, arr, cp, planes, nthreads=16)
        for f in X.RESULT_DTYPE.names:
            out[f][r::n_ref] = res[f]
    return out


def first_mismatch(got, exp):
    bad = np.nonzero(got != exp)[0]
    return f"{len(bad)} mismatches, first {bad[0]}: gpu {got[bad[0]]} oracle {exp[bad[0]]}" if len(bad) else ""


def test_2160p_four_refs_range64_graph_replays():
    w, h, R, n_ref = 3840, 2160, 64, 4
    refs = [synth.smooth_field(w, h, seed=501 + 101 * r) for r in range(n_ref)]
    cur = synth.multi_ref_current(refs, 32, 48, seed=508)
    pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, seed=504)]
    area = {}
    for (_, _, pw, ph, _, _) in pus:
        k = "small" if pw * ph <= 256 else ("large" if pw * ph == 4096 else "mid")
        area[k] = area.get(k, 0) + pw * ph
    assert area["small"] > area["mid"] > area["large"] > 0          # the 60 / 30 / 10 % area mix
    d_refs = []
    for r in refs:
        d = X.DeviceRef(w, h, frac=2)
        d.load(r)
        d_refs.append(d)
    dcur = device_cur(cur)
    plan = X.Plan(X.me_params(w, h, R, LAM), X.make_pus(pus), n_ref, X.PLAN_BATCHED)
    s = X.Stream()
    print(f"2160p: {len(pus)} PUs x {n_ref} refs, graph run 1", flush=True)
    plan.run_graph(dcur, d_refs, s.h)
    s.sync()
    got = plan.fetch()
    print("2160p: oracle", flush=True)
    exp = oracle_search(w, h, R, pus, cur, oracle_refs(refs))
    assert not first_mismatch(got, exp), first_mismatch(got, exp)
    for rep in range(2):
        plan.results.zero(s.h)
        plan.run_graph(dcur, d_refs, s.h)
        s.sync()
        assert np.array_equal(plan.fetch(), exp), f"replay {rep}"


STREAMS = [(1280, 720)] * 6 + [(1920, 1080)] * 6 + [(3840, 2160)] * 4


def serve(specs, seqs, pus, n_frames, slots):
    srv = X.Server(64 * GIB, max_inflight=slots, search_range=64, n_ref=2)
    sids = [srv.open_stream(w, h, X.make_pus(pus[i])) for i, (w, h) in enumerate(specs)]
    jobs = [(i, f, srv.submit(sids[i], seqs[i][f])) for f in range(n_frames) for i in range(len(specs))]
    out = {i: [None] * n_frames for i in range(len(specs))}
    for i, f, j in jobs:
        out[i][f] = srv.wait(j)[0]
    st = srv.stats()
    srv.close()
    return out, st


def test_sixteen_streams_alone_mix_and_oracle():
    n_frames = 3
    seqs = [synth.moving_sequence(w, h, n_frames, seed=600 + 10 * i) for i, (w, h) in enumerate(STREAMS)]
    pus = [[(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, 601 + 10 * i)]
           for i, (w, h) in enumerate(STREAMS)]
    print("16 streams in the mix", flush=True)
    mix, st = serve(STREAMS, seqs, pus, n_frames, slots=8)
    assert st["completed"] == 16 * n_frames and st["rejected"] == 0
    for i in range(16):
        assert len(mix[i][0]) == 0 and len(mix[i][1]) == len(pus[i]) and len(mix[i][2]) == 2 * len(pus[i])
    for i in (0, 6, 12):
        print(f"stream {i} alone", flush=True)
        alone, _ = serve([STREAMS[i]], [seqs[i]], [pus[i]], n_frames, slots=1)
        for f in range(n_frames):
            assert np.array_equal(alone[0][f], mix[i][f]), f"stream {i} frame {f}: alone and mix differ"
    # the oracle: frame 2 (references frames 1 and 0) of a 720p stream, frame 1 of a 1080p
    # stream, and the first CTU row of frame 2 of a 2160p stream
    for i, f, rows in ((1, 2, None), (7, 1, None), (13, 2, 1)):
        w, h = STREAMS[i]
        sel = [k for k, p in enumerate(pus[i]) if rows is None or p[1] < 64 * rows]
        print(f"oracle: stream {i} ({w}x{h}) frame {f}, {len(sel)} PUs", flush=True)
        exp = oracle_search(w, h, 64, [pus[i][k] for k in sel], seqs[i][f],
                            oracle_refs([seqs[i][f - 1 - r] for r in range(min(f, 2))]))
        k_ref = min(f, 2)
        got = mix[i][f].reshape(len(pus[i]), k_ref)[sel].reshape(-1)
        assert not first_mismatch(got, exp), first_mismatch(got, exp)
