"""GPU parity of the frame-level ME job (integer full search + sub-pel
refinement) and of the interpolated reference planes, against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu

LAM = 4 * 65536


def oracle_ref(img, frac=2):
    h, w = img.shape
    p = O.Plane(w, h)
    p.set_picture(img)
    p.expand_border()
    return O.interp_planes(p, frac) if frac else [p] + [None] * 15


def oracle_cur(img):
    h, w = img.shape
    c = O.Plane(w, h)
    c.set_picture(img)
    return c


def device_cur(img):
    h, w = img.shape
    d = X.DevicePlane(w, h, pad=0, stride=((w + 63) // 64) * 64)
    d.upload_picture(img)
    return d


def oracle_results(prm, pus, cur, planes_by_ref, n_ref):
    out = np.zeros(len(pus) * n_ref, dtype=X.RESULT_DTYPE)
    arr = np.array([(p[0], p[1], p[2], p[3], p[4], p[5]) for p in pus], dtype=O.PU_DTYPE)
    cp = oracle_cur(cur)
    for r in range(n_ref):
        res = O.me_search_batch(prm, arr, cp, planes_by_ref[r], nthreads=8)
        for f in X.RESULT_DTYPE.names:
            out[f][r::n_ref] = res[f]
    return out


def compare(got, exp, pus, n_ref):
    bad = np.nonzero(got != exp)[0]
    if len(bad):
        i = bad[0]
        raise AssertionError(f"{len(bad)} mismatches; first inst {i} pu {pus[i // n_ref]} ref {i % n_ref}: "
                             f"gpu {got[i]} oracle {exp[i]}")


@pytest.mark.parametrize("w,h", [(64, 48), (200, 120), (1920 // 4, 1080 // 4)])
def test_interpolated_planes_match_oracle(w, h):
    img = synth.smooth_field(w, h, seed=w + h)
    img[0, :] = 255
    img[:, -1] = 0
    planes = oracle_ref(img, 2)
    dref = X.DeviceRef(w, h, frac=2)
    dref.load(img)
    pad = X.XME_PAD
    got0 = dref.planes[0].download_padded()
    assert np.array_equal(got0, planes[0].view()[: h + 2 * pad, : w + 2 * pad])
    for k in range(1, 16):
        got = dref.planes[k].download_padded()
        exp = planes[k].view()[: h + 2 * pad, : w + 2 * pad]
        assert np.array_equal(got, exp), f"plane {k}"


def test_half_pel_planes_only():
    w, h = 96, 64
    img = synth.smooth_field(w, h, seed=3)
    planes = oracle_ref(img, 1)
    dref = X.DeviceRef(w, h, frac=1)
    dref.load(img)
    for k in (2, 8, 10):
        assert np.array_equal(dref.planes[k].download_padded(), planes[k].view()[: h + 160, : w + 160])


def run_case(w, h, pus, R, n_ref=1, mode=X.PLAN_BATCHED, seed=0, prm_kw=None, cost_map=False):
    prm_kw = prm_kw or {}
    refs = [synth.smooth_field(w, h, seed=seed * 10 + r) for r in range(n_ref)]
    cur = synth.multi_ref_current(refs, 16, min(R, 24), seed=seed + 99)
    frac = prm_kw.get("use_fractional_me", 2)
    o_planes = [oracle_ref(r, frac) for r in refs]
    d_refs = []
    for r in refs:
        d = X.DeviceRef(w, h, frac=frac)
        d.load(r)
        d_refs.append(d)
    dcur = device_cur(cur)
    prm_o = O.me_param(w, h, R, LAM, **prm_kw)
    prm_d = X.me_params(w, h, R, LAM, **prm_kw)
    plan = X.Plan(prm_d, X.make_pus(pus), n_ref, mode)
    cm = X.DeviceBuffer(len(pus) * n_ref * 4 * R * R * 4) if cost_map else None
    plan.run(dcur, d_refs, cost_map=cm)
    got = plan.fetch()
    exp = oracle_results(prm_o, pus, cur, o_planes, n_ref)
    compare(got, exp, pus, n_ref)
    if cost_map:
        maps = cm.download(len(pus) * n_ref * 4 * R * R, np.int32).reshape(len(pus) * n_ref, 2 * R, 2 * R)
        for i in range(0, len(pus) * n_ref, max(1, len(pus) * n_ref // 7)):
            p = pus[i // n_ref]
            _, exp_map = O.me_search(prm_o, tuple(p), oracle_cur(cur), o_planes[i % n_ref], cost_map=True)
            assert np.array_equal(maps[i], exp_map), f"cost map of inst {i}"
    return plan, dcur, d_refs, exp


@pytest.mark.parametrize("part", [1, 2, 3, 4, 19, 20, 23, 24])
def test_fs_subpel_every_config1_shape(part):
    pw, ph = X.PU_W[part], X.PU_H[part]
    w, h = 256, 192
    pus = [(x, y, pw, ph, 0, 0) for (x, y, _, _) in synth.shape_tiling(w, h, pw, ph)]
    run_case(w, h, pus, 16, seed=part)


def test_cost_maps_match_every_candidate():
    w, h = 192, 128
    pus = [(x, y, 16, 16, 0, 0) for (x, y, _, _) in synth.shape_tiling(w, h, 16, 16)]
    pus += [(0, 0, 8, 8, 0, 0), (184, 120, 8, 8, 0, 0), (64, 64, 32, 8, 0, 0)]
    run_case(w, h, pus, 32, seed=5, cost_map=True, prm_kw=dict(use_fractional_me=0))


def test_all_25_shapes_with_mvps_and_two_refs():
    w, h = 256, 160
    rng = np.random.default_rng(1)
    pus = []
    for part in range(25):
        pw, ph = X.PU_W[part], X.PU_H[part]
        for _ in range(3):
            x = int(rng.integers(0, (w - pw) // 4 + 1)) * 4
            y = int(rng.integers(0, h - ph + 1))
            pus.append((x, y, pw, ph, int(rng.integers(-60, 61)), int(rng.integers(-60, 61))))
    run_case(w, h, pus, 12, n_ref=2, seed=7)


@pytest.mark.parametrize("kw", [dict(enable_pmvr=1), dict(use_fast_sub_me=1), dict(use_hadamard=0),
                                dict(use_fractional_me=1), dict(use_fractional_me=0)])
def test_subpel_switches(kw):
    w, h = 192, 128
    pus = [(x, y, 16, 16, 6, -5) for (x, y, _, _) in synth.shape_tiling(w, h, 16, 16)]
    run_case(w, h, pus, 16, seed=11, prm_kw=kw)


@pytest.mark.parametrize("mvp,bcost2,pmvr,exp", [((-237, -242), 64, 0, (-234, -235, 48)),
                                                   ((-237, -242), 64, 1, (-236, -236, 40)),
                                                   ((-237, -236), 56, 0, (-234, -235, 32)),
                                                   ((-237, -236), 56, 1, (-234, -236, 24))])
def test_pmvr_flat_frame_known_answers(mvp, bcost2, pmvr, exp):
    """The hand-worked flat-frame cases of the oracle's PMVR test, on the GPU: every SAD is
    0, so the costs are 4 * MV bits and the answers are fixed literals."""
    w, h = 128, 96
    flat = np.full((h, w), 77, np.uint8)
    dref = X.DeviceRef(w, h, frac=2)
    dref.load(flat)
    plan = X.Plan(X.me_params(w, h, 8, LAM, use_hadamard=0, enable_pmvr=pmvr), X.make_pus([(0, 0, 16, 16) + mvp]), 1)
    plan.run(device_cur(flat), [dref])
    r = plan.fetch()[0]
    assert (r["status"], r["bmv2_x"], r["bmv2_y"], r["bcost2"]) == (0, -58, -58, bcost2)
    assert (r["bmv_x"], r["bmv_y"], r["cost"], r["mvcost"]) == exp + (exp[2],)


def test_frame_edges_range64():
    w, h = 320, 192
    pus = [(0, 0, 16, 16, 0, 0), (304, 176, 16, 16, 0, 0), (0, 128, 64, 64, 0, 0), (256, 0, 64, 64, 0, 0),
           (152, 96, 8, 8, 0, 0), (256, 128, 64, 16, 0, 0), (0, 64, 16, 64, 0, 0)]
    run_case(w, h, pus, 64, seed=13)


def test_plain_equals_batched_and_graph_replays():
    w, h = 256, 192
    pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, 3)]
    plan, dcur, drefs, exp = run_case(w, h, pus, 16, n_ref=2, seed=17)
    prm_d = X.me_params(w, h, 16, LAM)
    plain = X.Plan(prm_d, X.make_pus(pus), 2, X.PLAN_PLAIN)
    plain.run(dcur, drefs)
    assert np.array_equal(plain.fetch(), exp)
    s = X.Stream()
    for rep in range(4):
        plan.results.zero(s.h)
        plan.run_graph(dcur, drefs, s.h)
        s.sync()
        assert np.array_equal(plan.fetch(), exp), f"replay {rep}"
        plain.run(dcur, drefs, stream=s.h)   # eager work between replays
        s.sync()
        assert np.array_equal(plain.fetch(), exp), f"eager run after replay {rep}"


def test_split_launches_and_workgroup_cap_keep_results():
    """The search split into many small launches, and with a capped grid, gives
    the same results, also over graph replays (the captured graph is rebuilt)."""
    w, h = 256, 192
    pus = [(x, y, pw, ph, int(v) % 9 - 4, int(v) % 7 - 3)
           for v, (x, y, pw, ph) in enumerate(synth.mixed_pus(w, h, 5))]
    plan, dcur, drefs, exp = run_case(w, h, pus, 16, n_ref=2, seed=23)
    s = X.Stream()
    for split, cap in ((2_000_000, 0), (50_000_000, 7), (0, 0)):
        X.check(X.lib().xme_plan_set_launch_split(plan.h, split), "split")
        X.check(X.lib().xme_plan_set_max_workgroups(plan.h, cap), "cap")
        for rep in range(3):
            plan.results.zero(s.h)
            plan.run_graph(dcur, drefs, s.h)
            s.sync()
            assert np.array_equal(plan.fetch(), exp), (split, cap, rep)


def test_spread_search_centres_split_windows():
    """A random MVP per PU spreads the search centres of a CTU, so its window
    is split into candidate-row ranges (8 lanes per candidate row: bands of
    8 x 16 = 128 candidate rows, the tile shape where splitting matters).
    The split job equals the oracle; its speed is measured by the bench, not
    asserted here."""
    w, h, R = 640, 384, 64
    ref = synth.smooth_field(w, h, seed=31)
    cur, _ = synth.displaced(ref, 16, 16, seed=32)
    dref = X.DeviceRef(w, h, frac=2)
    dref.load(ref)
    dcur = device_cur(cur)
    rng = np.random.default_rng(33)
    pus = [(x, y, 8, 8) + tuple(int(v) for v in rng.integers(-160, 161, 2))
           for (x, y, _, _) in synth.shape_tiling(w, h, 8, 8)]
    saved = X.tune_get(1, R)
    X.tune_set(1, R, 16, 8)
    try:
        plan = X.Plan(X.me_params(w, h, R, LAM, use_fractional_me=0), X.make_pus(pus), 1)
        plan.run(dcur, [dref])
        got = plan.fetch()
    finally:
        X.tune_set(1, R, *saved)
    exp = oracle_results(O.me_param(w, h, R, LAM, use_fractional_me=0), pus, cur, [oracle_ref(ref, 0)], 1)
    compare(got, exp, pus, 1)


def test_no_legal_candidate_status():
    w, h = 64, 64
    pus = [(16, 16, 16, 16, 0, 0)]
    run_case(w, h, pus, 4, seed=2, prm_kw=dict(mv_range_y=(40, 41)))
