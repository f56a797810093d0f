"""GPU parity of the DIA and HEX integer searches (followed by the sub-pel
refinement) against the oracle: every PU shape, MV predictors, several
references, MV windows narrowed so that walks reach their edges, graph
replays, and a 1080p frame of mixed shapes with a random MVP per PU."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_me import LAM, compare, device_cur, oracle_results, run_case
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu

METHODS = [pytest.param(X.ME_DIA, id="dia"), pytest.param(X.ME_HEX, id="hex")]


@pytest.mark.parametrize("method", METHODS)
def test_all_25_shapes_with_mvps_two_refs(method):
    w, h = 256, 160
    rng = np.random.default_rng(41)
    pus = []
    for part in range(25):
        pw, ph = X.PU_W[part], X.PU_H[part]
        for _ in range(4):
            x = int(rng.integers(0, (w - pw) // 4 + 1)) * 4
            y = int(rng.integers(0, h - ph + 1))
            pus.append((x, y, pw, ph, int(rng.integers(-60, 61)), int(rng.integers(-60, 61))))
    run_case(w, h, pus, 16, n_ref=2, seed=43, prm_kw=dict(me_method=method))


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("kw", [dict(use_fractional_me=0), dict(use_hadamard=0), dict(enable_pmvr=1)])
def test_switches(method, kw):
    w, h = 192, 128
    pus = [(x, y, pw, ph, 5, -3) for (x, y, pw, ph) in synth.mixed_pus(w, h, 7)]
    run_case(w, h, pus, 16, seed=47, prm_kw=dict(me_method=method, **kw))


@pytest.mark.parametrize("method", METHODS)
def test_walks_reach_a_narrow_mv_window(method):
    """MV range +-10 qpel: windows of [-4, 4] full-pel points, while the
    current picture moves by up to 16 pixels; picture-corner PUs too."""
    w, h = 192, 128
    pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, 9)]
    pus += [(0, 0, 16, 16, -40, -40), (176, 112, 16, 16, 40, 40), (0, 64, 64, 64, 0, 0)]
    run_case(w, h, pus, 16, seed=53,
             prm_kw=dict(me_method=method, mv_range_x=(-40, 40), mv_range_y=(-40, 40)))


@pytest.mark.parametrize("method", METHODS)
def test_plain_equals_batched_graph_replays_and_counts(method):
    w, h = 256, 192
    pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, 3)]
    plan, dcur, drefs, exp = run_case(w, h, pus, 32, n_ref=2, seed=59, prm_kw=dict(me_method=method))
    prm_d = X.me_params(w, h, 32, LAM, me_method=method)
    assert plan.info()[1:] == (0, 0)
    one = plan.search_count()
    assert one[0] > 0 and one[1] > one[0]
    plain = X.Plan(prm_d, X.make_pus(pus), 2, X.PLAN_PLAIN)
    plain.run(dcur, drefs)
    assert np.array_equal(plain.fetch(), exp)
    s = X.Stream()
    for rep in range(3):
        plan.results.zero(s.h)
        plan.run_graph(dcur, drefs, s.h)
        s.sync()
        assert np.array_equal(plan.fetch(), exp), f"replay {rep}"
    # the same points are evaluated on every run
    assert plan.search_count() == one


def test_cost_map_and_unknown_method_are_rejected():
    prm = X.me_params(64, 64, 8, LAM, me_method=X.ME_HEX)
    plan = X.Plan(prm, X.make_pus([(16, 16, 16, 16, 0, 0)]), 1)
    ref = X.DeviceRef(64, 64, frac=2)
    ref.load(synth.smooth_field(64, 64, seed=1))
    cur = device_cur(synth.smooth_field(64, 64, seed=2))
    cm = X.DeviceBuffer(4 * 16 * 16)
    with pytest.raises(RuntimeError):
        plan.run(cur, [ref], cost_map=cm)
    with pytest.raises(RuntimeError):
        X.Plan(X.me_params(64, 64, 8, LAM, me_method=3), X.make_pus([(16, 16, 16, 16, 0, 0)]), 1)


@pytest.mark.parametrize("method", METHODS)
def test_random_mvps_mixed_shapes_1080p_range64(method):
    W, H = 1920, 1080
    ref = synth.smooth_field(W, H, seed=500)
    cur, _ = synth.displaced(ref, 16, 24, seed=501)
    rng = np.random.default_rng(502)
    pus = [(x, y, pw, ph, int(rng.integers(-64, 65)), int(rng.integers(-64, 65)))
           for (x, y, pw, ph) in synth.mixed_pus(W, 1024, 502)]
    dref = X.DeviceRef(W, H, frac=2)
    dref.load(ref)
    dcur = device_cur(cur)
    plan = X.Plan(X.me_params(W, H, 64, LAM, me_method=method), X.make_pus(pus), 1)
    plan.run(dcur, [dref])
    got = plan.fetch()
    p = O.Plane(W, H)
    p.set_picture(ref)
    p.expand_border()
    exp = oracle_results(O.me_param(W, H, 64, LAM, me_method=method), pus, cur, [O.interp_planes(p, 2)], 1)
    compare(got, exp, pus, 1)
