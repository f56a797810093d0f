"""GPU parity of the DIA, HEX, UMH and TZ integer searches (followed by the
sub-pel refinement) against the oracle: every PU shape, MV predictors, several
references, MV windows narrowed so that walks reach their edges, graph
replays, a 1080p frame of mixed shapes with a random MVP per PU, and for TZ
motion up to 48 pixels (rings to 64, the raster step offset by the diamond's
centre, the while loop) and picture-edge PUs whose raster reads leave the
padded reference (status 2)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_me import LAM, compare, device_cur, oracle_results, run_case
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu

METHODS = [pytest.param(X.ME_DIA, id="dia"), pytest.param(X.ME_HEX, id="hex"), pytest.param(X.ME_TZ, id="tz")]


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
        X.Plan(X.me_params(64, 64, 8, LAM, me_method=5), X.make_pus([(16, 16, 16, 16, 0, 0)]), 1)


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


# ---------------------------------------------------------------------------
# UMH: the same kernels with the per-(PU, reference) encoder context

def random_umh_ctx(n_pu, n_ref, seed, zero_frac=0.3):
    """Seeded contexts: predictor MVs (some (0,0)), ref_idx = the reference
    index, predicted SADs and betas spread so that every early-termination
    route is taken somewhere."""
    rng = np.random.default_rng(seed)
    n = n_pu * n_ref
    c = np.zeros(n, dtype=X.UMH_CTX_DTYPE)
    for f in ("mvp1", "mvp2", "mvp3"):
        for a in ("_x", "_y"):
            c[f + a] = rng.integers(-96, 97, n)
    for f in ("mvp1", "mvp3"):
        z = rng.random(n) < zero_frac
        c[f + "_x"][z] = 0
        c[f + "_y"][z] = 0
    c["ref_idx"] = np.arange(n) % n_ref
    c["pred_sad"] = rng.integers(0, 6000, n)
    c["pred_sad_uplayer"] = rng.integers(0, 6000, n)
    c["beta2"] = rng.uniform(-1.0, 0.6, n)
    c["beta3"] = rng.uniform(-1.0, 0.2, n)
    return c


def run_umh_case(w, h, pus, R, n_ref, seed, level, prm_kw=None, ctx=None, d_states=None):
    """The UMH plan against the oracle, the plan's context set from ctx (host)
    or, when given, from device-resident xme_umh_state records."""
    prm_kw = dict(prm_kw or {})
    refs = [synth.smooth_field(w, h, seed=seed * 10 + r) for r in range(n_ref)]
    cur = synth.multi_ref_current(refs, 16, min(R, 24), seed=seed + 99)
    d_refs = []
    for r in refs:
        d = X.DeviceRef(w, h, frac=2)
        d.load(r)
        d_refs.append(d)
    dcur = device_cur(cur)
    ctx = random_umh_ctx(len(pus), n_ref, seed + 5) if ctx is None else ctx
    plan = X.Plan(X.me_params(w, h, R, LAM, me_method=X.ME_UMH, umh_big_hex_level=level, **prm_kw),
                  X.make_pus(pus), n_ref)
    if d_states is None:
        plan.set_umh_context(ctx)
    else:
        plan.set_umh_states(d_states)
    plan.run(dcur, d_refs)
    got = plan.fetch()
    prm_o = O.me_param(w, h, R, LAM, me_method=O.ME_UMH, umh_big_hex_level=level, **prm_kw)
    arr = np.array([tuple(p) for p in pus], dtype=O.PU_DTYPE)
    cp = O.Plane(w, h)
    cp.set_picture(cur)
    exp = np.zeros(len(pus) * n_ref, dtype=X.RESULT_DTYPE)
    for r in range(n_ref):
        p = O.Plane(w, h)
        p.set_picture(refs[r])
        p.expand_border()
        res = O.me_search_batch(prm_o, arr, cp, O.interp_planes(p, 2), nthreads=8,
                                ctxs=np.ascontiguousarray(ctx[r::n_ref]).view(O.UMH_CTX_DTYPE))
        for f in X.RESULT_DTYPE.names:
            exp[f][r::n_ref] = res[f]
    compare(got, exp, pus, n_ref)
    return plan, dcur, d_refs, exp


@pytest.mark.parametrize("level", [0, 1, 2])
def test_umh_all_25_shapes_two_refs(level):
    w, h = 256, 160
    rng = np.random.default_rng(61 + level)
    pus = []
    for part in range(25):
        pw, ph = X.PU_W[part], X.PU_H[part]
        for _ in range(4):
            x = int(rng.integers(0, (w - pw) // 4 + 1)) * 4
            y = int(rng.integers(0, h - ph + 1))
            pus.append((x, y, pw, ph, int(rng.integers(-60, 61)), int(rng.integers(-60, 61))))
    run_umh_case(w, h, pus, 16, 2, 67 + level, level)


def test_umh_no_early_termination_and_switches():
    """betas of -1 (bound 0): every instance runs the cross, the grid, the
    big hexagons, the hexagon and the diamond; sub-pel switches on top."""
    w, h = 192, 128
    pus = [(x, y, pw, ph, 6, -3) for (x, y, pw, ph) in synth.mixed_pus(w, h, 11)]
    ctx = random_umh_ctx(len(pus), 1, 71)
    ctx["beta2"] = -1.0
    ctx["beta3"] = -1.0
    for kw in (dict(), dict(use_fractional_me=0), dict(use_hadamard=0), dict(enable_pmvr=1)):
        run_umh_case(w, h, pus, 32, 1, 73, 2, kw, ctx)


def test_umh_narrow_mv_window_and_edges():
    w, h = 192, 128
    pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, 13)]
    pus += [(0, 0, 16, 16, -40, -40), (176, 112, 16, 16, 40, 40), (0, 64, 64, 64, 0, 0)]
    run_umh_case(w, h, pus, 16, 2, 79, 2, dict(mv_range_x=(-40, 40), mv_range_y=(-40, 40)))


def test_umh_graph_replays_and_context_updates():
    w, h = 256, 192
    pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, 17)]
    plan, dcur, drefs, exp = run_umh_case(w, h, pus, 32, 2, 83, 1)
    s = X.Stream()
    for rep in range(3):
        plan.results.zero(s.h)
        plan.run_graph(dcur, drefs, s.h)
        s.sync()
        assert np.array_equal(plan.fetch(), exp), f"replay {rep}"
    # a new context changes the results of the replayed graph like a plain run
    ctx2 = random_umh_ctx(len(pus), 2, 89)
    plan.set_umh_context(ctx2, s.h)
    plan.run_graph(dcur, drefs, s.h)
    s.sync()
    got_graph = plan.fetch()
    plan.run(dcur, drefs)
    assert np.array_equal(plan.fetch(), got_graph)


def test_umh_needs_its_context():
    prm = X.me_params(64, 64, 8, LAM, me_method=X.ME_UMH)
    plan = X.Plan(prm, X.make_pus([(16, 16, 16, 16, 0, 0)]), 1)
    ref = X.DeviceRef(64, 64, frac=2)
    ref.load(synth.smooth_field(64, 64, seed=1))
    cur = device_cur(synth.smooth_field(64, 64, seed=2))
    with pytest.raises(RuntimeError):
        plan.run(cur, [ref])
    with pytest.raises(RuntimeError):
        X.Plan(X.me_params(64, 64, 8, LAM, me_method=X.ME_UMH, umh_big_hex_level=3),
               X.make_pus([(16, 16, 16, 16, 0, 0)]), 1)


def test_umh_random_mvps_mixed_shapes_1080p_range64():
    W, H = 1920, 1080
    ref = synth.smooth_field(W, H, seed=600)
    cur, _ = synth.displaced(ref, 16, 24, seed=601)
    rng = np.random.default_rng(602)
    pus = [(x, y, pw, ph, int(rng.integers(-64, 65)), int(rng.integers(-64, 65)))
           for (x, y, pw, ph) in synth.mixed_pus(W, 1024, 602)]
    ctx = random_umh_ctx(len(pus), 1, 603)
    dref = X.DeviceRef(W, H, frac=2)
    dref.load(ref)
    dcur = device_cur(cur)
    plan = X.Plan(X.me_params(W, H, 64, LAM, me_method=X.ME_UMH, umh_big_hex_level=2), X.make_pus(pus), 1)
    plan.set_umh_context(ctx)
    plan.run(dcur, [dref])
    got = plan.fetch()
    p = O.Plane(W, H)
    p.set_picture(ref)
    p.expand_border()
    cp = O.Plane(W, H)
    cp.set_picture(cur)
    exp = O.me_search_batch(O.me_param(W, H, 64, LAM, me_method=O.ME_UMH, umh_big_hex_level=2),
                            np.array([tuple(q) for q in pus], dtype=O.PU_DTYPE), cp, O.interp_planes(p, 2),
                            nthreads=16, ctxs=ctx.view(O.UMH_CTX_DTYPE))
    compare(got, exp.astype(X.RESULT_DTYPE), pus, 1)


@pytest.mark.parametrize("method", [X.ME_DIA, X.ME_HEX, X.ME_UMH])
def test_no_legal_candidate_status(method):
    """An MV range that leaves no legal candidate (me.c:919-921): status 1 and
    XME_MAX_DISTORTION, as the oracle returns, for every pattern search."""
    w, h = 64, 64
    pus = [(16, 16, 16, 16, 0, 0), (32, 8, 8, 8, 0, 0)]
    kw = dict(mv_range_y=(40, 41))
    if method == X.ME_UMH:
        ctx = random_umh_ctx(len(pus), 1, 97)
        run_umh_case(w, h, pus, 4, 1, 2, 2, kw, ctx)
    else:
        run_case(w, h, pus, 4, seed=2, prm_kw=dict(me_method=method, **kw))


def umh_single(w, h, cur, ref, pu, R, **prm_kw):
    """One PU's UMH search (no early termination, integer only) on the GPU and in the oracle."""
    ctx = np.zeros(1, dtype=X.UMH_CTX_DTYPE)
    ctx["beta2"] = -1.0
    ctx["beta3"] = -1.0
    dref = X.DeviceRef(w, h, frac=0)
    dref.load(ref)
    plan = X.Plan(X.me_params(w, h, R, LAM, me_method=X.ME_UMH, use_fractional_me=0, **prm_kw), X.make_pus([pu]), 1)
    plan.set_umh_context(ctx)
    plan.run(device_cur(cur), [dref])
    got = plan.fetch()
    p = O.Plane(w, h)
    p.set_picture(ref)
    p.expand_border()
    cp = O.Plane(w, h)
    cp.set_picture(cur)
    exp = O.me_search_batch(O.me_param(w, h, R, LAM, me_method=O.ME_UMH, use_fractional_me=0, **prm_kw),
                            np.array([pu], dtype=O.PU_DTYPE), cp, [p] + [None] * 15,
                            ctxs=ctx.view(O.UMH_CTX_DTYPE))
    compare(got, exp, [pu], 1)
    return int(got["bmv2_x"][0]), int(got["bmv2_y"][0]), int(got["bcost2"][0])


def test_umh_diamond_unchecked_at_the_window_edge():
    """The oracle's hand-worked case: on a flat frame the first diamond around the clipped
    MVP candidate (-58, 0) takes (-59, 0), outside the window (cost 16)."""
    flat = np.full((96, 128), 60, np.uint8)
    assert umh_single(128, 96, flat, flat, (0, 0, 16, 16, -237, 0), 8) == (-59, 0, 16)


def test_umh_diamond_absolute_points_when_the_window_misses_an_offset():
    """The oracle's hand-worked case: with the window ending at 0 on both axes the diamond
    evaluates the absolute MV (-1, 0), where the current block matches exactly (cost 112)."""
    w, h = 128, 96
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 256, (h, w), dtype=np.uint8)
    blk = ref[32:48, 47:63].copy()
    ref[12:28, 38:54] = np.clip(blk.astype(int) + rng.choice([-3, 3], blk.shape), 0, 255)
    cur = rng.integers(0, 256, (h, w), dtype=np.uint8)
    cur[32:48, 48:64] = blk
    assert umh_single(w, h, cur, ref, (48, 32, 16, 16, -40, -80), 8, mv_range_x=(-8192, 24),
                      mv_range_y=(-2048, 24)) == (-1, 0, 112)


# ---------------------------------------------------------------------------
# TZ: long walks, the offset raster step and reads outside the reference

def tz_case(seed, pus, W=256, H=192, R=64, mv=48, n_ref=1):
    """The TZ plan against the oracle on a smooth picture whose 16x16 tiles move by
    up to `mv` pixels; returns the oracle's results."""
    refs = [synth.smooth_field(W, H, seed=seed + 7 * r) for r in range(n_ref)]
    cur, _ = synth.displaced(refs[0], 16, mv, seed=seed + 1)
    d_refs = []
    for r in refs:
        d = X.DeviceRef(W, H, frac=2)
        d.load(r)
        d_refs.append(d)
    dcur = device_cur(cur)
    plan = X.Plan(X.me_params(W, H, R, LAM, me_method=X.ME_TZ), X.make_pus(pus), n_ref)
    plan.run(dcur, d_refs)
    got = plan.fetch()
    planes = []
    for r in refs:
        p = O.Plane(W, H)
        p.set_picture(r)
        p.expand_border()
        planes.append(O.interp_planes(p, 2))
    exp = oracle_results(O.me_param(W, H, R, LAM, me_method=O.ME_TZ), pus, cur, planes, n_ref)
    compare(got, exp, pus, n_ref)
    return exp


@pytest.mark.parametrize("seed", [0, 4])
def test_tz_long_walks_raster_and_while_loop(seed):
    """Motion up to 48 pixels over a mixed tiling, two references: about a quarter
    of the PUs take the raster step (offset by the diamond's centre) and half run
    the while loop's pattern searches (counted on the oracle's routes)."""
    W, H = 256, 192
    pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(W, H, seed)]
    tz_case(seed, pus, n_ref=2)
    # route coverage of the same inputs, reference 0
    ref = synth.smooth_field(W, H, seed=seed)
    cur, _ = synth.displaced(ref, 16, 48, seed=seed + 1)
    rp = O.Plane(W, H)
    rp.set_picture(ref)
    rp.expand_border()
    cp = O.Plane(W, H)
    cp.set_picture(cur)
    prm = O.me_param(W, H, 64, LAM, me_method=O.ME_TZ, use_fractional_me=0)
    res = O.me_search_batch(prm, np.array([p for p in pus], dtype=O.PU_DTYPE), cp, [rp] + [None] * 15)
    raster = loops = 0
    pad = rp.view()
    for i, pu in enumerate(pus):
        x, y, pw, ph = pu[:4]
        # the start: the only MV candidate (0,0), SAD + MV cost (mvbits(0) = 1 per component)
        blk = pad[O.PAD + y:O.PAD + y + ph, O.PAD + x:O.PAD + x + pw].astype(int)
        c0 = int(np.abs(blk - cur[y:y + ph, x:x + pw].astype(int)).sum()) + ((LAM * 2) >> 16)
        _, _, _, route = O.pattern_trace(prm, pu, cp, rp, (0, 0), c0)
        raster += (route & 6) == 6
        loops += (route >> 8) >= 2
    assert raster >= len(pus) // 8 and loops >= len(pus) // 3, (raster, loops, len(pus))
    assert (res["status"] == 0).all()


def test_tz_reads_outside_the_reference_are_status_2():
    """PUs along the picture's edges with MVPs of 30-57 pixels pointing out of it:
    their diamond centres lie far out, and some offset raster reads leave the
    padded reference; GPU and oracle skip the same points and both give status 2
    there, equal results elsewhere."""
    W, H = 256, 192
    n2 = 0
    for seed in (1, 4):
        rng = np.random.default_rng(seed)
        pus = []
        for y in range(0, H - 16, 16):
            pus.append((0, y, 16, 16, int(rng.integers(-230, -120)), int(rng.integers(-60, 60))))
            pus.append((W - 16, y, 16, 16, int(rng.integers(120, 230)), int(rng.integers(-60, 60))))
        for x in range(0, W - 16, 16):
            pus.append((x, 0, 16, 16, int(rng.integers(-60, 60)), int(rng.integers(-230, -120))))
        exp = tz_case(seed, pus)
        n2 += int((exp["status"] == 2).sum())
    assert n2 >= 2
