"""Pins the oracle's DIA, HEX (encoder/me.c:1101-1154) and UMH
(me.c:1030-1099) integer searches.  For DIA and HEX:

* Hand-worked paths on a flat frame, where every SAD is 0 and a point's cost
  is its MV bits alone, so each step can be worked out on paper from the
  hexagon / diamond tables and the strict '<' (equal-cost points are not
  taken).
* On textured frames, checked against an independent numpy cost:
  - every accepted move is a hexagon move (HEX stage) or a diamond move;
  - the cost falls strictly along the path;
  - each stage that ends on its own ends in a local minimum: no point of the
    full hexagon (diamond) around its last point is strictly cheaper.  A
    wrong table entry or a wrong index update of the half patterns leaves a
    cheaper neighbour unvisited and fails this.
* The integer result of the whole search is the last point of the path.
* Where the current picture is the reference moved by a small whole-pixel
  shift, both searches find the shift, as the full search does.

For UMH, pictures built so that one stage alone can find the match: the
grid's repeated (0, 2), points on the cross and on the big hexagons, the
predictor points; early-termination routes forced through the context
(including the strict double bound); walks that fall and end in diamond
minima.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref
from tests.test_oracle_interp import smooth_frame

HEX = [(-1, -2), (1, -2), (-2, 0), (2, 0), (-1, 2), (1, 2)]
DIA = [(0, -1), (-1, 0), (1, 0), (0, 1)]


def plane_of(img):
    h, w = img.shape
    p = O.Plane(w, h)
    p.set_picture(img)
    p.expand_border()
    return p


def cur_plane(img):
    h, w = img.shape
    c = O.Plane(w, h)
    c.set_picture(img)
    return c


def cost_fn(cur, ref, pu, mvp, lam):
    """SAD against the border-clamped reference + lambda * MV bits (numpy)."""
    x, y, w, h = pu
    ref_pad = npref.clamp_picture(ref, O.PAD)
    blk = cur[y:y + h, x:x + w].astype(np.int64)

    def bits(v):
        return 1 if v == 0 else npref.se_len(v)

    def cost(mx, my):
        r = ref_pad[O.PAD + y + my:O.PAD + y + my + h, O.PAD + x + mx:O.PAD + x + mx + w].astype(np.int64)
        return int(np.abs(blk - r).sum()) + ((lam * (bits(4 * mx - mvp[0]) + bits(4 * my - mvp[1]))) >> 16)
    return cost


def run_path(method, cur, ref, pu, mvp, lam, start, R=16):
    w, h = cur.shape[1], cur.shape[0]
    prm = O.me_param(w, h, R, lam, use_fractional_me=0, me_method=method)
    cost = cost_fn(cur, ref, pu, mvp, lam)
    return O.me_pattern_path(prm, pu + mvp, cur_plane(cur), plane_of(ref), start, cost(*start)), cost


@pytest.mark.parametrize("method,start,expected", [
    # HEX from (3, 6): hexagon -> (1, 6) [bits 7 + 11], half hexagon -> (0, 4) [1 + 11],
    # the next half hexagon finds nothing below 12; diamond -> (0, 3) [1 + 9], and
    # the half diamond's (0, 2) costs 10 as well, which a strict '<' does not take
    (O.ME_HEX, (3, 6), [(3, 6), (1, 6), (0, 4), (0, 3)]),
    # DIA from (2, 5): diamond -> (1, 5) [7 + 11]; half diamond -> (0, 5) [1 + 11];
    # then (0, 4) and (0, 6) cost 12 too: stop
    (O.ME_DIA, (2, 5), [(2, 5), (1, 5), (0, 5)]),
    # DIA from (3, 6): every diamond point costs >= 20, no move
    (O.ME_DIA, (3, 6), [(3, 6)]),
    # DIA from (1, 1) [14]: (1, 0) and (0, 1) both cost 8; the diamond lists
    # (0, -1) before (-1, 0), so (1, 0) is taken; the half diamond then reaches
    # (0, 0) [2], around which every point costs 8
    (O.ME_DIA, (1, 1), [(1, 1), (1, 0), (0, 0)]),
])
def test_hand_worked_paths_on_a_flat_frame(method, start, expected):
    flat = np.full((64, 64), 90, np.uint8)
    path, _ = run_path(method, flat, flat, (16, 16, 16, 16), (0, 0), 65536, start)
    assert path == expected


def check_path(method, path, cost, legal, R):
    # a stage only continues from a point inside the MV window
    assert all(legal(*p) for p in path[:-1]), path
    costs = [cost(*p) for p in path]
    assert all(b < a for a, b in zip(costs, costs[1:])), (path, costs)
    moves = [(b[0] - a[0], b[1] - a[1]) for a, b in zip(path, path[1:])]
    n_hex = 0
    if method == O.ME_HEX:
        while n_hex < len(moves) and moves[n_hex] in HEX:
            n_hex += 1
    assert all(m in DIA for m in moves[n_hex:]), moves
    if method == O.ME_HEX:
        # the hexagon stage ended on its own unless its point left the MV
        # window or it ran its R + 1 steps
        c = path[n_hex]
        if legal(*c) and n_hex < R + 1:
            assert all(cost(c[0] + dx, c[1] + dy) >= costs[n_hex] for dx, dy in HEX), (path, c)
    c = path[-1]
    n_dia = len(moves) - n_hex
    if legal(*c) and n_dia < R + 1:
        assert all(cost(c[0] + dx, c[1] + dy) >= costs[-1] for dx, dy in DIA), (path, c)


@pytest.mark.parametrize("method", [O.ME_DIA, O.ME_HEX])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_paths_fall_and_end_in_local_minima(method, seed):
    w, h = 160, 112
    rng = np.random.default_rng(seed)
    ref = smooth_frame(w, h, seed)
    cur = np.roll(ref, (int(rng.integers(-6, 7)), int(rng.integers(-6, 7))), axis=(0, 1))
    cur = np.clip(cur.astype(int) + rng.integers(-4, 5, cur.shape), 0, 255).astype(np.uint8)
    lam = 4 * 65536
    R = 8
    for _ in range(12):
        pw, ph = [(8, 8), (16, 16), (16, 8), (8, 32), (32, 16)][int(rng.integers(0, 5))]
        pu = (int(rng.integers(0, (w - pw) // 4 + 1)) * 4, int(rng.integers(0, h - ph + 1)), pw, ph)
        mvp = (int(rng.integers(-20, 21)), int(rng.integers(-20, 21)))
        (xmin, xmax), (ymin, ymax) = npref.mv_window(w, h, *pu)
        start = (int(np.clip((mvp[0] + 2) >> 2, xmin, xmax)), int(np.clip((mvp[1] + 2) >> 2, ymin, ymax)))
        path, cost = run_path(method, cur, ref, pu, mvp, lam, start, R)
        check_path(method, path, cost, lambda mx, my: xmin <= mx <= xmax and ymin <= my <= ymax, R)


@pytest.mark.parametrize("method", [O.ME_DIA, O.ME_HEX])
def test_search_result_is_the_path_end_and_edges_stop_it(method):
    """PUs at the picture corners, where the hexagon reaches past the MV
    window: the search's integer result is the path's last point."""
    w, h = 96, 64
    ref = smooth_frame(w, h, 9)
    cur = np.roll(ref, (5, -7), axis=(0, 1))
    lam = 4 * 65536
    planes = [plane_of(ref)] + [None] * 15
    for pu, mvp in [((0, 0, 16, 16), (-300, -300)), ((80, 48, 16, 16), (300, 300)), ((40, 24, 8, 8), (0, 0)),
                    ((0, 48, 8, 16), (-270, 60))]:
        prm = O.me_param(w, h, 16, lam, use_fractional_me=0, me_method=method)
        res, _ = O.me_search(prm, pu + mvp, cur_plane(cur), planes)
        exp_cost, exp_mv = None, None
        (xmin, xmax), (ymin, ymax) = npref.mv_window(w, h, *pu)
        cost = cost_fn(cur, ref, pu, mvp, lam)
        best = None
        for q in (mvp, (0, 0)):
            c = (min(max((q[0] + 2) >> 2, xmin), xmax), min(max((q[1] + 2) >> 2, ymin), ymax))
            if best is None or cost(*c) < best[0]:
                best = (cost(*c), c)
        path, _ = run_path(method, cur, ref, pu, mvp, lam, best[1])
        exp_cost, exp_mv = cost(*path[-1]), path[-1]
        assert (res["bcost2"], res["bmv2_x"], res["bmv2_y"]) == (exp_cost, exp_mv[0], exp_mv[1])
        assert (res["bmv_x"], res["bmv_y"]) == (4 * exp_mv[0], 4 * exp_mv[1])


@pytest.mark.parametrize("method", [O.ME_DIA, O.ME_HEX])
@pytest.mark.parametrize("shift", [(1, 0), (-2, 1), (2, 2)])
def test_small_shift_found_like_full_search(method, shift):
    w, h = 128, 96
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ref = np.clip(np.rint(128 + 50 * np.sin(xx / 11.0 + 0.3) * np.cos(yy / 9.0 + 1.1)), 0, 255).astype(np.uint8)
    # pred = ref[p + mv]: the current picture is the reference moved by -shift
    cur = np.roll(ref, (-shift[1], -shift[0]), axis=(0, 1))
    planes = [plane_of(ref)] + [None] * 15
    lam = 4 * 65536
    for pu in [(32, 32, 16, 16), (64, 40, 16, 16), (48, 24, 32, 32)]:
        prm = O.me_param(w, h, 8, lam, use_fractional_me=0, me_method=method)
        res, _ = O.me_search(prm, pu + (0, 0), cur_plane(cur), planes)
        fs_cost, fs_mv = npref.fs_brute(cur, npref.clamp_picture(ref, O.PAD), O.PAD, pu, 8, lam)
        assert fs_mv == shift
        assert (res["bmv2_x"], res["bmv2_y"], res["bcost2"]) == (shift[0], shift[1], fs_cost)


def low_freq(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    f = 128 + 60 * np.sin(xx / 23.0 + rng.uniform(0, 6)) * np.cos(yy / 19.0 + rng.uniform(0, 6))
    return np.clip(np.rint(f + rng.normal(0, 1, (h, w))), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("method", [O.ME_DIA, O.ME_HEX])
def test_long_walks_turn_every_way(method):
    """Shifts of up to 14 pixels on a low-frequency field: walks of many steps
    that turn in every direction, so that every entry of the index tables of
    the half patterns is used."""
    w, h = 192, 160
    lam = 4 * 65536
    R = 24
    turns = set()
    for seed in range(6):
        rng = np.random.default_rng(100 + seed)
        ref = low_freq(w, h, seed)
        shift = (int(rng.integers(-14, 15)), int(rng.integers(-14, 15)))
        cur = np.roll(ref, (-shift[1], -shift[0]), axis=(0, 1))
        for _ in range(6):
            pu = (int(rng.integers(6, 30)) * 4, int(rng.integers(32, h - 48)), 16, 16)
            (xmin, xmax), (ymin, ymax) = npref.mv_window(w, h, *pu)
            path, cost = run_path(method, cur, ref, pu, (0, 0), lam, (0, 0), R)
            check_path(method, path, cost, lambda mx, my: xmin <= mx <= xmax and ymin <= my <= ymax, R)
            moves = [(b[0] - a[0], b[1] - a[1]) for a, b in zip(path, path[1:])]
            turns.update(zip(moves, moves[1:]))
    pat = HEX if method == O.ME_HEX else DIA
    # consecutive moves of the same pattern: every move is followed by some
    # other move at least once
    firsts = {a for a, b in turns if a in pat and b in pat}
    assert firsts == set(pat), sorted(turns)


@pytest.mark.parametrize("method", [O.ME_DIA, O.ME_HEX])
def test_walks_stop_where_they_leave_the_mv_window(method):
    """An MV range of +-10 qpel leaves the window [-4, 4] (full-pel, after the
    6-pixel margin) while the picture moved by 9 pixels: the walks reach the
    window edge; the point past it ends the walk."""
    w, h = 160, 128
    ref = low_freq(w, h, 7)
    lam = 4 * 65536
    left = 0
    for shift in [(9, 0), (-9, 3), (2, 9), (-7, -8)]:
        cur = np.roll(ref, (-shift[1], -shift[0]), axis=(0, 1))
        for pu in [(48, 40, 16, 16), (80, 64, 8, 8), (64, 48, 32, 16)]:
            prm = O.me_param(w, h, 16, lam, use_fractional_me=0, me_method=method,
                             mv_range_x=(-40, 40), mv_range_y=(-40, 40))
            cost = cost_fn(cur, ref, pu, (0, 0), lam)
            path = O.me_pattern_path(prm, pu + (0, 0), cur_plane(cur), plane_of(ref), (0, 0), cost(0, 0))
            check_path(method, path, cost, lambda mx, my: -4 <= mx <= 4 and -4 <= my <= 4, 16)
            left += not (-4 <= path[-1][0] <= 4 and -4 <= path[-1][1] <= 4)
    assert left > 0


# ---------------------------------------------------------------------------
# UMH (me.c:1030-1099, then the hexagon or the diamond)

def umh_ctx(**kw):
    c = np.zeros(1, dtype=O.UMH_CTX_DTYPE)
    c["beta2"] = -1.0          # beta + 1 = 0: no early termination
    c["beta3"] = -1.0
    for k, v in kw.items():
        c[k] = v
    return c


def umh_search(cur, ref, pu, mvp, lam, R, ctx, level=0):
    w, h = cur.shape[1], cur.shape[0]
    prm = O.me_param(w, h, R, lam, use_fractional_me=0, me_method=O.ME_UMH, umh_big_hex_level=level)
    res = O.me_search_batch(prm, np.array([pu + mvp], dtype=O.PU_DTYPE), cur_plane(cur),
                            [plane_of(ref)] + [None] * 15, ctxs=ctx)
    return (int(res["bmv2_x"][0]), int(res["bmv2_y"][0])), int(res["bcost2"][0])


def umh_route(cur, ref, pu, mvp, lam, R, ctx, level=0):
    w, h = cur.shape[1], cur.shape[0]
    prm = O.me_param(w, h, R, lam, use_fractional_me=0, me_method=O.ME_UMH, umh_big_hex_level=level)
    cost = cost_fn(cur, ref, pu, mvp, lam)
    (xmin, xmax), (ymin, ymax) = npref.mv_window(w, h, *pu)
    # the MV candidates (clipped MVP, clipped (0, 0)); the walk starts at the
    # cheaper (the first on a tie), UMH's first diamond goes around the MVP
    p0 = (min(max((mvp[0] + 2) >> 2, xmin), xmax), min(max((mvp[1] + 2) >> 2, ymin), ymax))
    z = (min(max(0, xmin), xmax), min(max(0, ymin), ymax))
    start = z if cost(*z) < cost(*p0) else p0
    path, route = O.me_pattern_path(prm, pu + mvp, cur_plane(cur), plane_of(ref), start, cost(*start), ctx=ctx,
                                    p0=p0, with_route=True)
    return path, route, cost


def test_umh_grid_repeats_0_2_and_never_tries_0_minus_2():
    """Rows repeat with period 4 and the current picture is the reference
    moved by 2 rows, so (0, 2) and (0, -2) match exactly (SAD 0) and cost the
    same MV bits.  No earlier point matches: the diamonds and the vertical
    cross step by odd rows.  The grid lists (0, 2) at entry 10, where (0, -2)
    would complete it, and (0, 2) again at entry 21, so the search ends at
    (0, 2) [bits 1 + 9]; a grid with (0, -2) at entry 10 would take (0, -2)
    first and keep it."""
    w, h = 96, 96
    rows = np.array([20, 200, 90, 150], np.uint8)
    ref = np.repeat(rows[np.arange(h) % 4][:, None], w, axis=1)
    cur = np.roll(ref, -2, axis=0)
    mv, cost = umh_search(cur, ref, (32, 32, 16, 16), (0, 0), 65536, 8, umh_ctx())
    assert (mv, cost) == ((0, 2), 10)


# lambda 250 on 8x8 PUs of a noise texture: a point one step off (0, 0) costs
# 6 more bits (1500), more than the SAD (about 5400) varies between
# neighbouring points, so the diamonds before the cross keep (0, 0) as the
# centre of the cross and the big hexagons (each test asserts it), while an
# exact match (SAD 0, at most 14 bits) still beats every other point
LAM_HEAVY = 250 * 65536


def centre_stays(cur, ref, pu):
    cost = cost_fn(cur, ref, pu, (0, 0), LAM_HEAVY)
    return all(cost(dx, dy) >= cost(0, 0) for dx, dy in DIA)


@pytest.mark.parametrize("shift", [(7, 0), (-5, 0), (0, 3), (0, -3), (-7, 0)])
def test_umh_cross_finds_odd_horizontal_and_short_vertical_shifts(shift):
    """A noise texture moved by a whole-pixel shift that lies on the
    unsymmetrical cross around (0, 0) (odd steps up to R across, up to R/2
    down and up): only that point matches, and the search ends there."""
    w, h = 128, 96
    rng = np.random.default_rng(5)
    ref = rng.integers(0, 256, (h, w)).astype(np.uint8)
    cur = np.roll(ref, (-shift[1], -shift[0]), axis=(0, 1))
    assert centre_stays(cur, ref, (48, 40, 8, 8))
    mv, _ = umh_search(cur, ref, (48, 40, 8, 8), (0, 0), LAM_HEAVY, 8, umh_ctx())
    assert mv == shift


def test_umh_vertical_cross_stops_at_half_the_range():
    """(0, 5) lies past R/2 = 4 on the vertical arm at R = 8 (and off the grid
    and the small patterns): nothing leads there.  At R = 12 the arm reaches
    it."""
    w, h = 128, 96
    rng = np.random.default_rng(6)
    ref = rng.integers(0, 256, (h, w)).astype(np.uint8)
    cur = np.roll(ref, -5, axis=0)
    assert centre_stays(cur, ref, (48, 40, 8, 8))
    mv, _ = umh_search(cur, ref, (48, 40, 8, 8), (0, 0), LAM_HEAVY, 8, umh_ctx())
    assert mv != (0, 5)
    mv, _ = umh_search(cur, ref, (48, 40, 8, 8), (0, 0), LAM_HEAVY, 12, umh_ctx())
    assert mv == (0, 5)


@pytest.mark.parametrize("level,shift,found", [
    (2, (8, 4), True),      # HEX4 entry (4, 2) at level 2
    (1, (8, 4), False),     # FAST_HEX4 has no (4, 2)
    (1, (-4, 6), True),     # FAST_HEX4 entry (-2, 3) at level 2
    (0, (-4, 6), False),    # no big hexagons
    (2, (0, -12), True),    # HEX4 entry (0, -4) at level 3 (R = 16: levels 1..4)
])
def test_umh_big_hexagons(level, shift, found):
    """Matches off the cross and the grid, at a big-hexagon point around
    (0, 0) (the heavy lambda keeps every earlier centre at (0, 0))."""
    w, h = 160, 128
    rng = np.random.default_rng(7)
    ref = rng.integers(0, 256, (h, w)).astype(np.uint8)
    cur = np.roll(ref, (-shift[1], -shift[0]), axis=(0, 1))
    assert centre_stays(cur, ref, (64, 56, 8, 8))
    mv, _ = umh_search(cur, ref, (64, 56, 8, 8), (0, 0), LAM_HEAVY, 16, umh_ctx(), level)
    assert (mv == shift) == found


@pytest.mark.parametrize("kw,route", [
    (dict(), 0 | 2 << 4),
    (dict(pred_sad_uplayer=10 ** 6, beta3=1.0), 1 | 3 << 4),
    (dict(pred_sad_uplayer=10 ** 6, beta2=1.0), 1 | 2 << 4),
    (dict(pred_sad=10 ** 6, beta3=1.0), 2 | 3 << 4),
    (dict(pred_sad=10 ** 6, beta2=1.0), 2 | 2 << 4),
])
def test_umh_early_termination_routes(kw, route):
    """beta + 1 scales pred_sad_uplayer at the first check and pred_sad at the
    later ones; beta3 jumps to the diamond, beta2 to the hexagon."""
    w, h = 128, 96
    ref = smooth_frame(w, h, 3)
    cur = np.roll(ref, (2, -3), axis=(0, 1))
    _, r, _ = umh_route(cur, ref, (48, 40, 16, 16), (0, 0), 4 * 65536, 8, umh_ctx(**kw))
    assert r == route


def test_umh_early_termination_is_a_strict_double_compare():
    """On a flat frame with MVP (0, 0) the best after the predictor points is
    (0, 0) at 2 bits (lambda factor 65536).  pred_sad_uplayer 2 with beta3 0
    gives the bound 2 * (0 + 1) = 2: 2 < 2 fails; beta3 1e-9 gives
    2.000000002, which it passes."""
    flat = np.full((64, 64), 90, np.uint8)
    _, r, _ = umh_route(flat, flat, (16, 16, 16, 16), (0, 0), 65536, 8, umh_ctx(pred_sad_uplayer=2, beta3=0.0))
    assert r & 15 != 1
    _, r, _ = umh_route(flat, flat, (16, 16, 16, 16), (0, 0), 65536, 8, umh_ctx(pred_sad_uplayer=2, beta3=1e-9))
    assert r == 1 | 3 << 4


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("seed", [1, 2])
def test_umh_paths_fall_and_end_in_diamond_minima(level, seed):
    w, h = 160, 112
    rng = np.random.default_rng(seed)
    ref = low_freq(w, h, seed)
    cur = np.roll(ref, (int(rng.integers(-9, 10)), int(rng.integers(-9, 10))), axis=(0, 1))
    cur = np.clip(cur.astype(int) + rng.integers(-3, 4, cur.shape), 0, 255).astype(np.uint8)
    lam = 4 * 65536
    for _ in range(8):
        pw, ph = [(8, 8), (16, 16), (32, 16), (16, 32)][int(rng.integers(0, 4))]
        pu = (int(rng.integers(8, (w - pw) // 4 - 8)) * 4, int(rng.integers(24, h - ph - 24)), pw, ph)
        mvp = (int(rng.integers(-20, 21)), int(rng.integers(-20, 21)))
        ctx = umh_ctx(mvp1_x=int(rng.integers(-30, 31)), mvp1_y=int(rng.integers(-30, 31)),
                      mvp3_x=int(rng.integers(-30, 31)), ref_idx=1, mvp2_x=8, mvp2_y=-12)
        path, _, cost = umh_route(cur, ref, pu, mvp, lam, 16, ctx, level)
        (xmin, xmax), (ymin, ymax) = npref.mv_window(w, h, *pu)
        legal = lambda mx, my: xmin <= mx <= xmax and ymin <= my <= ymax
        assert all(legal(*p) for p in path[:-1])
        costs = [cost(*p) for p in path]
        assert all(b < a for a, b in zip(costs, costs[1:])), (path, costs)
        c = path[-1]
        if legal(*c):
            assert all(cost(c[0] + dx, c[1] + dy) >= costs[-1] for dx, dy in DIA), (path, c)
        res = umh_search(cur, ref, pu, mvp, lam, 16, ctx, level)
        assert res == (c, costs[-1])


@pytest.mark.parametrize("shift", [(1, 0), (-2, 1), (2, 2), (-3, -1)])
def test_umh_small_shift_found_like_full_search(shift):
    w, h = 128, 96
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ref = np.clip(np.rint(128 + 50 * np.sin(xx / 11.0 + 0.3) * np.cos(yy / 9.0 + 1.1)), 0, 255).astype(np.uint8)
    cur = np.roll(ref, (-shift[1], -shift[0]), axis=(0, 1))
    lam = 4 * 65536
    for pu in [(32, 32, 16, 16), (64, 40, 16, 16)]:
        fs_cost, fs_mv = npref.fs_brute(cur, npref.clamp_picture(ref, O.PAD), O.PAD, pu, 8, lam)
        assert fs_mv == shift
        for level in (0, 2):
            assert umh_search(cur, ref, pu, (0, 0), lam, 8, umh_ctx(), level) == (shift, fs_cost)


@pytest.mark.parametrize("ref_idx,found", [(0, False), (1, True)])
def test_umh_tries_mvp2_only_for_later_references(ref_idx, found):
    """A noise texture matching at (5, 7) only, off every pattern around
    (0, 0); mvp2 = (20, 28) qpel rounds to it and is tried when ref_idx > 0
    (me.c:1048-1050)."""
    w, h = 128, 96
    rng = np.random.default_rng(8)
    ref = rng.integers(0, 256, (h, w)).astype(np.uint8)
    cur = np.roll(ref, (-7, -5), axis=(0, 1))
    assert centre_stays(cur, ref, (48, 40, 8, 8))
    mv, _ = umh_search(cur, ref, (48, 40, 8, 8), (0, 0), LAM_HEAVY, 8,
                       umh_ctx(ref_idx=ref_idx, mvp2_x=20, mvp2_y=28))
    assert (mv == (5, 7)) == found


@pytest.mark.parametrize("mvp,first,second_diamond", [((0, 0), (1, 0), False), ((0, 4), (1, 1), True)])
def test_umh_second_diamond_needs_a_nonzero_mvp(mvp, first, second_diamond):
    """A horizontal ramp moved by 3 pixels.  The first diamond goes around
    the MVP candidate (IPEL of the MVP: (0, 0), or (0, 1) for MVP (0, 4) qpel,
    which is cheaper than (0, 0) there) and moves the best one step right.
    Only with an MVP candidate other than (0, 0) does a second diamond run
    before the first early termination (me.c:1033-1036), one more step right.
    The bound pred_sad_uplayer = cost(first), beta3 = 0, tells the two apart:
    only a cost below cost(first) jumps."""
    w, h = 128, 96
    ref = np.repeat((2 * np.arange(w)).astype(np.uint8)[None, :], h, axis=0)
    cur = np.roll(ref, -3, axis=1)
    pu, lam = (48, 40, 16, 16), 4 * 65536
    cost = cost_fn(cur, ref, pu, mvp, lam)
    nxt = (first[0] + 1, first[1])
    assert cost(*nxt) < cost(*first) < min(cost(0, 0), cost((mvp[0] + 2) >> 2, (mvp[1] + 2) >> 2))
    _, route, _ = umh_route(cur, ref, pu, mvp, lam, 8, umh_ctx(pred_sad_uplayer=cost(*first), beta3=0.0))
    assert (route == 1 | 3 << 4) == second_diamond


def test_umh_gpu_test_contexts_take_every_route():
    """The seeded contexts of the UMH GPU parity test (all 25 shapes, two
    references, big-hexagon level 2) send searches through the whole prefix
    and out at each of the four early-termination points, to the hexagon and
    to the diamond: the GPU test compares every route with the oracle."""
    from tests.test_gpu_pattern import random_umh_ctx
    from xavs2_amd import synth
    w, h, R, level, seed = 256, 160, 16, 2, 69
    rng = np.random.default_rng(61 + level)
    pus = []
    for part in range(25):
        pw, ph = O.PU_W[part], O.PU_H[part]
        for _ in range(4):
            x = int(rng.integers(0, (w - pw) // 4 + 1)) * 4
            y = int(rng.integers(0, h - ph + 1))
            pus.append((x, y, pw, ph, int(rng.integers(-60, 61)), int(rng.integers(-60, 61))))
    refs = [synth.smooth_field(w, h, seed=seed * 10 + r) for r in range(2)]
    cur = synth.multi_ref_current(refs, 16, min(R, 24), seed=seed + 99)
    ctx = random_umh_ctx(len(pus), 2, seed + 5)
    routes = set()
    for r in range(2):
        for i, pu in enumerate(pus):
            c = np.ascontiguousarray(ctx[i * 2 + r:i * 2 + r + 1]).view(O.UMH_CTX_DTYPE)
            _, route, _ = umh_route(cur, refs[r], pu[:4], pu[4:], 4 * 65536, R, c, level)
            routes.add((route & 15, route >> 4))
    assert (0, 2) in routes
    for where in (1, 2, 3, 4):
        assert any(k == where for k, _ in routes), (where, sorted(routes))
    assert {t for _, t in routes} == {2, 3}


def test_umh_first_diamond_is_unchecked_at_the_window_edge():
    """DIA_ITER (me.c:328-334) tests its relative offsets (0,-1), (-1,0), (1,0), (0,1)
    against the window (CHECK_MV_RANGE_X4, me.h:51-56), so the points around a centre on
    the window edge are evaluated unchecked.  Flat frame, costs 4 * MV bits (mvbits 1, 3, 5,
    7, 9 for |MVD| 0, 1, 2-3, 4-7, 8-15); 16x16 PU at (0, 0) of 128x96: the window starts at
    -58.  MVP (-237, 0) qpel: mvc[0] = clip(IPEL) = (-58, 0), MVD 5, cost 4 * (7 + 1) = 32.
    The first diamond takes (-59, 0), one pixel outside the window (MVD 1: cost 16); the
    second diamond around it finds nothing cheaper ((-60, 0): MVD -3, cost 24); cross and
    grid points are checked (those cheaper than 16 are all outside); the hexagon around
    (-59, 0) finds nothing below 16, and the diamond needs a legal centre.  Result (-59, 0),
    16.  A diamond that checked its points would stay at (-58, 0) and the hexagon would
    then take (-60, 0) (cost 24): the HEX method, which has no DIA_ITER, does exactly that."""
    w, h = 128, 96
    flat = np.full((h, w), 60, np.uint8)
    pu, mvp = (0, 0, 16, 16), (-237, 0)
    assert umh_search(flat, flat, pu, mvp, 4 * 65536, 8, umh_ctx()) == ((-59, 0), 16)
    prm = O.me_param(w, h, 8, 4 * 65536, use_fractional_me=0, me_method=O.ME_HEX)
    res = O.me_search_batch(prm, np.array([pu + mvp], dtype=O.PU_DTYPE), cur_plane(flat),
                            [plane_of(flat)] + [None] * 15)
    assert (int(res["bmv2_x"][0]), int(res["bmv2_y"][0]), int(res["bcost2"][0])) == (-60, 0, 24)


def test_umh_diamond_takes_absolute_points_when_the_window_misses_an_offset():
    """When the window does not hold all four unit offsets, ME_COST_IPEL_X4's else branch
    evaluates ME_COST_IPEL(0,-1), (-1,0), (1,0), (0,1): absolute MVs, range-checked, whatever
    the centre.  MV range x and y limited to 24 qpel: the window of the 16x16 PU at (48, 32)
    ends at ((24 >> 2) - 6) = 0 on both axes, so (1, 0) and (0, 1) are outside.  The
    current PU is the reference at (-1, 0) exactly; MVP (-40, -80): mvc[0] = (-10, -20),
    where a noisy copy of the same block sits (SAD several hundred).  The first diamond,
    centred at (-10, -20), evaluates the absolute (-1, 0): SAD 0, cost 4 * (mvbits(36) +
    mvbits(80)) = 4 * (13 + 15) = 112, which no other point of a uniform-noise picture
    beats.  A diamond around (-10, -20) never reaches (-1, 0) (nor do the cross and grid
    around it)."""
    w, h = 128, 96
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 256, (h, w), dtype=np.uint8)
    blk = ref[32:48, 47:63].copy()
    ref[12:28, 38:54] = np.clip(blk.astype(int) + rng.choice([-3, 3], blk.shape), 0, 255)
    cur = rng.integers(0, 256, (h, w), dtype=np.uint8)
    cur[32:48, 48:64] = blk
    prm = O.me_param(w, h, 8, 4 * 65536, use_fractional_me=0, me_method=O.ME_UMH,
                     mv_range_x=(-8192, 24), mv_range_y=(-2048, 24))
    res = O.me_search_batch(prm, np.array([(48, 32, 16, 16, -40, -80)], dtype=O.PU_DTYPE), cur_plane(cur),
                            [plane_of(ref)] + [None] * 15, ctxs=umh_ctx())
    assert (int(res["bmv2_x"][0]), int(res["bmv2_y"][0]), int(res["bcost2"][0])) == (-1, 0, 112)
    # the same picture with the offsets inside the window: the diamond stays around (-10, -20)
    prm = O.me_param(w, h, 8, 4 * 65536, use_fractional_me=0, me_method=O.ME_UMH)
    res = O.me_search_batch(prm, np.array([(48, 32, 16, 16, -40, -80)], dtype=O.PU_DTYPE), cur_plane(cur),
                            [plane_of(ref)] + [None] * 15, ctxs=umh_ctx())
    assert (int(res["bmv2_x"][0]), int(res["bmv2_y"][0])) != (-1, 0)
