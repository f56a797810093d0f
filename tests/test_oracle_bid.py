"""Pins the oracle's bi-directional refinement (xavs2_me_search_bid,
encoder/me.c:1278-1392):

* a known answer: reference 2 at the backward MV reproduces the current PU,
  so the target clip(2 * cur - pred) is the current PU itself, and reference
  1 at the forward start reproduces it too: the start stays with the two MV
  rates (against mvp1 and mvp2) as its cost;
* the target's formula and clip: with a constant backward prediction c, a
  reference 1 that holds clip(2 * cur - c) at the start scores 0, and one
  that holds cur does not;
* the extra halving of fpel_cmp and the MV rates, through an independent
  numpy evaluation of the 9 half-pel and 8 quarter-pel Spiral points in order;
* the backward prediction of a plane that is not built comes from mc_luma:
  half-pel-only planes give the same results as full ones;
* a backward MV outside the window leaves every point without a cost.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref
from tests.test_oracle_sym import cur_plane, planes_of, texture

LAM = 4 * 65536
SPIRAL = [(0, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (1, 1)]


def bits(v):
    return 1 if v == 0 else npref.se_len(v)


def rate(mx, my, px, py):
    return (LAM * (bits(mx - px) + bits(my - py))) >> 16


def job(x, y, w, h, fwd, bwd, mvp1=(0, 0), mvp2=(0, 0)):
    return dict(pix_x=x, pix_y=y, w=w, h=h, fwd_x=fwd[0], fwd_y=fwd[1], bwd_x=bwd[0], bwd_y=bwd[1],
                mvp1_x=mvp1[0], mvp1_y=mvp1[1], mvp2_x=mvp2[0], mvp2_y=mvp2[1])


@pytest.mark.parametrize("d,e", [((3, -2), (-1, 4)), ((-4, 1), (2, 2))])
def test_known_answer_both_references_reproduce_the_pu(d, e):
    w, h = 96, 80
    cur = texture(w, h, 3)
    p1 = planes_of(np.roll(cur, (d[1], d[0]), axis=(0, 1)))    # ref1[p + d] = cur[p]
    p2 = planes_of(np.roll(cur, (e[1], e[0]), axis=(0, 1)))    # ref2[p + e] = cur[p]
    prm = O.me_param(w, h, 16, LAM)
    fwd, bwd = (4 * d[0], 4 * d[1]), (4 * e[0], 4 * e[1])
    mvp1, mvp2 = (3, -5), (-7, 2)
    cost, mx, my, mvcost = O.me_search_bid(prm, job(32, 24, 16, 16, fwd, bwd, mvp1, mvp2), cur_plane(cur), p1, p2)
    exp = rate(*fwd, *mvp1) + rate(*bwd, *mvp2)
    assert (cost, mx, my, mvcost) == (exp, fwd[0], fwd[1], exp)


def test_target_is_twice_cur_minus_prediction_clipped():
    w, h = 96, 80
    cur = texture(w, h, 4)
    c = 90
    target = np.clip(2 * cur.astype(np.int32) - c, 0, 255).astype(np.uint8)
    assert np.any(2 * cur.astype(np.int32) - c > 255) or np.any(2 * cur.astype(np.int32) - c < 0)
    p2 = planes_of(np.full((h, w), c, np.uint8))
    d = (2, 1)
    prm = O.me_param(w, h, 16, LAM)
    j = job(32, 24, 16, 16, (4 * d[0], 4 * d[1]), (0, 0))
    good = O.me_search_bid(prm, j, cur_plane(cur), planes_of(np.roll(target, (d[1], d[0]), axis=(0, 1))), p2)
    assert good[:3] == (rate(8, 4, 0, 0) + rate(0, 0, 0, 0), 8, 4)
    bad = O.me_search_bid(prm, j, cur_plane(cur), planes_of(np.roll(cur, (d[1], d[0]), axis=(0, 1))), p2)
    assert bad[0] > good[0]


def numpy_bid(had, frac, pmvr, j, cur, ref1, ref2):
    """me.c:1278-1392 restated with numpy planes, target and SATD / SAD."""
    w, h = cur.shape[1], cur.shape[0]
    pad = O.PAD
    pl1 = [npref.interp_plane_padded(ref1, k, pad) for k in range(16)]
    pl2 = [npref.interp_plane_padded(ref2, k, pad) for k in range(16)]
    x, y, bw, bh = j["pix_x"], j["pix_y"], j["w"], j["h"]
    org = cur[y:y + bh, x:x + bw].astype(np.int64)

    def legal(mx, my):
        return (-64 - x) * 4 <= mx <= (w + 64 - x - bw) * 4 and (-64 - y) * 4 <= my <= (h + 64 - y - bh) * 4

    def block(pl, mx, my):
        return pl[(my & 3) * 4 + (mx & 3)][pad + y + (my >> 2):pad + y + (my >> 2) + bh,
                                           pad + x + (mx >> 2):pad + x + (mx >> 2) + bw].astype(np.int64)

    bx2, by2 = j["bwd_x"], j["bwd_y"]
    target = np.clip(2 * org - block(pl2, bx2, by2), 0, 255)
    bid = rate(bx2, by2, j["mvp2_x"], j["mvp2_y"])

    def cost(mx, my):
        if not (legal(mx, my) and legal(bx2, by2)):
            return O.MAX_DISTORTION
        b = block(pl1, mx, my)
        d = npref.satd(target, b) if had else int(np.abs(target - b).sum())
        return (d >> 1) + rate(mx, my, j["mvp1_x"], j["mvp1_y"]) + bid

    best, bx, by = O.MAX_DISTORTION, j["fwd_x"], j["fwd_y"]
    for pos in range(0 if had else 1, 9):
        mx, my = j["fwd_x"] + 2 * SPIRAL[pos][0], j["fwd_y"] + 2 * SPIRAL[pos][1]
        c = cost(mx, my)
        if c < best:
            best, bx, by = c, mx, my
    if frac >= 2:
        qx, qy = bx, by
        cx, cy = (j["mvp1_x"] >> 1) << 1, (j["mvp1_y"] >> 1) << 1
        for pos in range(1, 9):
            sx, sy = SPIRAL[pos]
            if pmvr:
                if abs(qx - cx) > 2 or abs(qy - cy) > 2:
                    mx, my = qx + 2 * sx, qy + 2 * sy
                    if abs(mx - cx) <= 2 and abs(my - cy) <= 2:
                        continue
                else:
                    mx, my = qx + sx, qy + sy
                    if abs(mx - cx) > 2 or abs(my - cy) > 2:
                        continue
            else:
                mx, my = qx + sx, qy + sy
            c = cost(mx, my)
            if c < best:
                best, bx, by = c, mx, my
    return best, bx, by, rate(bx, by, j["mvp1_x"], j["mvp1_y"]) + bid


@pytest.mark.parametrize("had,pmvr", [(1, 0), (0, 0), (1, 1)])
@pytest.mark.parametrize("seed", [1, 2])
def test_matches_numpy_on_random_textures(had, pmvr, seed):
    w, h = 80, 64
    rng = np.random.default_rng(seed)
    cur, r1, r2 = texture(w, h, seed), texture(w, h, seed + 10), texture(w, h, seed + 20)
    p1, p2 = planes_of(r1), planes_of(r2)
    prm = O.me_param(w, h, 16, LAM, use_hadamard=had, enable_pmvr=pmvr)
    for _ in range(6):
        bw, bh = [(8, 8), (16, 16), (16, 8), (4, 8)][int(rng.integers(0, 4))]
        fwd = (int(rng.integers(-40, 41)), int(rng.integers(-40, 41)))
        j = job(int(rng.integers(0, (w - bw) // 4 + 1)) * 4, int(rng.integers(0, h - bh + 1)), bw, bh, fwd,
                (int(rng.integers(-40, 41)), int(rng.integers(-40, 41))),
                mvp1=(fwd[0] + int(rng.integers(-6, 7)), fwd[1] + int(rng.integers(-6, 7))),
                mvp2=(int(rng.integers(-20, 21)), int(rng.integers(-20, 21))))
        got = O.me_search_bid(prm, j, cur_plane(cur), p1, p2)
        assert got == numpy_bid(had, 2, pmvr, j, cur, r1, r2), j


def test_backward_prediction_without_its_plane_uses_mc_luma():
    w, h = 96, 80
    rng = np.random.default_rng(9)
    cur, r1, r2 = texture(w, h, 5), texture(w, h, 6), texture(w, h, 7)
    p1, full2 = planes_of(r1), planes_of(r2)
    half2 = [p if k in (0, 2, 8, 10) else None for k, p in enumerate(planes_of(r2))]
    prm = O.me_param(w, h, 16, LAM, use_fractional_me=1)
    n_missing = 0
    for _ in range(24):
        bwd = (int(rng.integers(-30, 31)), int(rng.integers(-30, 31)))
        n_missing += half2[(bwd[1] & 3) * 4 + (bwd[0] & 3)] is None
        j = job(int(rng.integers(0, 20)) * 4, int(rng.integers(0, 60)), 16, 16,
                (2 * int(rng.integers(-10, 11)), 2 * int(rng.integers(-10, 11))), bwd)
        assert O.me_search_bid(prm, j, cur_plane(cur), p1, half2) == O.me_search_bid(prm, j, cur_plane(cur), p1,
                                                                                      full2), j
    assert n_missing > 10


def test_backward_mv_outside_the_window():
    w, h = 96, 80
    cur = texture(w, h, 8)
    p1, p2 = planes_of(texture(w, h, 9)), planes_of(texture(w, h, 10))
    prm = O.me_param(w, h, 16, LAM)
    cost, mx, my, mvcost = O.me_search_bid(prm, job(0, 24, 16, 16, (4, 4), (-4 * 70, 0)), cur_plane(cur), p1, p2)
    assert (cost, mx, my) == (O.MAX_DISTORTION, 4, 4)
    # the returned MV cost still holds both rates (me.c:1388)
    assert mvcost == rate(4, 4, 0, 0) + rate(-280, 0, 0, 0)
