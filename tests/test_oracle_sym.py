"""Pins the oracle's symmetric bi-prediction refinement (xavs2_me_search_sym,
encoder/me.c:1190-1262, frame sequences):

* scale_mv_skip (predict.h:148-151) on hand-worked values, including the
  truncating MULTI / dist_src and the flooring shift of negative MVs;
* a known answer: the current picture is the average of two references moved
  by +d and -d, so in a B slice with equal distances the start MV 4d predicts
  it exactly (SATD 0) and stays; in a P slice the second MV is not mirrored;
* with SAD as fpel_cmp the start point itself is never evaluated (the
  half-pel loop starts at Spiral[1]);
* half-pel-only planes (use_fractional_me 1) give no cost at quarter
  positions, and MVs outside the window none at all;
* an independent numpy evaluation (separable FIR planes, numpy SATD, numpy
  average) of the Spiral points in order on random textures.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref

LAM = 4 * 65536
SPIRAL = [(0, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (1, 1)]


@pytest.mark.parametrize("mv,dd,ds,expected", [
    (5, 2, 1, 10),       # (5*2*16384 + 8192) >> 14 = 172032 >> 14 = 10
    (-5, 2, 1, -10),     # (-163840 + 8192) >> 14 = floor(-9.5) = -10
    (7, 1, 3, 2),        # 16384 / 3 = 5461: (38227 + 8192) >> 14 = 2
    (-7, 1, 3, -2),      # (-38227 + 8192) >> 14 = floor(-1.83) = -2
    (6, 1, 2, 3),        # (6*8192 + 8192) >> 14 = 57344 >> 14 = 3
    (1, 1, 2, 1),        # (8192 + 8192) >> 14 = 16384 >> 14 = 1
    (-1, 1, 2, 0),       # (-8192 + 8192) >> 14 = 0
    (3, 1, 6, 0),        # 16384 / 6 truncates to 2730: (8190 + 8192) >> 14 = 0 (exact 1/2 would round to 1)
    (2053, 1, 5, 410),   # 16384 / 5 = 3276: (6725628 + 8192) >> 14 = 410 (exact division: 411)
])
def test_scale_mv_skip_hand_values(mv, dd, ds, expected):
    assert O.scale_mv_skip(mv, dd, ds) == expected


def planes_of(img, frac=2):
    h, w = img.shape
    p = O.Plane(w, h)
    p.set_picture(img)
    p.expand_border()
    return O.interp_planes(p, frac) if frac else [p] + [None] * 15


def cur_plane(img):
    h, w = img.shape
    c = O.Plane(w, h)
    c.set_picture(img)
    return c


def texture(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    f = 128 + 50 * np.sin(xx / 5.3 + rng.uniform(0, 6)) * np.cos(yy / 4.1 + rng.uniform(0, 6))
    return np.clip(np.rint(f + rng.normal(0, 6, (h, w))), 0, 255).astype(np.uint8)


def job(x, y, w, h, mv, mvp=(0, 0), mvp1=(0, 0), dist=(1, 1), b_slice=1):
    return dict(pix_x=x, pix_y=y, w=w, h=h, mv_x=mv[0], mv_y=mv[1], mvp_x=mvp[0], mvp_y=mvp[1],
                mvp1_x=mvp1[0], mvp1_y=mvp1[1], dist_fwd=dist[0], dist_bwd=dist[1], b_slice=b_slice)


def bits(v):
    return 1 if v == 0 else npref.se_len(v)


@pytest.mark.parametrize("d", [(3, -2), (-4, 1)])
def test_mirrored_shift_known_answer(d):
    """ref1 holds cur's pixels at p + d and ref2 at p - d, so the average of
    the two predictions at MVs 4d and -4d is cur itself: SATD 0, and the
    start MV stays with its MV bits as the cost."""
    w, h = 96, 80
    cur = texture(w, h, 3)
    ref1 = np.roll(cur, (d[1], d[0]), axis=(0, 1))       # ref1[p + d] = cur[p]
    ref2 = np.roll(cur, (-d[1], -d[0]), axis=(0, 1))     # ref2[p - d] = cur[p]
    p1, p2 = planes_of(ref1), planes_of(ref2)
    prm = O.me_param(w, h, 16, LAM)
    mv = (4 * d[0], 4 * d[1])
    cost, mx, my, mvcost = O.me_search_sym(prm, job(32, 24, 16, 16, mv), cur_plane(cur), p1, p2)
    exp_bits = (LAM * (bits(mv[0]) + bits(mv[1]))) >> 16
    assert (cost, mx, my, mvcost) == (exp_bits, mv[0], mv[1], exp_bits)
    # P slice: the second MV is +d, where ref2 does not hold cur
    cost_p, _, _, _ = O.me_search_sym(prm, job(32, 24, 16, 16, mv, b_slice=0), cur_plane(cur), p1, p2)
    assert cost_p > exp_bits


def test_sad_mode_never_evaluates_the_start():
    w, h = 96, 80
    d = (3, -2)
    cur = texture(w, h, 4)
    p1 = planes_of(np.roll(cur, (d[1], d[0]), axis=(0, 1)))
    p2 = planes_of(np.roll(cur, (-d[1], -d[0]), axis=(0, 1)))
    prm = O.me_param(w, h, 16, LAM, use_hadamard=0, use_fractional_me=1)
    mv = (4 * d[0], 4 * d[1])
    cost, mx, my, _ = O.me_search_sym(prm, job(32, 24, 16, 16, mv), cur_plane(cur), p1, p2)
    # the exact start (cost = bits only) is not among Spiral[1..4] * 2
    assert (mx, my) != mv
    assert (mx - mv[0], my - mv[1]) in [(2 * a, 2 * b) for a, b in SPIRAL[1:5]]
    assert cost > (LAM * (bits(mv[0]) + bits(mv[1]))) >> 16


def test_half_pel_planes_only_and_window():
    w, h = 96, 80
    cur = texture(w, h, 5)
    # use_fractional_me 1 builds planes 2, 8 and 10 only (frame.c:343-350):
    # the others are absent (NULL), as the encoder holds them
    p1, p2 = planes_of(texture(w, h, 6), 1), planes_of(texture(w, h, 7), 1)
    for pl in (p1, p2):
        for k in range(1, 16):
            if k not in (2, 8, 10):
                pl[k] = None
    prm = O.me_param(w, h, 16, LAM, use_fractional_me=1)
    # a start on a quarter position: every half-pel point lands on a plane
    # that use_fractional_me 1 does not build
    cost, mx, my, _ = O.me_search_sym(prm, job(32, 24, 16, 16, (5, 3)), cur_plane(cur), p1, p2)
    assert (cost, mx, my) == (O.MAX_DISTORTION, 5, 3)
    # a start outside the qpel window (more than 64 pixels left of the picture)
    prm2 = O.me_param(w, h, 16, LAM)
    p1f, p2f = planes_of(texture(w, h, 6)), planes_of(texture(w, h, 7))
    cost, mx, my, _ = O.me_search_sym(prm2, job(0, 24, 16, 16, (-4 * 70, 0)), cur_plane(cur), p1f, p2f)
    assert cost == O.MAX_DISTORTION and (mx, my) == (-4 * 70, 0)


def numpy_sym(prm_kw, j, cur, ref1, ref2):
    """The Spiral points of me.c:1210-1252 in order, costed with numpy planes,
    average and SATD / SAD."""
    w, h = cur.shape[1], cur.shape[0]
    had = prm_kw.get("use_hadamard", 1)
    pad = O.PAD
    pl1 = [npref.interp_plane_padded(ref1, k, pad) for k in range(16)]
    pl2 = [npref.interp_plane_padded(ref2, k, pad) for k in range(16)]
    x, y, bw, bh = j["pix_x"], j["pix_y"], j["w"], j["h"]
    org = cur[y:y + bh, x:x + bw].astype(np.int64)

    def legal(mx, my):
        lo_x, hi_x = (-64 - x) * 4, (w + 64 - x - bw) * 4
        lo_y, hi_y = (-64 - y) * 4, (h + 64 - y - bh) * 4
        return lo_x <= mx <= hi_x and lo_y <= my <= hi_y

    def cost(mx, my):
        sx = O.scale_mv_skip(mx, j["dist_bwd"], j["dist_fwd"])
        sy = O.scale_mv_skip(my, j["dist_bwd"], j["dist_fwd"])
        if j["b_slice"]:
            sx, sy = -sx, -sy
        if not (legal(mx, my) and legal(sx, sy)):
            return O.MAX_DISTORTION
        a = pl1[(my & 3) * 4 + (mx & 3)][pad + y + (my >> 2):pad + y + (my >> 2) + bh,
                                          pad + x + (mx >> 2):pad + x + (mx >> 2) + bw].astype(np.int64)
        b = pl2[(sy & 3) * 4 + (sx & 3)][pad + y + (sy >> 2):pad + y + (sy >> 2) + bh,
                                          pad + x + (sx >> 2):pad + x + (sx >> 2) + bw].astype(np.int64)
        pred = (a + b + 1) >> 1
        d = npref.satd(org, pred) if had else int(np.abs(org - pred).sum())
        return d + ((LAM * (bits(mx - j["mvp_x"]) + bits(my - j["mvp_y"]))) >> 16)

    best, bx, by = O.MAX_DISTORTION, j["mv_x"], j["mv_y"]
    for pos in range(0 if had else 1, 5):
        mx, my = j["mv_x"] + 2 * SPIRAL[pos][0], j["mv_y"] + 2 * SPIRAL[pos][1]
        c = cost(mx, my)
        if c < best:
            best, bx, by = c, mx, my
    qx, qy = bx, by
    for pos in range(1, 5):
        mx, my = qx + SPIRAL[pos][0], qy + SPIRAL[pos][1]
        c = cost(mx, my)
        if c < best:
            best, bx, by = c, mx, my
    return best, bx, by


@pytest.mark.parametrize("had", [0, 1])
@pytest.mark.parametrize("seed", [1, 2])
def test_matches_numpy_on_random_textures(had, seed):
    w, h = 80, 64
    rng = np.random.default_rng(seed)
    cur, r1, r2 = texture(w, h, seed), texture(w, h, seed + 10), texture(w, h, seed + 20)
    p1, p2 = planes_of(r1), planes_of(r2)
    prm = O.me_param(w, h, 16, LAM, use_hadamard=had)
    for _ in range(6):
        bw, bh = [(8, 8), (16, 16), (16, 8), (4, 8)][int(rng.integers(0, 4))]
        j = job(int(rng.integers(0, (w - bw) // 4 + 1)) * 4, int(rng.integers(0, h - bh + 1)), bw, bh,
                (int(rng.integers(-40, 41)), int(rng.integers(-40, 41))),
                mvp=(int(rng.integers(-20, 21)), int(rng.integers(-20, 21))),
                dist=(int(rng.integers(1, 5)), int(rng.integers(1, 5))), b_slice=int(rng.integers(0, 2)))
        cost, mx, my, _ = O.me_search_sym(prm, j, cur_plane(cur), p1, p2)
        assert (cost, mx, my) == numpy_sym(dict(use_hadamard=had), j, cur, r1, r2), j


def test_mirror_straddling_the_window():
    """A PU at the left edge in a B slice with distances 1 and 8: the mirror of
    a candidate m is -8m, and the window starts at -256 qpel, so the candidates
    with m <= 32 keep their mirror inside and the others lose their cost."""
    w, h = 80, 64
    cur, r1, r2 = texture(w, h, 31), texture(w, h, 32), texture(w, h, 33)
    prm = O.me_param(w, h, 16, LAM)
    p1, p2 = planes_of(r1), planes_of(r2)
    for mv in ((33, 0), (31, 2), (34, -1)):
        j = job(0, 24, 16, 16, mv, dist=(1, 8), b_slice=1)
        got = O.me_search_sym(prm, j, cur_plane(cur), p1, p2)
        assert got[:3] == numpy_sym(dict(), j, cur, r1, r2), mv
        assert got[1] <= 32
