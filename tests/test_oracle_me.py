"""Pins the oracle's integer full search and sub-pel refinement: brute force
lexicographic (cost, order) minima, engineered ties, window edges, PMVR."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref
from tests.test_oracle_interp import smooth_frame


def make_pair(w, h, seed, shift=(3, -2)):
    ref = smooth_frame(w, h, seed)
    cur = np.roll(ref, shift, axis=(0, 1))
    rng = np.random.default_rng(seed + 1)
    cur = np.clip(cur.astype(int) + rng.integers(-3, 4, cur.shape), 0, 255).astype(np.uint8)
    return cur, ref


def planes_of(ref, frac=2):
    h, w = ref.shape
    p = O.Plane(w, h); p.set_picture(ref); p.expand_border()
    return O.interp_planes(p, frac) if frac else [p] + [None] * 15


def cur_plane(cur):
    h, w = cur.shape
    c = O.Plane(w, h); c.set_picture(cur)
    return c


@pytest.mark.parametrize("pu,R,mvp", [((32, 32, 16, 16), 8, (0, 0)), ((0, 0, 8, 8), 12, (0, 0)),
                                     ((100, 56, 8, 16), 8, (9, -7)), ((40, 8, 32, 8), 10, (-30, 21)),
                                     ((112, 72, 16, 16), 16, (0, 0))])
def test_integer_fs_matches_brute_force(pu, R, mvp):
    w, h = 128, 96
    cur, ref = make_pair(w, h, 5)
    planes = planes_of(ref, 0)
    lam = 4 * 65536
    prm = O.me_param(w, h, R, lam, use_fractional_me=0)
    res, cmap = O.me_search(prm, pu + mvp, cur_plane(cur), planes, cost_map=True)
    exp_cost, (bx, by) = npref.fs_brute(cur, npref.clamp_picture(ref, O.PAD), O.PAD, pu, R, lam, mvp)
    assert res["status"] == 0
    assert (res["cost"], res["bmv2_x"], res["bmv2_y"]) == (exp_cost, bx, by)
    assert (res["bmv_x"], res["bmv_y"]) == (4 * bx, 4 * by)
    # the cost map holds every legal raster candidate, and its minimum is the result
    legal = cmap >= 0
    assert legal.sum() > 0
    assert cmap[legal].min() >= exp_cost


def test_flat_frame_tie_goes_to_candidate_list():
    # every SAD is 0; MV bits break the tie, the first strict minimum is the (0,0) candidate
    w, h = 96, 64
    cur = np.full((h, w), 90, np.uint8); ref = cur.copy()
    planes = planes_of(ref, 2)
    prm = O.me_param(w, h, 8, 4 * 65536)
    res, _ = O.me_search(prm, (32, 16, 16, 16, 0, 0), cur_plane(cur), planes)
    assert (res["bmv2_x"], res["bmv2_y"], res["bmv_x"], res["bmv_y"]) == (0, 0, 0, 0)
    assert res["bcost2"] == (4 * 65536 * 2) >> 16       # mvbits[0] = 1 per component


def test_period2_texture_raster_order_breaks_ties():
    # columns alternate 0/255: displacements with equal parity tie on SAD; lambda 0 removes the
    # MV cost, so the first raster position of the best class must win.
    w, h = 64, 48
    img = np.tile(np.array([0, 255], np.uint8), (h, w // 2))
    planes = planes_of(img, 0)
    prm = O.me_param(w, h, 4, 0, use_fractional_me=0)
    res, cmap = O.me_search(prm, (24, 16, 8, 8, 0, 0), cur_plane(img), planes, cost_map=True)
    # mvc (0,0) is evaluated first with cost 0 and nothing is strictly smaller
    assert (res["bmv2_x"], res["bmv2_y"], res["cost"]) == (0, 0, 0)
    zero = np.argwhere(cmap == 0)
    assert len(zero) == 4 * 8  # even horizontal displacements, every row
    # with a shifted MVP the candidate list starts elsewhere but the tie still resolves to it
    res2, _ = O.me_search(prm, (24, 16, 8, 8, 8, 4), cur_plane(img), planes)
    assert (res2["bmv2_x"], res2["bmv2_y"]) == (2, 1)


def test_window_clip_at_frame_corner():
    w, h = 128, 96
    cur, ref = make_pair(w, h, 9)
    planes = planes_of(ref, 0)
    prm = O.me_param(w, h, 64, 4 * 65536, use_fractional_me=0)
    res, cmap = O.me_search(prm, (0, 0, 16, 16, 0, 0), cur_plane(cur), planes, cost_map=True)
    (xmin, xmax), (ymin, ymax) = npref.mv_window(w, h, 0, 0, 16, 16)
    assert xmin == -58 and ymin == -58
    legal = cmap >= 0
    # columns i - 64 < -58 are illegal, columns above xmax too
    assert not legal[:, :64 - 58].any() and legal[64:70, 64 - 58].all()
    assert not legal[:, 64 + xmax + 1:].any()
    exp_cost, (bx, by) = npref.fs_brute(cur, npref.clamp_picture(ref, O.PAD), O.PAD, (0, 0, 16, 16), 64, 4 * 65536)
    assert (res["cost"], res["bmv2_x"], res["bmv2_y"]) == (exp_cost, bx, by)


def subpel_brute(cur, planes, pu, bmv_int, lam, mvp, hadamard=True, pmvr=False, fast=False, frac=2):
    """Sequential sub-pel refinement written from the Spiral tables: centre, 8 half-pel
    points around the integer best, 8 quarter-pel points around the half-pel best."""
    x, y, w, h = pu
    spiral = [(0, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (1, 1)]
    spiral2 = [(0, 0), (0, -1), (-1, -1), (-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0), (1, -1)]
    pat = spiral2 if fast else spiral
    step = 2 if fast else 1
    blk = cur[y:y + h, x:x + w]

    def bits(v):
        return 1 if v == 0 else npref.se_len(v)

    def cost(mx, my):
        p = planes[((my & 3) << 2) + (mx & 3)]
        o = p.pad
        r = p.view()[o + y + (my >> 2):o + y + (my >> 2) + h, o + x + (mx >> 2):o + x + (mx >> 2) + w]
        d = npref.satd(blk, r) if hadamard else npref.sad(blk, r)
        return d + ((lam * (bits(mx - mvp[0]) + bits(my - mvp[1]))) >> 16)

    cx, cy = 4 * bmv_int[0], 4 * bmv_int[1]
    best = (cost(cx, cy), cx, cy)
    for pos in range(1, 9, step):
        mx, my = cx + 2 * pat[pos][0], cy + 2 * pat[pos][1]
        c = cost(mx, my)
        if c < best[0]:
            best = (c, mx, my)
    hx, hy = best[1], best[2]
    ctr = ((mvp[0] >> 1) << 1, (mvp[1] >> 1) << 1)
    if frac > 1:
        for pos in range(1, 9, step):
            sx, sy = pat[pos]
            if pmvr:
                if abs(hx - ctr[0]) > 2 or abs(hy - ctr[1]) > 2:
                    mx, my = hx + 2 * sx, hy + 2 * sy
                    if abs(mx - ctr[0]) <= 2 and abs(my - ctr[1]) <= 2:
                        continue
                else:
                    mx, my = hx + sx, hy + sy
                    if abs(mx - ctr[0]) > 2 or abs(my - ctr[1]) > 2:
                        continue
            else:
                mx, my = hx + sx, hy + sy
            c = cost(mx, my)
            if c < best[0]:
                best = (c, mx, my)
    return best


@pytest.mark.parametrize("pmvr,fast,hadamard,frac", [(0, 0, 1, 2), (1, 0, 1, 2), (0, 1, 1, 2),
                                                     (0, 0, 0, 2), (0, 0, 1, 1), (1, 1, 0, 2)])
def test_subpel_refine_matches_sequential_spiral(pmvr, fast, hadamard, frac):
    w, h = 128, 96
    cur, ref = make_pair(w, h, 21, shift=(5, 7))
    planes = planes_of(ref, frac)
    lam = 4 * 65536
    cp = cur_plane(cur)
    for pu, mvp in [((48, 32, 16, 16), (0, 0)), ((16, 40, 8, 8), (6, -3)), ((64, 16, 32, 8), (-11, 2)),
                    ((8, 8, 16, 64), (1, 1)), ((96, 64, 12, 16), (0, 0))]:
        if pu[1] + pu[3] > h:
            continue
        prm = O.me_param(w, h, 16, lam, use_hadamard=hadamard, use_fractional_me=frac,
                         use_fast_sub_me=fast, enable_pmvr=pmvr)
        res, _ = O.me_search(prm, pu + mvp, cp, planes)
        start = (res["bmv2_x"], res["bmv2_y"])
        exp = subpel_brute(cur, planes, pu, start, lam, mvp, hadamard, pmvr, fast, frac)
        # without Hadamard the start cost is the integer cost, which equals the SAD of the
        # full-pel plane at (4bmx, 4bmy) plus the same MV cost, so one formula covers both
        assert (res["cost"], res["bmv_x"], res["bmv_y"]) == exp, (pu, mvp)


def test_no_legal_candidate_returns_max_distortion():
    # a vertical MV range that excludes 0 after clipping: every candidate fails CHECK_MV_RANGE
    w, h = 64, 64
    cur, ref = make_pair(w, h, 2)
    planes = planes_of(ref, 0)
    prm = O.me_param(w, h, 4, 4 * 65536, use_fractional_me=0, mv_range_y=(40, 41))
    res, _ = O.me_search(prm, (16, 16, 16, 16, 0, 0), cur_plane(cur), planes)
    assert res["status"] == 1 and res["cost"] == O.MAX_DISTORTION


def test_batch_equals_single_calls():
    w, h = 128, 96
    cur, ref = make_pair(w, h, 4)
    planes = planes_of(ref, 2)
    prm = O.me_param(w, h, 8, 4 * 65536)
    pus = np.array([(x, y, 16, 16, 0, 0) for y in range(0, 96, 16) for x in range(0, 128, 16)],
                   dtype=O.PU_DTYPE)
    batch = O.me_search_batch(prm, pus, cur_plane(cur), planes, nthreads=3)
    for k, p in enumerate(pus):
        single, _ = O.me_search(prm, tuple(int(v) for v in p), cur_plane(cur), planes)
        assert single["cost"] == batch[k]["cost"] and single["bmv_x"] == batch[k]["bmv_x"]


def test_packed_mv_range_check_equals_box_test():
    """CHECK_MV_RANGE packs (mx, my) into one word; for the MV boxes an encoder
    forms (|MV| below 8192 full-pel) it must agree with the plain box test, which
    the GPU search uses for bands that lie wholly inside the box."""
    L = O.lib()
    L.xo_check_mv_range.argtypes = [O.C.c_int] * 6
    rng = np.random.default_rng(0)
    boxes = [(-58, -58, 1913, 1073), (-122, -70, 58, 30), (-8192 // 4 + 6, -2048 // 4 + 6, 2041, 505),
             (-200, -300, -100, -250), (100, 40, 300, 41)]
    for (x0, y0, x1, y1) in boxes:
        xs = np.unique(np.concatenate([np.arange(x0 - 3, x0 + 4), np.arange(x1 - 3, x1 + 4),
                                       rng.integers(x0 - 500, x1 + 500, 60)]))
        ys = np.unique(np.concatenate([np.arange(y0 - 3, y0 + 4), np.arange(y1 - 3, y1 + 4),
                                       rng.integers(y0 - 500, y1 + 500, 60)]))
        for mx in xs:
            for my in ys:
                box = x0 <= mx <= x1 and y0 <= my <= y1
                assert bool(L.xo_check_mv_range(int(mx), int(my), x0, y0, x1, y1)) == box, (mx, my, x0, y0, x1, y1)


# pmvr_adapt_mv (me.c:412-424, TH_PMVR = 2, defines.h:227), worked by hand from the text:
# (centre, half-pel best, Spiral step) -> (quarter-pel point, skipped).
PMVR_KNOWN = [
    # near branch: |mv - ctr| <= 2 on both axes, one qpel step, skipped when it leaves the +-2 box
    ((0, 0), (0, 0), (1, 0), (1, 0), 0),
    ((0, 0), (2, 0), (1, 0), (3, 0), 1),          # |2| is not > 2: near branch; 3 > 2: skipped
    ((0, 0), (2, 0), (-1, 0), (1, 0), 0),
    ((0, 0), (1, 1), (1, 1), (2, 2), 0),          # lands on the box corner: kept
    ((0, 0), (2, -2), (0, -1), (2, -3), 1),
    ((0, 0), (-2, 2), (-1, 1), (-3, 3), 1),
    ((4, -6), (6, -4), (1, 1), (7, -3), 1),
    ((4, -6), (6, -4), (-1, -1), (5, -5), 0),
    # far branch: |mv - ctr| > 2 on an axis, two qpel steps, skipped when it lands in the box
    ((0, 0), (3, 0), (1, 0), (5, 0), 0),
    ((0, 0), (3, 0), (-1, 0), (1, 0), 1),
    ((0, 0), (0, 3), (0, -1), (0, 1), 1),
    ((0, 0), (3, 0), (-1, 1), (1, 2), 1),         # both axes within 2: skipped
    ((0, 0), (3, 1), (-1, 1), (1, 3), 0),         # y lands at 3: kept
    ((0, 0), (4, 0), (-1, 0), (2, 0), 1),         # exactly 2 from the centre: inside
    ((0, 0), (-3, -3), (1, 1), (-1, -1), 1),
    ((-238, -242), (-234, -242), (-1, 0), (-236, -242), 1),
    ((-238, -242), (-234, -244), (-1, -1), (-236, -246), 0),
]


@pytest.mark.parametrize("ctr,mv,step,point,skip", PMVR_KNOWN)
def test_pmvr_adapt_known_answers(ctr, mv, step, point, skip):
    import ctypes as C
    mx, my = C.c_int(99999), C.c_int(99999)
    s = O.lib().xo_pmvr_adapt(C.byref(mx), C.byref(my), ctr[0], ctr[1], mv[0], mv[1], step[0], step[1])
    assert (s, mx.value, my.value) == (skip, point[0], point[1])


# (MVP, integer cost, PMVR, expected (bmv_x, bmv_y, cost)); see the test's docstring.
PMVR_FLAT_CASES = [
    ((-237, -242), 64, 0, (-234, -235, 48)),
    ((-237, -242), 64, 1, (-236, -236, 40)),
    ((-237, -236), 56, 0, (-234, -235, 32)),
    ((-237, -236), 56, 1, (-234, -236, 24)),
]


@pytest.mark.parametrize("mvp,bcost2,pmvr,exp", PMVR_FLAT_CASES)
def test_pmvr_quarter_pel_stage_known_answer(mvp, bcost2, pmvr, exp):
    """A flat frame (every SAD 0, interpolated planes flat too) and SAD costs: the cost is
    4 * (mvbits(mx - pmx) + mvbits(my - pmy)), worked by hand (mvbits 1, 3, 5, 7, 9 for
    |MVD| 0, 1, 2-3, 4-7, 8-15).  PU 16x16 at (0, 0) of a 128x96 frame: the window starts at
    -58 (md_inter.c:913-926) and both MVPs lie outside it, so the integer best is (-58, -58).

    MVP (-237, -242): integer MVD (5, 10), cost 4 * (7 + 9) = 64; half-pel best (-234, -234),
    MVD (3, 8), cost 56.  The PMVR centre is ((-237 >> 1) << 1, (-242 >> 1) << 1) =
    (-238, -242); the half-pel best is 4 qpel from it in x, so PMVR steps two qpel and
    keeps (-236, -236) (MVD (1, 6), cost 40), which the plain +-1 pattern cannot reach
    (its best is (-234, -235), MVD (3, 7), cost 48).

    MVP (-237, -236): integer MVD (5, 4), cost 56; half-pel best (-234, -234), MVD (3, 2),
    cost 40.  Centre (-238, -236) (the shift floors: -237 >> 1 = -119): x is 4 away, so two
    qpel steps; (-236, -236) and (-236, -234) land inside the +-2 box and are skipped, and
    (-234, -236) wins (MVD (3, 0), cost 24).  A centre truncated towards zero (-236, -236)
    would take the near branch and end at (-234, -235), cost 32, the plain pattern's answer."""
    w, h = 128, 96
    flat = np.full((h, w), 77, np.uint8)
    planes = planes_of(flat, 2)
    prm = O.me_param(w, h, 8, 4 * 65536, use_hadamard=0, enable_pmvr=pmvr)
    res, _ = O.me_search(prm, (0, 0, 16, 16) + mvp, cur_plane(flat), planes)
    assert (res["status"], res["bmv2_x"], res["bmv2_y"], res["bcost2"]) == (0, -58, -58, bcost2)
    assert (res["bmv_x"], res["bmv_y"], res["cost"]) == exp
    assert res["mvcost"] == exp[2]
