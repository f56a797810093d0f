"""Pins the oracle's callers of the two bi-prediction refinements,
pred_inter_search_bi (encoder/md_inter.c:1036-1128) and
pred_inter_search_dual (md_inter.c:1133-1219), frame sequences, every
reference row reconstructed:

* each gate on hand-worked boundaries: check_mvd's +-4096 (and +-(64 << 2)
  vertically with frame threads, md_inter.c:68-76), check_mv_range's box of
  the PU reaching 1 << lcu_level pixels past the picture (predict.h:68-78),
  check_mv_range_sym's mirrored MV (negated in B slices, md_inter.c:83-110);
* a failed gate leaves the search out: MAX_DISTORTION, the MV as it came in
  (full-pel best x 4), no MV cost;
* the known answer of the symmetric refinement (the picture is the average
  of two references moved by +d and -d) reached through the caller, and the
  bid refinement through it, equal to the refinements called alone with the
  caller's inputs;
* the stored MVs' own checks after the search (md_inter.c:1107-1122);
* dual: REF_COST(ref_idx) = (lambda * REF_BITS[ref_idx + 1]) >> 16 added per
  reference, the first strict minimum over references, the second MV scaled
  by scale_mv_skip(fst, d_snd, d_fst), a failing p_me->mvp gate voiding all.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_sym import cur_plane, planes_of, texture

LAM = 4 * 65536
MAX = O.MAX_DISTORTION
W, H = 128, 96


@pytest.fixture(scope="module")
def avg_case():
    """ref1 moved by +d and ref2 by -d average to the picture: in a B slice with
    equal distances, the symmetric MV 4d predicts it exactly."""
    d = (3, -2)
    cur = texture(W, H, 5)
    r1 = np.roll(cur, (d[1], d[0]), axis=(0, 1))       # r1[p + d] = cur[p]
    r2 = np.roll(cur, (-d[1], -d[0]), axis=(0, 1))     # r2[p - d] = cur[p]
    return d, cur_plane(cur), planes_of(r1), planes_of(r2)


def bi_job(**kw):
    j = dict(pix_x=48, pix_y=32, w=16, h=16, fwd_imv_x=3, fwd_imv_y=-2, bwd_x=-12, bwd_y=8, mvp_x=0, mvp_y=0,
             mvp1_x=0, mvp1_y=0, mvp2_x=0, mvp2_y=0)
    j.update(kw)
    return j


def run_bi(case, job, dist=(2, 2), frame_threads=1, **prm_kw):
    _, cp, p1, p2 = case
    prm = O.me_param(W, H, 32, LAM, **prm_kw)
    return O.pred_inter_search_bi(prm, job, dist[0], dist[1], cp, p1, p2, lcu_level=6,
                                  frame_threads=frame_threads)


def test_known_answer_through_the_caller(avg_case):
    d = avg_case[0]
    r = run_bi(avg_case, bi_job())
    # the symmetric MV starts at 4d and stays: SATD 0, cost = its MV rate against mvp
    prm = O.me_param(W, H, 32, LAM)
    exp_rate = (LAM * (O.mvbits(32, 4 * d[0]) + O.mvbits(32, 4 * d[1]))) >> 16
    assert (r["sym_mv_x"], r["sym_mv_y"]) == (4 * d[0], 4 * d[1])
    assert r["sym_cost"] == exp_rate and r["sym_mvcost"] == exp_rate
    # bid: the same as the refinement called alone with the caller's inputs
    _, cp, p1, p2 = avg_case
    alone = O.me_search_bid(prm, dict(pix_x=48, pix_y=32, w=16, h=16, fwd_x=12, fwd_y=-8, bwd_x=-12, bwd_y=8,
                                      mvp1_x=0, mvp1_y=0, mvp2_x=0, mvp2_y=0), cp, p1, p2)
    assert (r["bid_cost"], r["fwd_x"], r["fwd_y"], r["bid_mvcost"]) == alone
    assert (r["bwd_x"], r["bwd_y"]) == (-12, 8)


@pytest.mark.parametrize("field,value,sym_ran,bid_ran,bid_max", [
    # PU (48, 32) 16x16, lcu_level 6: the box is x in [-448, 512], y in [-384, 448] (qpel)
    ("mvp1_x", 448, True, True, False),    # check_mv_range_sym(mvp1): mirror -448 inside
    ("mvp1_x", 452, False, True, False),   # mirror -452 < -448: the symmetric search is gated
    ("mvp1_x", 4095, False, True, False),  # bid: check_mvd(mvp1) holds (< 4096), so does fwd - mvp1
    ("mvp1_x", 4096, False, False, True),  # bid: check_mvd(mvp1) fails
    ("mvp1_x", -4096, False, True, True),  # bid runs (>= -4096), then fwd - mvp1 = 4108 fails the stored check
    ("mvp1_x", -4097, False, False, True),
    ("mvp2_y", 4096, True, False, True),   # only bid gates on mvp2
    ("mvp_x", 9000, True, True, False),    # p_me->mvp only prices the symmetric MV
])
def test_gates_on_the_predictors(avg_case, field, value, sym_ran, bid_ran, bid_max):
    r = run_bi(avg_case, bi_job(**{field: value}))
    if sym_ran:
        assert r["sym_mvcost"] != -1
    else:
        assert (r["sym_cost"], r["sym_mv_x"], r["sym_mv_y"], r["sym_mvcost"]) == (MAX, 12, -8, -1)
    assert (r["bid_mvcost"] != -1) == bid_ran
    assert (r["bid_cost"] == MAX) == bid_max
    if not bid_ran:
        assert (r["fwd_x"], r["fwd_y"]) == (12, -8)


def test_frame_threads_narrow_the_vertical_mvd():
    """With frame threads the vertical MVD must lie in [-(64 << 2), 64 << 2):
    mvp1.y = 255 passes, 256 fails (lcu_level 6)."""
    case = None
    d = (0, 0)
    img = texture(W, H, 9)
    cp, p1 = cur_plane(img), planes_of(img)
    prm = O.me_param(W, H, 32, LAM)
    for y, ok in ((255, True), (256, False), (-256, True), (-257, False)):
        r = O.pred_inter_search_bi(prm, bi_job(fwd_imv_x=0, fwd_imv_y=0, bwd_x=0, bwd_y=0, mvp1_y=y), 2, 2, cp, p1,
                                   p1, lcu_level=6, frame_threads=2)
        assert (r["bid_mvcost"] != -1) == ok, y
        r1 = O.pred_inter_search_bi(prm, bi_job(fwd_imv_x=0, fwd_imv_y=0, bwd_x=0, bwd_y=0, mvp1_y=y), 2, 2, cp, p1,
                                    p1, lcu_level=6, frame_threads=1)
        assert r1["bid_mvcost"] != -1


def test_mv_range_box_and_the_mirror():
    """PU at (0, 0), 16x16, lcu_level 6: MVs down to -(0 + 64) * 4 = -256 qpel.  A
    full-pel best of -64 starts at -256 (inside), -65 at -260 (outside: both
    searches gated).  In a B slice the mirror -scale_mv_skip(mv, d_bwd, d_fwd)
    is checked too: with d_fwd 1, d_bwd 4, mv 80 mirrors to -320 < -256, so the
    symmetric search is gated while bid is not."""
    img = texture(W, H, 11)
    cp, p1 = cur_plane(img), planes_of(img)
    prm = O.me_param(W, H, 32, LAM)
    job = bi_job(pix_x=0, pix_y=0, bwd_x=0, bwd_y=0, fwd_imv_y=0)
    r = O.pred_inter_search_bi(prm, dict(job, fwd_imv_x=-64), 2, 2, cp, p1, p1)
    assert r["bid_mvcost"] != -1
    r = O.pred_inter_search_bi(prm, dict(job, fwd_imv_x=-65), 2, 2, cp, p1, p1)
    assert r["bid_mvcost"] == -1 and r["sym_mvcost"] == -1 and (r["fwd_x"], r["sym_mv_x"]) == (-260, -260)
    assert O.scale_mv_skip(80, 4, 1) == 320
    r = O.pred_inter_search_bi(prm, dict(job, fwd_imv_x=20), 1, 4, cp, p1, p1)
    assert r["sym_mvcost"] == -1 and r["sym_cost"] == MAX and r["bid_mvcost"] != -1
    r = O.pred_inter_search_bi(prm, dict(job, fwd_imv_x=16), 1, 4, cp, p1, p1)     # mirror -256: inside
    assert r["sym_mvcost"] != -1


def test_stored_mv_checks_after_the_search():
    """A start MV on the box's right edge (max_x = ((128 - (112 + 16)) + 64) * 4 =
    256) on a flat picture: the refinement's costs are the MV rates, smallest
    toward mvp (0, 0), so the search stays at 256 and passes.  Start 1 qpel
    outside: gated (MAX).  The mirror of a B slice with equal distances is
    -256, inside."""
    flat = np.full((H, W), 77, np.uint8)
    cp, p1 = cur_plane(flat), planes_of(flat)
    prm = O.me_param(W, H, 32, LAM)
    job = bi_job(pix_x=112, pix_y=32, fwd_imv_x=64, fwd_imv_y=0, bwd_x=0, bwd_y=0)
    r = O.pred_inter_search_bi(prm, job, 2, 2, cp, p1, p1)
    assert r["sym_cost"] != MAX and r["sym_mv_x"] <= 256
    assert r["bid_cost"] != MAX and r["fwd_x"] <= 256


def dual_case(n_ref=2):
    img = texture(W, H, 21)
    return cur_plane(img), [planes_of(img) for _ in range(n_ref)]


def test_dual_ref_cost_and_first_minimum():
    """Identical references: the symmetric searches of (0, 1) and (1, 0) score the
    same, so the totals differ by REF_COST: (LAM * 1) >> 16 = 4 for ref 0 and
    (LAM * 3) >> 16 = 12 for ref 1; ref 0 wins."""
    cp, planes = dual_case()
    prm = O.me_param(W, H, 32, LAM)
    job = dict(pix_x=48, pix_y=32, w=16, h=16, mvp_x=0, mvp_y=0, mvp1=[(0, 0), (0, 0)], imv=[(0, 0), (0, 0)])
    cost, fr, sr, mvc, fst, snd = O.pred_inter_search_dual(prm, job, [2, 4], cp, planes)
    alone = O.me_search_sym(prm, dict(pix_x=48, pix_y=32, w=16, h=16, mv_x=0, mv_y=0, mvp_x=0, mvp_y=0, mvp1_x=0,
                                      mvp1_y=0, dist_fwd=2, dist_bwd=4, b_slice=0), cp, planes[0], planes[1])
    assert (cost, fr, sr) == (alone[0] + 4, 0, 1) and mvc == alone[3]
    # ref 1 alone (ref 0 gated by a huge mvp1) pays 12
    job2 = dict(job, mvp1=[(5000, 0), (0, 0)])
    cost2, fr2, sr2, _, _, _ = O.pred_inter_search_dual(prm, job2, [2, 4], cp, planes)
    alone1 = O.me_search_sym(prm, dict(pix_x=48, pix_y=32, w=16, h=16, mv_x=0, mv_y=0, mvp_x=0, mvp_y=0, mvp1_x=0,
                                       mvp1_y=0, dist_fwd=4, dist_bwd=2, b_slice=0), cp, planes[1], planes[0])
    assert (cost2, fr2, sr2) == (alone1[0] + 12, 1, 0)


def test_dual_second_mv_and_mvp_gate():
    """The second MV is scale_mv_skip(fst, d_snd, d_fst): fst (12, -8) with
    distances 2 and 6 gives (36, -24) (exact), and with 6 and 2 gives
    ((12 * 2 * 2730 + 8192) >> 14, ..) = (4, -3) (16384 / 6 truncates).
    A p_me->mvp beyond check_mvd voids every reference."""
    flat = np.full((H, W), 90, np.uint8)
    cp, planes = cur_plane(flat), [planes_of(flat)] * 2
    prm = O.me_param(W, H, 32, LAM, use_fractional_me=0)
    job = dict(pix_x=48, pix_y=32, w=16, h=16, mvp_x=0, mvp_y=0, mvp1=[(12, -8), (12, -8)], imv=[(3, -2), (3, -2)])
    _, _, _, _, fst, snd = O.pred_inter_search_dual(prm, job, [2, 6], cp, planes)
    assert fst == [(12, -8), (12, -8)]
    assert snd[0] == (36, -24) and snd[1] == (4, -3)
    assert O.scale_mv_skip(12, 2, 6) == 4 and O.scale_mv_skip(-8, 2, 6) == -3
    cost, fr, sr, mvc, _, _ = O.pred_inter_search_dual(prm, dict(job, mvp_x=4096), [2, 6], cp, planes)
    assert (cost, fr, sr, mvc) == (MAX, -1, -1, -1)


def test_mvp_prices_the_symmetric_mv_and_mvp1_checks_it(avg_case):
    """p_me->mvp = (9000, 0) prices the symmetric MV (12, -8):
    (LAM * (bits(12 - 9000) + bits(-8))) >> 16 is its MV cost, while the stored
    MV's check uses mvp1 (the local mvp of md_inter.c:1063), so the cost stays
    valid: against mvp, 12 - 9000 would fail check_mvd."""
    r = run_bi(avg_case, bi_job(mvp_x=9000))
    rate = (LAM * (O.mvbits(32, 12 - 9000) + O.mvbits(32, -8))) >> 16
    assert (r["sym_mv_x"], r["sym_mv_y"], r["sym_mvcost"]) == (12, -8, rate)
    assert r["sym_cost"] == rate


def test_backward_mv_gate_and_its_stored_check(avg_case):
    """The backward MV is gated by check_mv_range before the bid search
    (md_inter.c:1092-1096) and checked again, against mvp2, after it
    (:1113-1116).  PU (48, 32) 16x16: the box's right edge is x = 512 qpel.
    bwd_x 516 leaves the search out (no MV cost, fwd as it came).  With
    mvp2_x = -4000 and bwd_x = 200 the gate (check_mvd(mvp2)) passes and the
    search runs, but 200 - (-4000) = 4200 fails check_mvd afterwards: the
    search's MV cost is kept, its cost is voided."""
    r = run_bi(avg_case, bi_job(bwd_x=516))
    assert r["bid_mvcost"] == -1 and r["bid_cost"] == MAX and (r["fwd_x"], r["fwd_y"]) == (12, -8)
    r = run_bi(avg_case, bi_job(bwd_x=512))
    assert r["bid_mvcost"] != -1
    r = run_bi(avg_case, bi_job(bwd_x=200, mvp2_x=-4000))
    assert r["bid_mvcost"] != -1 and r["bid_cost"] == MAX
    r = run_bi(avg_case, bi_job(bwd_x=200, mvp2_x=-3800))
    assert r["bid_cost"] != MAX


def test_x4_of_the_full_pel_best_wraps_in_int16(avg_case):
    """fwd_mv.x <<= 2 shifts an int16 component of mv_t (md_inter.c:1068):
    4 * 16387 = 65548 wraps to 12, the same start as a full-pel best of 3."""
    assert run_bi(avg_case, bi_job(fwd_imv_x=16387)) == run_bi(avg_case, bi_job(fwd_imv_x=3))


def test_dual_mirror_keeps_its_sign_in_f_slices():
    """F slices mirror without the B-slice negation (md_inter.c:97-101): with
    reference 0 the picture moved by d at distance 1 and reference 1 moved by
    2d at distance 2, the symmetric MV 4d averages two exact predictions
    (its mirror is +8d), so reference 0 costs its MV rate plus REF_COST(0)."""
    d = (3, -2)
    img = texture(W, H, 31)
    cp = cur_plane(img)
    r0 = planes_of(np.roll(img, (d[1], d[0]), axis=(0, 1)))
    r1 = planes_of(np.roll(img, (2 * d[1], 2 * d[0]), axis=(0, 1)))
    prm = O.me_param(W, H, 32, LAM)
    job = dict(pix_x=48, pix_y=32, w=16, h=16, mvp_x=0, mvp_y=0, mvp1=[(0, 0), (0, 0)], imv=[d, (0, 0)])
    cost, fr, _, mvc, fst, snd = O.pred_inter_search_dual(prm, job, [1, 2], cp, [r0, r1])
    rate = (LAM * (O.mvbits(32, 4 * d[0]) + O.mvbits(32, 4 * d[1]))) >> 16
    assert fst[0] == (4 * d[0], 4 * d[1]) and snd[0] == (8 * d[0], 8 * d[1])
    assert (cost, fr, mvc) == (rate + 4, 0, rate)
    # PU at x = 0: the box starts at -(0 + 64) * 4 = -256 qpel.  A start of
    # 4 * 40 = 160 mirrors to +320 (inside); negated, -320 would gate
    # reference 0 (reference 1 is voided here by its mvp1)
    job = dict(pix_x=0, pix_y=32, w=16, h=16, mvp_x=0, mvp_y=0, mvp1=[(160, 0), (5000, 0)], imv=[(40, 0), (0, 0)])
    cost, fr, _, _, _, _ = O.pred_inter_search_dual(prm, job, [1, 2], cp, [r0, r1])
    assert fr == 0 and cost < MAX


def test_dual_first_minimum_among_equal_totals():
    """References 1 and 2 identical, at one distance, with the same inputs,
    and reference 0 voided by an mvp1 beyond check_mvd: the totals of 1 and 2
    are equal (REF_BITS[2] = REF_BITS[3] = 3), so the first strict minimum
    keeps reference 1."""
    img = texture(W, H, 41)
    cp = cur_plane(img)
    a = planes_of(np.roll(img, (1, 2), axis=(0, 1)))
    b = planes_of(np.roll(img, (-2, 1), axis=(0, 1)))
    prm = O.me_param(W, H, 32, LAM)
    for x, y in ((16, 16), (48, 32), (96, 64)):
        job = dict(pix_x=x, pix_y=y, w=16, h=16, mvp_x=0, mvp_y=0, mvp1=[(5000, 0), (0, 0), (0, 0)],
                   imv=[(0, 0), (-2, 1), (-2, 1)])
        cost, fr, sr, _, _, _ = O.pred_inter_search_dual(prm, job, [2, 3, 3], cp, [a, b, b])
        assert (fr, sr) == (1, 0) and cost < MAX
