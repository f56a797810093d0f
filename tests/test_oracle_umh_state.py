"""Pins the oracle's UMH encoder-state producer (get_mvp_default_sad,
md_inter.c:527-608, and fast_me_prepare_info, md_inter.c:640-718) and
xavs2_me_init_umh_threshold (me.c:647-665) on cases worked by hand from the
text: which neighbour's min cost the MVP's y median picks, the upper-layer
mode of each mode, the nearest reference's float scaling and its B-slice
negation, the co-located MV, the pred_sad rule, the betas and the int32
square of a large pred_sad."""
import numpy as np
import pytest

from oracle import oracle as O

DPOC = [2, 4, 6, 8]
MULTI = [16384 // d for d in DPOC]
L, U, UL, UR = (4, -8), (12, 6), (-20, 30), (7, -3)
P_SLICE, B_SLICE = 1, 2


def mvp_job(ref_idx=0, refs=(0, 0, 0, 0), mvs=(L, U, UL, UR), ur=1, cb=(0, 0, 16, 16), bwd=0):
    return dict(ref_idx=ref_idx, bwd_2nd=bwd, cb=cb, ur_available=ur, ref=list(refs), mv=[list(m) for m in mvs])


def bsize_p33():
    return O.umh_bsize(33, P_SLICE)


def test_umh_bsize_known_answer():
    """QP 33, P slice: qp_per 5, qp_rem 3, q_bits 20, qp_const 2^20 / 6 = 174762,
    threshold_4x4 = (1048576 - 174762) / 9362 = 93; quantize_step = 93 / (4 * 5.61f) in
    float; nRx2N = 256 * step, then x4 per level up (nLx2N, 2NxnD), 2NxnU, Nx2N and
    2NxN = 4 * 2NxnU, 2Nx2N = 4 * 2NxN; SKIP 0."""
    b = O.umh_bsize(33, P_SLICE)
    step = float(np.float32(93) / (np.float32(4) * np.float32(5.61)))
    assert abs(step - 93 / 22.44) < 1e-5
    assert b[7] == 256 * step and b[6] == 4 * b[7] and b[5] == 4 * b[7] and b[4] == 16 * b[7]
    assert b[3] == 64 * b[7] and b[2] == 64 * b[7] and b[1] == 256 * b[7] and b[0] == 0
    # I slices round with 1/3: threshold_4x4 = (1048576 - 349525) / 9362 = 74
    bi = O.umh_bsize(33, 0)
    assert bi[7] == 256 * float(np.float32(74) / (np.float32(4) * np.float32(5.61)))


def test_median_y_picks_the_neighbour_cost_and_upper_layer_2nxn():
    """P slice, PRED_2NxN, ref 0, all neighbours on ref 0 (scaling by 2 * 8192 >> 14 is
    the identity).  MVP x: 4, 12, 7 -> closest pair (c, a): 5; y: -8, 6, -3 -> b alone
    positive: (c + a) / 2 = -5, taken from c, so pred_sad_space = nb_sad[2].  2NxN's
    upper layer is 2Nx2N: mvp1 = best_up[2], pred_sad_uplayer = 3001 / 2 = 1500, and
    with ref 0 of a P slice and a mode other than 2Nx2N, pred_sad = 1500.  Co-located
    (10, -7) on a reference at distance 4, current distance 2: (int16)(20 / 4.0f) = 5,
    (int16)(-14 / 4.0f) = -3 (truncation).  beta = bsize[2] / 1500^2 - alpha[2]."""
    j = O.umh_job(mvp_job(), mode=2, nb_sad=(111, 222, 333), mincost_up=(900, 800, 3001),
                  best_up=((1, 2), (3, 4), (-5, 6)), col=(1, (10, -7), 4))
    b = bsize_p33()
    o = O.umh_prepare(j, P_SLICE, DPOC, MULTI, b)
    assert (o.status, o.pmv_x, o.pmv_y, o.pred_sad_space) == (0, 5, -5, 333)
    c = o.ctx
    assert (c.mvp1_x, c.mvp1_y, c.pred_sad_uplayer, c.pred_sad, c.ref_idx) == (-5, 6, 1500, 1500, 0)
    assert (c.mvp2_x, c.mvp2_y, c.mvp3_x, c.mvp3_y) == (0, 0, 5, -3)
    assert c.beta2 == b[2] / 2250000 - float(np.float32(0.01))
    assert c.beta3 == b[2] / 2250000 - float(np.float32(0.07))


@pytest.mark.parametrize("mode,up", [(4, 0), (5, 0), (6, 1), (7, 1), (3, 2), (1, None)])
def test_upper_layer_of_each_mode(mode, up):
    """2NxnU / 2NxnD take 2NxN's result, nLx2N / nRx2N take Nx2N's, 2NxN / Nx2N take
    2Nx2N's, and 2Nx2N has none (mvp1 (0, 0), uplayer 0)."""
    best = ((11, 12), (21, 22), (31, 32))
    j = O.umh_job(mvp_job(), mode=mode, mincost_up=(501, 701, 901), best_up=best)
    o = O.umh_prepare(j, P_SLICE, DPOC, MULTI, bsize_p33())
    if up is None:
        assert (o.ctx.mvp1_x, o.ctx.mvp1_y, o.ctx.pred_sad_uplayer) == (0, 0, 0)
    else:
        assert (o.ctx.mvp1_x, o.ctx.mvp1_y) == best[up]
        assert o.ctx.pred_sad_uplayer == (501, 701, 901)[up] // 2


def test_nearest_reference_scaling_and_p_slice_pred_sad():
    """P slice, ref 2 (distance 6), the nearest reference ref 1 (distance 4) found
    (7, -9): mvp2 = (int16)(42 / 4.0f) = 10 and (int16)(-54 / 4.0f) = -13; pred_sad is
    the nearest reference's cost 4000 whatever the mode; no co-located MV."""
    j = O.umh_job(mvp_job(ref_idx=2), mode=1, nb_sad=(5, 6, 7), mincost_prev=4000, best_prev=(7, -9))
    o = O.umh_prepare(j, P_SLICE, DPOC, MULTI, bsize_p33())
    c = o.ctx
    assert (c.mvp2_x, c.mvp2_y, c.pred_sad, c.ref_idx, c.mvp3_x, c.mvp3_y) == (10, -13, 4000, 2, 0, 0)


def test_b_slice_forward_reference_negates_the_backward_mv():
    """B slice, ref_idx 1 (B_FWD): mvp2 is the negated B_BWD result; pred_sad comes
    from the mode rule (2Nx2N: pred_sad_space), not the reference's cost."""
    j = O.umh_job(mvp_job(ref_idx=1, refs=(0, -1, -1, -1)), mode=1, nb_sad=(77, 88, 99), mincost_prev=5000,
                  best_prev=(13, -4))
    o = O.umh_prepare(j, B_SLICE, [2, 6, 0, 0], [8192, 2731, 0, 0], bsize_p33())
    c = o.ctx
    # only the left neighbour: MVP_LEFT, sad of the left block
    assert o.pred_sad_space == 77
    assert (c.mvp2_x, c.mvp2_y, c.pred_sad) == (-13, 4, 77)


def test_pred_sad_square_wraps_in_int32():
    """pred_sad * pred_sad is an int (dist_t) product: 50000^2 = 2.5e9 wraps to
    2.5e9 - 2^32 = -1794967296."""
    j = O.umh_job(mvp_job(ref_idx=1), mode=1, mincost_prev=50000, best_prev=(0, 0))
    b = bsize_p33()
    o = O.umh_prepare(j, P_SLICE, DPOC, MULTI, b)
    assert o.ctx.pred_sad == 50000
    assert o.ctx.beta2 == b[1] / -1794967296 - float(np.float32(0.01))


def test_zero_pred_sad_gives_zero_betas_and_col_without_reference():
    j = O.umh_job(mvp_job(), mode=1, nb_sad=(0, 0, 0), col=(-1, (9, 9), 2))
    o = O.umh_prepare(j, P_SLICE, DPOC, MULTI, bsize_p33())
    assert (o.ctx.pred_sad, o.ctx.beta2, o.ctx.beta3, o.ctx.mvp3_x, o.ctx.mvp3_y) == (0, 0.0, 0.0, 0, 0)


def test_refused_indices_and_zero_distances():
    assert O.umh_prepare(O.umh_job(mvp_job(ref_idx=4), mode=1), P_SLICE, DPOC, MULTI, bsize_p33()).status == -2
    assert O.umh_prepare(O.umh_job(mvp_job(), mode=8), P_SLICE, DPOC, MULTI, bsize_p33()).status == -2
    j = O.umh_job(mvp_job(ref_idx=1), mode=1, mincost_prev=10, best_prev=(1, 1))
    assert O.umh_prepare(j, P_SLICE, [0, 4, 6, 8], MULTI, bsize_p33()).status == -3
