"""Pins the oracle's MV predictor (get_mvp_default, encoder/md_inter.c:470-524)
on hand-worked cases: each branch of the MVP type rule (:333-368), the
top-left standing in for an unavailable top-right, each sign rule and the
closest-pair rule of the median (:372-401) with its tie order and C's
truncating division, the neighbour scaling scale_mv_default (:405-410) with
its sign-symmetric rounding, the B-slice distances of the list (B_BWD = 0,
B_FWD = 1), invalid neighbours counting as (0, 0), and the overflow flag."""
import pytest

from oracle import oracle as O

# P-slice distances: ref 0 at 2, ref 1 at 4; ref_dpoc_multi = 16384 / dist
DPOC = [2, 4, 6, 8]
MULTI = [16384 // d for d in DPOC]


def job(ref_idx=0, cb=(0, 0, 16, 16), refs=(0, 0, 0, 0), mvs=((0, 0),) * 4, ur=1, bwd=0):
    return dict(ref_idx=ref_idx, bwd_2nd=bwd, cb=cb, ur_available=ur, ref=list(refs), mv=[list(m) for m in mvs])


def mvp(j, b_slice=0, dpoc=DPOC, multi=MULTI):
    rc, x, y = O.get_mvp_default(j, b_slice, dpoc, multi)
    assert rc == 0
    return x, y


L, U, UL, UR = (4, -8), (12, 6), (-20, 30), (7, -3)


@pytest.mark.parametrize("refs,expected", [
    ((0, -1, -1, -1), L),          # only the left is valid
    ((-1, 0, -1, -1), U),          # only the top
    ((-1, -1, -1, 0), UR),         # only the top-right
])
def test_single_valid_neighbour_is_taken(refs, expected):
    assert mvp(job(refs=refs, mvs=(L, U, UL, UR))) == expected


def test_unavailable_top_right_uses_top_left():
    # L and U invalid; the top-right is unavailable, so the top-left (valid) is c
    assert mvp(job(refs=(-1, -1, 0, -1), mvs=(L, U, UL, UR), ur=0)) == UL
    # the same neighbours with the top-right available: its invalid ref makes a median of zeros
    assert mvp(job(refs=(-1, -1, 0, -1), mvs=(L, U, UL, UR), ur=1)) == (0, 0)


@pytest.mark.parametrize("cb,refs,expected", [
    ((0, 0, 8, 16), (0, 1, 1, 1), L),       # narrow PU at the CU's left: left if it uses ref_idx
    ((8, 0, 8, 16), (1, 1, 1, 0), UR),      # narrow PU on the right: top-right if it uses ref_idx
    ((0, 0, 16, 8), (1, 0, 1, 1), U),       # flat PU at the top: top
    ((0, 8, 16, 8), (0, 1, 1, 1), L),       # flat PU below: left
])
def test_rectangular_pu_rules(cb, refs, expected):
    # neighbours on ref 1 are scaled by dist(ref 0) / dist(ref 1) = 2 / 4
    got = mvp(job(ref_idx=0, cb=cb, refs=refs, mvs=(L, U, UL, UR)))
    assert got == expected


def test_rectangular_pu_falls_back_to_the_median():
    # narrow PU at the left whose left neighbour is on another reference: median of
    # a = L, b = U, c = UR, all on ref 0 (no scaling)
    got = mvp(job(ref_idx=1, cb=(0, 0, 8, 16), refs=(0, 0, 0, 0), mvs=(L, U, UL, UR)))
    # ref_idx 1: dist 4, neighbours on ref 0 (multi 8192): every MV doubles,
    # a = (8, -16), b = (24, 12), c = (14, -6)
    # x: 8, 24, 14 all positive: |a-b| 16, |b-c| 10, |c-a| 6 -> (c + a) / 2 = 11
    # y: -16 < 0 with 12 > 0 and -6 < 0: b alone positive -> (c + a) / 2 = -11
    assert got == (11, -11)


@pytest.mark.parametrize("a,b,c,expected", [
    # each sign rule against a closest pair that would give 0
    (-1, 20, 2, 11),     # a alone negative: (b + c) / 2
    (1, -20, -2, -11),   # a alone positive
    (20, -1, 2, 11),     # b alone negative: (c + a) / 2
    (-20, 1, -2, -11),   # b alone positive
    (20, 2, -1, 11),     # c alone negative: (a + b) / 2
    (-20, -2, 1, -11),   # c alone positive
    (1, 10, 4, 2),       # closest pair c, a: (4 + 1) / 2 truncates to 2
    (0, 2, 4, 1),        # tie |a-b| = |b-c|: a, b wins
    (-3, 0, -10, -1),    # (a + b) / 2 = -3 / 2 truncates toward zero
])
def test_median_rules(a, b, c, expected):
    # square PU, all three on the searched reference: no scaling (2 * 8192 = 16384)
    got = mvp(job(refs=(0, 0, 0, 0), mvs=((a, 0), (b, 0), (0, 0), (c, 0))))
    assert got == (expected, 0)


@pytest.mark.parametrize("v,expected", [(5, 10), (-5, -10), (3, 6), (-3, -6), (0, 0)])
def test_scaling_is_sign_symmetric(v, expected):
    # target ref 1 (dist 4), left neighbour on ref 0 (multi 8192):
    # (|v| * 4 * 8192 + 8192) >> 14 = 2|v| + 0.5 floored, sign restored
    got = mvp(job(ref_idx=1, refs=(0, -1, -1, -1), mvs=((v, v), U, UL, UR)))
    assert got == (expected, expected)


def test_b_slice_uses_the_lists_distance():
    dpoc = [2, 6, 0, 0]                      # B_BWD = 0 at 2, B_FWD = 1 at 6
    multi = [8192, 2731, 0, 0]
    j = job(refs=(0, -1, -1, -1), mvs=((10, -10), U, UL, UR))
    # forward list (bwd_2nd 0) uses index 1: (10 * 6 * 2731 + 8192) >> 14 = 10
    assert mvp(j, 1, dpoc, multi) == (10, -10)
    # backward list uses index 0: (10 * 2 * 8192 + 8192) >> 14 = 10
    assert mvp(dict(j, bwd_2nd=1), 1, dpoc, multi) == (10, -10)
    # a forward distance of 3 with the backward's multi halves it: (10 * 3 * 2731 + 8192) >> 14 = 5
    assert mvp(j, 1, [2, 3, 0, 0], multi) == (5, -5)


def test_invalid_neighbours_count_as_zero_in_the_median():
    # left invalid; top and top-right valid: square PU -> median of (0, U, UR)
    got = mvp(job(refs=(-1, 0, 0, 0), mvs=((99, 99), (12, 6), UL, (8, -3))))
    # x: 0, 12, 8: |a-b| 12, |b-c| 4, |c-a| 8 -> (12 + 8) / 2 = 10
    # y: 0, 6, -3: c alone negative?  a = 0 is neither, so the closest pair:
    #    |a-b| 6, |b-c| 9, |c-a| 3 -> (-3 + 0) / 2 = -1
    assert got == (10, -1)


def test_scaling_overflow_is_flagged():
    rc, _, _ = O.get_mvp_default(job(ref_idx=0, refs=(1, -1, -1, -1), mvs=((30000, 0), U, UL, UR)), 0,
                                 [512] * 4, [16384] * 4)
    assert rc == -1


@pytest.mark.parametrize("v,expected", [(5, 3), (-5, -3), (3, 2), (-3, -2), (1, 1), (-1, -1)])
def test_scaling_rounds_exact_halves_away_from_zero(v, expected):
    # target ref 0 at distance 1, left neighbour on ref 1 (multi 8192): the scaled
    # value is v / 2, an exact half for odd v; sign * ((|v| * 8192 + 8192) >> 14)
    # rounds it away from zero, where a floor of the signed product would not
    dpoc = [1, 2, 3, 4]
    multi = [16384, 8192, 5461, 4096]
    got = mvp(job(ref_idx=0, refs=(1, -1, -1, -1), mvs=((v, -v), U, UL, UR)), 0, dpoc, multi)
    assert got == (expected, -expected)


def test_left_with_a_valid_top_right_is_a_median():
    # left and top-right valid, top invalid: none of the single-neighbour rules
    # applies, so a square PU takes the median of (L, (0, 0), UR)
    got = mvp(job(refs=(0, -1, -1, 0), mvs=(L, U, UL, UR)))
    # x: 4, 0, 7: |a-b| 4, |b-c| 7, |c-a| 3 -> (7 + 4) / 2 = 5
    # y: -8, 0, -3: |a-b| 8, |b-c| 3, |c-a| 5 -> (0 - 3) / 2 = -1
    assert got == (5, -1)


@pytest.mark.parametrize("kw", [dict(ref_idx=4), dict(refs=(4, 0, 0, 0)), dict(refs=(0, -2, 0, 0)),
                                dict(refs=(0, 0, 0, 4), ur=1), dict(refs=(0, 0, 4, 0), ur=0)])
def test_reference_indices_outside_the_lists_are_refused(kw):
    # the lists hold MAX_REFS = 4 entries (defines.h:364) and predict.h:100 asserts blkref < MAX_REFS
    rc, _, _ = O.get_mvp_default(job(**kw), 0, DPOC, MULTI)
    assert rc == -2


def test_unused_neighbour_index_is_not_checked():
    # with the top-right available the top-left is not read
    rc, _, _ = O.get_mvp_default(job(refs=(0, 0, 7, 0), ur=1), 0, DPOC, MULTI)
    assert rc == 0
