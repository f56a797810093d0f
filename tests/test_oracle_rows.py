"""Pins the oracle's row-wise reference-plane production, as an encoder
that finishes a picture LCU row by LCU row runs it (encoder/slice.c:535-543,
one slice): xavs2_frame_expand_border_lcurow (common/frame.c:537-575) and
interpolate_lcu_row (common/mc.c:748-785).

* the row windows on hand-worked values from the reference's constants
  (UP_SHIFT 4; PAD_OFFSET 4, MC_OFFSET 8), for luma and 4:2:0 chroma;
* the windows of all LCU rows tile [0, h) and [-4, h + 4) without overlap;
* run in row order on a picture whose rows below the current LCU row are
  not final yet (garbage) and whose pads start as garbage, every row leaves
  all planes equal to the whole-frame expansion and interpolation on every
  row above its window's end: the windows only ever read rows that the same
  or an earlier call has made final.
"""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("h,level,expect", [
    (1080, 6, {0: (-4, 56), 1: (56, 120), 15: (952, 1016), 16: (1016, 1084)}),
    (64, 6, {0: (-4, 68)}),
    (24, 3, {0: (-4, 0), 1: (0, 8), 2: (8, 28)}),
    (2160, 5, {0: (-4, 24), 1: (24, 56), 67: (2136, 2164)}),
])
def test_interpolation_row_windows(h, level, expect):
    for row, win in expect.items():
        assert O.lcu_row_window(h, level, row) == win


@pytest.mark.parametrize("h,row_log2,expect", [
    (1080, 6, {0: (0, 60), 1: (60, 124), 16: (1020, 1080)}),     # luma, 64x64 LCUs
    (540, 5, {0: (0, 28), 1: (28, 60), 16: (508, 540)}),         # its 4:2:0 chroma
    (64, 6, {0: (0, 64)}),
])
def test_border_row_windows(h, row_log2, expect):
    for row, win in expect.items():
        assert O.border_row_window(h, row_log2, row) == win


@pytest.mark.parametrize("h,level", [(1080, 6), (720, 6), (544, 5), (200, 4), (72, 3), (64, 6), (8, 3)])
def test_windows_tile_the_rows(h, level):
    rows = (h + (1 << level) - 1) >> level
    iw = [O.lcu_row_window(h, level, r) for r in range(rows)]
    bw = [O.border_row_window(h, level, r) for r in range(rows)]
    assert iw[0][0] == -4 and iw[-1][1] == h + 4
    assert bw[0][0] == 0 and bw[-1][1] == h
    for k in range(rows - 1):
        assert iw[k][1] == iw[k + 1][0] and bw[k][1] == bw[k + 1][0]


@pytest.mark.parametrize("w,h,level,frac", [(64, 48, 4, 2), (96, 72, 3, 2), (128, 64, 6, 2), (80, 56, 4, 1)])
def test_rows_in_order_reach_the_frame_planes(w, h, level, frac):
    rng = np.random.default_rng(w + h + level)
    img = rng.integers(0, 256, (h, w), dtype=np.uint8)
    # the whole-frame result
    full = O.Plane(w, h)
    full.set_picture(img)
    full.expand_border()
    exp = O.interp_planes(full, frac, lcu_level=level)
    # row by row on planes that start as garbage
    src = O.Plane(w, h)
    src.view()[...] = 0xA5
    planes = [src] + [O.Plane(w, h, src.pad, src.stride) for _ in range(15)]
    for p in planes[1:]:
        p.view()[...] = 0x5A
    if frac == 1:
        planes = [p if k in (0, 2, 8, 10) else None for k, p in enumerate(planes)]
    pad = src.pad
    rows = (h + (1 << level) - 1) >> level
    for r in range(rows):
        # rows of LCU row r become final; rows below stay garbage
        y_end = min(h, (r + 1) << level)
        src.picture()[(r << level):y_end] = img[(r << level):y_end]
        O.expand_border_lcu_row(src, level, r)
        O.interpolate_lcu_row(planes, level, r, frac)
        _, b1 = O.border_row_window(h, level, r)
        top = b1 + pad if r < rows - 1 else h + 2 * pad
        assert np.array_equal(src.view()[:top, : w + 2 * pad], exp[0].view()[:top, : w + 2 * pad]), r
        _, i1 = O.lcu_row_window(h, level, r)
        top = i1 + pad if r < rows - 1 else h + 2 * pad
        for k in range(1, 16):
            if planes[k] is None:
                continue
            assert np.array_equal(planes[k].view()[:top, : w + 2 * pad], exp[k].view()[:top, : w + 2 * pad]), (r, k)
