"""GPU parity of the row-wise reference-plane producer: xme_expand_border_lcu_row
(xavs2_frame_expand_border_lcurow, common/frame.c:537-575) and
xme_interpolate_lcu_row (interpolate_lcu_row, common/mc.c:748-785), run in
row order as the encoder runs them (encoder/slice.c:535-543), against the
oracle's whole-frame planes:

* the picture's rows become final one LCU row at a time (rows below stay
  garbage, every pad and plane starts as garbage); after each row, every
  plane equals the whole-frame result on all rows above the row's window
  end, and nothing below the window has been written;
* LCU levels 3 to 6, use_fractional_me 1 and 2, pictures of one and of many
  LCU rows;
* argument errors.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from xavs2_amd import lib as X

pytestmark = pytest.mark.gpu

GARBAGE_SRC, GARBAGE_PL = 0xA5, 0x5A


@pytest.mark.parametrize("w,h,level,frac", [(64, 48, 4, 2), (200, 120, 6, 2), (192, 136, 3, 2), (96, 72, 5, 1),
                                            (64, 64, 6, 2)])
def test_rows_in_order_match_the_frame_planes(w, h, level, frac):
    rng = np.random.default_rng(w * h + level)
    img = rng.integers(0, 256, (h, w), dtype=np.uint8)
    full = O.Plane(w, h)
    full.set_picture(img)
    full.expand_border()
    exp = O.interp_planes(full, frac, lcu_level=level)
    pad = X.XME_PAD
    ref = X.DeviceRef(w, h, frac=frac)
    shape = (h + 2 * pad, w + 2 * pad)
    ref.planes[0].upload_padded(np.full(shape, GARBAGE_SRC, np.uint8))
    for p in ref.planes[1:]:
        if p is not None:
            p.upload_padded(np.full(shape, GARBAGE_PL, np.uint8))
    rows = (h + (1 << level) - 1) >> level
    for r in range(rows):
        # LCU row r of the picture becomes final
        host = ref.planes[0].download_padded().copy()
        y_end = min(h, (r + 1) << level)
        host[pad + (r << level): pad + y_end, pad: pad + w] = img[(r << level): y_end]
        ref.planes[0].upload_padded(host)
        ref.planes[0].expand_border_lcu_row(level, r)
        ref.interpolate_lcu_row(level, r)
        _, b1 = O.border_row_window(h, level, r)
        _, i1 = O.lcu_row_window(h, level, r)
        last = r == rows - 1
        got0 = ref.planes[0].download_padded()
        top = h + 2 * pad if last else b1 + pad
        assert np.array_equal(got0[:top], exp[0].view()[:top, : w + 2 * pad]), f"full-pel plane after row {r}"
        top = h + 2 * pad if last else i1 + pad
        for k in range(1, 16):
            if ref.planes[k] is None:
                continue
            got = ref.planes[k].download_padded()
            ev = exp[k].view()[:, : w + 2 * pad]
            bad = np.argwhere(got[:top] != ev[:top])
            assert len(bad) == 0, f"plane {k} after row {r}: first mismatch at (y, x) = {tuple(bad[0] - pad)}"
            assert np.all(got[top:] == GARBAGE_PL), f"plane {k} written below row {r}'s window"


def test_argument_errors():
    w, h = 64, 48
    ref = X.DeviceRef(w, h, frac=2)
    L = X.lib()
    arr = (C.c_void_p * 16)(*[p.origin for p in ref.planes])
    p0 = ref.planes[0]
    assert L.xme_interpolate_lcu_row(arr, ref.stride, w, h, 2, 4, 0, None) == X.XME_OK
    for lvl, row in ((2, 0), (7, 0), (4, -1), (4, 3), (6, 1)):
        assert L.xme_interpolate_lcu_row(arr, ref.stride, w, h, 2, lvl, row, None) == X.XME_ERR_ARG, (lvl, row)
    assert L.xme_expand_border_lcu_row(C.c_void_p(p0.origin), p0.stride, w, h, p0.pad, 4, 2, None) == X.XME_OK
    for lg, row, pad in ((1, 0, 80), (8, 0, 80), (4, 3, 80), (4, 0, 82)):
        assert L.xme_expand_border_lcu_row(C.c_void_p(p0.origin), p0.stride, w, h, pad, lg, row,
                                           None) == X.XME_ERR_ARG, (lg, row, pad)
    assert L.xme_stream_sync(None) == X.XME_OK
