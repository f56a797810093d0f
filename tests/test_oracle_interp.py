"""Pins the oracle's frame interpolation, block filters and chroma MC against the
separable-FIR closed form (with the border clamp model) in numpy."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import npref


def smooth_frame(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    f = 128 + 40 * np.sin(xx / 7.0 + rng.uniform(0, 6)) * np.cos(yy / 5.0 + rng.uniform(0, 6))
    f += rng.normal(0, 12, (h, w))
    return np.clip(np.rint(f), 0, 255).astype(np.uint8)


def build(img, frac=2, lcu_level=6):
    h, w = img.shape
    p = O.Plane(w, h)
    p.set_picture(img)
    p.expand_border()
    return p, O.interp_planes(p, frac, lcu_level)


@pytest.mark.parametrize("wh,lcu", [((64, 48), 6), ((200, 120), 6), ((72, 40), 5), ((40, 24), 4)])
def test_planes_match_closed_form_with_border_clamp(wh, lcu):
    w, h = wh
    img = smooth_frame(w, h, 11 + w)
    img[0, :] = 0; img[-1, :] = 255; img[:, 0] = 255; img[:, -1] = 0   # hard edges at the border
    ref, planes = build(img, 2, lcu)
    pad = O.PAD
    full = ref.view()[0:h + 2 * pad, 0:w + 2 * pad]
    assert np.array_equal(full, npref.clamp_picture(img, pad))
    for k in range(1, 16):
        got = planes[k].view()[0:h + 2 * pad, 0:w + 2 * pad]
        exp = npref.interp_plane_padded(img, k, pad)
        assert np.array_equal(got, exp), f"plane {k} mismatch"


def test_half_pel_only_builds_b_h_j():
    img = smooth_frame(48, 32, 5)
    _, planes_q = build(img, 2)
    ref, planes_h = build(img, 1)
    for k in (2, 8, 10):
        assert np.array_equal(planes_h[k].view(), planes_q[k].view())
    for k in (1, 3, 4, 5, 6, 7, 9, 11, 12, 13, 14, 15):
        assert not planes_h[k].view().any()


def test_flat_frame_gives_flat_planes():
    # the taps of every phase sum to 64, so a flat picture stays flat after rounding
    for v in (0, 1, 77, 254, 255):
        img = np.full((24, 32), v, np.uint8)
        _, planes = build(img, 2)
        for k in range(1, 16):
            assert np.all(planes[k].view()[:24 + 160, :32 + 160] == v)


def test_taps_structure():
    assert all(sum(r) == 64 for r in npref.LUMA_TAPS)
    assert all(sum(r) == 64 for r in npref.CHROMA_TAPS)
    assert list(npref.LUMA_TAPS[3]) == list(npref.LUMA_TAPS[1][::-1])
    assert list(npref.LUMA_TAPS[2]) == list(npref.LUMA_TAPS[2][::-1])
    for k in range(1, 8):
        assert list(npref.CHROMA_TAPS[8 - k]) == list(npref.CHROMA_TAPS[k][::-1])


@pytest.mark.parametrize("w,h", [(4, 4), (8, 8), (16, 12), (64, 64), (8, 32), (24, 32)])
def test_luma_block_filters(w, h):
    rng = np.random.default_rng(w * 100 + h)
    src = rng.integers(0, 256, (h + 20, w + 20), dtype=np.uint8)
    L = O.lib()
    base = O.C.c_void_p(src.ctypes.data + 8 * src.strides[0] + 8)
    for dx in range(4):
        for dy in range(4):
            if dx == 0 and dy == 0:
                continue
            out = np.zeros((h, w), np.uint8)
            if dy == 0:
                L.xo_luma_block_hor(O.ptr(out), w, base, src.strides[0], w, h, dx)
            elif dx == 0:
                L.xo_luma_block_ver(O.ptr(out), w, base, src.strides[0], w, h, dy)
            else:
                L.xo_luma_block_ext(O.ptr(out), w, base, src.strides[0], w, h, dx, dy)
            assert np.array_equal(out, npref.block_filter_luma(src, 8, 8, w, h, dx, dy)), (dx, dy)


def test_block_filters_agree_with_planes_inside_picture():
    img = smooth_frame(96, 80, 3)
    ref, planes = build(img, 2)
    out = np.zeros((16, 16), np.uint8)
    planes_no = [ref] + [None] * 15
    for qx, qy in [(4 * 20 + 1, 4 * 30 + 2), (4 * 40 + 3, 4 * 10 + 3), (4 * 33 + 2, 4 * 44 + 1)]:
        exp = np.zeros((16, 16), np.uint8)
        O.lib().xo_mc_luma(O.ptr(exp), 16, qx, qy, 16, 16, O.plane_ptrs(planes), ref.stride)
        O.lib().xo_mc_luma(O.ptr(out), 16, qx, qy, 16, 16, O.plane_ptrs(planes_no), ref.stride)
        assert np.array_equal(out, exp)


@pytest.mark.parametrize("w,h", [(2, 2), (4, 4), (8, 4), (16, 16), (32, 32), (4, 16), (32, 8)])
def test_chroma_mc_all_phases(w, h):
    rng = np.random.default_rng(w + 7 * h)
    img = rng.integers(0, 256, (48, 56), dtype=np.uint8)
    pu = O.Plane(56, 48, pad=40); pu.set_picture(img); pu.expand_border()
    pv = O.Plane(56, 48, pad=40); pv.set_picture(255 - img); pv.expand_border()
    for fx in range(8):
        for fy in range(8):
            qx, qy = 8 * 10 + fx, 8 * 12 + fy
            du, dv = O.mc_chroma(qx, qy, w, h, pu, pv)
            xu = npref.block_filter_chroma(pu.view(), 40 + 10, 40 + 12, w, h, fx, fy)
            xv = npref.block_filter_chroma(pv.view(), 40 + 10, 40 + 12, w, h, fx, fy)
            assert np.array_equal(du, xu) and np.array_equal(dv, xv), (fx, fy)


def test_get_mc_pos_clamps_far_mvs():
    L = O.lib()
    assert L.xo_get_mc_pos(1920, 16, 100, 37) == 4 * 100 + 37
    assert L.xo_get_mc_pos(1920, 16, 100, -4 * 200 + 3) == 4 * (-16 - 8) + ((-800 + 3) & 7)
    assert L.xo_get_mc_pos(1920, 16, 1900, 4 * 40 + 1) == 4 * (1920 + 4) + 1


def test_get_mc_pos_keeps_the_eighth_pel_phase_when_clamping():
    """cu_get_mc_pos (mc.h:52-66) keeps mv & 7 when it clamps: the bit of value 4 survives
    (chroma reads it as half a chroma pixel).  Every value worked by hand."""
    L = O.lib()
    # left: -795 >> 2 = -199, 100 - 199 < -16 - 8: (-24 << 2) + (-795 & 7 = 5)
    assert L.xo_get_mc_pos(1920, 16, 100, -795) == -91
    assert L.xo_get_mc_pos(1920, 16, 100, -793) == -89          # -793 & 7 = 7
    # top: -396 >> 2 = -99 < -16: (-16 << 2) + (-396 & 7 = 4)
    assert L.xo_get_mc_pos(1080, 8, 0, -396) == -60
    # exactly at the left limit (-24 is not < -24): not clamped; one qpel further is
    assert L.xo_get_mc_pos(1920, 16, 0, -96) == -96
    assert L.xo_get_mc_pos(1920, 16, 0, -97) == -89             # -97 >> 2 = -25: -96 + (-97 & 7 = 7)
    # right: 166 >> 2 = 41, 1941 > 1924: (1924 << 2) + (166 & 7 = 6)
    assert L.xo_get_mc_pos(1920, 16, 1900, 166) == 7702
    # bottom: 127 >> 2 = 31, 1101 > 1084: (1084 << 2) + 7
    assert L.xo_get_mc_pos(1080, 32, 1070, 127) == 4343


def test_chroma_mc_at_a_clamped_position():
    """rdo_get_pred_inter (rdo.c:1616-1659) passes the clamped luma qpel position to
    mc_chroma, which splits it into a chroma integer position (>> 3) and a 1/8-pel phase
    (& 7).  Luma 112x96 (chroma 56x48, pad 40), 16x16 PU at (0, 0), MV (-795, -394):
    positions (-91, -90) by hand, i.e. chroma integer (-12, -12), phases (5, 6)."""
    L = O.lib()
    qx = L.xo_get_mc_pos(112, 16, 0, -795)
    qy = L.xo_get_mc_pos(96, 16, 0, -394)
    assert (qx, qy) == (-91, -90)
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (48, 56), dtype=np.uint8)
    pu = O.Plane(56, 48, pad=40); pu.set_picture(img); pu.expand_border()
    pv = O.Plane(56, 48, pad=40); pv.set_picture(255 - img); pv.expand_border()
    du, dv = O.mc_chroma(qx, qy, 8, 8, pu, pv)
    assert np.array_equal(du, npref.block_filter_chroma(pu.view(), 40 - 12, 40 - 12, 8, 8, 5, 6))
    assert np.array_equal(dv, npref.block_filter_chroma(pv.view(), 40 - 12, 40 - 12, 8, 8, 5, 6))
