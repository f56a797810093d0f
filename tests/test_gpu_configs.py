"""Bit-exact parity with the oracle at the full sizes of the 1080p workloads:

* integer full search, 1920x1080, 16x16 PUs, [-32, 32)^2, lambda factor
  4 * 65536, MVP (0, 0), current picture = the reference with every 16x16 block
  displaced by an integer MV in [-24, 24]^2 plus noise;
* sub-pel refinement, 1920x1080, the current picture built from quarter-pel
  true MVs applied through the AVS2 8-tap planes, PU shapes 8x8, 16x16, 32x32,
  64x64, 64x16, 16x64, 32x8, 8x32, full search [-32, 32)^2 then half- and
  quarter-pel Spiral points with SATD, and chroma MC of the 960x540 planes at
  the chosen MVs;
* full search [-64, 64)^2 + sub-pel over mixed PU shapes, 1920x1024, with a
  random MVP per PU (split search windows).
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu

LAM = 4 * 65536
W, H = 1920, 1080
THREADS = min(16, os.cpu_count() or 1)


def oracle_planes(img):
    p = O.Plane(W, H)
    p.set_picture(img)
    p.expand_border()
    return O.interp_planes(p, 2)


def gpu_run(prm, pus, cur, dref):
    plan = X.Plan(prm, X.make_pus(pus), 1)
    dcur = X.DevicePlane(W, H, pad=0, stride=((W + 63) // 64) * 64)
    dcur.upload_picture(cur)
    plan.run(dcur, [dref])
    return plan.fetch()


def oracle_run(prm_o, pus, cur, planes):
    arr = np.array([tuple(p) for p in pus], dtype=O.PU_DTYPE)
    cp = O.Plane(W, H)
    cp.set_picture(cur)
    return O.me_search_batch(prm_o, arr, cp, planes, nthreads=THREADS)


def assert_same(got, exp, pus):
    for f in X.RESULT_DTYPE.names:
        bad = np.nonzero(got[f] != exp[f])[0]
        assert len(bad) == 0, f"{f}: {len(bad)} mismatches, first PU {pus[bad[0]]}: {got[bad[0]]} vs {exp[bad[0]]}"


def test_integer_full_search_1080p_16x16():
    ref = synth.smooth_field(W, H, seed=100)
    cur, _ = synth.displaced(ref, 16, 24, seed=101)
    pus = [(x, y, 16, 16, 0, 0) for (x, y, _, _) in synth.shape_tiling(W, 1072, 16, 16)]
    assert len(pus) == 8040
    dref = X.DeviceRef(W, H, frac=2)
    dref.load(ref)
    got = gpu_run(X.me_params(W, H, 32, LAM, use_fractional_me=0), pus, cur, dref)
    exp = oracle_run(O.me_param(W, H, 32, LAM, use_fractional_me=0), pus, cur, oracle_planes(ref))
    assert_same(got, exp, pus)


def test_random_mvps_mixed_shapes_1080p_range64():
    """Every PU its own MVP (+-16 pixels): search centres spread over a CTU, so
    CTU windows exceed their LDS share and are split into candidate-row ranges,
    and the short bands of those ranges exercise the lanes pulled back inside
    an item.  Mixed shapes 8x8..64x64, R = 64, sub-pel refinement on."""
    ref = synth.smooth_field(W, H, seed=400)
    cur, _ = synth.displaced(ref, 16, 24, seed=401)
    rng = np.random.default_rng(402)
    pus = [(x, y, pw, ph, int(rng.integers(-64, 65)), int(rng.integers(-64, 65)))
           for (x, y, pw, ph) in synth.mixed_pus(W, 1024, 402)]
    dref = X.DeviceRef(W, H, frac=2)
    dref.load(ref)
    got = gpu_run(X.me_params(W, H, 64, LAM), pus, cur, dref)
    exp = oracle_run(O.me_param(W, H, 64, LAM), pus, cur, oracle_planes(ref))
    assert_same(got, exp, pus)


def qpel_current(planes, block, mv_range, seed):
    """Each block x block tile copies the interpolated plane of a seeded
    quarter-pel MV (pred = plane[(my & 3) * 4 + (mx & 3)] at (x + mx >> 2, y + my >> 2))."""
    rng = np.random.default_rng(seed)
    cur = np.empty((H, W), np.uint8)
    pad = planes[0].pad
    for by in range(0, H, block):
        for bx in range(0, W, block):
            mx, my = (int(v) for v in rng.integers(-4 * mv_range, 4 * mv_range + 1, size=2))
            pl = planes[((my & 3) << 2) + (mx & 3)].view()
            hh, ww = min(block, H - by), min(block, W - bx)
            y0, x0 = pad + by + (my >> 2), pad + bx + (mx >> 2)
            cur[by:by + hh, bx:bx + ww] = pl[y0:y0 + hh, x0:x0 + ww]
    noise = rng.normal(0.0, 3.0, cur.shape)
    return np.clip(np.rint(cur + noise), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("pw,ph", [(8, 8), (16, 16), (32, 32), (64, 64), (64, 16), (16, 64), (32, 8), (8, 32)])
def test_subpel_refinement_1080p(pw, ph):
    ref = synth.smooth_field(W, H, seed=200)
    planes = oracle_planes(ref)
    cur = qpel_current(planes, 16, 24, seed=201 + pw + ph)
    pus = [(x, y, pw, ph, 0, 0) for (x, y, _, _) in synth.shape_tiling(W, H, pw, ph)]
    dref = X.DeviceRef(W, H, frac=2)
    dref.load(ref)
    got = gpu_run(X.me_params(W, H, 32, LAM), pus, cur, dref)
    exp = oracle_run(O.me_param(W, H, 32, LAM), pus, cur, planes)
    assert_same(got, exp, pus)
    if (pw, ph) == (16, 16):
        chroma_at_final_mvs(pus, got)


def chroma_at_final_mvs(pus, res):
    cw, ch = W // 2, H // 2
    u = synth.smooth_field(cw, ch, seed=300)
    v = synth.smooth_field(cw, ch, seed=301)
    pu_o, pv_o = O.Plane(cw, ch, pad=40), O.Plane(cw, ch, pad=40)
    for p, img in ((pu_o, u), (pv_o, v)):
        p.set_picture(img)
        p.expand_border()
    du, dv = X.DevicePlane(cw, ch, pad=40), X.DevicePlane(cw, ch, pad=40)
    du.upload_padded(pu_o.view()[: ch + 80, : cw + 80])
    dv.upload_padded(pv_o.view()[: ch + 80, : cw + 80])
    blocks = np.array([(x, y, w, h, int(r["bmv_x"]), int(r["bmv_y"])) for (x, y, w, h, _, _), r in zip(pus, res)],
                      np.int32)
    n = len(blocks)
    off = np.arange(n, dtype=np.int64) * 64
    d_blocks, d_off = X.upload(blocks.ravel()), X.upload(off)
    ou, ov = X.DeviceBuffer(n * 64), X.DeviceBuffer(n * 64)
    X.check(X.lib().xme_mc_chroma_batch(C.c_void_p(du.origin), C.c_void_p(dv.origin), du.stride, d_blocks.at(), n,
                                        W, H, ou.at(), ov.at(), 8, d_off.at(), None), "chroma")
    gu = ou.download(n * 64, np.uint8).reshape(n, 8, 8)
    gv = ov.download(n * 64, np.uint8).reshape(n, 8, 8)
    for i in range(0, n, 7):
        x, y, w, h, mx, my = (int(t) for t in blocks[i])
        qx = O.lib().xo_get_mc_pos(W, w, x, mx)
        qy = O.lib().xo_get_mc_pos(H, h, y, my)
        eu, ev = O.mc_chroma(qx, qy, w // 2, h // 2, pu_o, pv_o)
        assert np.array_equal(gu[i], eu) and np.array_equal(gv[i], ev), (i, blocks[i])
