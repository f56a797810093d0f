"""GPU parity of the batched and per-call pixel primitives (SAD, SATD, SAD_x3/x4)
and of luma/chroma motion compensation, against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from xavs2_amd import lib as X
from xavs2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("satd", [False, True])
def test_batched_sad_satd_all_sizes(satd):
    rng = np.random.default_rng(5)
    fenc = rng.integers(0, 256, (96, 128), dtype=np.uint8)
    fref = rng.integers(0, 256, (160, 200), dtype=np.uint8)
    fenc[:64, :64] = 255
    d_fenc, d_fref = X.upload(fenc), X.upload(fref)
    for part in range(25):
        w, h = X.PU_W[part], X.PU_H[part]
        n = 40
        o1 = np.array([(int(rng.integers(0, 96 - h + 1)) * 128 + int(rng.integers(0, 128 - w + 1))) for _ in range(n)],
                      dtype=np.int64)
        o2 = np.array([(int(rng.integers(0, 160 - h + 1)) * 200 + int(rng.integers(0, 200 - w + 1))) for _ in range(n)],
                      dtype=np.int64)
        o2[0] = 0
        d1, d2, out = X.upload(o1), X.upload(o2), X.DeviceBuffer(4 * n)
        f = X.lib().xme_satd_batch if satd else X.lib().xme_sad_batch
        X.check(f(part, d_fenc.at(), 128, d_fref.at(), 200, d1.at(), d2.at(), n, out.at(), None), "batch")
        got = out.download(n, np.int32)
        ofn = O.satd if satd else O.sad
        for i in range(n):
            y1, x1 = divmod(int(o1[i]), 128)
            y2, x2 = divmod(int(o2[i]), 200)
            exp = ofn(w, h, fenc[y1:y1 + h, x1:x1 + w], fref[y2:y2 + h, x2:x2 + w])
            assert got[i] == exp, (part, i)


def test_batched_sad_x3_x4():
    rng = np.random.default_rng(6)
    fenc = rng.integers(0, 256, (64, 64), dtype=np.uint8)
    fref = rng.integers(0, 256, (200, 256), dtype=np.uint8)
    d_fenc, d_fref = X.upload(fenc), X.upload(fref)
    for part in (0, 2, 4, 13, 14, 19, 24):
        w, h = X.PU_W[part], X.PU_H[part]
        for nr, fn in ((3, X.lib().xme_sad_x3_batch), (4, X.lib().xme_sad_x4_batch)):
            n = 9
            o1 = np.zeros(n, np.int64)
            pos = [(int(rng.integers(0, 200 - h)), int(rng.integers(0, 256 - w))) for _ in range(n * nr)]
            o2 = np.array([y * 256 + x for y, x in pos], np.int64)
            d1, d2, out = X.upload(o1), X.upload(o2), X.DeviceBuffer(4 * n * nr)
            X.check(fn(part, d_fenc.at(), d_fref.at(), 256, d1.at(), d2.at(), n, out.at(), None), "xN")
            got = out.download(n * nr, np.int32)
            for i in range(n * nr):
                y, x = pos[i]
                assert got[i] == O.sad(w, h, fenc[:h, :w], fref[y:y + h, x:x + w])


def test_per_call_tables_match_oracle():
    """The exact-signature entry points an encoder installs into g_funcs."""
    L = X.lib()
    sad_t = (C.c_void_p * 25)()
    satd_t = (C.c_void_p * 25)()
    x3_t = (C.c_void_p * 25)()
    x4_t = (C.c_void_p * 25)()
    for g, t in ((L.xme_get_sad_table, sad_t), (L.xme_get_satd_table, satd_t), (L.xme_get_sad_x3_table, x3_t),
                 (L.xme_get_sad_x4_table, x4_t)):
        assert g(t) == 0
    cmp_t = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_ssize_t, C.c_void_p, C.c_ssize_t)
    x3_ft = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_ssize_t, C.c_void_p)
    x4_ft = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_ssize_t, C.c_void_p)
    rng = np.random.default_rng(8)
    fenc = rng.integers(0, 256, (64, 64), dtype=np.uint8)          # FENC_STRIDE layout
    ref = rng.integers(0, 256, (160, 240), dtype=np.uint8)
    for part in range(25):
        w, h = X.PU_W[part], X.PU_H[part]
        sad = cmp_t(sad_t[part])
        satd = cmp_t(satd_t[part])
        p2 = C.c_void_p(ref.ctypes.data + 7 * 240 + 13)
        assert sad(O.ptr(fenc), 64, p2, 240) == O.sad(w, h, fenc[:h, :w], ref[7:7 + h, 13:13 + w])
        assert satd(O.ptr(fenc), 64, p2, 240) == O.satd(w, h, fenc[:h, :w], ref[7:7 + h, 13:13 + w])
        offs = [(3, 5), (40, 90), (17, 9), (30, 100)]
        ptrs = [C.c_void_p(ref.ctypes.data + y * 240 + x) for y, x in offs]
        r3 = np.zeros(3, np.int32)
        r4 = np.zeros(4, np.int32)
        x3_ft(x3_t[part])(O.ptr(fenc), *ptrs[:3], 240, O.ptr(r3))
        x4_ft(x4_t[part])(O.ptr(fenc), *ptrs, 240, O.ptr(r4))
        exp = [O.sad(w, h, fenc[:h, :w], ref[y:y + h, x:x + w]) for y, x in offs]
        assert list(r3) == exp[:3] and list(r4) == exp


def test_mc_luma_and_chroma_batches():
    w, h = 192, 128
    img = synth.smooth_field(w, h, seed=31)
    dref = X.DeviceRef(w, h, frac=2)
    dref.load(img)
    refp = O.Plane(w, h)
    refp.set_picture(img)
    refp.expand_border()
    planes = O.interp_planes(refp, 2)
    cu = synth.smooth_field(w // 2, h // 2, seed=32)
    cv = synth.smooth_field(w // 2, h // 2, seed=33)
    pu, pv = O.Plane(w // 2, h // 2, pad=40), O.Plane(w // 2, h // 2, pad=40)
    pu.set_picture(cu)
    pu.expand_border()
    pv.set_picture(cv)
    pv.expand_border()
    du, dv = X.DevicePlane(w // 2, h // 2, pad=40), X.DevicePlane(w // 2, h // 2, pad=40)
    du.upload_padded(pu.view()[: h // 2 + 80, : w // 2 + 80])
    dv.upload_padded(pv.view()[: h // 2 + 80, : w // 2 + 80])
    rng = np.random.default_rng(2)
    blocks = []
    for part in (1, 2, 3, 4, 19, 20, 23, 24, 0, 13):
        bw, bh = X.PU_W[part], X.PU_H[part]
        for _ in range(6):
            x = int(rng.integers(0, (w - bw) // 4 + 1)) * 4
            y = int(rng.integers(0, h - bh + 1))
            blocks.append((x, y, bw, bh, int(rng.integers(-120, 121)), int(rng.integers(-120, 121))))
    blocks.append((0, 0, 16, 16, -4 * 200 + 3, 4 * 300 + 5))          # clamped far outside by get_mv_for_mc
    # clamped with the phase bit of value 4 set on each side (cu_get_mc_pos keeps mv & 7)
    blocks.append((0, 0, 16, 16, -795, -394))
    blocks.append((w - 16, h - 16, 16, 16, 166, 4 * 40 + 6))
    blocks.append((8, 4, 8, 8, -793, 4 * 50 + 7))
    n = len(blocks)
    barr = np.array(blocks, np.int32).ravel()
    off = np.arange(n, dtype=np.int64) * 64 * 64
    d_blocks, d_off = X.upload(barr), X.upload(off)
    # luma: from the planes, and from the full-pel plane only (block filters)
    out = X.DeviceBuffer(n * 64 * 64)
    for use_planes in (True, False):
        rd = dref.desc()
        if not use_planes:
            for k in range(1, 16):
                rd.planes[k] = None
        X.check(X.lib().xme_mc_luma_batch(C.byref(rd), d_blocks.at(), n, w, h, out.at(), 64, d_off.at(), None), "luma")
        got = out.download(n * 64 * 64, np.uint8).reshape(n, 64, 64)
        for i, (x, y, bw, bh, mx, my) in enumerate(blocks):
            qx = O.lib().xo_get_mc_pos(w, bw, x, mx)
            qy = O.lib().xo_get_mc_pos(h, bh, y, my)
            exp = np.zeros((bh, bw), np.uint8)
            O.lib().xo_mc_luma(O.ptr(exp), bw, qx, qy, bw, bh, O.plane_ptrs(planes if use_planes else
                                                                          [planes[0]] + [None] * 15), refp.stride)
            assert np.array_equal(got[i, :bh, :bw], exp), (i, use_planes)
    ou, ov = X.DeviceBuffer(n * 64 * 64), X.DeviceBuffer(n * 64 * 64)
    X.check(X.lib().xme_mc_chroma_batch(C.c_void_p(du.origin), C.c_void_p(dv.origin), du.stride, d_blocks.at(), n, w, h,
                                        ou.at(), ov.at(), 64, d_off.at(), None), "chroma")
    gu = ou.download(n * 64 * 64, np.uint8).reshape(n, 64, 64)
    gv = ov.download(n * 64 * 64, np.uint8).reshape(n, 64, 64)
    for i, (x, y, bw, bh, mx, my) in enumerate(blocks):
        qx = O.lib().xo_get_mc_pos(w, bw, x, mx)
        qy = O.lib().xo_get_mc_pos(h, bh, y, my)
        eu = np.zeros((bh // 2, bw // 2), np.uint8)
        ev = np.zeros((bh // 2, bw // 2), np.uint8)
        O.lib().xo_mc_chroma(O.ptr(eu), O.ptr(ev), bw // 2, qx, qy, bw // 2, bh // 2, pu.origin_ptr(), pv.origin_ptr(),
                             pu.stride)
        assert np.array_equal(gu[i, :bh // 2, :bw // 2], eu), i
        assert np.array_equal(gv[i, :bh // 2, :bw // 2], ev), i


def test_mc_bi_batches():
    """xme_mc_luma_bi_batch / xme_mc_chroma_bi_batch: the rounded average (a + b + 1) >> 1
    (pixel.c:844-856) of the oracle's mc_luma / mc_chroma from two references at two MVs
    (rdo.c:1616-1666), luma from the planes and from the full-pel planes alone, chroma with
    widths of 2 and 6 (the reference's pixf.average branch) beside the others."""
    w, h = 192, 128
    imgs = [synth.smooth_field(w, h, seed=41), synth.smooth_field(w, h, seed=42)]
    drefs, oplanes = [], []
    for img in imgs:
        d = X.DeviceRef(w, h, frac=2)
        d.load(img)
        drefs.append(d)
        rp = O.Plane(w, h)
        rp.set_picture(img)
        rp.expand_border()
        oplanes.append((rp, O.interp_planes(rp, 2)))
    cps, dps = [], []
    for seed in (43, 44, 45, 46):
        c = synth.smooth_field(w // 2, h // 2, seed=seed)
        op = O.Plane(w // 2, h // 2, pad=40)
        op.set_picture(c)
        op.expand_border()
        dp = X.DevicePlane(w // 2, h // 2, pad=40)
        dp.upload_padded(op.view()[: h // 2 + 80, : w // 2 + 80])
        cps.append(op)
        dps.append(dp)
    rng = np.random.default_rng(47)
    blocks = []
    for bw, bh in ((8, 8), (16, 16), (32, 32), (64, 64), (32, 8), (8, 32), (4, 8), (12, 16), (16, 12), (64, 16)):
        for _ in range(5):
            x = int(rng.integers(0, (w - bw) // 4 + 1)) * 4
            y = int(rng.integers(0, h - bh + 1))
            mv = [int(rng.integers(-120, 121)) for _ in range(4)]
            if rng.random() < 0.2:
                mv[2], mv[3] = 4 * int(rng.integers(-300, 300)) + 5, 4 * int(rng.integers(-300, 300)) + 3
            blocks.append((x, y, bw, bh, *mv))
    n = len(blocks)
    barr = np.array(blocks, np.int32).ravel()
    off = np.arange(n, dtype=np.int64) * 64 * 64
    d_blocks, d_off = X.upload(barr), X.upload(off)

    def mc_pos(size, blk, pos, mv):
        return O.lib().xo_get_mc_pos(size, blk, pos, mv)

    out = X.DeviceBuffer(n * 64 * 64)
    for use_planes in (True, False):
        rd = [d.desc() for d in drefs]
        if not use_planes:
            for r in rd:
                for k in range(1, 16):
                    r.planes[k] = None
        X.check(X.lib().xme_mc_luma_bi_batch(C.byref(rd[0]), C.byref(rd[1]), d_blocks.at(), n, w, h, out.at(), 64,
                                             d_off.at(), None), "luma bi")
        got = out.download(n * 64 * 64, np.uint8).reshape(n, 64, 64)
        for i, (x, y, bw, bh, m1x, m1y, m2x, m2y) in enumerate(blocks):
            preds = []
            for (rp, planes), (mx, my) in zip(oplanes, ((m1x, m1y), (m2x, m2y))):
                e = np.zeros((bh, bw), np.uint8)
                O.lib().xo_mc_luma(O.ptr(e), bw, mc_pos(w, bw, x, mx), mc_pos(h, bh, y, my), bw, bh,
                                   O.plane_ptrs(planes), rp.stride)
                preds.append(e.astype(np.int32))
            exp = ((preds[0] + preds[1] + 1) >> 1).astype(np.uint8)
            assert np.array_equal(got[i, :bh, :bw], exp), (i, use_planes)
    ou, ov = X.DeviceBuffer(n * 64 * 64), X.DeviceBuffer(n * 64 * 64)
    X.check(X.lib().xme_mc_chroma_bi_batch(C.c_void_p(dps[0].origin), C.c_void_p(dps[1].origin),
                                           C.c_void_p(dps[2].origin), C.c_void_p(dps[3].origin), dps[0].stride,
                                           d_blocks.at(), n, w, h, ou.at(), ov.at(), 64, d_off.at(), None), "chroma bi")
    gu = ou.download(n * 64 * 64, np.uint8).reshape(n, 64, 64)
    gv = ov.download(n * 64 * 64, np.uint8).reshape(n, 64, 64)
    for i, (x, y, bw, bh, m1x, m1y, m2x, m2y) in enumerate(blocks):
        u1, v1 = O.mc_chroma(mc_pos(w, bw, x, m1x), mc_pos(h, bh, y, m1y), bw // 2, bh // 2, cps[0], cps[1])
        u2, v2 = O.mc_chroma(mc_pos(w, bw, x, m2x), mc_pos(h, bh, y, m2y), bw // 2, bh // 2, cps[2], cps[3])
        eu = ((u1.astype(np.int32) + u2 + 1) >> 1).astype(np.uint8)
        ev = ((v1.astype(np.int32) + v2 + 1) >> 1).astype(np.uint8)
        assert np.array_equal(gu[i, :bh // 2, :bw // 2], eu), i
        assert np.array_equal(gv[i, :bh // 2, :bw // 2], ev), i
    # argument checks
    assert X.lib().xme_mc_chroma_bi_batch(None, None, None, None, 64, None, 1, w, h, None, None, 64, None,
                                          None) == X.XME_ERR_ARG
    assert X.lib().xme_mc_chroma_bi_batch(None, None, None, None, 64, None, 0, w, h, None, None, 64, None,
                                          None) == X.XME_OK
