"""ctypes binding of oracle/liboracle.so (the CPU restatement, the checker).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PU_W = [4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 12, 16, 4, 32, 24, 32, 8, 64, 48, 64, 16]
PU_H = [4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 12, 16, 4, 16, 24, 32, 8, 32, 48, 64, 16, 64]
PAD = 80
MAX_DISTORTION = 1 << 30
FENC_STRIDE = 64


class MeParam(C.Structure):
    _fields_ = [("frame_w", C.c_int), ("frame_h", C.c_int), ("search_range", C.c_int),
                ("lambda_factor", C.c_int), ("min_mv_range", C.c_int * 2),
                ("max_mv_range", C.c_int * 2), ("use_hadamard", C.c_int),
                ("use_fractional_me", C.c_int), ("use_fast_sub_me", C.c_int),
                ("enable_pmvr", C.c_int), ("me_method", C.c_int), ("umh_big_hex_level", C.c_int)]


# h->param->me_method (common/common.h:145-151)
ME_FS, ME_DIA, ME_HEX, ME_UMH, ME_TZ = 0, 1, 2, 3, 4

# xo_umh_ctx_t: per-(PU, reference) state read by the UMH search
UMH_CTX_DTYPE = np.dtype([("mvp1_x", "<i4"), ("mvp1_y", "<i4"), ("mvp2_x", "<i4"), ("mvp2_y", "<i4"),
                          ("mvp3_x", "<i4"), ("mvp3_y", "<i4"), ("ref_idx", "<i4"), ("pred_sad", "<i4"),
                          ("pred_sad_uplayer", "<i4"), ("pad", "<i4"), ("beta2", "<f8"), ("beta3", "<f8")])


class SymJob(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("pix_x", "pix_y", "w", "h", "mv_x", "mv_y", "mvp_x", "mvp_y",
                                       "mvp1_x", "mvp1_y", "dist_fwd", "dist_bwd", "b_slice")]


class SymResult(C.Structure):
    _fields_ = [("cost", C.c_int32), ("mv_x", C.c_int32), ("mv_y", C.c_int32), ("mvcost", C.c_int32)]


class BidJob(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("pix_x", "pix_y", "w", "h", "fwd_x", "fwd_y", "bwd_x", "bwd_y",
                                       "mvp1_x", "mvp1_y", "mvp2_x", "mvp2_y")]


class GateParam(C.Structure):
    _fields_ = [("lcu_level", C.c_int), ("frame_threads", C.c_int)]


class BiJob(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("pix_x", "pix_y", "w", "h", "fwd_imv_x", "fwd_imv_y", "bwd_x", "bwd_y",
                                       "mvp_x", "mvp_y", "mvp1_x", "mvp1_y", "mvp2_x", "mvp2_y")]


class BiResult(C.Structure):
    _fields_ = [(f, C.c_int32) for f in ("sym_cost", "sym_mv_x", "sym_mv_y", "sym_mvcost", "bid_cost", "fwd_x",
                                         "fwd_y", "bwd_x", "bwd_y", "bid_mvcost")]


class DualJob(C.Structure):
    _fields_ = [("pix_x", C.c_int), ("pix_y", C.c_int), ("w", C.c_int), ("h", C.c_int), ("mvp_x", C.c_int),
                ("mvp_y", C.c_int), ("mvp1", (C.c_int * 2) * 4), ("imv", (C.c_int * 2) * 4)]


class DualResult(C.Structure):
    _fields_ = [("cost", C.c_int32), ("fst_ref", C.c_int32), ("snd_ref", C.c_int32), ("mvcost", C.c_int32),
                ("fst", (C.c_int32 * 2) * 4), ("snd", (C.c_int32 * 2) * 4)]


class MvpJob(C.Structure):
    _fields_ = [("ref_idx", C.c_int), ("bwd_2nd", C.c_int), ("cb_x", C.c_int), ("cb_y", C.c_int),
                ("cb_w", C.c_int), ("cb_h", C.c_int), ("ur_available", C.c_int), ("ref", C.c_int * 4),
                ("mv", (C.c_int * 2) * 4)]


class UmhCtx(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("mvp1_x", "mvp1_y", "mvp2_x", "mvp2_y", "mvp3_x", "mvp3_y", "ref_idx",
                                       "pred_sad", "pred_sad_uplayer", "pad")] + \
               [("beta2", C.c_double), ("beta3", C.c_double)]


class UmhJob(C.Structure):
    _fields_ = [("mvp", MvpJob), ("nb_sad", C.c_int * 3), ("mode", C.c_int), ("mincost_up", C.c_int * 3),
                ("best_up", (C.c_int * 2) * 3), ("mincost_prev", C.c_int), ("best_prev", C.c_int * 2),
                ("col_valid", C.c_int), ("col_ref", C.c_int), ("col_mv", C.c_int * 2), ("col_dpoc", C.c_int)]


class UmhOut(C.Structure):
    _fields_ = [("pmv_x", C.c_int), ("pmv_y", C.c_int), ("pred_sad_space", C.c_int), ("status", C.c_int),
                ("ctx", UmhCtx)]


class Pu(C.Structure):
    _fields_ = [("pix_x", C.c_int), ("pix_y", C.c_int), ("w", C.c_int), ("h", C.c_int),
                ("mvp_x", C.c_int), ("mvp_y", C.c_int)]


class MeResult(C.Structure):
    _fields_ = [("cost", C.c_int32), ("bcost2", C.c_int32), ("bmv_x", C.c_int16),
                ("bmv_y", C.c_int16), ("bmv2_x", C.c_int16), ("bmv2_y", C.c_int16),
                ("mvcost", C.c_int32), ("status", C.c_int32)]


ME_RESULT_DTYPE = np.dtype([("cost", "<i4"), ("bcost2", "<i4"), ("bmv_x", "<i2"), ("bmv_y", "<i2"),
                            ("bmv2_x", "<i2"), ("bmv2_y", "<i2"), ("mvcost", "<i4"), ("status", "<i4")])
PU_DTYPE = np.dtype([("pix_x", "<i4"), ("pix_y", "<i4"), ("w", "<i4"), ("h", "<i4"),
                     ("mvp_x", "<i4"), ("mvp_y", "<i4")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        ip = C.c_ssize_t
        L.xo_sad.restype = C.c_int32
        L.xo_sad.argtypes = [C.c_int, C.c_int, P, ip, P, ip]
        L.xo_satd.restype = C.c_int32
        L.xo_satd.argtypes = [C.c_int, C.c_int, P, ip, P, ip]
        L.xo_sad_x3.argtypes = [C.c_int, C.c_int, P, P, P, P, ip, P]
        L.xo_sad_x4.argtypes = [C.c_int, C.c_int, P, P, P, P, P, ip, P]
        L.xo_avg.argtypes = [C.c_int, C.c_int, P, ip, P, ip, P, ip]
        for f in (L.xo_sad_batch, L.xo_satd_batch):
            f.argtypes = [C.c_int, C.c_int, P, ip, P, ip, P, P, C.c_int, P]
        L.xo_mvbits_max_mvd.argtypes = [C.c_int]
        L.xo_mvbits.argtypes = [C.c_int, C.c_int]
        L.xo_lambda_factor.argtypes = [C.c_double]
        L.xo_me_search.restype = C.c_int32
        L.xo_me_search.argtypes = [C.POINTER(MeParam), C.POINTER(Pu), P, ip, P, ip, P, P]
        L.xo_me_search_batch.argtypes = [C.POINTER(MeParam), P, C.c_int, P, ip, P, ip, P, C.c_int]
        L.xo_me_pattern_path.argtypes = [C.POINTER(MeParam), C.POINTER(Pu), P, P, ip, P, ip, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, P, C.c_int, P]
        L.xo_pattern_landscape.argtypes = [C.POINTER(MeParam), C.POINTER(Pu), P, P, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_int, P, C.c_int, P, P, P, P]
        L.xo_me_pattern_trace.argtypes = [C.POINTER(MeParam), C.POINTER(Pu), P, P, ip, P, ip, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, P, C.c_int, P, P, P, P]
        L.xo_scale_mv_skip.argtypes = [C.c_int, C.c_int, C.c_int]
        L.xo_pred_inter_search_bi.argtypes = [C.POINTER(MeParam), C.POINTER(GateParam), C.POINTER(BiJob), C.c_int,
                                              C.c_int, P, ip, P, P, ip, C.POINTER(BiResult)]
        L.xo_pred_inter_search_dual.argtypes = [C.POINTER(MeParam), C.POINTER(GateParam), C.POINTER(DualJob),
                                                C.c_int, P, P, ip, P, ip, C.POINTER(DualResult)]
        L.xo_me_search_sym.restype = C.c_int32
        L.xo_me_search_sym.argtypes = [C.POINTER(MeParam), C.POINTER(SymJob), P, ip, P, P, ip, C.POINTER(SymResult)]
        L.xo_me_search_bid.restype = C.c_int32
        L.xo_me_search_bid.argtypes = [C.POINTER(MeParam), C.POINTER(BidJob), P, ip, P, P, ip, C.POINTER(SymResult)]
        L.xo_me_search_ctx.restype = C.c_int32
        L.xo_me_search_ctx.argtypes = [C.POINTER(MeParam), C.POINTER(Pu), P, P, ip, P, ip, P]
        L.xo_me_search_batch_ctx.argtypes = [C.POINTER(MeParam), P, P, C.c_int, P, ip, P, ip, P, C.c_int]
        L.xo_expand_border.argtypes = [P, ip, C.c_int, C.c_int, C.c_int]
        L.xo_expand_border_lcu_row.argtypes = [P, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_interpolate_lcu_row.argtypes = [P, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_lcu_row_window.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.xo_border_row_window.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.xo_interpolate_frame.argtypes = [P, ip, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_mc_luma.argtypes = [P, ip, C.c_int, C.c_int, C.c_int, C.c_int, P, ip]
        L.xo_mc_chroma.argtypes = [P, P, ip, C.c_int, C.c_int, C.c_int, C.c_int, P, P, ip]
        for f in (L.xo_luma_block_hor, L.xo_luma_block_ver, L.xo_chroma_block_hor, L.xo_chroma_block_ver):
            f.argtypes = [P, ip, P, ip, C.c_int, C.c_int, C.c_int]
        for f in (L.xo_luma_block_ext, L.xo_chroma_block_ext):
            f.argtypes = [P, ip, P, ip, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_dct.argtypes = [C.c_int, P, P, C.c_int]
        L.xo_idct.argtypes = [C.c_int, P, P, C.c_int]
        L.xo_quant.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_dequant.argtypes = [P, C.c_int, C.c_int, C.c_int]
        L.xo_quant_params.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P]
        L.xo_dequant_params.argtypes = [C.c_int, C.c_int, P, P]
        L.xo_residual_batch.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P, C.c_int]
        L.xo_get_mc_pos.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_pmvr_adapt.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.xo_umh_prepare.argtypes = [C.POINTER(UmhJob), C.c_int, P, P, P, C.POINTER(UmhOut)]
        L.xo_umh_bsize.argtypes = [C.c_int, C.c_int, P]
        for f in (L.xo_dct_part, L.xo_idct_part, L.xo_dct_half_part):
            f.argtypes = [C.c_int, C.c_int, P, P, C.c_int]
        L.xo_wavelet_fwd.argtypes = [C.c_int, C.c_int, P, P]
        L.xo_wavelet_inv.argtypes = [C.c_int, C.c_int, P, P]
        L.xo_intpl_luma_hor.argtypes = [P, ip, P, ip, P, ip, C.c_int, C.c_int, P]
        L.xo_intpl_luma_ver.argtypes = [P, ip, P, ip, C.c_int, C.c_int, P]
        L.xo_intpl_luma_ext.argtypes = [P, ip, P, ip, C.c_int, C.c_int, P]
        L.xo_intpl_luma_block_ext.argtypes = [P, ip, P, ip, C.c_int, C.c_int, P, P]
        L.xo_intpl_chroma_block_hor.argtypes = [P, ip, P, ip, C.c_int, C.c_int, P]
        L.xo_intpl_chroma_block_ver.argtypes = [P, ip, P, ip, C.c_int, C.c_int, P]
        L.xo_intpl_chroma_block_ext.argtypes = [P, ip, P, ip, C.c_int, C.c_int, P, P]
        L.xo_get_mvp_default.argtypes = [C.POINTER(MvpJob), C.c_int, P, P, P, P]
        L.xo_inter_recon_batch.argtypes = [C.c_int, C.c_int, P, ip, P, ip, C.c_int, C.c_int, P, C.c_int, P, P, ip,
                                           P, P]
        L.xo_inter_recon_bi_batch.argtypes = [C.c_int, C.c_int, P, ip, P, P, ip, C.c_int, C.c_int, P, P, C.c_int, P,
                                              P, ip, P, P]
        L.xo_recon_pred_batch.argtypes = [C.c_int, C.c_int, P, ip, P, ip, P, C.c_int, P, P, ip, P, P]
        L.xo_chroma_qp.argtypes = [C.c_int]
        L.xo_tmat.restype = C.POINTER(C.c_int16)
        L.xo_tmat.argtypes = [C.c_int]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# padded frames: a plane is stored as a 2-D array with PAD pixels on each side;
# the origin (0, 0) of the picture is at [PAD, PAD].

class Plane:
    """A padded 8-bit plane with the picture origin at (pad, pad)."""

    def __init__(self, w, h, pad=PAD, stride=None, fill=0):
        self.w, self.h, self.pad = w, h, pad
        self.stride = stride if stride is not None else ((w + 2 * pad + 63) // 64) * 64
        self.rows = h + 2 * pad + 1
        self.buf = np.full(self.rows * self.stride + 64, fill, dtype=np.uint8)

    @property
    def origin(self):
        return self.pad * self.stride + self.pad

    def origin_ptr(self):
        return C.c_void_p(self.buf.ctypes.data + self.origin)

    def view(self):
        """2-D view of the full padded area (rows x stride)."""
        return self.buf[: self.rows * self.stride].reshape(self.rows, self.stride)

    def picture(self):
        v = self.view()
        return v[self.pad:self.pad + self.h, self.pad:self.pad + self.w]

    def set_picture(self, img):
        self.picture()[...] = img

    def expand_border(self):
        lib().xo_expand_border(self.origin_ptr(), self.stride, self.w, self.h, self.pad)


def sad(w, h, a, b):
    """a, b: 2-D uint8 arrays (blocks)."""
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    return lib().xo_sad(w, h, ptr(a), a.strides[0], ptr(b), b.strides[0])


def satd(w, h, a, b):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    return lib().xo_satd(w, h, ptr(a), a.strides[0], ptr(b), b.strides[0])


def mvbits(search_range, v):
    return lib().xo_mvbits(search_range, v)


def lambda_factor(lam):
    return lib().xo_lambda_factor(lam)


def me_param(frame_w, frame_h, search_range, lambda_factor, use_hadamard=1, use_fractional_me=2,
             use_fast_sub_me=0, enable_pmvr=0, mv_range_x=(-8192, 8191), mv_range_y=(-2048, 2047),
             me_method=ME_FS, umh_big_hex_level=0):
    p = MeParam()
    p.me_method = me_method
    p.umh_big_hex_level = umh_big_hex_level
    p.frame_w, p.frame_h, p.search_range, p.lambda_factor = frame_w, frame_h, search_range, lambda_factor
    p.min_mv_range[0], p.max_mv_range[0] = mv_range_x
    p.min_mv_range[1], p.max_mv_range[1] = mv_range_y
    p.use_hadamard, p.use_fractional_me = use_hadamard, use_fractional_me
    p.use_fast_sub_me, p.enable_pmvr = use_fast_sub_me, enable_pmvr
    return p


def interp_planes(ref_plane, use_fractional_me=2, lcu_level=6):
    """Build the 15 interpolated planes of a border-expanded reference plane.
    Returns a list of 16 Plane objects (index 0 is ref_plane itself)."""
    planes = [ref_plane] + [Plane(ref_plane.w, ref_plane.h, ref_plane.pad, ref_plane.stride)
                            for _ in range(15)]
    arr = (C.c_void_p * 16)(*[p.origin_ptr().value for p in planes])
    rc = lib().xo_interpolate_frame(arr, ref_plane.stride, ref_plane.w, ref_plane.h, lcu_level,
                                    use_fractional_me)
    assert rc == 0
    return planes


def lcu_row_window(h, lcu_level, lcu_y):
    """Rows [y0, y1) interpolate_lcu_row produces for LCU row lcu_y (one slice)."""
    y0, y1 = C.c_int(), C.c_int()
    lib().xo_lcu_row_window(h, lcu_level, lcu_y, C.byref(y0), C.byref(y1))
    return y0.value, y1.value


def border_row_window(h, row_log2, lcu_y):
    """Rows [y0, y1) xavs2_frame_expand_border_lcurow expands for LCU row lcu_y."""
    y0, y1 = C.c_int(), C.c_int()
    lib().xo_border_row_window(h, row_log2, lcu_y, C.byref(y0), C.byref(y1))
    return y0.value, y1.value


def expand_border_lcu_row(plane, row_log2, lcu_y):
    lib().xo_expand_border_lcu_row(plane.origin_ptr(), plane.stride, plane.w, plane.h, plane.pad, row_log2, lcu_y)


def interpolate_lcu_row(planes, lcu_level, lcu_y, use_fractional_me=2):
    """interpolate_lcu_row on a list of 16 Planes (index 0 the full-pel one)."""
    p0 = planes[0]
    arr = (C.c_void_p * 16)(*[(p.origin_ptr().value if p is not None else None) for p in planes])
    assert lib().xo_interpolate_lcu_row(arr, p0.stride, p0.w, p0.h, lcu_level, lcu_y, use_fractional_me) == 0


def plane_ptrs(planes):
    return (C.c_void_p * 16)(*[(p.origin_ptr().value if p is not None else None) for p in planes])


def me_search(prm, pu_tuple, cur_plane, planes, cost_map=False):
    """pu_tuple = (x, y, w, h, mvp_x, mvp_y).  Returns (result dict, cost_map or None)."""
    pu = Pu(*pu_tuple)
    res = MeResult()
    R = prm.search_range
    cm = np.zeros(4 * R * R, dtype=np.int32) if cost_map else None
    lib().xo_me_search(C.byref(prm), C.byref(pu), cur_plane.origin_ptr(), cur_plane.stride,
                       plane_ptrs(planes), planes[0].stride, C.byref(res),
                       ptr(cm) if cm is not None else None)
    out = {f: getattr(res, f) for f, _ in MeResult._fields_}
    return out, (cm.reshape(2 * R, 2 * R) if cm is not None else None)


def me_pattern_path(prm, pu_tuple, cur_plane, ref_plane, start, start_cost, max_pts=4096, ctx=None, p0=None,
                    with_route=False):
    """Points of the DIA / HEX / UMH integer stage from `start` (full-pel)
    with cost `start_cost`, the best after each step that moved it: a list of
    (x, y), the start first, the integer best last.  ctx: a UMH_CTX_DTYPE
    record; p0: the first MV candidate (default: start).  with_route also
    returns the UMH route code (see xo_me_pattern_path)."""
    path = np.zeros(2 * max_pts, dtype=np.int32)
    route = C.c_int(0)
    c = np.ascontiguousarray(ctx, dtype=UMH_CTX_DTYPE) if ctx is not None else None
    p0 = start if p0 is None else p0
    n = lib().xo_me_pattern_path(C.byref(prm), C.byref(Pu(*pu_tuple)), ptr(c) if c is not None else None,
                                 cur_plane.origin_ptr(), cur_plane.stride, ref_plane.origin_ptr(), ref_plane.stride,
                                 start[0], start[1], start_cost, p0[0], p0[1], ptr(path), max_pts, C.byref(route))
    assert n <= max_pts
    pts = [(int(path[2 * k]), int(path[2 * k + 1])) for k in range(n)]
    return (pts, route.value) if with_route else pts


def pattern_landscape(prm, pu_tuple, costs, start, start_cost=None, p0=None, ctx=None, max_trace=4096):
    """The integer stage of prm.me_method over a cost table instead of
    pictures: costs is a (2 span + 1)^2 int array indexed [y + span, x + span].
    Returns (trace of evaluated points, (bx, by), bcost, route)."""
    costs = np.ascontiguousarray(costs, dtype=np.int32)
    span = (costs.shape[0] - 1) // 2
    assert costs.shape == (2 * span + 1, 2 * span + 1)
    if start_cost is None:
        start_cost = int(costs[start[1] + span, start[0] + span])
    p0 = start if p0 is None else p0
    trace = np.zeros(2 * max_trace, dtype=np.int32)
    bx, by, bc, route = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    c = np.ascontiguousarray(ctx, dtype=UMH_CTX_DTYPE) if ctx is not None else None
    n = lib().xo_pattern_landscape(C.byref(prm), C.byref(Pu(*pu_tuple)), ptr(c) if c is not None else None,
                                   ptr(costs), span, start[0], start[1], start_cost, p0[0], p0[1], ptr(trace),
                                   max_trace, C.byref(bx), C.byref(by), C.byref(bc), C.byref(route))
    assert n <= max_trace
    pts = [(int(trace[2 * k]), int(trace[2 * k + 1])) for k in range(n)]
    return pts, (bx.value, by.value), bc.value, route.value


def pattern_trace(prm, pu_tuple, cur_plane, ref_plane, start, start_cost, p0=None, ctx=None, max_trace=8192):
    """The integer stage of prm.me_method on pictures: (evaluated points in
    order, (bx, by), bcost, route)."""
    p0 = start if p0 is None else p0
    trace = np.zeros(2 * max_trace, dtype=np.int32)
    bx, by, bc, route = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    c = np.ascontiguousarray(ctx, dtype=UMH_CTX_DTYPE) if ctx is not None else None
    n = lib().xo_me_pattern_trace(C.byref(prm), C.byref(Pu(*pu_tuple)), ptr(c) if c is not None else None,
                                  cur_plane.origin_ptr(), cur_plane.stride, ref_plane.origin_ptr(), ref_plane.stride,
                                  start[0], start[1], start_cost, p0[0], p0[1], ptr(trace), max_trace, C.byref(bx),
                                  C.byref(by), C.byref(bc), C.byref(route))
    assert n <= max_trace
    pts = [(int(trace[2 * k]), int(trace[2 * k + 1])) for k in range(n)]
    return pts, (bx.value, by.value), bc.value, route.value


def me_search_sym(prm, job, cur_plane, planes1, planes2):
    """xavs2_me_search_sym for one job: a dict with the SymJob fields.
    Returns (cost, mv_x, mv_y, mvcost)."""
    j = SymJob(**job)
    r = SymResult()
    lib().xo_me_search_sym(C.byref(prm), C.byref(j), cur_plane.origin_ptr(), cur_plane.stride,
                           plane_ptrs(planes1), plane_ptrs(planes2), planes1[0].stride, C.byref(r))
    return r.cost, r.mv_x, r.mv_y, r.mvcost


def me_search_bid(prm, job, cur_plane, planes1, planes2):
    """xavs2_me_search_bid for one job: a dict with the BidJob fields.
    Returns (cost, fwd_x, fwd_y, mvcost)."""
    j = BidJob(**job)
    r = SymResult()
    lib().xo_me_search_bid(C.byref(prm), C.byref(j), cur_plane.origin_ptr(), cur_plane.stride,
                           plane_ptrs(planes1), plane_ptrs(planes2), planes1[0].stride, C.byref(r))
    return r.cost, r.mv_x, r.mv_y, r.mvcost


def pred_inter_search_bi(prm, job, dist_fwd, dist_bwd, cur_plane, planes_fwd, planes_bwd, lcu_level=6,
                         frame_threads=1):
    """pred_inter_search_bi for one PU: job is a dict of BiJob fields.
    Returns a dict of the BiResult fields."""
    g = GateParam(lcu_level, frame_threads)
    r = BiResult()
    lib().xo_pred_inter_search_bi(C.byref(prm), C.byref(g), C.byref(BiJob(**job)), dist_fwd, dist_bwd,
                                  cur_plane.origin_ptr(), cur_plane.stride, plane_ptrs(planes_fwd),
                                  plane_ptrs(planes_bwd), planes_fwd[0].stride, C.byref(r))
    return {f: getattr(r, f) for f, _ in BiResult._fields_}


def pred_inter_search_dual(prm, job, dist, cur_plane, planes_by_ref, lcu_level=6, frame_threads=1):
    """pred_inter_search_dual for one PU: job = dict(pix_x, pix_y, w, h, mvp_x,
    mvp_y, mvp1=[(x, y)] * n_ref, imv=[(x, y)] * n_ref); planes_by_ref: one
    16-plane list per reference.  Returns (cost, fst_ref, snd_ref, mvcost, fst,
    snd) with fst / snd lists of n_ref (x, y)."""
    n_ref = len(planes_by_ref)
    j = DualJob()
    for f in ("pix_x", "pix_y", "w", "h", "mvp_x", "mvp_y"):
        setattr(j, f, job[f])
    for r in range(n_ref):
        j.mvp1[r][0], j.mvp1[r][1] = job["mvp1"][r]
        j.imv[r][0], j.imv[r][1] = job["imv"][r]
    d = (C.c_int * 4)(*(list(dist) + [0] * (4 - len(dist))))
    tabs = [plane_ptrs(p) for p in planes_by_ref]
    arr = (C.c_void_p * 4)(*[C.cast(t, C.c_void_p) for t in tabs] + [None] * (4 - n_ref))
    res = DualResult()
    lib().xo_pred_inter_search_dual(C.byref(prm), C.byref(GateParam(lcu_level, frame_threads)), C.byref(j),
                                    n_ref, d, cur_plane.origin_ptr(), cur_plane.stride, arr,
                                    planes_by_ref[0][0].stride, C.byref(res))
    return (res.cost, res.fst_ref, res.snd_ref, res.mvcost, [tuple(res.fst[r]) for r in range(n_ref)],
            [tuple(res.snd[r]) for r in range(n_ref)])


def get_mvp_default(job, b_slice, ref_dpoc, ref_dpoc_multi):
    """get_mvp_default for one PU: job has ref_idx, bwd_2nd, cb (x, y, w, h),
    ur_available, ref (4 neighbour ref_idx: left, top, top-left, top-right)
    and mv (4 neighbour MVs).  Returns (status, pmv_x, pmv_y)."""
    j = MvpJob()
    j.ref_idx, j.bwd_2nd = job["ref_idx"], job["bwd_2nd"]
    j.cb_x, j.cb_y, j.cb_w, j.cb_h = job["cb"]
    j.ur_available = job["ur_available"]
    for k in range(4):
        j.ref[k] = job["ref"][k]
        j.mv[k][0], j.mv[k][1] = job["mv"][k]
    d = np.ascontiguousarray(ref_dpoc, dtype=np.int32)
    m = np.ascontiguousarray(ref_dpoc_multi, dtype=np.int32)
    assert d.size == 4 and m.size == 4, "ref_dpoc tables hold MAX_REFS = 4 entries"
    x, y = C.c_int(), C.c_int()
    rc = lib().xo_get_mvp_default(C.byref(j), b_slice, ptr(d), ptr(m), C.byref(x), C.byref(y))
    return rc, x.value, y.value


def umh_bsize(i_qp, slice_type):
    out = np.zeros(8, np.float64)
    lib().xo_umh_bsize(i_qp, slice_type, ptr(out))
    return out


def umh_job(mvp_job, mode, nb_sad=(0, 0, 0), mincost_up=(0, 0, 0), best_up=((0, 0),) * 3, mincost_prev=0,
            best_prev=(0, 0), col=None):
    """A UmhJob record; mvp_job as for get_mvp_default; col = (col_ref, (mv_x, mv_y), col_dpoc) or None."""
    j = UmhJob()
    m = j.mvp
    m.ref_idx, m.bwd_2nd = mvp_job["ref_idx"], mvp_job["bwd_2nd"]
    m.cb_x, m.cb_y, m.cb_w, m.cb_h = mvp_job["cb"]
    m.ur_available = mvp_job["ur_available"]
    for k in range(4):
        m.ref[k] = mvp_job["ref"][k]
        m.mv[k][0], m.mv[k][1] = mvp_job["mv"][k]
    for k in range(3):
        j.nb_sad[k] = nb_sad[k]
        j.mincost_up[k] = mincost_up[k]
        j.best_up[k][0], j.best_up[k][1] = best_up[k]
    j.mode, j.mincost_prev = mode, mincost_prev
    j.best_prev[0], j.best_prev[1] = best_prev
    if col is not None:
        j.col_valid = 1
        j.col_ref, (j.col_mv[0], j.col_mv[1]), j.col_dpoc = col
    return j


def umh_prepare(job, slice_type, ref_dpoc, ref_dpoc_multi, bsize):
    out = UmhOut()
    d = np.ascontiguousarray(ref_dpoc, dtype=np.int32)
    m = np.ascontiguousarray(ref_dpoc_multi, dtype=np.int32)
    b = np.ascontiguousarray(bsize, dtype=np.float64)
    lib().xo_umh_prepare(C.byref(job), slice_type, ptr(d), ptr(m), ptr(b), C.byref(out))
    return out


def scale_mv_skip(mv, dist_dst, dist_src):
    return lib().xo_scale_mv_skip(mv, dist_dst, dist_src)


def me_search_batch(prm, pus, cur_plane, planes, nthreads=1, ctxs=None):
    """pus: structured array of PU_DTYPE; ctxs: UMH_CTX_DTYPE array, one per
    PU, or None.  Returns structured array of ME_RESULT_DTYPE."""
    pus = np.ascontiguousarray(pus, dtype=PU_DTYPE)
    res = np.zeros(len(pus), dtype=ME_RESULT_DTYPE)
    c = np.ascontiguousarray(ctxs, dtype=UMH_CTX_DTYPE) if ctxs is not None else None
    assert c is None or len(c) == len(pus)
    lib().xo_me_search_batch_ctx(C.byref(prm), ptr(pus), ptr(c) if c is not None else None, len(pus),
                                 cur_plane.origin_ptr(), cur_plane.stride, plane_ptrs(planes), planes[0].stride,
                                 ptr(res), nthreads)
    return res


def tmat(n):
    p = lib().xo_tmat(n)
    return np.ctypeslib.as_array(p, shape=(n * n,)).reshape(n, n).copy()


def dct(n, blk):
    blk = np.ascontiguousarray(blk, dtype=np.int16)
    out = np.zeros((n, n), dtype=np.int16)
    lib().xo_dct(n, ptr(blk), ptr(out), n)
    return out


def idct(n, coef):
    coef = np.ascontiguousarray(coef, dtype=np.int16)
    out = np.zeros((n, n), dtype=np.int16)
    lib().xo_idct(n, ptr(coef), ptr(out), n)
    return out


def quant_params(log2n, qp, b_intra):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    lib().xo_quant_params(log2n, qp, b_intra, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def dequant_params(log2n, qp):
    a, b = C.c_int(), C.c_int()
    lib().xo_dequant_params(log2n, qp, C.byref(a), C.byref(b))
    return a.value, b.value


def residual_batch(n, qp, b_intra, blocks):
    """blocks: int16 array (nblk, n, n).  Returns (level, recon, nnz)."""
    blocks = np.ascontiguousarray(blocks, dtype=np.int16)
    nblk = blocks.shape[0]
    level = np.zeros_like(blocks)
    recon = np.zeros_like(blocks)
    nnz = np.zeros(nblk, dtype=np.int32)
    lib().xo_residual_batch(n, qp, b_intra, ptr(blocks), ptr(level), ptr(recon), ptr(nnz), nblk)
    return level, recon, nnz


TU_DCT_HALF = 0x100


def inter_recon_batch(n, qp, cur_plane, planes, frame_w, frame_h, tus, planes2=None, mv2=None):
    """Luma inter reconstruction of n x n TUs; tus: int32 (ntu, 8) records
    {tu_x, tu_y, pu_x, pu_y, pu_w, pu_h, mv_x, mv_y}; with planes2 and mv2
    (int32 (ntu, 2)) the TUs are bi-predicted; n = 32 | TU_DCT_HALF uses
    dct_half; n = 64 transforms the wavelet's low band and keeps 32x32 levels.
    Returns (level (ntu, n, n), rec (frame_h, frame_w) picture with the TUs
    written, nnz, ssd)."""
    tus = np.ascontiguousarray(tus, dtype=np.int32).reshape(-1, 8)
    ntu = len(tus)
    ns = n & ~TU_DCT_HALF
    nl = min(ns, 32)                 # a 64x64 TU keeps the 32x32 levels of its wavelet low band
    level = np.zeros((ntu, nl, nl), np.int16)
    rec = np.zeros((frame_h, frame_w), np.uint8)
    nnz = np.zeros(ntu, np.int32)
    ssd = np.zeros(ntu, np.int32)
    if planes2 is None:
        lib().xo_inter_recon_batch(n, qp, cur_plane.origin_ptr(), cur_plane.stride, plane_ptrs(planes),
                                   planes[0].stride, frame_w, frame_h, ptr(tus), ntu, ptr(level), ptr(rec), frame_w,
                                   ptr(nnz), ptr(ssd))
    else:
        m2 = np.ascontiguousarray(mv2, dtype=np.int32).reshape(-1, 2)
        lib().xo_inter_recon_bi_batch(n, qp, cur_plane.origin_ptr(), cur_plane.stride, plane_ptrs(planes),
                                      plane_ptrs(planes2), planes[0].stride, frame_w, frame_h, ptr(tus), ptr(m2), ntu,
                                      ptr(level), ptr(rec), frame_w, ptr(nnz), ptr(ssd))
    return level, rec, nnz, ssd


def recon_pred_batch(n, qp, cur, pred, blocks):
    """Reconstruction of n x n blocks whose prediction is the picture pred
    at the block's position (chroma of inter CUs); cur, pred: uint8 (h, w)
    planes; blocks: int32 (nblk, 2) {x, y}.  Returns (level (nblk, n, n),
    rec (h, w) with the blocks written, nnz, ssd)."""
    cur = np.ascontiguousarray(cur, dtype=np.uint8)
    pred = np.ascontiguousarray(pred, dtype=np.uint8)
    assert cur.shape == pred.shape
    blocks = np.ascontiguousarray(blocks, dtype=np.int32).reshape(-1, 2)
    nblk = len(blocks)
    ns = n & ~TU_DCT_HALF
    nl = min(ns, 32)
    h, w = cur.shape
    assert nblk == 0 or (blocks.min() >= 0 and (blocks[:, 0] + ns).max() <= w and (blocks[:, 1] + ns).max() <= h)
    level = np.zeros((nblk, nl, nl), np.int16)
    rec = np.zeros((h, w), np.uint8)
    nnz = np.zeros(nblk, np.int32)
    ssd = np.zeros(nblk, np.int32)
    lib().xo_recon_pred_batch(n, qp, ptr(cur), w, ptr(pred), w, ptr(blocks), nblk, ptr(level), ptr(rec), w,
                              ptr(nnz), ptr(ssd))
    return level, rec, nnz, ssd


def chroma_qp(luma_qp):
    return lib().xo_chroma_qp(int(luma_qp))


def mc_chroma(qx, qy, w, h, pu, pv):
    """pu, pv: Plane objects (chroma, pad 40).  Returns (u_block, v_block)."""
    du = np.zeros((h, w), dtype=np.uint8)
    dv = np.zeros((h, w), dtype=np.uint8)
    lib().xo_mc_chroma(ptr(du), ptr(dv), w, qx, qy, w, h, pu.origin_ptr(), pv.origin_ptr(), pu.stride)
    return du, dv
