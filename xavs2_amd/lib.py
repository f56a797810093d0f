"""ctypes mirror of include/xavs2_amd.h (libxavs2_amd.so).

The library is the product; this module only loads it and converts numpy
arrays.  Loading fails loudly when the shared object has not been built: there
is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxavs2_amd.so")

XME_OK = 0
XME_ERR_ARG, XME_ERR_HIP, XME_ERR_NOMEM, XME_ERR_BUDGET, XME_ERR_STATE = -1, -2, -3, -4, -5
XME_PAD = 80
XME_MAX_REFS = 8
XME_MVP_REFS = 4
XME_TU_DCT_HALF = 0x100
MAX_DISTORTION = 1 << 30
PLAN_BATCHED, PLAN_PLAIN = 0, 1


class IntplFuncs(C.Structure):
    """xme_intpl_funcs: the interpolation members of g_funcs, in order."""
    _fields_ = [(n, C.c_void_p) for n in (
        "intpl_luma_hor", "intpl_luma_ver", "intpl_luma_ext", "intpl_luma_ver_x3", "intpl_luma_hor_x3",
        "intpl_luma_ext_x3", "intpl_luma_block_hor", "intpl_luma_block_ver", "intpl_luma_block_ext",
        "intpl_chroma_block_hor", "intpl_chroma_block_ver", "intpl_chroma_block_ext")]


PU_W = [4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 12, 16, 4, 32, 24, 32, 8, 64, 48, 64, 16]
PU_H = [4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 12, 16, 4, 16, 24, 32, 8, 32, 48, 64, 16, 64]


def part_index(w, h):
    """enum LumaPU index of a w x h block (PART_INDEX, pixel.h:178)."""
    for i in range(25):
        if PU_W[i] == w and PU_H[i] == h:
            return i
    raise ValueError(f"{w}x{h} is not a LumaPU size")


class XmeError(RuntimeError):
    pass


class MeParams(C.Structure):
    _fields_ = [("frame_w", C.c_int32), ("frame_h", C.c_int32), ("search_range", C.c_int32),
                ("lambda_factor", C.c_int32), ("min_mv_range", C.c_int32 * 2), ("max_mv_range", C.c_int32 * 2),
                ("use_hadamard", C.c_int32), ("use_fractional_me", C.c_int32), ("use_fast_sub_me", C.c_int32),
                ("enable_pmvr", C.c_int32), ("me_method", C.c_int32), ("umh_big_hex_level", C.c_int32)]


# integer search methods (XME_ME_*, h->param->me_method)
ME_FS, ME_DIA, ME_HEX, ME_UMH, ME_TZ = 0, 1, 2, 3, 4

# xme_umh_ctx: per-(PU, reference) encoder state of the UMH search
UMH_CTX_DTYPE = np.dtype([("mvp1_x", "<i4"), ("mvp1_y", "<i4"), ("mvp2_x", "<i4"), ("mvp2_y", "<i4"),
                          ("mvp3_x", "<i4"), ("mvp3_y", "<i4"), ("ref_idx", "<i4"), ("pred_sad", "<i4"),
                          ("pred_sad_uplayer", "<i4"), ("reserved", "<i4"), ("beta2", "<f8"), ("beta3", "<f8")])


PU_DTYPE = np.dtype([("x", "<i2"), ("y", "<i2"), ("w", "u1"), ("h", "u1"), ("r0", "u1"), ("r1", "u1"),
                     ("mvp_x", "<i2"), ("mvp_y", "<i2")])
RESULT_DTYPE = np.dtype([("cost", "<i4"), ("bcost2", "<i4"), ("bmv_x", "<i2"), ("bmv_y", "<i2"),
                         ("bmv2_x", "<i2"), ("bmv2_y", "<i2"), ("mvcost", "<i4"), ("status", "<i4")])


# xme_sym_job / xme_sym_result: the symmetric bi-prediction batch
SYM_JOB_DTYPE = np.dtype([("x", "<i2"), ("y", "<i2"), ("w", "u1"), ("h", "u1"), ("r0", "u1"), ("r1", "u1"),
                          ("mv_x", "<i2"), ("mv_y", "<i2"), ("mvp_x", "<i2"), ("mvp_y", "<i2"),
                          ("mvp1_x", "<i2"), ("mvp1_y", "<i2"), ("dist_fwd", "<i2"), ("dist_bwd", "<i2"),
                          ("b_slice", "<i4")])
BID_JOB_DTYPE = np.dtype([("x", "<i2"), ("y", "<i2"), ("w", "u1"), ("h", "u1"), ("r0", "u1"), ("r1", "u1"),
                          ("fwd_x", "<i2"), ("fwd_y", "<i2"), ("bwd_x", "<i2"), ("bwd_y", "<i2"),
                          ("mvp1_x", "<i2"), ("mvp1_y", "<i2"), ("mvp2_x", "<i2"), ("mvp2_y", "<i2")])
MVP_JOB_DTYPE = np.dtype([("ref_idx", "i1"), ("bwd_2nd", "i1"), ("ur_available", "i1"), ("reserved", "i1"),
                          ("cb_x", "i1"), ("cb_y", "i1"), ("cb_w", "u1"), ("cb_h", "u1"), ("ref", "i1", (4,)),
                          ("mv", "<i2", (4, 2))])
MVP_RESULT_DTYPE = np.dtype([("mvp_x", "<i2"), ("mvp_y", "<i2"), ("status", "<i4")])
# xme_umh_job / xme_umh_state: the UMH search's encoder-state batch
UMH_JOB_DTYPE = np.dtype([("mvp", MVP_JOB_DTYPE), ("nb_sad", "<i4", (3,)), ("mode", "<i4"),
                          ("mincost_up", "<i4", (3,)), ("best_up", "<i2", (3, 2)), ("mincost_prev", "<i4"),
                          ("best_prev", "<i2", (2,)), ("col_mv", "<i2", (2,)), ("col_valid", "i1"),
                          ("col_ref", "i1"), ("reserved", "<i2"), ("col_dpoc", "<i4")])
UMH_STATE_DTYPE = np.dtype([("mvp_x", "<i2"), ("mvp_y", "<i2"), ("pred_sad_space", "<i4"), ("status", "<i4"),
                            ("reserved", "<i4"), ("ctx", UMH_CTX_DTYPE)])
SYM_RESULT_DTYPE = np.dtype([("cost", "<i4"), ("mv_x", "<i2"), ("mv_y", "<i2"), ("mvcost", "<i4"),
                             ("status", "<i4")])
# xme_bi_job / xme_bi_result and xme_dual_job / xme_dual_result: the callers
# pred_inter_search_bi and pred_inter_search_dual
BI_JOB_DTYPE = np.dtype([("x", "<i2"), ("y", "<i2"), ("w", "u1"), ("h", "u1"), ("r0", "u1"), ("r1", "u1"),
                         ("fwd_imv_x", "<i2"), ("fwd_imv_y", "<i2"), ("bwd_x", "<i2"), ("bwd_y", "<i2"),
                         ("mvp_x", "<i2"), ("mvp_y", "<i2"), ("mvp1_x", "<i2"), ("mvp1_y", "<i2"),
                         ("mvp2_x", "<i2"), ("mvp2_y", "<i2")])
BI_RESULT_DTYPE = np.dtype([("sym_cost", "<i4"), ("sym_mv_x", "<i2"), ("sym_mv_y", "<i2"), ("sym_mvcost", "<i4"),
                            ("bid_cost", "<i4"), ("fwd_x", "<i2"), ("fwd_y", "<i2"), ("bwd_x", "<i2"),
                            ("bwd_y", "<i2"), ("bid_mvcost", "<i4"), ("status", "<i4")])
DUAL_JOB_DTYPE = np.dtype([("x", "<i2"), ("y", "<i2"), ("w", "u1"), ("h", "u1"), ("r0", "u1"), ("r1", "u1"),
                           ("mvp_x", "<i2"), ("mvp_y", "<i2"), ("mvp1", "<i2", (4, 2)), ("imv", "<i2", (4, 2))])
DUAL_RESULT_DTYPE = np.dtype([("cost", "<i4"), ("fst_ref", "i1"), ("snd_ref", "i1"), ("reserved", "<i2"),
                              ("mvcost", "<i4"), ("fst", "<i2", (4, 2)), ("snd", "<i2", (4, 2)), ("status", "<i4")])


class RefDesc(C.Structure):
    _fields_ = [("planes", C.c_void_p * 16), ("stride", C.c_ssize_t)]


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_int64), ("max_inflight", C.c_int32), ("search_range", C.c_int32),
                ("lambda_factor", C.c_int32), ("n_ref", C.c_int32), ("use_hadamard", C.c_int32),
                ("use_fractional_me", C.c_int32), ("use_fast_sub_me", C.c_int32), ("enable_pmvr", C.c_int32),
                ("workgroup_share_pct", C.c_int32), ("launch_split_absdiff", C.c_int64),
                ("me_method", C.c_int32), ("reserved", C.c_int32), ("full_share_below_absdiff", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("throughput_fps", C.c_double), ("resident_bytes", C.c_int64), ("inflight_bytes", C.c_int64),
                ("peak_bytes", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise XmeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P, I, Z = C.c_void_p, C.c_int, C.c_ssize_t
        L.xme_last_error.restype = C.c_char_p
        L.xme_version.restype = C.c_char_p
        L.xme_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.xme_free.argtypes = [P]
        L.xme_memcpy_h2d.argtypes = [P, P, C.c_size_t, P]
        L.xme_memcpy_d2h.argtypes = [P, P, C.c_size_t, P]
        L.xme_memcpy_d2d.argtypes = [P, P, C.c_size_t, P]
        L.xme_memset.argtypes = [P, I, C.c_size_t, P]
        L.xme_stream_create.argtypes = [C.POINTER(C.c_void_p)]
        L.xme_stream_destroy.argtypes = [P]
        L.xme_stream_sync.argtypes = [P]
        L.xme_device_count.argtypes = [C.POINTER(C.c_int)]
        L.xme_plan_create.argtypes = [C.POINTER(MeParams), P, I, I, I, C.POINTER(C.c_void_p)]
        L.xme_plan_destroy.argtypes = [P]
        L.xme_plan_set_max_workgroups.argtypes = [P, I]
        L.xme_plan_workgroups.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.xme_plan_set_launch_split.argtypes = [P, C.c_int64]
        L.xme_plan_info.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.xme_plan_search_count.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.xme_plan_set_umh_context.argtypes = [P, P, P]
        L.xme_me_run.argtypes = [P, P, Z, P, P, P, P]
        L.xme_me_run_graph.argtypes = [P, P, Z, P, P, P]
        L.xme_me_search_static.argtypes = [C.POINTER(MeParams), P, I, I, P, Z, P, P, P]
        L.xme_me_time.argtypes = [P, P, Z, P, P, P, I, C.POINTER(C.c_double), C.POINTER(C.c_double), P]
        L.xme_copy_kernel.argtypes = [P, P, C.c_size_t, P]
        L.xme_tune.argtypes = [C.POINTER(MeParams), P, I, P]
        L.xme_tune_get.argtypes = [I, I, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.xme_tune_set.argtypes = [I, I, I, I]
        L.xme_expand_border.argtypes = [P, Z, I, I, I, P]
        L.xme_interpolate_frame.argtypes = [P, Z, I, I, I, P]
        L.xme_expand_border_lcu_row.argtypes = [P, Z, I, I, I, I, I, P]
        L.xme_interpolate_lcu_row.argtypes = [P, Z, I, I, I, I, I, P]
        L.xme_sad_batch.argtypes = [I, P, Z, P, Z, P, P, I, P, P]
        L.xme_satd_batch.argtypes = [I, P, Z, P, Z, P, P, I, P, P]
        L.xme_sad_x3_batch.argtypes = [I, P, P, Z, P, P, I, P, P]
        L.xme_sad_x4_batch.argtypes = [I, P, P, Z, P, P, I, P, P]
        L.xme_mc_chroma_batch.argtypes = [P, P, Z, P, I, I, I, P, P, Z, P, P]
        L.xme_mc_chroma_bi_batch.argtypes = [P, P, P, P, Z, P, I, I, I, P, P, Z, P, P]
        L.xme_mc_luma_batch.argtypes = [P, P, I, I, I, P, Z, P, P]
        L.xme_mc_luma_bi_batch.argtypes = [C.POINTER(RefDesc), C.POINTER(RefDesc), P, I, I, I, P, Z, P, P]
        L.xme_me_search_sym_batch.argtypes = [C.POINTER(MeParams), P, I, P, Z, C.POINTER(RefDesc),
                                              C.POINTER(RefDesc), P, P]
        L.xme_me_search_bid_batch.argtypes = [C.POINTER(MeParams), P, I, P, Z, C.POINTER(RefDesc),
                                              C.POINTER(RefDesc), P, P]
        L.xme_pred_inter_search_bi_batch.argtypes = [C.POINTER(MeParams), I, I, P, I, I, I, P, Z,
                                                     C.POINTER(RefDesc), C.POINTER(RefDesc), P, P]
        L.xme_pred_inter_search_dual_batch.argtypes = [C.POINTER(MeParams), I, I, P, I, I, C.POINTER(C.c_int),
                                                       P, Z, C.POINTER(RefDesc), P, P]
        L.xme_dct_batch.argtypes = [I, P, P, I, P]
        L.xme_idct_batch.argtypes = [I, P, P, I, P]
        L.xme_quant_batch.argtypes = [I, P, I, I, I, P, P]
        L.xme_dequant_batch.argtypes = [I, P, I, I, P]
        L.xme_residual_batch.argtypes = [I, I, I, P, P, P, P, I, P]
        L.xme_get_mvp_batch.argtypes = [P, I, I, P, P, P, P]
        L.xme_umh_prepare_batch.argtypes = [P, I, I, P, P, P, P, P]
        L.xme_umh_bsize.argtypes = [I, I, P]
        L.xme_plan_set_umh_states.argtypes = [P, P, P]
        L.xme_inter_recon_batch.argtypes = [I, I, P, Z, C.POINTER(RefDesc), I, I, P, I, P, P, P, P, Z, P, P, P]
        L.xme_inter_recon_bi_batch.argtypes = [I, I, P, Z, C.POINTER(RefDesc), C.POINTER(RefDesc), I, I, P, P, I,
                                               P, P, P, P, Z, P, P, P]
        L.xme_recon_pred_batch.argtypes = [I, I, P, Z, P, Z, I, I, P, I, P, P, Z, P, P, P]
        L.xme_chroma_qp.argtypes = [I]
        L.xme_recon_chroma_batch.argtypes = [I, I, P, P, Z, P, P, Z, I, I, P, I, P, P, P, Z, P, P, P]
        L.xme_get_sad_table.argtypes = [P]
        L.xme_get_dct_table.argtypes = [P, P, P]
        L.xme_get_quant_funcs.argtypes = [P, P]
        L.xme_get_avg_table.argtypes = [P]
        L.xme_get_intpl_funcs.argtypes = [P]
        L.xme_get_satd_table.argtypes = [P]
        L.xme_get_sad_x3_table.argtypes = [P]
        L.xme_get_sad_x4_table.argtypes = [P]
        L.xme_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]
        L.xme_server_destroy.argtypes = [P]
        L.xme_server_open_stream.argtypes = [P, I, I, P, I, C.POINTER(C.c_int)]
        L.xme_server_stats_reset.argtypes = [P]
        L.xme_server_close_stream.argtypes = [P, I]
        L.xme_server_submit.argtypes = [P, I, P, Z, C.POINTER(C.c_int64)]
        L.xme_server_job_size.argtypes = [P, C.c_int64, C.POINTER(C.c_int)]
        L.xme_server_wait.argtypes = [P, C.c_int64, P, C.POINTER(C.c_double)]
        L.xme_server_stats_get.argtypes = [P, C.POINTER(ServerStats)]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != XME_OK:
        raise XmeError(f"{what} failed ({rc}): {lib().xme_last_error().decode()}")
    return rc


def ptr(a):
    return C.c_void_p(a.ctypes.data)


class DeviceBuffer:
    """A device allocation owned by Python."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().xme_malloc(C.byref(p), max(self.nbytes, 1)), "xme_malloc")
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.xme_free(C.c_void_p(self.ptr))
            self.ptr = None

    def at(self, offset=0):
        return C.c_void_p(self.ptr + int(offset))

    def upload(self, arr, offset=0, stream=None):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes + offset <= self.nbytes
        check(lib().xme_memcpy_h2d(self.at(offset), ptr(arr), arr.nbytes, stream), "h2d")

    def download(self, shape, dtype, offset=0, stream=None):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes + offset <= self.nbytes
        check(lib().xme_memcpy_d2h(ptr(out), self.at(offset), out.nbytes, stream), "d2h")
        return out

    def zero(self, stream=None):
        check(lib().xme_memset(self.at(0), 0, self.nbytes, stream), "memset")


def upload(arr):
    arr = np.ascontiguousarray(arr)
    b = DeviceBuffer(arr.nbytes)
    b.upload(arr)
    return b


class DevicePlane:
    """A padded 8-bit plane in device memory, laid out like xavs2's frame planes:
    the picture origin sits `pad` rows and columns into the allocation."""

    def __init__(self, w, h, pad=XME_PAD, stride=None):
        self.w, self.h, self.pad = w, h, pad
        self.stride = stride if stride is not None else ((w + 2 * pad + 255) // 256) * 256
        self.rows = h + 2 * pad + 1
        self.buf = DeviceBuffer(self.rows * self.stride + 256)

    @property
    def origin(self):
        return self.buf.ptr + self.pad * self.stride + self.pad

    def upload_padded(self, padded):
        """padded: (h + 2 pad) x (w + 2 pad) uint8 array (full padded area)."""
        full = np.zeros((self.rows, self.stride), np.uint8)
        full[:padded.shape[0], :padded.shape[1]] = padded
        self.buf.upload(full.ravel())

    def upload_picture(self, img):
        full = np.zeros((self.rows, self.stride), np.uint8)
        full[self.pad:self.pad + self.h, self.pad:self.pad + self.w] = img
        self.buf.upload(full.ravel())

    def download_padded(self):
        full = self.buf.download(self.rows * self.stride, np.uint8)[: self.rows * self.stride]
        return full.reshape(self.rows, self.stride)[: self.h + 2 * self.pad, : self.w + 2 * self.pad]

    def expand_border(self, stream=None):
        check(lib().xme_expand_border(C.c_void_p(self.origin), self.stride, self.w, self.h, self.pad, stream),
              "xme_expand_border")

    def expand_border_lcu_row(self, row_log2, lcu_y, stream=None):
        check(lib().xme_expand_border_lcu_row(C.c_void_p(self.origin), self.stride, self.w, self.h, self.pad,
                                              row_log2, lcu_y, stream), "xme_expand_border_lcu_row")


class DeviceRef:
    """A reference picture: full-pel plane plus its 15 interpolated planes."""

    def __init__(self, w, h, frac=2):
        self.w, self.h, self.frac = w, h, frac
        self.planes = [DevicePlane(w, h)]
        stride = self.planes[0].stride
        for k in range(1, 16):
            need = frac == 2 or (frac == 1 and k in (2, 8, 10))
            self.planes.append(DevicePlane(w, h, stride=stride) if need else None)

    @property
    def stride(self):
        return self.planes[0].stride

    def load(self, img, stream=None):
        """Upload a picture, expand its border and build the interpolated planes on the GPU."""
        self.planes[0].upload_picture(img)
        self.planes[0].expand_border(stream)
        if self.frac:
            arr = (C.c_void_p * 16)(*[(p.origin if p is not None else None) for p in self.planes])
            check(lib().xme_interpolate_frame(arr, self.stride, self.w, self.h, self.frac, stream),
                  "xme_interpolate_frame")

    def interpolate_lcu_row(self, lcu_level, lcu_y, stream=None):
        """interpolate_lcu_row for one LCU row of the planes (the full-pel
        plane's rows must be final and border-expanded down to 4 rows below)."""
        arr = (C.c_void_p * 16)(*[(p.origin if p is not None else None) for p in self.planes])
        check(lib().xme_interpolate_lcu_row(arr, self.stride, self.w, self.h, self.frac, lcu_level, lcu_y, stream),
              "xme_interpolate_lcu_row")

    def desc(self):
        d = RefDesc()
        for k in range(16):
            d.planes[k] = self.planes[k].origin if self.planes[k] is not None else None
        d.stride = self.stride
        return d


def me_params(frame_w, frame_h, search_range, lambda_factor, use_hadamard=1, use_fractional_me=2,
              use_fast_sub_me=0, enable_pmvr=0, mv_range_x=(-8192, 8191), mv_range_y=(-2048, 2047),
              me_method=ME_FS, umh_big_hex_level=0):
    p = MeParams()
    p.me_method = me_method
    p.umh_big_hex_level = umh_big_hex_level
    p.frame_w, p.frame_h, p.search_range, p.lambda_factor = frame_w, frame_h, search_range, lambda_factor
    p.min_mv_range[0], p.max_mv_range[0] = mv_range_x
    p.min_mv_range[1], p.max_mv_range[1] = mv_range_y
    p.use_hadamard, p.use_fractional_me = use_hadamard, use_fractional_me
    p.use_fast_sub_me, p.enable_pmvr = use_fast_sub_me, enable_pmvr
    return p


def me_search_sym_batch(params, jobs, cur, ref1, ref2, stream=None):
    """xme_me_search_sym_batch on host job records (SYM_JOB_DTYPE): `cur` is a
    DevicePlane holding the current picture, ref1 / ref2 DeviceRefs of the same
    stride.  Returns the SYM_RESULT_DTYPE records (blocking)."""
    return _bipred_batch("xme_me_search_sym_batch", SYM_JOB_DTYPE, params, jobs, cur, ref1, ref2, stream)


def me_search_bid_batch(params, jobs, cur, ref1, ref2, stream=None):
    """xme_me_search_bid_batch on host job records (BID_JOB_DTYPE), as
    me_search_sym_batch."""
    return _bipred_batch("xme_me_search_bid_batch", BID_JOB_DTYPE, params, jobs, cur, ref1, ref2, stream)


def pred_inter_search_bi_batch(params, jobs, dist_fwd, dist_bwd, cur, ref_fwd, ref_bwd, lcu_level=6,
                               frame_threads=1, stream=None):
    """xme_pred_inter_search_bi_batch on host job records (BI_JOB_DTYPE):
    `cur` a DevicePlane, ref_fwd / ref_bwd DeviceRefs of one stride.  Returns
    the BI_RESULT_DTYPE records (blocking)."""
    name = "xme_pred_inter_search_bi_batch"
    jobs = np.ascontiguousarray(jobs, dtype=BI_JOB_DTYPE)
    n = len(jobs)
    d_jobs = upload(jobs) if n else DeviceBuffer(1)
    d_res = DeviceBuffer(max(n, 1) * BI_RESULT_DTYPE.itemsize)
    r1, r2 = ref_fwd.desc(), ref_bwd.desc()
    check(lib().xme_pred_inter_search_bi_batch(C.byref(params), lcu_level, frame_threads, C.c_void_p(d_jobs.ptr),
                                               n, dist_fwd, dist_bwd, C.c_void_p(cur.origin), cur.stride,
                                               C.byref(r1), C.byref(r2), C.c_void_p(d_res.ptr), stream), name)
    check(lib().xme_stream_sync(stream), "xme_stream_sync")
    return d_res.download(n, BI_RESULT_DTYPE)


def pred_inter_search_dual_batch(params, jobs, dist, cur, refs, lcu_level=6, frame_threads=1, stream=None):
    """xme_pred_inter_search_dual_batch on host job records (DUAL_JOB_DTYPE):
    refs a list of 2..4 DeviceRefs of one stride, dist their distances.
    Returns the DUAL_RESULT_DTYPE records (blocking)."""
    name = "xme_pred_inter_search_dual_batch"
    jobs = np.ascontiguousarray(jobs, dtype=DUAL_JOB_DTYPE)
    n = len(jobs)
    d_jobs = upload(jobs) if n else DeviceBuffer(1)
    d_res = DeviceBuffer(max(n, 1) * DUAL_RESULT_DTYPE.itemsize)
    descs = (RefDesc * 4)(*[r.desc() for r in refs])
    dists = (C.c_int * 4)(*(list(dist) + [0] * (4 - len(dist))))
    check(lib().xme_pred_inter_search_dual_batch(C.byref(params), lcu_level, frame_threads, C.c_void_p(d_jobs.ptr),
                                                 n, len(refs), dists, C.c_void_p(cur.origin), cur.stride, descs,
                                                 C.c_void_p(d_res.ptr), stream), name)
    check(lib().xme_stream_sync(stream), "xme_stream_sync")
    return d_res.download(n, DUAL_RESULT_DTYPE)


def _bipred_batch(name, dtype, params, jobs, cur, ref1, ref2, stream):
    jobs = np.ascontiguousarray(jobs, dtype=dtype)
    n = len(jobs)
    d_jobs = upload(jobs) if n else DeviceBuffer(1)
    d_res = DeviceBuffer(max(n, 1) * SYM_RESULT_DTYPE.itemsize)
    r1, r2 = ref1.desc(), ref2.desc()
    check(getattr(lib(), name)(C.byref(params), C.c_void_p(d_jobs.ptr), n, C.c_void_p(cur.origin), cur.stride,
                               C.byref(r1), C.byref(r2), C.c_void_p(d_res.ptr), stream), name)
    check(lib().xme_stream_sync(stream), "xme_stream_sync")
    return d_res.download(n, SYM_RESULT_DTYPE)


def make_pus(tuples):
    """tuples: iterable of (x, y, w, h, mvp_x, mvp_y)."""
    a = np.zeros(len(tuples), dtype=PU_DTYPE)
    for i, t in enumerate(tuples):
        a[i] = (t[0], t[1], t[2], t[3], 0, 0, t[4], t[5])
    return a


class Plan:
    def __init__(self, params, pus, n_ref, mode=PLAN_BATCHED):
        self.params = params
        self.pus = np.ascontiguousarray(pus, dtype=PU_DTYPE)
        self.n_ref = n_ref
        h = C.c_void_p()
        check(lib().xme_plan_create(C.byref(params), ptr(self.pus), len(self.pus), n_ref, mode, C.byref(h)),
              "xme_plan_create")
        self.h = h
        self.n = len(self.pus) * n_ref
        self.results = DeviceBuffer(self.n * RESULT_DTYPE.itemsize)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.xme_plan_destroy(self.h)
            self.h = None

    def info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().xme_plan_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "xme_plan_info")
        return a.value, b.value, c.value

    def search_count(self):
        """(points, byte absdiffs) evaluated by the last run of a pattern-search plan."""
        a, b = C.c_int64(), C.c_int64()
        check(lib().xme_plan_search_count(self.h, C.byref(a), C.byref(b)), "xme_plan_search_count")
        return a.value, b.value

    def set_umh_context(self, ctx, stream=None):
        """ctx: UMH_CTX_DTYPE array of n_pu * n_ref entries (pu * n_ref + r)."""
        self._umh = np.ascontiguousarray(ctx, dtype=UMH_CTX_DTYPE)
        assert len(self._umh) == self.n
        check(lib().xme_plan_set_umh_context(self.h, ptr(self._umh), stream), "xme_plan_set_umh_context")
        check(lib().xme_stream_sync(stream), "sync")

    def set_umh_states(self, d_states, stream=None):
        """d_states: device address (int or c_void_p) of n_pu * n_ref xme_umh_state records."""
        d = d_states if isinstance(d_states, C.c_void_p) else C.c_void_p(d_states)
        check(lib().xme_plan_set_umh_states(self.h, d, stream), "xme_plan_set_umh_states")

    def _refs(self, refs):
        arr = (RefDesc * len(refs))()
        for i, r in enumerate(refs):
            arr[i] = r.desc()
        return arr

    def run(self, cur_plane, refs, cost_map=None, stream=None):
        self._ref_arr = self._refs(refs)
        check(lib().xme_me_run(self.h, C.c_void_p(cur_plane.origin), cur_plane.stride, self._ref_arr,
                               C.c_void_p(self.results.ptr), C.c_void_p(cost_map.ptr) if cost_map else None,
                               stream), "xme_me_run")

    def run_graph(self, cur_plane, refs, stream):
        self._ref_arr = self._refs(refs)
        check(lib().xme_me_run_graph(self.h, C.c_void_p(cur_plane.origin), cur_plane.stride, self._ref_arr,
                                     C.c_void_p(self.results.ptr), stream), "xme_me_run_graph")

    def fetch(self, stream=None):
        return self.results.download(self.n, RESULT_DTYPE, stream=stream)

    def time(self, cur_plane, refs, stream, reps):
        """Runs the job reps times; returns (mean search-kernel ms, mean job ms, per-run job ms)."""
        self._ref_arr = self._refs(refs)
        a, b = C.c_double(), C.c_double()
        per = np.zeros(reps, dtype=np.float64)
        check(lib().xme_me_time(self.h, C.c_void_p(cur_plane.origin), cur_plane.stride, self._ref_arr,
                                C.c_void_p(self.results.ptr), stream, reps, C.byref(a), C.byref(b), ptr(per)),
              "xme_me_time")
        return a.value, b.value, per


def tune(params, parts, stream=None):
    arr = np.ascontiguousarray(parts, dtype=np.int32)
    n = lib().xme_tune(C.byref(params), ptr(arr), len(arr), stream)
    if n < 0:
        check(n, "xme_tune")
    out = {}
    for p in parts:
        k, lc = C.c_int(), C.c_int()
        lib().xme_tune_get(int(p), params.search_range, C.byref(k), C.byref(lc))
        out[int(p)] = (k.value, lc.value)
    return out


def tune_set(part, search_range, k_rows, lanes_per_row):
    check(lib().xme_tune_set(int(part), int(search_range), int(k_rows), int(lanes_per_row)), "xme_tune_set")


def tune_get(part, search_range):
    k, lc = C.c_int(), C.c_int()
    check(lib().xme_tune_get(int(part), int(search_range), C.byref(k), C.byref(lc)), "xme_tune_get")
    return k.value, lc.value


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib().xme_stream_create(C.byref(s)), "xme_stream_create")
        self.h = s

    def sync(self):
        check(lib().xme_stream_sync(self.h), "xme_stream_sync")

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.xme_stream_destroy(self.h)
            self.h = None


class Server:
    """The continuous-batching frame-ME server (xme_server_*)."""

    def __init__(self, memory_budget, max_inflight=4, search_range=64, lambda_factor=4 * 65536, n_ref=2,
                 use_hadamard=1, use_fractional_me=2, use_fast_sub_me=0, enable_pmvr=0, workgroup_share_pct=0,
                 launch_split_absdiff=0, me_method=0, full_share_below_absdiff=0):
        cfg = ServerConfig()
        cfg.workgroup_share_pct = workgroup_share_pct
        cfg.full_share_below_absdiff = int(full_share_below_absdiff)
        cfg.launch_split_absdiff = launch_split_absdiff
        cfg.me_method = me_method
        cfg.memory_budget, cfg.max_inflight, cfg.search_range = int(memory_budget), max_inflight, search_range
        cfg.lambda_factor, cfg.n_ref = lambda_factor, n_ref
        cfg.use_hadamard, cfg.use_fractional_me = use_hadamard, use_fractional_me
        cfg.use_fast_sub_me, cfg.enable_pmvr = use_fast_sub_me, enable_pmvr
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().xme_server_create(C.byref(cfg), C.byref(h)), "xme_server_create")
        self.h = h

    def open_stream(self, w, h, pus):
        """pus: structured array of PU_DTYPE (see make_pus).  Raises XmeError
        (code XME_ERR_BUDGET = -4) when the stream does not fit the budget."""
        pus = np.ascontiguousarray(pus, dtype=PU_DTYPE)
        sid = C.c_int()
        check(lib().xme_server_open_stream(self.h, w, h, ptr(pus), len(pus), C.byref(sid)), "xme_server_open_stream")
        return sid.value

    def close_stream(self, sid):
        check(lib().xme_server_close_stream(self.h, sid), "xme_server_close_stream")

    def submit(self, sid, luma):
        """Returns the job id; returns None when the frame is rejected for the budget."""
        luma = np.ascontiguousarray(luma, dtype=np.uint8)
        jid = C.c_int64()
        rc = lib().xme_server_submit(self.h, sid, ptr(luma), luma.strides[0], C.byref(jid))
        if rc == XME_ERR_BUDGET:
            return None
        check(rc, "xme_server_submit")
        return jid.value

    def wait(self, jid):
        """Returns (results structured array, latency in ms)."""
        n = C.c_int()
        check(lib().xme_server_job_size(self.h, jid, C.byref(n)), "xme_server_job_size")
        res = np.zeros(n.value, dtype=RESULT_DTYPE)
        lat = C.c_double()
        check(lib().xme_server_wait(self.h, jid, ptr(res) if n.value else None, C.byref(lat)), "xme_server_wait")
        return res, lat.value

    def stats(self):
        s = ServerStats()
        check(lib().xme_server_stats_get(self.h, C.byref(s)), "xme_server_stats_get")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def reset_stats(self):
        check(lib().xme_server_stats_reset(self.h), "xme_server_stats_reset")

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.xme_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
