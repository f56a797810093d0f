// Internal definitions shared by the HIP kernels and the host runtime of
// libxavs2_amd.so.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/xavs2_amd.h"

#define XME_WAVE 64

// ---------------------------------------------------------------------------
// Device-side descriptors.

// One PU of a plan (16 bytes).
struct DevPu {
    int16_t x, y;
    uint8_t w, h, part, pad0;
    int16_t mvp_x, mvp_y;
    int32_t pad1;
};

// Search parameters of a plan (copied into kernel arguments).
struct DevParams {
    int32_t frame_w, frame_h, R, lambda;
    int32_t min_mv[2], max_mv[2];
    int32_t hadamard, frac, fast, pmvr;
    int32_t n_pu, n_ref;
    int32_t method;            // XME_ME_FS, _DIA, _HEX, _UMH or _TZ
    int32_t big_hex;           // UMH big-hexagon level (0, 1, 2)
    int32_t max_mvd;           // largest |MV difference| of the reference's mvbits table
};

// Per-frame pointers, read by the kernels from device memory so that a captured
// graph stays valid while the caller changes buffers between frames.
struct FrameDesc {
    const uint8_t *cur;
    int64_t i_cur;
    const uint8_t *planes[XME_MAX_REFS][16];
    int64_t stride[XME_MAX_REFS];
    xme_result *results;
    int32_t *cost_map;
};

// Per (PU, reference) state produced by the candidate-list pass.
struct DevInst {
    int16_t omx, omy;          // full-search centre (best MV candidate)
    int16_t mvc[2][2];         // candidate list (md_inter.c:957-959)
    int8_t n_mvc, status;      // status 1: no legal candidate (me.c:919-921)
    int16_t flags;             // bit 0: a TZ raster read left the padded plane (undefined in the reference)
};

static_assert(sizeof(DevPu) == 16 && sizeof(DevInst) == 16, "the search kernel moves these as uint4");

// A workgroup task of the integer search: a set of items that share one
// reference window staged in LDS.
struct FsTask {
    int32_t item_begin, item_end;
    int32_t ref;
    int16_t wx0, wy0;          // window origin (picture coordinates, wx0 % 4 == 0)
    int16_t ww, wh;            // window size in bytes / rows (ww % 4 == 0)
    int16_t cx0, cy0;          // current-picture region staged in LDS
    int16_t cw, ch;            // (cw % 4 == 0)
};

// A wave work item: one band of candidate rows/columns for a group of
// (PU, reference) instances of one PU shape and one MVP, listed at
// item_insts[list .. list + count).  Instance ids are pu * n_ref + ref.
struct FsItem {
    int32_t list;
    int16_t count;
    uint16_t cost;             // relative work (host-side ordering only)
    int16_t row0, nrows;       // candidate rows [row0, row0 + nrows) of [0, 2R)
    int16_t col0, ncols;       // candidate columns [col0, col0 + ncols) of [0, 2R)
    uint8_t K, LC;             // rows per lane, lanes per candidate row
    uint8_t pad[2];
};
// Instances of one item: one per lane when the search kernel fetches them.
#define XME_MAX_GROUP 64
static_assert(XME_MAX_GROUP <= 64, "an item's instances are fetched one per lane");

// Order of a candidate in xavs2_me_search: the MV candidate list comes first
// (me.c:913-917), then the raster rows j, columns i (me.c:1159-1163).
#define XME_ORDER_RASTER 2

// Key of a candidate for the lexicographic (cost, order) minimum, which is
// the first strict minimum in evaluation order (COPY3_IF_LT, me.h:70-75).
__host__ __device__ inline uint64_t xme_key(uint32_t cost, uint32_t order)
{
    return ((uint64_t)cost << 32) | order;
}
#define XME_KEY_NONE (((uint64_t)XME_MAX_DISTORTION) << 32)

// ---------------------------------------------------------------------------
// Shared integer helpers (host and device).

__host__ __device__ inline int xme_clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }

// h->mvbits[v] (me.c:614-641): 1 for v = 0, else 2*floor(log2|v|) + 3, the
// length of the signed Exp-Golomb code.  Valid for |v| within the table.
__device__ inline int xme_mvbits(int v)
{
    unsigned a = (unsigned)(v < 0 ? -v : v);
    return a == 0 ? 1 : 2 * (31 - __clz(a)) + 3;
}

// Largest |v| of the reference's mvbits table for search range R
// (me.c:592-612).
int xme_mvbits_max_mvd(int R);

// WEIGHTED_COST (common/common.h:87) with the reference's int arithmetic.
__host__ __device__ inline int xme_weighted_cost(int lambda, int bits)
{
    return (int)((uint32_t)lambda * (uint32_t)bits) >> 16;
}

// CHECK_MV_RANGE (me.h:46-48) with the packed bounds of xavs2_me_search
// (me.c:896-897), restated bit for bit.
__host__ __device__ inline uint32_t xme_pack2(int mx, int my)
{
    return ((uint32_t)mx << 16) | ((uint32_t)my & 0x7FFFu);
}

struct MvWindow {
    int min_fpel[2], max_fpel[2];
    uint32_t swar_min, swar_max;
};

// MV limits of pred_inter_search_single (md_inter.c:910-926).
__host__ __device__ inline MvWindow xme_mv_window(const DevParams &p, int x, int y, int w, int h)
{
    MvWindow m;
    int lo = (-64 - x) * 4, hi = (p.frame_w + 64 - x - w) * 4;
    lo = xme_clip3(p.min_mv[0], p.max_mv[0], lo);
    hi = xme_clip3(p.min_mv[0], p.max_mv[0], hi);
    m.min_fpel[0] = (lo >> 2) + 6;
    m.max_fpel[0] = (hi >> 2) - 6;
    lo = (-64 - y) * 4;
    hi = (p.frame_h + 64 - y - h) * 4;
    lo = xme_clip3(p.min_mv[1], p.max_mv[1], lo);
    hi = xme_clip3(p.min_mv[1], p.max_mv[1], hi);
    m.min_fpel[1] = (lo >> 2) + 6;
    m.max_fpel[1] = (hi >> 2) - 6;
    m.swar_min = xme_pack2(-m.min_fpel[0], -m.min_fpel[1]);
    m.swar_max = xme_pack2(m.max_fpel[0], m.max_fpel[1]) | 0x8000u;
    return m;
}

__host__ __device__ inline bool xme_mv_legal(const MvWindow &m, int mx, int my)
{
    uint32_t v = xme_pack2(mx, my);
    return !(((v + m.swar_min) | (m.swar_max - v)) & 0x80004000u);
}

// ---------------------------------------------------------------------------
// The residual path of transform_kernels.hip for other kernels' hosts: the
// fused DCT -> quant -> dequant -> IDCT on int16 blocks (half: dct_half for
// N = 32), and the fused luma inter reconstruction (one launch per TU size;
// resid / rres may be null; every plane of `planes` (and `planes2` when bi)
// must exist).
int xme_residual_run(int n, int qp, int b_intra, int half, const int16_t *in, int16_t *level, int16_t *recon,
                     int32_t *nnz, int nblk, hipStream_t s);
int xme_recon_fused(int n, int qp, int half, int bi, const uint8_t *const planes[16], const uint8_t *const planes2[16],
                    int64_t i_ref, const uint8_t *cur, int64_t i_cur, uint8_t *rec, int64_t i_rec, const int32_t *tus,
                    const int32_t *mv2, int ntu, int fw, int fh, int16_t *resid, int16_t *level, int16_t *rres,
                    int32_t *nnz, int32_t *ssd, hipStream_t s);

// ---------------------------------------------------------------------------
// Host-side helpers (defined in runtime.cpp).
void xme_set_error(const char *fmt, ...);
int xme_hip_fail(hipError_t e, const char *what);
#define XME_HIP(call)                                   \
    do {                                                \
        hipError_t e__ = (call);                        \
        if (e__ != hipSuccess)                          \
            return xme_hip_fail(e__, #call);            \
    } while (0)

// Launchers implemented in the .hip files.
int xme_launch_mvc(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, DevInst *d_inst,
                   uint64_t *d_best, int n_inst, hipStream_t s);
int xme_launch_fs(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, const DevInst *d_inst,
                  const FsTask *d_tasks, int n_tasks, const FsItem *d_items, const int32_t *d_item_insts,
                  uint64_t *d_best, uint32_t *d_counters, size_t lds_bytes, int grid, bool persistent, int kmax,
                  hipStream_t s);
// Rows per lane of the narrow-PU search path (8-pixel-wide PUs of <= 16 rows),
// which has a kernel variant of its own.
#define XME_NARROW_K 16
// LDS bytes of a search task beyond its staged window and current pixels
// (per-wave MV-cost term tables of 4 waves x 8 x 4 x 64 dwords).
#define XME_FS_TERM_BYTES (4 * 8 * 4 * 64 * 4)
// Sub-pel refinement of every instance.  `d_list` holds PU indices: the
// first n_small PUs have at most 64 4x4 tiles (xme_subpel_large false), the
// next n_large more; each part gets a kernel variant of its own.
bool xme_subpel_large(int w, int h);
int xme_launch_subpel(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, const DevInst *d_inst,
                      const uint64_t *d_best, const int32_t *d_list, int n_small, int n_large, hipStream_t s);
// Pattern searches (XME_ME_DIA, XME_ME_HEX) of every instance.  The six PU
// size classes of xme_pattern_class (0: <= 16 pixel quads ... 5: > 256) run
// as wave ranges, largest class first: range k holds class 5 - k with
// 1, 1, 2, 4, 8, 16 instance ids per wave (d_list back to back, padded with
// -1 to whole workgroups of 4 waves); wave_end[k] is the first wave after
// range k, a multiple of 4.  d_points[id]
// receives the points instance id evaluated.
int xme_pattern_class(int w, int h);
int xme_launch_pattern(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, DevInst *d_inst,
                       uint64_t *d_best, const int32_t *d_list, const int wave_end[6], uint32_t *d_points,
                       const xme_umh_ctx *d_umh, hipStream_t s);
int xme_fs_max_lds(void);
int xme_fs_prepare(void);
// Device bytes a plan holds (work lists, per-instance state, descriptor).
size_t xme_plan_device_bytes(const xme_plan *pl);
