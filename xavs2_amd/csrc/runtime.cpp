// Host runtime of libxavs2_amd.so: errors, device memory, and the frame-level
// motion-estimation plan (work lists for the search kernels, HIP graph replay).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

#include "xme_internal.h"
#include "tune.h"

static thread_local char g_err[512];

void xme_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int xme_hip_fail(hipError_t e, const char *what)
{
    xme_set_error("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? XME_ERR_NOMEM : XME_ERR_HIP;
}

extern "C" const char *xme_last_error(void) { return g_err; }
extern "C" const char *xme_version(void) { return "xavs2_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// device, memory, streams

extern "C" int xme_device_count(int *count)
{
    XME_HIP(hipGetDeviceCount(count));
    return XME_OK;
}

extern "C" int xme_set_device(int device)
{
    XME_HIP(hipSetDevice(device));
    return XME_OK;
}

extern "C" int xme_malloc(void **p, size_t bytes)
{
    if (!p)
        return XME_ERR_ARG;
    XME_HIP(hipMalloc(p, bytes ? bytes : 1));
    return XME_OK;
}

extern "C" int xme_free(void *p)
{
    if (p)
        XME_HIP(hipFree(p));
    return XME_OK;
}

extern "C" int xme_host_alloc(void **p, size_t bytes)
{
    XME_HIP(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
    return XME_OK;
}

extern "C" int xme_host_free(void *p)
{
    if (p)
        XME_HIP(hipHostFree(p));
    return XME_OK;
}

extern "C" int xme_memcpy_h2d(void *d, const void *h, size_t n, void *s)
{
    XME_HIP(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, (hipStream_t)s));
    XME_HIP(hipStreamSynchronize((hipStream_t)s));
    return XME_OK;
}

extern "C" int xme_memcpy_d2h(void *h, const void *d, size_t n, void *s)
{
    XME_HIP(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, (hipStream_t)s));
    XME_HIP(hipStreamSynchronize((hipStream_t)s));
    return XME_OK;
}

extern "C" int xme_memcpy_d2d(void *d, const void *src, size_t n, void *s)
{
    XME_HIP(hipMemcpyAsync(d, src, n, hipMemcpyDeviceToDevice, (hipStream_t)s));
    return XME_OK;
}

extern "C" int xme_memset(void *d, int v, size_t n, void *s)
{
    XME_HIP(hipMemsetAsync(d, v, n, (hipStream_t)s));
    return XME_OK;
}

extern "C" int xme_stream_create(void **s)
{
    hipStream_t st;
    XME_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = st;
    return XME_OK;
}

extern "C" int xme_stream_destroy(void *s)
{
    if (s)
        XME_HIP(hipStreamDestroy((hipStream_t)s));
    return XME_OK;
}

extern "C" int xme_stream_sync(void *s)
{
    XME_HIP(hipStreamSynchronize((hipStream_t)s));
    return XME_OK;
}

extern "C" int xme_device_sync(void)
{
    XME_HIP(hipDeviceSynchronize());
    return XME_OK;
}

// ---------------------------------------------------------------------------
// plan

static const uint8_t kPuW[XME_NUM_PU] = { 4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32,
                                          16, 12, 16, 4, 32, 24, 32, 8, 64, 48, 64, 16 };
static const uint8_t kPuH[XME_NUM_PU] = { 4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64,
                                          12, 16, 4, 16, 24, 32, 8, 32, 48, 64, 16, 64 };

int xme_part_index(int w, int h)
{
    for (int i = 0; i < XME_NUM_PU; i++)
        if (kPuW[i] == w && kPuH[i] == h)
            return i;
    return -1;
}

int xme_pu_w(int part) { return kPuW[part]; }
bool xme_narrow_shape(int part) { return kPuW[part] * kPuH[part] <= 256; }
int xme_pu_h(int part) { return kPuH[part]; }

// Largest |v| of the reference's mvbits table (me.c:592-612).
int xme_mvbits_max_mvd(int R)
{
    int me_range = R > 256 ? R : 256;
    int subpel_num = 4 * (2 * me_range + 3);
    int max_bits = 5 + 2 * (int)std::ceil(std::log((double)subpel_num + 1) / std::log(2.0) + 1e-10);
    return (1 << (max_bits >> 1)) - 1;
}

struct DescSlot {
    FrameDesc *host;   // pinned
    hipEvent_t ev;
    bool used;
};

struct LaunchGroup {       // plain path: one launch per PU shape
    int task_begin, task_end;
    size_t lds;
    int variant;           // kernel variant: the plan's K bound, or XME_NARROW_K
};

struct xme_plan {
    int device;
    int mode;
    DevParams p;
    int n_pu, n_ref, n_inst;
    int64_t n_cand, n_absdiff;
    DevPu *d_pus = nullptr;
    DevInst *d_inst = nullptr;
    uint64_t *d_best = nullptr;
    FsTask *d_tasks = nullptr;
    FsItem *d_items = nullptr;
    int32_t *d_item_insts = nullptr;
    int32_t *d_sp_list = nullptr;      // sub-pel PU order: small PUs, then large ones
    int n_sp_small = 0, n_sp_large = 0;
    uint32_t *d_counters = nullptr;
    FrameDesc *d_desc = nullptr;
    int n_tasks = 0;
    size_t lds_bytes = 0;
    int grid = 0;
    int grid_full = 0;         // occupancy-sized grid, before any cap
    // batched plans: tasks [n_tasks - n_tasks_narrow, n_tasks) hold the
    // narrow-PU items and run in a second persistent launch
    int n_tasks_narrow = 0;
    size_t lds_narrow = 0;
    int grid_narrow = 0, grid_narrow_full = 0;
    // absdiffs per task (upper bound: every candidate of every item) and the
    // launches the batched search is split into: {first task, task count}
    std::vector<int64_t> task_work;
    std::vector<std::pair<int, int>> chunks_gen, chunks_nar;
    int kmax = 8;              // rows-per-lane bound of the search kernel variant
    std::vector<LaunchGroup> groups;
    DescSlot slots[8];
    int slot_next = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t graph_stream = nullptr;
    size_t dev_bytes = 0;
    // pattern-search plans: instance ids per wave and class bounds, counters
    int32_t *d_pat_list = nullptr;
    int pat_wave_end[6] = { 0, 0, 0, 0, 0, 0 };
    uint32_t *d_pat_points = nullptr;     // points evaluated per instance, last run
    xme_umh_ctx *d_umh = nullptr;         // UMH plans: per-instance encoder state
    bool umh_set = false;
    std::vector<int32_t> inst_area;        // PU area per instance
    // device buffers carved from caller-given memory instead of hipMalloc
    // (xme_me_search_static): nothing of it is freed with the plan
    uint8_t *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    hipStream_t arena_stream = nullptr;
    FrameDesc arena_desc;                  // host copy of the frame descriptor (pageable)
};

namespace {

struct Rect {
    int x0, y0, x1, y1;   // half-open
    void add(const Rect &o)
    {
        x0 = std::min(x0, o.x0);
        y0 = std::min(y0, o.y0);
        x1 = std::max(x1, o.x1);
        y1 = std::max(y1, o.y1);
    }
};

struct HostItem {
    FsItem it;
    std::vector<int> insts;   // instance ids pu * n_ref + ref (one shape, one MVP)
};

// The possible full-search centres of a PU: the clipped MVP and the clipped
// (0,0) (md_inter.c:957-959).  Which one wins depends on the pixels; the
// window covers both.
static int possible_centres(const DevParams &p, const DevPu &pu, int c[2][2])
{
    MvWindow w = xme_mv_window(p, pu.x, pu.y, pu.w, pu.h);
    const int q[2][2] = { { pu.mvp_x, pu.mvp_y }, { 0, 0 } };
    int n = 0;
    for (int k = 0; k < 2; k++) {
        int x = xme_clip3(w.min_fpel[0], w.max_fpel[0], (q[k][0] + 2) >> 2);
        int y = xme_clip3(w.min_fpel[1], w.max_fpel[1], (q[k][1] + 2) >> 2);
        bool dup = false;
        for (int j = 0; j < n; j++)
            dup |= c[j][0] == x && c[j][1] == y;
        if (!dup) {
            c[n][0] = x;
            c[n][1] = y;
            n++;
        }
    }
    return n;
}

// LDS read footprint of one item (picture coordinates), see fs_strip.
// Window bytes an item's lanes read at search centre (cx, cy): the item's
// candidate columns (at least 4) and rows (at least K; lanes beyond them are
// pulled back inside, see lane_col/lane_row in fs_kernels.hip), plus the PU.
static Rect item_footprint(const DevParams &p, const DevPu &pu, const FsItem &it, int cx, int cy)
{
    const int R = p.R;
    Rect r;
    const int bx = pu.x + cx - R, by = pu.y + cy - R;
    r.x0 = bx + it.col0 - 3;
    r.x1 = bx + it.col0 + std::max((int)it.ncols, 4) + pu.w + 8;
    r.y0 = by + it.row0;
    r.y1 = by + it.row0 + std::max((int)it.nrows, (int)it.K) + pu.h - 1;
    return r;
}

static int window_pitch_host(int ww)
{
    int dw = ww >> 2;
    return (dw | 1) << 2;
}

// Items for a group of instances sharing PU shape and MVP: one item per band
// of candidate rows/columns, covering every instance of the group.
// Splits the batched search into launches of at most max_work absdiffs each
// (0: one launch per kernel variant).  Tasks keep their order.
static void split_launches(xme_plan *pl, int64_t max_work)
{
    const int n_gen = pl->n_tasks - pl->n_tasks_narrow;
    auto split = [&](int b, int e, std::vector<std::pair<int, int>> &out) {
        out.clear();
        int first = b;
        int64_t sum = 0;
        for (int t = b; t < e; t++) {
            sum += pl->task_work[t];
            if (max_work > 0 && sum >= max_work && t + 1 < e) {
                out.push_back({ first, t + 1 - first });
                first = t + 1;
                sum = 0;
            }
        }
        if (e > first)
            out.push_back({ first, e - first });
    };
    split(0, n_gen, pl->chunks_gen);
    split(n_gen, pl->n_tasks, pl->chunks_nar);
}

// Whether a PU shape's items go to the narrow-PU kernel variant.
static bool uses_narrow(const DevParams &p, int part)
{
    return xme_narrow_shape(part) && xme_fs_kmax() == 8 && xme_tune_lookup(part, p.R).K == XME_NARROW_K;
}

// Tile shape of the items of a PU shape over candidate rows [row_lo, row_hi):
// K rows per lane and LC lanes per candidate row (the tuned choice, with a
// band no taller than the rows asked for when a window is split into row
// ranges: more lanes per row, fewer row groups).  Returns the item count.
static int band_shape(const DevParams &p, int part, int row_lo, int row_hi, int &K, int &LC)
{
    const TuneChoice tc = xme_tune_lookup(part, p.R);
    if (uses_narrow(p, part)) {
        K = XME_NARROW_K;                         // the short-PU kernel variant
    } else {
        const int Kt = std::min(std::min(tc.K, 8), xme_fs_kmax());
        K = (xme_pu_h(part) % Kt == 0) ? Kt : 4;
    }
    LC = tc.LC;
    while (LC < 64 && (64 / LC) * K > row_hi - row_lo)
        LC *= 2;
    const int band_rows = (64 / LC) * K, band_cols = 4 * LC, R2 = 2 * p.R;
    return ((row_hi - row_lo + band_rows - 1) / band_rows) * ((R2 + band_cols - 1) / band_cols);
}

static void items_for(const DevParams &p, const std::vector<DevPu> &pus, const std::vector<int> &group, int part,
                      int row_lo, int row_hi, std::vector<HostItem> &out)
{
    const DevPu &pu = pus[group[0] / p.n_ref];
    int K, LC;
    band_shape(p, part, row_lo, row_hi, K, LC);
    const int band_rows = (64 / LC) * K, band_cols = 4 * LC;
    const int R2 = 2 * p.R;
    for (int r0 = row_lo; r0 < row_hi; r0 += band_rows)
        for (int c0 = 0; c0 < R2; c0 += band_cols) {
            HostItem h;
            h.it.list = 0;
            h.it.count = (int16_t)group.size();
            h.it.row0 = (int16_t)r0;
            h.it.nrows = (int16_t)std::min(band_rows, row_hi - r0);
            h.it.col0 = (int16_t)c0;
            h.it.ncols = (int16_t)std::min(band_cols, R2 - c0);
            h.it.K = (uint8_t)K;
            h.it.LC = (uint8_t)LC;
            h.it.pad[0] = h.it.pad[1] = 0;
            h.it.cost = (uint16_t)std::min<int64_t>(65535, ((int64_t)pu.w * pu.h * group.size()) >> 4);
            h.insts = group;
            out.push_back(h);
        }
}

// Split instances of one shape into groups whose items carry similar work
// (about 4096 pixels of PU area per item: enough calls per item to hide the
// latency of fetching the next item's instances), with at least 4 items per
// shape so that each of the 4 waves of a task's workgroup gets work
// (items_per_group(part): the bands one group yields).  MVPs may differ inside
// a group: a wave keeps its cached MV costs while consecutive instances share
// search centre and MVP, so instances are kept in PU order, MVPs together.
template <typename ItemsPerGroup>
static std::vector<std::vector<int>> make_groups(const std::vector<DevPu> &pus, const std::vector<int> &insts,
                                                 int n_ref, int max_group, ItemsPerGroup items_per_group)
{
    std::map<int, std::vector<int>> by_key;
    std::vector<int> sorted(insts);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) {
        const DevPu &pa = pus[a / n_ref], &pb = pus[b / n_ref];
        return std::make_pair(pa.mvp_x, pa.mvp_y) < std::make_pair(pb.mvp_x, pb.mvp_y);
    });
    for (int id : sorted)
        by_key[pus[id / n_ref].part].push_back(id);
    std::vector<std::vector<int>> out;
    for (auto &kv : by_key) {
        const DevPu &pu = pus[kv.second[0] / n_ref];
        const int n = (int)kv.second.size();
        const int min_groups = (4 + items_per_group(kv.first) - 1) / items_per_group(kv.first);
        const int g = std::max(1, std::min(std::min(max_group, 4096 / (pu.w * pu.h)),
                                           (n + min_groups - 1) / min_groups));
        for (size_t i = 0; i < kv.second.size(); i += g)
            out.emplace_back(kv.second.begin() + i, kv.second.begin() + std::min(kv.second.size(), i + g));
    }
    return out;
}

// Build a task from items of one reference.  Returns the LDS bytes it needs.
static size_t make_task(const DevParams &p, const std::vector<DevPu> &pus, const std::vector<HostItem> &its,
                        int ref, FsTask &t)
{
    Rect win = { 1 << 30, 1 << 30, -(1 << 30), -(1 << 30) };
    Rect cur = win;
    for (const HostItem &h : its) {
        for (int id : h.insts) {
            const DevPu &pu = pus[id / p.n_ref];
            int c[2][2];
            int n = possible_centres(p, pu, c);
            for (int k = 0; k < n; k++)
                win.add(item_footprint(p, pu, h.it, c[k][0], c[k][1]));
            cur.add(Rect{ pu.x, pu.y, pu.x + pu.w, pu.y + pu.h });
        }
    }
    t.ref = ref;
    const int wx0 = (int)std::floor(win.x0 / 4.0) * 4;
    t.wx0 = (int16_t)wx0;
    t.wy0 = (int16_t)win.y0;
    t.ww = (int16_t)(((win.x1 - wx0) + 3) & ~3);
    t.wh = (int16_t)(win.y1 - win.y0);
    t.cx0 = (int16_t)cur.x0;
    t.cy0 = (int16_t)cur.y0;
    t.cw = (int16_t)(((cur.x1 - cur.x0) + 3) & ~3);
    t.ch = (int16_t)(cur.y1 - cur.y0);
    const size_t staged = (size_t)window_pitch_host(t.ww) * t.wh + (size_t)t.cw * t.ch;
    return ((staged + 15) & ~(size_t)15) + XME_FS_TERM_BYTES;
}

} // namespace

static void plan_free(xme_plan *pl)
{
    if (pl->arena) {
        delete pl;
        return;
    }
    hipFree(pl->d_pus);
    hipFree(pl->d_inst);
    hipFree(pl->d_best);
    hipFree(pl->d_tasks);
    hipFree(pl->d_items);
    hipFree(pl->d_item_insts);
    hipFree(pl->d_sp_list);
    hipFree(pl->d_pat_list);
    hipFree(pl->d_pat_points);
    hipFree(pl->d_umh);
    hipFree(pl->d_counters);
    hipFree(pl->d_desc);
    for (auto &s : pl->slots) {
        if (s.host)
            hipHostFree(s.host);
        if (s.ev)
            hipEventDestroy(s.ev);
    }
    if (pl->graph_exec)
        hipGraphExecDestroy(pl->graph_exec);
    if (pl->graph)
        hipGraphDestroy(pl->graph);
    delete pl;
}

extern "C" int xme_plan_destroy(xme_plan *pl)
{
    if (pl)
        plan_free(pl);
    return XME_OK;
}

// Device memory of a plan: hipMalloc, or (arena plans) the next 256-byte
// aligned block of the arena.
static hipError_t plan_dev_alloc(xme_plan *pl, void **d, size_t n)
{
    n = n ? n : 16;
    if (!pl->arena) {
        pl->dev_bytes += n;
        return hipMalloc(d, n);
    }
    const size_t at = (pl->arena_used + 255) & ~(size_t)255;
    if (at + n > pl->arena_cap) {
        *d = nullptr;
        pl->arena_used = pl->arena_cap + 1;    // marks the overflow for the caller
        return hipErrorOutOfMemory;
    }
    *d = pl->arena + at;
    pl->arena_used = at + n;
    pl->dev_bytes = pl->arena_used;
    return hipSuccess;
}

static hipError_t plan_upload(xme_plan *pl, void *d, const void *h, size_t n)
{
    if (!n)
        return hipSuccess;
    return pl->arena ? hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, pl->arena_stream)
                     : hipMemcpy(d, h, n, hipMemcpyHostToDevice);
}

static hipError_t plan_zero(xme_plan *pl, void *d, size_t n)
{
    return pl->arena ? hipMemsetAsync(d, 0, n, pl->arena_stream) : hipMemset(d, 0, n);
}

static int plan_create_impl(const xme_me_params *prm, const xme_pu *pus_in, int n_pu, int n_ref, int mode,
                            uint8_t *arena, size_t arena_cap, hipStream_t arena_stream, xme_plan **out);

extern "C" int xme_plan_create(const xme_me_params *prm, const xme_pu *pus_in, int n_pu, int n_ref, int mode,
                               xme_plan **out)
{
    return plan_create_impl(prm, pus_in, n_pu, n_ref, mode, nullptr, 0, nullptr, out);
}

static int plan_create_impl(const xme_me_params *prm, const xme_pu *pus_in, int n_pu, int n_ref, int mode,
                            uint8_t *arena, size_t arena_cap, hipStream_t arena_stream, xme_plan **out)
{
    if (!prm || !pus_in || !out || n_pu <= 0 || n_ref <= 0 || n_ref > XME_MAX_REFS ||
        (mode != XME_PLAN_BATCHED && mode != XME_PLAN_PLAIN)) {
        xme_set_error("xme_plan_create: bad arguments");
        return XME_ERR_ARG;
    }
    if (prm->lambda_factor < 0 || prm->lambda_factor >= (1 << 24)) {
        xme_set_error("xme_plan_create: lambda_factor must lie in [0, 2^24) (sqrt(lambda) < 256)");
        return XME_ERR_ARG;
    }
    if (prm->search_range < 1 || prm->search_range > 128 || prm->frame_w < 8 || prm->frame_h < 8 ||
        prm->frame_w > 8192 || prm->frame_h > 8192 || prm->use_fractional_me < 0 || prm->use_fractional_me > 2) {
        xme_set_error("xme_plan_create: search range must be 1..128 and the frame 8..8192 pixels");
        return XME_ERR_ARG;
    }
    if (prm->me_method < XME_ME_FS || prm->me_method > XME_ME_TZ || prm->umh_big_hex_level < 0 ||
        prm->umh_big_hex_level > 2) {
        xme_set_error("xme_plan_create: me_method %d / umh_big_hex_level %d: full search (0), DIA (1), HEX (2), "
                      "UMH (3, big-hexagon level 0..2) and TZ (4) are provided", prm->me_method,
                      prm->umh_big_hex_level);
        return XME_ERR_ARG;
    }
    DevParams p;
    p.method = prm->me_method;
    p.big_hex = prm->umh_big_hex_level;
    p.frame_w = prm->frame_w;
    p.frame_h = prm->frame_h;
    p.R = prm->search_range;
    p.lambda = prm->lambda_factor;
    p.min_mv[0] = prm->min_mv_range[0];
    p.min_mv[1] = prm->min_mv_range[1];
    p.max_mv[0] = prm->max_mv_range[0];
    p.max_mv[1] = prm->max_mv_range[1];
    p.hadamard = prm->use_hadamard ? 1 : 0;
    p.frac = prm->use_fractional_me;
    p.fast = prm->use_fast_sub_me ? 1 : 0;
    p.pmvr = prm->enable_pmvr ? 1 : 0;
    p.n_pu = n_pu;
    p.n_ref = n_ref;

    const int max_mvd = xme_mvbits_max_mvd(p.R);
    p.max_mvd = max_mvd;
    std::vector<DevPu> pus(n_pu);
    std::vector<int> parts(n_pu);
    for (int i = 0; i < n_pu; i++) {
        const xme_pu &q = pus_in[i];
        int part = xme_part_index(q.w, q.h);
        if (part < 0 || (q.x & 3) || q.x < 0 || q.y < 0 || q.x + q.w > p.frame_w || q.y + q.h > p.frame_h) {
            xme_set_error("xme_plan_create: PU %d (%d,%d %dx%d) is not a LumaPU inside the frame at x %% 4 == 0",
                          i, q.x, q.y, q.w, q.h);
            return XME_ERR_ARG;
        }
        DevPu &d = pus[i];
        d.x = q.x;
        d.y = q.y;
        d.w = q.w;
        d.h = q.h;
        d.part = (uint8_t)part;
        d.pad0 = 0;
        d.mvp_x = q.mvp_x;
        d.mvp_y = q.mvp_y;
        d.pad1 = 0;
        parts[i] = part;
        MvWindow w = xme_mv_window(p, d.x, d.y, d.w, d.h);
        // every MV difference the search can form must stay inside mvbits
        // (the pattern searches evaluate points up to 2 pixels outside the
        // window)
        const int margin = p.method == XME_ME_FS ? 1 : 2;
        for (int a = 0; a < 2; a++) {
            const int pm = a ? d.mvp_y : d.mvp_x;
            const int lo = std::min(w.min_fpel[a], (pm + 2) >> 2) - margin;
            const int hi = std::max(w.max_fpel[a], (pm + 2) >> 2) + margin;
            if (std::abs(4 * lo - pm) + 3 > max_mvd || std::abs(4 * hi - pm) + 3 > max_mvd) {
                xme_set_error("xme_plan_create: PU %d: MV differences leave the mvbits table", i);
                return XME_ERR_ARG;
            }
        }
    }

    if (arena) {
        // what every plan holds at least: the PU records, and per instance its
        // state, best key and place in an item list
        const size_t floor_bytes = (size_t)n_pu * sizeof(DevPu) +
                                   (size_t)n_pu * n_ref * (sizeof(DevInst) + sizeof(uint64_t) + sizeof(int32_t));
        if (floor_bytes > arena_cap) {
            xme_set_error("xme_me_search_static: the job's plan needs at least %zu bytes, more than the %zu bytes of "
                          "static scratch (fewer PUs or references per call)", floor_bytes, arena_cap);
            return XME_ERR_BUDGET;
        }
    }
    xme_plan *pl = new xme_plan();
    hipGetDevice(&pl->device);
    pl->arena = arena;
    pl->arena_cap = arena_cap;
    pl->arena_stream = arena_stream;
    pl->mode = mode;
    pl->kmax = xme_fs_kmax();
    pl->p = p;
    pl->n_pu = n_pu;
    pl->n_ref = n_ref;
    pl->n_inst = n_pu * n_ref;

    // work accounting: legal integer candidates and absolute differences
    pl->n_cand = 0;
    pl->n_absdiff = 0;
    for (int i = 0; i < n_pu && p.method == XME_ME_FS; i++) {
        MvWindow w = xme_mv_window(p, pus[i].x, pus[i].y, pus[i].w, pus[i].h);
        int c[2][2];
        possible_centres(p, pus[i], c);
        // with the candidate-list winner unknown on the host, count around the
        // clipped (0,0) centre (exact whenever the MVP is (0,0))
        const int cx = c[0][0], cy = c[0][1];
        int64_t nx = 0, ny = 0;
        for (int k = -p.R; k < p.R; k++) {
            nx += (cx + k >= w.min_fpel[0] && cx + k <= w.max_fpel[0]);
            ny += (cy + k >= w.min_fpel[1] && cy + k <= w.max_fpel[1]);
        }
        pl->n_cand += nx * ny * n_ref;
        pl->n_absdiff += nx * ny * n_ref * pus[i].w * pus[i].h;
    }

    const size_t lds_cap = (size_t)xme_fs_max_lds();
    // search windows are split until this many workgroups share a CU's LDS
    const size_t lds_share = lds_cap / (size_t)xme_fs_wg_per_cu();
    std::vector<FsTask> tasks;
    std::vector<FsItem> items;
    std::vector<int32_t> item_insts;
    std::vector<int> task_part;
    const int R2 = 2 * p.R;

    auto emit_task = [&](const std::vector<HostItem> &its, int ref, int part) -> size_t {
        FsTask t;
        size_t lds = make_task(p, pus, its, ref, t);
        t.item_begin = (int)items.size();
        // larger PUs first inside a task: the last items to finish are small
        std::vector<HostItem> sorted = its;
        std::stable_sort(sorted.begin(), sorted.end(),
                         [](const HostItem &a, const HostItem &b) { return a.it.cost > b.it.cost; });
        for (const HostItem &h : sorted) {
            FsItem it = h.it;
            it.list = (int32_t)item_insts.size();
            it.count = (int16_t)h.insts.size();
            for (int id : h.insts)
                item_insts.push_back(id);
            items.push_back(it);
        }
        t.item_end = (int)items.size();
        tasks.push_back(t);
        task_part.push_back(part);
        return lds;
    };

    std::vector<int32_t> pat_list;
    if (p.method != XME_ME_FS) {
        // pattern searches: one launch in either mode, the instances of each
        // PU-size class packed 64 / S per wave, the largest class first.
        // A class lists its instances reference-major in PU order and is
        // padded to whole workgroups of 4 waves: the kernel hands each XCD a
        // contiguous run of a class's workgroups, so one L2 sees neighbouring
        // PUs of one reference plane.
        static const int per_wave[6] = { 1, 1, 2, 4, 8, 16 };
        int waves = 0;
        std::vector<int32_t> cls[6];
        for (int r = 0; r < n_ref; r++)
            for (int i = 0; i < n_pu; i++)
                cls[5 - xme_pattern_class(pus[i].w, pus[i].h)].push_back(i * n_ref + r);
        for (int k = 0; k < 6; k++) {
            for (int32_t id : cls[k])
                pat_list.push_back(id);
            const int nw = ((((int)cls[k].size() + per_wave[k] - 1) / per_wave[k]) + 3) & ~3;
            pat_list.resize(pat_list.size() + (size_t)(nw * per_wave[k] - (int)cls[k].size()), -1);
            waves += nw;
            pl->pat_wave_end[k] = waves;
        }
    } else if (mode == XME_PLAN_BATCHED) {
        // group PUs by 64x64 CTU; split the candidate rows until the window fits.
        // Pass 0 builds the generic tasks, pass 1 the narrow-PU tasks.
        std::map<std::pair<int, int>, std::vector<int>> ctus;
        for (int i = 0; i < n_pu; i++)
            ctus[{ pus[i].y >> 6, pus[i].x >> 6 }].push_back(i);
        size_t lds_pass[2] = { 0, 0 };
        int begin_pass[2] = { 0, 0 };
        for (int pass = 0; pass < 2; pass++) {
            begin_pass[pass] = (int)tasks.size();
            for (auto &kv : ctus) {
                std::vector<int> members;
                for (int i : kv.second)
                    if (uses_narrow(p, pus[i].part) == (pass == 1))
                        members.push_back(i);
                if (members.empty())
                    continue;
                for (int r = 0; r < n_ref; r++) {
                    int splits = 1;
                    for (;;) {
                        bool ok = true;
                        std::vector<std::vector<HostItem>> groups(splits);
                        const int rows_per = (R2 + splits - 1) / splits;
                        std::vector<int> insts;
                        for (int i : members)
                            insts.push_back(i * n_ref + r);
                        const auto inst_groups = make_groups(pus, insts, n_ref, XME_MAX_GROUP, [&](int part) {
                            int K, LC;
                            return band_shape(p, part, 0, rows_per, K, LC);
                        });
                        for (int s = 0; s < splits; s++) {
                            for (const auto &ig : inst_groups)
                                items_for(p, pus, ig, pus[ig[0] / n_ref].part, s * rows_per,
                                          std::min(R2, (s + 1) * rows_per), groups[s]);
                            FsTask t;
                            if (!groups[s].empty() && make_task(p, pus, groups[s], r, t) > lds_share)
                                ok = false;
                        }
                        if (ok || splits >= R2 / 4) {
                            for (auto &g : groups)
                                if (!g.empty())
                                    lds_pass[pass] = std::max(lds_pass[pass], emit_task(g, r, -1));
                            break;
                        }
                        splits *= 2;
                    }
                }
            }
            // longest tasks first (all CTU tasks have similar work; edges are smaller)
            auto work = [&](const FsTask &t) {
                int64_t w = 0;
                for (int k = t.item_begin; k < t.item_end; k++)
                    w += items[k].cost;
                return w;
            };
            std::stable_sort(tasks.begin() + begin_pass[pass], tasks.end(),
                             [&](const FsTask &a, const FsTask &b) { return work(a) > work(b); });
        }
        pl->lds_bytes = lds_pass[0];
        pl->lds_narrow = lds_pass[1];
        pl->n_tasks_narrow = (int)tasks.size() - begin_pass[1];
    } else {
        // plain full search: every (PU, ref) is its own task; one launch per PU shape
        for (int part = 0; part < XME_NUM_PU; part++) {
            LaunchGroup g;
            g.task_begin = (int)tasks.size();
            g.lds = 0;
            g.variant = uses_narrow(p, part) ? XME_NARROW_K : xme_fs_kmax();
            for (int i = 0; i < n_pu; i++) {
                if (parts[i] != part)
                    continue;
                for (int r = 0; r < n_ref; r++) {
                    int splits = 1;
                    for (;;) {
                        const int rows_per = (R2 + splits - 1) / splits;
                        std::vector<std::vector<HostItem>> groups(splits);
                        bool ok = true;
                        for (int s = 0; s < splits; s++) {
                            items_for(p, pus, std::vector<int>{ i * n_ref + r }, part, s * rows_per,
                                      std::min(R2, (s + 1) * rows_per), groups[s]);
                            FsTask t;
                            if (make_task(p, pus, groups[s], r, t) > lds_share)
                                ok = false;
                        }
                        if (ok || splits >= R2 / 4) {
                            for (auto &gg : groups)
                                if (!gg.empty())
                                    g.lds = std::max(g.lds, emit_task(gg, r, part));
                            break;
                        }
                        splits *= 2;
                    }
                }
            }
            g.task_end = (int)tasks.size();
            if (g.task_end > g.task_begin)
                pl->groups.push_back(g);
        }
    }
    pl->n_tasks = (int)tasks.size();
    pl->task_work.resize(tasks.size());
    for (size_t ti = 0; ti < tasks.size(); ti++) {
        int64_t w = 0;
        for (int k = tasks[ti].item_begin; k < tasks[ti].item_end; k++) {
            const FsItem &it = items[k];
            const DevPu &pu = pus[item_insts[it.list] / n_ref];
            w += (int64_t)it.count * pu.w * pu.h * it.nrows * it.ncols;
        }
        pl->task_work[ti] = w;
    }
    split_launches(pl, 0);

    for (const FsTask &t : tasks) {
        size_t need = (((size_t)window_pitch_host(t.ww) * t.wh + (size_t)t.cw * t.ch + 15) & ~(size_t)15) +
                      XME_FS_TERM_BYTES;
        if (need > lds_cap) {
            xme_set_error("xme_plan_create: a search window needs %zu bytes of LDS", need);
            plan_free(pl);
            return XME_ERR_ARG;
        }
    }

    hipError_t e = hipSuccess;
    auto up = [&](void **d, const void *h, size_t n) {
        if (e != hipSuccess)
            return;
        e = plan_dev_alloc(pl, d, n);
        if (e == hipSuccess)
            e = plan_upload(pl, *d, h, n);
    };
    up((void **)&pl->d_pus, pus.data(), pus.size() * sizeof(DevPu));
    up((void **)&pl->d_tasks, tasks.data(), tasks.size() * sizeof(FsTask));
    up((void **)&pl->d_items, items.data(), items.size() * sizeof(FsItem));
    up((void **)&pl->d_item_insts, item_insts.data(), item_insts.size() * sizeof(int32_t));
    up((void **)&pl->d_pat_list, pat_list.data(), pat_list.size() * sizeof(int32_t));
    if (e == hipSuccess && p.method != XME_ME_FS) {
        e = plan_dev_alloc(pl, (void **)&pl->d_pat_points, sizeof(uint32_t) * pl->n_inst);
        if (e == hipSuccess)
            e = plan_zero(pl, pl->d_pat_points, sizeof(uint32_t) * pl->n_inst);
        pl->inst_area.resize(pl->n_inst);
        for (int i = 0; i < pl->n_inst; i++)
            pl->inst_area[i] = pus[i / n_ref].w * pus[i / n_ref].h;
    }
    if (e == hipSuccess && p.method == XME_ME_UMH)
        e = plan_dev_alloc(pl, (void **)&pl->d_umh, sizeof(xme_umh_ctx) * pl->n_inst);
    std::vector<int32_t> sp_list;      // lives until the arena uploads have been consumed
    {
        sp_list.reserve(pus.size());
        for (int large = 0; large < 2; large++)
            for (size_t i = 0; i < pus.size(); i++)
                if (xme_subpel_large(pus[i].w, pus[i].h) == (large == 1))
                    sp_list.push_back((int32_t)i);
        pl->n_sp_large = 0;
        for (const DevPu &q : pus)
            pl->n_sp_large += xme_subpel_large(q.w, q.h) ? 1 : 0;
        pl->n_sp_small = (int)pus.size() - pl->n_sp_large;
        up((void **)&pl->d_sp_list, sp_list.data(), sp_list.size() * sizeof(int32_t));
    }
    if (e == hipSuccess)
        e = plan_dev_alloc(pl, (void **)&pl->d_inst, sizeof(DevInst) * pl->n_inst);
    if (e == hipSuccess)
        e = plan_dev_alloc(pl, (void **)&pl->d_best, sizeof(uint64_t) * pl->n_inst);
    // the search launches' queue counters start at zero (the last workgroup
    // of a launch resets them)
    if (e == hipSuccess)
        e = plan_dev_alloc(pl, (void **)&pl->d_counters, 64);
    if (e == hipSuccess)
        e = plan_zero(pl, pl->d_counters, 64);
    if (e == hipSuccess)
        e = plan_dev_alloc(pl, (void **)&pl->d_desc, sizeof(FrameDesc));
    for (auto &s : pl->slots) {
        s.host = nullptr;
        s.ev = nullptr;
        s.used = false;
        if (pl->arena)
            continue;              // the descriptor goes up from arena_desc
        if (e == hipSuccess)
            e = hipHostMalloc((void **)&s.host, sizeof(FrameDesc), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&s.ev, hipEventDisableTiming);
    }
    // arena uploads are asynchronous copies from the host lists above, which
    // end with this call
    if (e == hipSuccess && pl->arena)
        e = hipStreamSynchronize(arena_stream);
    if (e != hipSuccess) {
        int rc;
        if (pl->arena)
            (void)hipStreamSynchronize(arena_stream);   // copies already queued read the host lists
        if (pl->arena && pl->arena_used > pl->arena_cap) {
            xme_set_error("xme_me_search_static: the job's plan needs more than the %zu bytes of static scratch "
                          "(fewer PUs or references per call)", pl->arena_cap);
            rc = XME_ERR_BUDGET;
        } else {
            rc = xme_hip_fail(e, "xme_plan_create");
        }
        plan_free(pl);
        return rc;
    }
    if (mode == XME_PLAN_BATCHED && p.method == XME_ME_FS) {
        int prc = xme_fs_prepare();
        if (prc) {
            plan_free(pl);
            return prc;
        }
        int cus = 0;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
        auto fill = [&](int variant, size_t lds, int n) {
            int per_cu = 0;
            hipError_t oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, xme_fs_kernel_ptr(true, variant),
                                                                          256, lds);
            if (oe != hipSuccess || per_cu < 1)
                per_cu = 1;
            return std::max(1, std::min(per_cu * cus, n));
        };
        pl->grid_full = fill(pl->kmax, pl->lds_bytes, pl->n_tasks - pl->n_tasks_narrow);
        pl->grid = pl->grid_full;
        pl->grid_narrow_full = fill(XME_NARROW_K, pl->lds_narrow, pl->n_tasks_narrow);
        pl->grid_narrow = pl->grid_narrow_full;
    }
    *out = pl;
    return XME_OK;
}

extern "C" int xme_plan_info(const xme_plan *pl, int64_t *n_inst, int64_t *n_cand, int64_t *n_absdiff)
{
    if (!pl)
        return XME_ERR_ARG;
    if (n_inst)
        *n_inst = pl->n_inst;
    if (n_cand)
        *n_cand = pl->n_cand;
    if (n_absdiff)
        *n_absdiff = pl->n_absdiff;
    return XME_OK;
}

extern "C" int xme_plan_set_umh_context(xme_plan *pl, const xme_umh_ctx *ctx, void *stream)
{
    if (!pl || !ctx || pl->p.method != XME_ME_UMH) {
        xme_set_error("xme_plan_set_umh_context: needs a UMH plan and a context array");
        return XME_ERR_ARG;
    }
    XME_HIP(hipMemcpyAsync(pl->d_umh, ctx, sizeof(xme_umh_ctx) * pl->n_inst, hipMemcpyHostToDevice,
                           (hipStream_t)stream));
    pl->umh_set = true;
    return XME_OK;
}

extern "C" int xme_plan_set_umh_states(xme_plan *pl, const xme_umh_state *d_states, void *stream)
{
    if (!pl || !d_states || pl->p.method != XME_ME_UMH) {
        xme_set_error("xme_plan_set_umh_states: needs a UMH plan and a state array");
        return XME_ERR_ARG;
    }
    XME_HIP(hipMemcpy2DAsync(pl->d_umh, sizeof(xme_umh_ctx), &d_states->ctx, sizeof(xme_umh_state),
                             sizeof(xme_umh_ctx), (size_t)pl->n_inst, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    pl->umh_set = true;
    return XME_OK;
}

extern "C" int xme_plan_search_count(const xme_plan *pl, int64_t *n_points, int64_t *n_absdiff)
{
    if (!pl)
        return XME_ERR_ARG;
    int64_t pts = 0, ad = 0;
    if (pl->d_pat_points) {
        std::vector<uint32_t> h(pl->n_inst);
        XME_HIP(hipMemcpy(h.data(), pl->d_pat_points, sizeof(uint32_t) * pl->n_inst, hipMemcpyDeviceToHost));
        for (int i = 0; i < pl->n_inst; i++) {
            pts += h[i];
            ad += (int64_t)h[i] * pl->inst_area[i];
        }
    }
    if (n_points)
        *n_points = pts;
    if (n_absdiff)
        *n_absdiff = ad;
    return XME_OK;
}

extern "C" int xme_plan_set_max_workgroups(xme_plan *pl, int max_wg)
{
    if (!pl || max_wg < 0) {
        xme_set_error("xme_plan_set_max_workgroups: bad arguments");
        return XME_ERR_ARG;
    }
    if (pl->mode != XME_PLAN_BATCHED)
        return XME_OK;
    const int g = max_wg == 0 ? pl->grid_full : std::max(1, std::min(max_wg, pl->grid_full));
    const int gn = max_wg == 0 ? pl->grid_narrow_full : std::max(1, std::min(max_wg, pl->grid_narrow_full));
    if ((g != pl->grid || gn != pl->grid_narrow) && pl->graph_exec) {
        // the captured launch holds the old grid
        hipGraphExecDestroy(pl->graph_exec);
        hipGraphDestroy(pl->graph);
        pl->graph_exec = nullptr;
        pl->graph = nullptr;
    }
    pl->grid = g;
    pl->grid_narrow = gn;
    return XME_OK;
}

extern "C" int xme_plan_set_launch_split(xme_plan *pl, int64_t max_absdiff_per_launch)
{
    if (!pl || max_absdiff_per_launch < 0) {
        xme_set_error("xme_plan_set_launch_split: bad arguments");
        return XME_ERR_ARG;
    }
    if (pl->mode != XME_PLAN_BATCHED)
        return XME_OK;
    split_launches(pl, max_absdiff_per_launch);
    if (pl->graph_exec) {
        // the captured graph holds the old launches
        hipGraphExecDestroy(pl->graph_exec);
        hipGraphDestroy(pl->graph);
        pl->graph_exec = nullptr;
        pl->graph = nullptr;
    }
    return XME_OK;
}

extern "C" int xme_plan_workgroups(const xme_plan *pl, int *grid, int *grid_full)
{
    if (!pl)
        return XME_ERR_ARG;
    if (grid)
        *grid = pl->grid;
    if (grid_full)
        *grid_full = pl->grid_full;
    return XME_OK;
}

size_t xme_plan_device_bytes(const xme_plan *pl) { return pl ? pl->dev_bytes : 0; }

int xme_plan_debug(const xme_plan *pl, int *n_tasks, int64_t *lds_bytes, int *grid)
{
    *n_tasks = pl->n_tasks;
    *lds_bytes = (int64_t)pl->lds_bytes;
    *grid = pl->grid;
    return XME_OK;
}

static int fill_desc(xme_plan *pl, const pel_t *cur, intptr_t i_cur, const xme_ref_desc *refs, xme_result *res,
                     int32_t *cost_map, hipStream_t s)
{
    if (!cur || !refs || !res || (reinterpret_cast<uintptr_t>(cur) & 3) || (i_cur & 3)) {
        xme_set_error("xme_me_run: the current plane must be 4-byte aligned with a stride multiple of 4");
        return XME_ERR_ARG;
    }
    for (int r = 0; r < pl->n_ref; r++) {
        if (!refs[r].planes[0] || (reinterpret_cast<uintptr_t>(refs[r].planes[0]) & 3) || (refs[r].stride & 3) ||
            refs[r].stride < pl->p.frame_w + 2 * XME_PAD) {
            xme_set_error("xme_me_run: reference %d needs a 4-byte aligned padded plane (stride >= width + %d)", r,
                          2 * XME_PAD);
            return XME_ERR_ARG;
        }
        for (int k = 1; k < 16; k++) {
            const bool need = pl->p.frac == 2 || (pl->p.frac == 1 && (k == 2 || k == 8 || k == 10));
            if (need && (!refs[r].planes[k] || (reinterpret_cast<uintptr_t>(refs[r].planes[k]) & 3))) {
                xme_set_error("xme_me_run: reference %d lacks interpolated plane %d", r, k);
                return XME_ERR_ARG;
            }
        }
    }
    DescSlot &slot = pl->slots[pl->slot_next];
    if (!pl->arena) {
        pl->slot_next = (pl->slot_next + 1) % 8;
        if (slot.used)
            XME_HIP(hipEventSynchronize(slot.ev));
    }
    FrameDesc *d = pl->arena ? &pl->arena_desc : slot.host;
    memset(d, 0, sizeof(*d));
    d->cur = cur;
    d->i_cur = i_cur;
    for (int r = 0; r < pl->n_ref; r++) {
        for (int k = 0; k < 16; k++)
            d->planes[r][k] = refs[r].planes[k];
        d->stride[r] = refs[r].stride;
    }
    d->results = res;
    d->cost_map = cost_map;
    XME_HIP(hipMemcpyAsync(pl->d_desc, d, sizeof(FrameDesc), hipMemcpyHostToDevice, s));
    if (!pl->arena) {
        XME_HIP(hipEventRecord(slot.ev, s));
        slot.used = true;
    }
    return XME_OK;
}

// The integer search: batched plans run the generic tasks and then the
// narrow-PU tasks as two persistent launches (each with its own queue
// counters); plain plans launch once per PU shape.
static int launch_search(xme_plan *pl, hipStream_t s)
{
    int rc = XME_OK;
    if (pl->p.method != XME_ME_FS) {
        if (pl->p.method == XME_ME_UMH && !pl->umh_set) {
            xme_set_error("xme_me_run: a UMH plan needs xme_plan_set_umh_context before it runs");
            return XME_ERR_STATE;
        }
        return xme_launch_pattern(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_best, pl->d_pat_list,
                                  pl->pat_wave_end, pl->d_pat_points, pl->d_umh, s);
    }
    if (pl->mode == XME_PLAN_BATCHED) {
        // launches of one variant run in stream order and share its queue
        // counters (the last workgroup of a launch resets them)
        for (const auto &c : pl->chunks_gen) {
            rc = xme_launch_fs(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_tasks + c.first, c.second, pl->d_items,
                               pl->d_item_insts, pl->d_best, pl->d_counters, pl->lds_bytes,
                               std::min(pl->grid, c.second), true, pl->kmax, s);
            if (rc)
                return rc;
        }
        for (const auto &c : pl->chunks_nar) {
            rc = xme_launch_fs(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_tasks + c.first, c.second, pl->d_items,
                               pl->d_item_insts, pl->d_best, pl->d_counters + 4, pl->lds_narrow,
                               std::min(pl->grid_narrow, c.second), true, XME_NARROW_K, s);
            if (rc)
                return rc;
        }
        return XME_OK;
    }
    for (const LaunchGroup &g : pl->groups) {
        rc = xme_launch_fs(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_tasks + g.task_begin,
                           g.task_end - g.task_begin, pl->d_items, pl->d_item_insts, pl->d_best, pl->d_counters, g.lds, 0,
                           false, g.variant, s);
        if (rc)
            return rc;
    }
    return XME_OK;
}

static int launch_all(xme_plan *pl, hipStream_t s)
{
    int rc = xme_launch_mvc(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_best, pl->n_inst, s);
    if (rc)
        return rc;
    rc = launch_search(pl, s);
    if (rc)
        return rc;
    return xme_launch_subpel(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_best, pl->d_sp_list, pl->n_sp_small,
                             pl->n_sp_large, s);
}

extern "C" int xme_me_run(xme_plan *pl, const pel_t *cur, intptr_t i_cur, const xme_ref_desc *refs,
                          xme_result *res, int32_t *cost_map, void *stream)
{
    if (!pl)
        return XME_ERR_ARG;
    if (cost_map && pl->p.method != XME_ME_FS) {
        xme_set_error("xme_me_run: cost maps are defined for the full search only");
        return XME_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc = fill_desc(pl, cur, i_cur, refs, res, cost_map, s);
    if (rc)
        return rc;
    return launch_all(pl, s);
}

extern "C" int xme_me_run_graph(xme_plan *pl, const pel_t *cur, intptr_t i_cur, const xme_ref_desc *refs,
                                xme_result *res, void *stream)
{
    if (!pl)
        return XME_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (s == nullptr) {
        xme_set_error("xme_me_run_graph: needs a created stream (the null stream cannot be captured)");
        return XME_ERR_ARG;
    }
    int rc = fill_desc(pl, cur, i_cur, refs, res, nullptr, s);
    if (rc)
        return rc;
    if (!pl->graph_exec) {
        XME_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = launch_all(pl, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g)
                hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return xme_hip_fail(e, "hipStreamEndCapture");
        pl->graph = g;
        XME_HIP(hipGraphInstantiate(&pl->graph_exec, g, nullptr, nullptr, 0));
    }
    XME_HIP(hipGraphLaunch(pl->graph_exec, s));
    return XME_OK;
}

extern "C" int xme_me_time(xme_plan *pl, const pel_t *cur, intptr_t i_cur, const xme_ref_desc *refs,
                           xme_result *res, void *stream, int reps, double *search_ms, double *job_ms_mean,
                           double *job_ms)
{
    if (!pl || reps < 1)
        return XME_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    std::vector<hipEvent_t> ev(4 * reps, nullptr);
    for (auto &e : ev)
        XME_HIP(hipEventCreate(&e));
    int rc = XME_OK;
    for (int r = 0; r < reps && rc == XME_OK; r++) {
        rc = fill_desc(pl, cur, i_cur, refs, res, nullptr, s);
        if (rc)
            break;
        XME_HIP(hipEventRecord(ev[4 * r + 0], s));
        rc = xme_launch_mvc(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_best, pl->n_inst, s);
        if (rc)
            break;
        XME_HIP(hipEventRecord(ev[4 * r + 1], s));
        rc = launch_search(pl, s);
        if (rc)
            break;
        XME_HIP(hipEventRecord(ev[4 * r + 2], s));
        rc = xme_launch_subpel(pl->p, pl->d_desc, pl->d_pus, pl->d_inst, pl->d_best, pl->d_sp_list, pl->n_sp_small,
                               pl->n_sp_large, s);
        XME_HIP(hipEventRecord(ev[4 * r + 3], s));
    }
    if (rc == XME_OK)
        XME_HIP(hipStreamSynchronize(s));
    double sum_s = 0, sum_j = 0;
    for (int r = 0; r < reps && rc == XME_OK; r++) {
        float a = 0, b = 0;
        XME_HIP(hipEventElapsedTime(&a, ev[4 * r + 1], ev[4 * r + 2]));
        XME_HIP(hipEventElapsedTime(&b, ev[4 * r + 0], ev[4 * r + 3]));
        sum_s += a;
        sum_j += b;
        if (job_ms)
            job_ms[r] = b;
    }
    for (auto &e : ev)
        hipEventDestroy(e);
    if (rc)
        return rc;
    if (search_ms)
        *search_ms = sum_s / reps;
    if (job_ms_mean)
        *job_ms_mean = sum_j / reps;
    return XME_OK;
}

// ---------------------------------------------------------------------------
// Frame-level ME without device allocation.  The plan's device buffers (work
// lists, per-instance state and best keys, queue counters, frame descriptor)
// are carved from this file-scope array, one per GPU for the life of the
// process.  Host code never passes the array's name to HIP: its device
// address comes from hipGetSymbolAddress, once per device.

static __device__ __align__(256) uint8_t xme_scratch[XME_STATIC_SCRATCH_BYTES];

namespace {
std::mutex g_static_mu;                 // one static search per process at a time
uint8_t *g_scratch_ptr[64] = {};        // device address of xme_scratch, per device
}

extern "C" int xme_me_search_static(const xme_me_params *prm, const xme_pu *pus, int n_pu, int n_ref,
                                    const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *refs,
                                    xme_result *d_results, void *stream)
{
    if (!prm || !pus || !refs || !d_results || n_pu <= 0 || n_ref <= 0) {
        xme_set_error("xme_me_search_static: bad arguments");
        return XME_ERR_ARG;
    }
    if (prm->me_method == XME_ME_UMH) {
        xme_set_error("xme_me_search_static: the UMH search needs its per-PU context (xme_plan_set_umh_context)");
        return XME_ERR_ARG;
    }
    int dev = 0;
    XME_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        xme_set_error("xme_me_search_static: device %d out of range", dev);
        return XME_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(xme_scratch));
        if (e != hipSuccess || !ptr)
            return xme_hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(xme_scratch)");
        g_scratch_ptr[dev] = static_cast<uint8_t *>(ptr);
    }
    hipStream_t s = (hipStream_t)stream;
    xme_plan *pl = nullptr;
    int rc = plan_create_impl(prm, pus, n_pu, n_ref, XME_PLAN_BATCHED, g_scratch_ptr[dev], XME_STATIC_SCRATCH_BYTES,
                              s, &pl);
    if (rc)
        return rc;
    rc = fill_desc(pl, d_cur, i_cur, refs, d_results, nullptr, s);
    if (rc == XME_OK)
        rc = launch_all(pl, s);
    // the scratch is free for the next call, and the plan's host copy of the
    // frame descriptor may go, once the kernels have finished
    hipError_t e = hipStreamSynchronize(s);
    plan_free(pl);
    if (rc == XME_OK && e != hipSuccess)
        rc = xme_hip_fail(e, "xme_me_search_static");
    return rc;
}
