// Bi-prediction refinements for batches of PUs, gfx950: the symmetric one
// (xavs2_me_search_sym, source/encoder/me.c:1190-1262) and the
// bi-directional one (xavs2_me_search_bid, me.c:1278-1392).
//
// One wave per job.  A stage's candidates (the half-pel Spiral points around
// the start MV, then the quarter-pel points around the half-pel best) are
// independent: groups of P lanes take one candidate each, a lane a 4x4 tile
// of the PU (more than one when the PU has more than 64 tiles).  A lane
// forms the prediction of its tile as the rounded average of the two
// references' planes at the candidate MV and its mirrored, distance-scaled
// MV, then takes the tile's Hadamard sum (or SAD) against the current PU.
// The group's last lane holds the candidate's cost; the wave then applies the
// reference's sequential strict-'<' update in Spiral order from uniform
// registers.  The bi-directional search keeps the backward MV fixed and
// compares reference 1 against the target clip(2 * cur - pred_bwd), which
// each lane builds once per job for its tiles, in registers.
#include "xme_dpp.h"
#include "xme_satd.h"
#include "xme_internal.h"

namespace {

__constant__ int8_t c_bi_spiral[9][2] = { { 0, 0 }, { 0, -1 }, { 0, 1 }, { -1, -1 }, { 1, -1 },
                                         { -1, 0 }, { 1, 0 }, { -1, 1 }, { 1, 1 } };

typedef const __attribute__((address_space(1))) uint8_t *gbytes;
typedef const __attribute__((address_space(1))) uint32_t *gdwords;

struct SymPlanes {
    const uint8_t *p1[16];
    const uint8_t *p2[16];
};

struct SymParams {
    int frame_w, frame_h;
    int min_mv[2], max_mv[2];
    int lambda, hadamard, frac, pmvr, max_mvd;
};

constexpr int SYM_WAVES = 4;

// One wave per job, SYM_WAVES jobs per workgroup; the grid is a multiple of
// 8 workgroups.  Workgroup b runs on XCD b % 8, and each XCD takes one
// contiguous range of the jobs, so the plane lines that neighbouring PUs
// share meet in one L2.
__host__ __device__ inline int job_grid(int n) { return ((n + SYM_WAVES - 1) / SYM_WAVES + 7) & ~7; }

__device__ __forceinline__ int job_index()
{
    const int per_xcd = (int)gridDim.x >> 3;
    const int lb = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    return __builtin_amdgcn_readfirstlane(lb * SYM_WAVES + (int)(threadIdx.x >> 6));
}
constexpr int SYM_MAX_MV = 8191;     // |MV| bound of a valid job (qpel), see sym_job_ok

// Four bytes at any address; the dword after the aligned one is read, which
// the XME_PAD columns of right padding keep inside the plane.
__device__ __forceinline__ uint32_t sym_load_dw(gbytes p)
{
    const uintptr_t a = (uintptr_t)p;
    gdwords q = (gdwords)(a & ~(uintptr_t)3);
    return __builtin_amdgcn_alignbyte(q[1], q[0], (uint32_t)(a & 3));
}

// (a + b + 1) >> 1 on each of four bytes (g_funcs.pixf.avg, pixel.c:1007-1019).
__device__ __forceinline__ uint32_t avg4(uint32_t a, uint32_t b)
{
    return (a | b) - (((a ^ b) >> 1) & 0x7f7f7f7fu);
}

// Plane k of 16 without dynamic indexing of the kernel-argument array (a
// select chain over uniform registers).
__device__ __forceinline__ const uint8_t *pick_plane(const uint8_t *const (&pl)[16], int k)
{
    const uint8_t *r = pl[0];
#pragma unroll
    for (int q = 1; q < 16; q++)
        r = k == q ? pl[q] : r;
    return r;
}

// scale_mv_skip (source/common/predict.h:148-151).
__device__ __forceinline__ int scale_mv(int mv, int dist_dst, int dist_src)
{
    return (int16_t)((mv * dist_dst * (16384 / dist_src) + 8192) >> 14);
}

__device__ __forceinline__ int64_t scale_mv_wide(int mv, int dist_dst, int dist_src)
{
    return ((int64_t)mv * dist_dst * (16384 / dist_src) + 8192) >> 14;
}

// The qpel MV window of the job (md_inter.c:913-924) packed as in
// xavs2_me_search_sym (me.c:1207-1208), with the plain limits beside it.
struct SymWindow {
    int lo[2], hi[2];
    uint32_t mv_min, mv_max;
};

__device__ __forceinline__ bool sym_legal(const SymWindow &w, int mx, int my)
{
    // CHECK_MV_RANGE (me.h:46-48) as the reference evaluates it; the plain
    // compare keeps every read inside the padded planes for any input
    const uint32_t v = xme_pack2(mx, my);
    const bool packed = !(((v + w.mv_min) | (w.mv_max - v)) & 0x80004000u);
    return packed && mx >= w.lo[0] && mx <= w.hi[0] && my >= w.lo[1] && my <= w.hi[1];
}

// A job the kernel evaluates: a PU of 4x4 tiles inside the frame at a column
// that is a multiple of 4, distances that keep scale_mv_skip inside int, and
// MVs (both the candidates within 4 qpel of the start and their mirrors)
// within +-SYM_MAX_MV, whose differences to the MVP stay inside the mvbits
// table.  The reference's caller gates the search the same way
// (check_mv_range / check_mvd, md_inter.c:1177-1184).
__device__ __forceinline__ bool sym_job_ok(const SymParams &p, const xme_sym_job &j)
{
    if ((j.w & 3) || (j.h & 3) || j.w < 4 || j.h < 4 || j.w > 64 || j.h > 64)
        return false;
    if ((j.x & 3) || j.x < 0 || j.y < 0 || j.x + j.w > p.frame_w || j.y + j.h > p.frame_h)
        return false;
    if (j.dist_fwd == 0 || j.dist_fwd > 255 || j.dist_fwd < -255 || j.dist_bwd > 255 || j.dist_bwd < -255)
        return false;
    const int mv[2] = { j.mv_x, j.mv_y }, mvp[2] = { j.mvp_x, j.mvp_y };
#pragma unroll
    for (int a = 0; a < 2; a++) {
#pragma unroll
        for (int e = -4; e <= 4; e += 8) {
            const int v = mv[a] + e;
            if (v > SYM_MAX_MV || v < -SYM_MAX_MV)
                return false;
            if ((int64_t)(v < 0 ? -v : v) * (j.dist_bwd < 0 ? -j.dist_bwd : j.dist_bwd) *
                    (16384 / (j.dist_fwd < 0 ? -j.dist_fwd : j.dist_fwd)) + 8192 > 0x7fffffffLL)
                return false;
            const int64_t s = scale_mv_wide(v, j.dist_bwd, j.dist_fwd);
            if (s > SYM_MAX_MV || s < -SYM_MAX_MV)
                return false;
            const int d = v - mvp[a];
            if (d > p.max_mvd || d < -p.max_mvd)
                return false;
        }
    }
    return true;
}

struct SymJob {
    gbytes cur;                // PU origin in the current picture
    int64_t i_cur, i_ref, off; // off: PU offset in the reference planes
    int w, h, tw, ntiles;
    int mvp_x, mvp_y, dd, ds;
    bool neg, hadamard;
    int lambda;
    SymWindow win;
};

// ME_COST_QPEL_SYM (me.c:353-382) of candidate (mx, my), summed over the
// lane's tiles t0, t0 + P, ...; returns -1 when the candidate has no cost
// (an MV outside the window or a plane that is not built).
__device__ __forceinline__ int sym_tiles(const SymJob &c, const SymPlanes &pl, int mx, int my, int t0, int P)
{
    int sx = scale_mv(mx, c.dd, c.ds), sy = scale_mv(my, c.dd, c.ds);
    if (c.neg) {
        sx = -sx;
        sy = -sy;
    }
    if (!sym_legal(c.win, mx, my) || !sym_legal(c.win, sx, sy))
        return -1;
    const uint8_t *a = pick_plane(pl.p1, ((my & 3) << 2) + (mx & 3));
    const uint8_t *b = pick_plane(pl.p2, ((sy & 3) << 2) + (sx & 3));
    if (a == nullptr || b == nullptr)
        return -1;
    const gbytes ga = (gbytes)a + c.off + (int64_t)(my >> 2) * c.i_ref + (mx >> 2);
    const gbytes gb = (gbytes)b + c.off + (int64_t)(sy >> 2) * c.i_ref + (sx >> 2);
    int s = 0;
    for (int t = t0; t < c.ntiles; t += P) {
        const int ty = (t / c.tw) * 4, tx = (t % c.tw) * 4;
        uint32_t cr[4], rr[4];
#pragma unroll
        for (int y = 0; y < 4; y++) {
            cr[y] = *(gdwords)(c.cur + (ty + y) * c.i_cur + tx);
            const int64_t o = (int64_t)(ty + y) * c.i_ref + tx;
            rr[y] = avg4(sym_load_dw(ga + o), sym_load_dw(gb + o));
        }
        s += xme_tile_cost_rows(cr, rr, c.hadamard);
    }
    return s;
}

// Candidate k of a stage (quarter-pel units).  Stage 0: the start point
// alone (frac 0), else Spiral[pos0 + k] * 2 around the start (me.c:1227-1238).
// Stage 1: Spiral[1 + k] around the half-pel best, with the PMVR adaptation
// (pmvr_adapt_mv, me.c:412-424) when enabled; *skip marks a point the
// reference does not evaluate.
__device__ __forceinline__ void sym_cand(int stage, int pos0, int cx, int cy, bool pmvr, int ctr_x, int ctr_y, int k,
                                         int &mx, int &my, bool &skip)
{
    skip = false;
    if (stage == 0) {
        mx = cx + 2 * c_bi_spiral[pos0 + k][0];
        my = cy + 2 * c_bi_spiral[pos0 + k][1];
        return;
    }
    const int dx = c_bi_spiral[1 + k][0], dy = c_bi_spiral[1 + k][1];
    if (pmvr) {
        const int ax = cx - ctr_x, ay = cy - ctr_y;
        if (ax > 2 || ax < -2 || ay > 2 || ay < -2) {
            mx = cx + 2 * dx;
            my = cy + 2 * dy;
            const int bx = mx - ctr_x, by = my - ctr_y;
            skip = bx <= 2 && bx >= -2 && by <= 2 && by >= -2;
        } else {
            mx = cx + dx;
            my = cy + dy;
            const int bx = mx - ctr_x, by = my - ctr_y;
            skip = bx > 2 || bx < -2 || by > 2 || by < -2;
        }
        return;
    }
    mx = cx + dx;
    my = cy + dy;
}

// ME_COST_QPEL_SYM over the lane's tiles, and the candidate's cost from the
// group's distortion sum.
struct SymEval {
    const SymJob &c;
    const SymPlanes &pl;
    __device__ __forceinline__ int tiles(int mx, int my, int t0, int P) const { return sym_tiles(c, pl, mx, my, t0, P); }
    __device__ __forceinline__ int cost(int s, int mx, int my) const
    {
        return (c.hadamard ? (s >> 1) : s) +
               xme_weighted_cost(c.lambda, xme_mvbits(mx - c.mvp_x) + xme_mvbits(my - c.mvp_y));
    }
};

// Evaluates the stage's n <= 9 candidates and applies the strict-'<' update
// to (bcost, bx, by) in candidate order.
template <class Eval>
__device__ __forceinline__ void bi_stage(const Eval &ev, int stage, int pos0, int n, int cx, int cy, bool pmvr,
                                         int ctr_x, int ctr_y, int P, int &bcost, int &bx, int &by)
{
    const int lane = threadIdx.x & 63;
    const int groups = 64 / P, g = lane / P, t0 = lane % P;
    for (int k0 = 0; k0 < n; k0 += groups) {
        const int k = k0 + g;
        int mx = cx, my = cy;
        bool skip = true;
        if (k < n)
            sym_cand(stage, pos0, cx, cy, pmvr, ctr_x, ctr_y, k, mx, my, skip);
        int s = skip ? -1 : ev.tiles(mx, my, t0, P);
        const bool none = s < 0;         // the same in every lane of the group
        s = group_sum(none ? 0 : s, P);
        const int cost = none ? XME_MAX_DISTORTION : ev.cost(s, mx, my);
        const int ng = min(groups, n - k0);
        for (int gi = 0; gi < ng; gi++) {        // uniform: the reference's candidate order
            const int src = gi * P + P - 1;
            const int cc = __builtin_amdgcn_readlane(cost, src);
            if (cc < bcost) {
                bcost = cc;
                bx = __builtin_amdgcn_readlane(mx, src);
                by = __builtin_amdgcn_readlane(my, src);
            }
        }
    }
}

// A refinement's outcome, the same in every lane of the wave.
struct SearchOut {
    int cost, mv_x, mv_y, mvcost, status;
};

// xavs2_me_search_sym for job j, by the whole wave.
__device__ SearchOut sym_search(const SymParams &p, const SymPlanes &pl, int64_t i_ref, const uint8_t *cur,
                                int64_t i_cur, const xme_sym_job &j)
{
    if (!sym_job_ok(p, j))
        return SearchOut{ XME_MAX_DISTORTION, j.mv_x, j.mv_y, -1, 2 };
    SymJob c;
    c.cur = (gbytes)cur + (int64_t)j.y * i_cur + j.x;
    c.i_cur = i_cur;
    c.i_ref = i_ref;
    c.off = (int64_t)j.y * i_ref + j.x;
    c.w = j.w;
    c.h = j.h;
    c.tw = j.w >> 2;
    c.ntiles = c.tw * (j.h >> 2);
    c.mvp_x = j.mvp_x;
    c.mvp_y = j.mvp_y;
    c.dd = j.dist_bwd;
    c.ds = j.dist_fwd;
    c.neg = j.b_slice != 0;
    c.hadamard = p.hadamard != 0;
    c.lambda = p.lambda;
    {
        // p_me->mv_min / mv_max (md_inter.c:913-924)
        const int lo_x = xme_clip3(p.min_mv[0], p.max_mv[0], (-64 - j.x) * 4);
        const int hi_x = xme_clip3(p.min_mv[0], p.max_mv[0], (p.frame_w + 64 - j.x - j.w) * 4);
        const int lo_y = xme_clip3(p.min_mv[1], p.max_mv[1], (-64 - j.y) * 4);
        const int hi_y = xme_clip3(p.min_mv[1], p.max_mv[1], (p.frame_h + 64 - j.y - j.h) * 4);
        // the plain limits also stay within 64 pixels of the picture, where
        // the XME_PAD border holds every read, whatever min/max_mv_range say
        c.win.lo[0] = max(lo_x, (-64 - j.x) * 4);
        c.win.hi[0] = min(hi_x, (p.frame_w + 64 - j.x - j.w) * 4);
        c.win.lo[1] = max(lo_y, (-64 - j.y) * 4);
        c.win.hi[1] = min(hi_y, (p.frame_h + 64 - j.y - j.h) * 4);
        c.win.mv_min = xme_pack2(-lo_x, -lo_y);
        c.win.mv_max = xme_pack2(hi_x, hi_y) | 0x8000u;
    }
    int P = 1;
    while (P < c.ntiles && P < 64)
        P <<= 1;
    int bcost = XME_MAX_DISTORTION, bx = j.mv_x, by = j.mv_y;
    if (p.frac == 0) {
        // the start point alone; its cost is returned even when there is none
        // (me.c:1219-1226)
        bi_stage(SymEval{ c, pl }, 0, 0, 1, j.mv_x, j.mv_y, false, 0, 0, P, bcost, bx, by);
        return SearchOut{ bcost, j.mv_x, j.mv_y, -1, 0 };
    }
    // half-pel: Spiral[0..4] with SATD, Spiral[1..4] with SAD (me.c:1210-1211)
    const int pos0 = c.hadamard ? 0 : 1;
    const SymEval ev{ c, pl };
    bi_stage(ev, 0, pos0, 5 - pos0, j.mv_x, j.mv_y, false, 0, 0, P, bcost, bx, by);
    if (p.frac >= 2) {
        const int ctr_x = (j.mvp1_x >> 1) << 1, ctr_y = (j.mvp1_y >> 1) << 1;
        bi_stage(ev, 1, 1, 4, bx, by, p.pmvr != 0, ctr_x, ctr_y, P, bcost, bx, by);
    }
    return SearchOut{ bcost, (int16_t)bx, (int16_t)by,
                      xme_weighted_cost(p.lambda, xme_mvbits(bx - j.mvp_x) + xme_mvbits(by - j.mvp_y)), 0 };
}

__global__ void __launch_bounds__(256) sym_kernel(SymParams p, SymPlanes pl, int64_t i_ref, const uint8_t *cur,
                                                  int64_t i_cur, const xme_sym_job *__restrict__ jobs, int n,
                                                  xme_sym_result *__restrict__ out)
{
    const int ji = job_index();
    if (ji >= n)
        return;
    const SearchOut o = sym_search(p, pl, i_ref, cur, i_cur, jobs[ji]);
    if ((threadIdx.x & 63) == 0) {
        xme_sym_result res;
        res.cost = o.cost;
        res.mv_x = (int16_t)o.mv_x;
        res.mv_y = (int16_t)o.mv_y;
        res.mvcost = o.mvcost;
        res.status = o.status;
        out[ji] = res;
    }
}

// A bi-directional job the kernel evaluates: a PU of 4x4 tiles inside the
// frame at a column that is a multiple of 4, a forward MV (and the
// candidates within 4 qpel of it) and a backward MV within +-SYM_MAX_MV whose
// differences to mvp1 / mvp2 stay inside the mvbits table (the caller's
// check_mv_range / check_mvd gate, md_inter.c:1089-1092).
__device__ __forceinline__ bool bid_job_ok(const SymParams &p, const xme_bid_job &j)
{
    if ((j.w & 3) || (j.h & 3) || j.w < 4 || j.h < 4 || j.w > 64 || j.h > 64)
        return false;
    if ((j.x & 3) || j.x < 0 || j.y < 0 || j.x + j.w > p.frame_w || j.y + j.h > p.frame_h)
        return false;
    const int fwd[2] = { j.fwd_x, j.fwd_y }, bwd[2] = { j.bwd_x, j.bwd_y };
    const int mvp1[2] = { j.mvp1_x, j.mvp1_y }, mvp2[2] = { j.mvp2_x, j.mvp2_y };
#pragma unroll
    for (int a = 0; a < 2; a++) {
        if (bwd[a] > SYM_MAX_MV || bwd[a] < -SYM_MAX_MV || bwd[a] - mvp2[a] > p.max_mvd ||
            bwd[a] - mvp2[a] < -p.max_mvd)
            return false;
#pragma unroll
        for (int e = -4; e <= 4; e += 8) {
            const int v = fwd[a] + e;
            if (v > SYM_MAX_MV || v < -SYM_MAX_MV || v - mvp1[a] > p.max_mvd || v - mvp1[a] < -p.max_mvd)
                return false;
        }
    }
    return true;
}

// clip(2 * o - p) on each of four bytes: sub_ps then add_ps
// (pixel.c:780-806), whose int16 difference never wraps for 8-bit pixels.
__device__ __forceinline__ uint32_t target4(uint32_t o, uint32_t p)
{
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int v = 2 * xme_bytei(o, i) - xme_bytei(p, i);
        r |= (uint32_t)min(max(v, 0), 255) << (8 * i);
    }
    return r;
}

// ME_COST_QPEL_BID (me.c:388-398) over the lane's tiles t0, t0 + P, ...
// (at most 4: a PU has at most 256 tiles and P = 64 there) against the
// target clip(2 * cur - pred_bwd).  The target rows of the lane's first tile
// are built once per job (`tgt`); a PU of more than 64 tiles builds those of
// its other tiles per candidate, so that one small register file serves
// every PU size.
struct BidEval {
    const SymJob &c;             // cur, strides, offset, tiles, window, lambda, switches
    const uint8_t *const (&p1)[16];
    const uint32_t (&tgt)[4];
    gbytes gb;                   // the backward prediction at the PU origin
    bool bwd_legal;
    int mvp1_x, mvp1_y, bid_bits;
    __device__ __forceinline__ int tiles(int mx, int my, int t0, int P) const
    {
        if (!bwd_legal || !sym_legal(c.win, mx, my))
            return -1;
        const uint8_t *a = pick_plane(p1, ((my & 3) << 2) + (mx & 3));
        if (a == nullptr)           // the reference reads through a NULL plane here
            return -1;
        const gbytes ga = (gbytes)a + c.off + (int64_t)(my >> 2) * c.i_ref + (mx >> 2);
        int s = 0;
        if (t0 < c.ntiles) {
            const int ty = (t0 / c.tw) * 4, tx = (t0 % c.tw) * 4;
            uint32_t rr[4];
#pragma unroll
            for (int y = 0; y < 4; y++)
                rr[y] = sym_load_dw(ga + (int64_t)(ty + y) * c.i_ref + tx);
            s = xme_tile_cost_rows(tgt, rr, c.hadamard);
        }
        for (int t = t0 + P; t < c.ntiles; t += P) {
            const int ty = (t / c.tw) * 4, tx = (t % c.tw) * 4;
            uint32_t rr[4], tg[4];
#pragma unroll
            for (int y = 0; y < 4; y++) {
                rr[y] = sym_load_dw(ga + (int64_t)(ty + y) * c.i_ref + tx);
                tg[y] = target4(*(gdwords)(c.cur + (ty + y) * c.i_cur + tx),
                                sym_load_dw(gb + (int64_t)(ty + y) * c.i_ref + tx));
            }
            s += xme_tile_cost_rows(tg, rr, c.hadamard);
        }
        return s;
    }
    __device__ __forceinline__ int cost(int s, int mx, int my) const
    {
        // fpel_cmp (SATD = half the Hadamard sum, or SAD) halved once more
        const int d = c.hadamard ? ((s >> 1) >> 1) : (s >> 1);
        return d + xme_weighted_cost(c.lambda, xme_mvbits(mx - mvp1_x) + xme_mvbits(my - mvp1_y)) + bid_bits;
    }
};

// xavs2_me_search_bid for job j, by the whole wave.
__device__ SearchOut bid_search(const SymParams &p, const SymPlanes &pl, int64_t i_ref, const uint8_t *cur,
                                int64_t i_cur, const xme_bid_job &j)
{
    const int lane = threadIdx.x & 63;
    if (!bid_job_ok(p, j))
        return SearchOut{ XME_MAX_DISTORTION, j.fwd_x, j.fwd_y, -1, 2 };
    SymJob c;
    c.cur = (gbytes)cur + (int64_t)j.y * i_cur + j.x;
    c.i_cur = i_cur;
    c.i_ref = i_ref;
    c.off = (int64_t)j.y * i_ref + j.x;
    c.w = j.w;
    c.h = j.h;
    c.tw = j.w >> 2;
    c.ntiles = c.tw * (j.h >> 2);
    c.mvp_x = j.mvp1_x;
    c.mvp_y = j.mvp1_y;
    c.dd = 1;
    c.ds = 1;
    c.neg = false;
    c.hadamard = p.hadamard != 0;
    c.lambda = p.lambda;
    {
        // p_me->mv_min / mv_max (md_inter.c:913-924), packed as in
        // xavs2_me_search_bid (me.c:1300-1301)
        const int lo_x = xme_clip3(p.min_mv[0], p.max_mv[0], (-64 - j.x) * 4);
        const int hi_x = xme_clip3(p.min_mv[0], p.max_mv[0], (p.frame_w + 64 - j.x - j.w) * 4);
        const int lo_y = xme_clip3(p.min_mv[1], p.max_mv[1], (-64 - j.y) * 4);
        const int hi_y = xme_clip3(p.min_mv[1], p.max_mv[1], (p.frame_h + 64 - j.y - j.h) * 4);
        c.win.lo[0] = max(lo_x, (-64 - j.x) * 4);
        c.win.hi[0] = min(hi_x, (p.frame_w + 64 - j.x - j.w) * 4);
        c.win.lo[1] = max(lo_y, (-64 - j.y) * 4);
        c.win.hi[1] = min(hi_y, (p.frame_h + 64 - j.y - j.h) * 4);
        c.win.mv_min = xme_pack2(-lo_x, -lo_y);
        c.win.mv_max = xme_pack2(hi_x, hi_y) | 0x8000u;
    }
    int P = 1;
    while (P < c.ntiles && P < 64)
        P <<= 1;
    const int t0 = lane % P;
    const bool bwd_legal = sym_legal(c.win, j.bwd_x, j.bwd_y);
    const uint8_t *p2 = pick_plane(pl.p2, ((j.bwd_y & 3) << 2) + (j.bwd_x & 3));
    if (bwd_legal && p2 == nullptr) {
        // the reference predicts from mc_luma here (me.c:1329-1335): not provided
        return SearchOut{ XME_MAX_DISTORTION, j.fwd_x, j.fwd_y, -1, 2 };
    }
    // the target rows of the lane's first tile: clip(2 * cur - pred_bwd)
    // (me.c:1322-1337)
    const gbytes gb = bwd_legal ? (gbytes)p2 + c.off + (int64_t)(j.bwd_y >> 2) * i_ref + (j.bwd_x >> 2) : nullptr;
    uint32_t tgt[4] = { 0, 0, 0, 0 };
    if (bwd_legal && t0 < c.ntiles) {
        const int ty = (t0 / c.tw) * 4, tx = (t0 % c.tw) * 4;
#pragma unroll
        for (int y = 0; y < 4; y++)
            tgt[y] = target4(*(gdwords)(c.cur + (ty + y) * i_cur + tx), sym_load_dw(gb + (int64_t)(ty + y) * i_ref + tx));
    }
    const int bid_bits = xme_weighted_cost(p.lambda, xme_mvbits(j.bwd_x - j.mvp2_x) + xme_mvbits(j.bwd_y - j.mvp2_y));
    const BidEval ev{ c, pl.p1, tgt, gb, bwd_legal, j.mvp1_x, j.mvp1_y, bid_bits };
    int bcost = XME_MAX_DISTORTION, bx = j.fwd_x, by = j.fwd_y;
    if (p.frac == 0) {
        bi_stage(ev, 0, 0, 1, j.fwd_x, j.fwd_y, false, 0, 0, P, bcost, bx, by);
        return SearchOut{ bcost, j.fwd_x, j.fwd_y, -1, 0 };
    }
    // half-pel: Spiral[0..8] with SATD, Spiral[1..8] with SAD (me.c:1340-1349)
    const int pos0 = c.hadamard ? 0 : 1;
    bi_stage(ev, 0, pos0, 9 - pos0, j.fwd_x, j.fwd_y, false, 0, 0, P, bcost, bx, by);
    if (p.frac >= 2) {
        const int ctr_x = (j.mvp1_x >> 1) << 1, ctr_y = (j.mvp1_y >> 1) << 1;
        bi_stage(ev, 1, 1, 8, bx, by, p.pmvr != 0, ctr_x, ctr_y, P, bcost, bx, by);
    }
    return SearchOut{ bcost, (int16_t)bx, (int16_t)by,
                      xme_weighted_cost(p.lambda, xme_mvbits(bx - j.mvp1_x) + xme_mvbits(by - j.mvp1_y)) + bid_bits,
                      0 };
}

__global__ void __launch_bounds__(256) bid_kernel(SymParams p, SymPlanes pl, int64_t i_ref, const uint8_t *cur,
                                                  int64_t i_cur, const xme_bid_job *__restrict__ jobs, int n,
                                                  xme_sym_result *__restrict__ out)
{
    const int ji = job_index();
    if (ji >= n)
        return;
    const SearchOut o = bid_search(p, pl, i_ref, cur, i_cur, jobs[ji]);
    if ((threadIdx.x & 63) == 0) {
        xme_sym_result res;
        res.cost = o.cost;
        res.mv_x = (int16_t)o.mv_x;
        res.mv_y = (int16_t)o.mv_y;
        res.mvcost = o.mvcost;
        res.status = o.status;
        out[ji] = res;
    }
}

// ---------------------------------------------------------------------------
// The callers, pred_inter_search_bi (md_inter.c:1036-1128) and
// pred_inter_search_dual (md_inter.c:1133-1219): one wave per PU runs the
// gates on uniform registers and the refinements they let through.

struct GateParams {
    int lcu_level, frame_threads;
};

// check_mvd (md_inter.c:68-76).
__device__ __forceinline__ bool gate_mvd(const GateParams &g, int dx, int dy)
{
    const int vy = g.frame_threads > 1 ? ((1 << g.lcu_level) << 2) : 4096;
    return dx < 4096 && dx >= -4096 && dy < vy && dy >= -vy;
}

// The qpel box of check_mv_range and check_mv_range_sym (predict.h:70-79,
// md_inter.c:85-95): the PU may reach 1 << lcu_level pixels past the
// picture, inside min/max_mv_range.
struct GateBox {
    int min_x, min_y, max_x, max_y;
    __device__ bool holds(int x, int y) const { return x <= max_x && x >= min_x && y <= max_y && y >= min_y; }
};

__device__ __forceinline__ GateBox gate_box(const SymParams &p, const GateParams &g, int px, int py, int w, int h)
{
    const int bsize = 1 << g.lcu_level;
    GateBox b;
    b.min_x = max(-((px + bsize) << 2), p.min_mv[0]);
    b.min_y = max(-((py + bsize) << 2), p.min_mv[1]);
    b.max_x = min(((p.frame_w - (px + w)) + bsize) << 2, p.max_mv[0]);
    b.max_y = min(((p.frame_h - (py + h)) + bsize) << 2, p.max_mv[1]);
    return b;
}

// scale_mv_skip (predict.h:148-151); *undef is set when the reference's int
// product overflows (undefined there).
__device__ __forceinline__ int gate_scale(int mv, int dist_dst, int dist_src, bool &undef)
{
    const int64_t s = (int64_t)mv * dist_dst * (16384 / dist_src) + 8192;
    if (s > 0x7fffffffLL || s < -0x80000000LL)
        undef = true;
    return (int16_t)((int)s >> 14);
}

// check_mv_range_sym (md_inter.c:83-110), frame sequences.
__device__ __forceinline__ bool gate_range_sym(const GateBox &b, int mx, int my, int dist_fwd, int dist_bwd, bool b_slice,
                                               bool &undef)
{
    int bx = gate_scale(mx, dist_bwd, dist_fwd, undef), by = gate_scale(my, dist_bwd, dist_fwd, undef);
    if (b_slice) {
        bx = -bx;
        by = -by;
    }
    return b.holds(mx, my) && b.holds(bx, by);
}

// A PU the refinements take (their job checks repeat this).
__device__ __forceinline__ bool caller_pu_ok(const SymParams &p, int x, int y, int w, int h)
{
    return !((w & 3) || (h & 3) || w < 4 || h < 4 || w > 64 || h > 64 || (x & 3) || x < 0 || y < 0 ||
             x + w > p.frame_w || y + h > p.frame_h);
}

// pred_inter_search_bi in two launches, so that each has the register file
// of its own search.  The first (md_inter.c:1083-1090) runs the symmetric
// search's gates and search and stores the PU's result as if the
// bi-directional search had not run; the second runs that search's gates
// and search and the stored MVs' checks, and completes the result.
__global__ void __launch_bounds__(256) bi_sym_kernel(SymParams p, GateParams g, SymPlanes pl, int64_t i_ref,
                                                     const uint8_t *cur, int64_t i_cur,
                                                     const xme_bi_job *__restrict__ jobs, int n, int dist_fwd,
                                                     int dist_bwd, xme_bi_result *__restrict__ out)
{
    const int ji = job_index();
    if (ji >= n)
        return;
    const xme_bi_job j = jobs[ji];
    bool undef = !caller_pu_ok(p, j.x, j.y, j.w, j.h);
    const GateBox box = gate_box(p, g, j.x, j.y, j.w, j.h);
    // fwd_mv.x <<= 2 on the int16 components of mv_t (md_inter.c:1068-1069)
    int mv_x = (int16_t)(j.fwd_imv_x * 4), mv_y = (int16_t)(j.fwd_imv_y * 4);
    const int fwd_x = mv_x, fwd_y = mv_y;
    int cost = XME_MAX_DISTORTION, sym_mvcost = -1;
    bool ok = gate_range_sym(box, j.mvp1_x, j.mvp1_y, dist_fwd, dist_bwd, true, undef);
    ok &= gate_range_sym(box, mv_x, mv_y, dist_fwd, dist_bwd, true, undef);
    ok &= gate_mvd(g, j.mvp1_x, j.mvp1_y);
    if (ok && !undef) {
        xme_sym_job sj;
        sj.x = j.x;
        sj.y = j.y;
        sj.w = j.w;
        sj.h = j.h;
        sj.mv_x = (int16_t)mv_x;
        sj.mv_y = (int16_t)mv_y;
        sj.mvp_x = j.mvp_x;
        sj.mvp_y = j.mvp_y;
        sj.mvp1_x = j.mvp1_x;
        sj.mvp1_y = j.mvp1_y;
        sj.dist_fwd = (int16_t)dist_fwd;
        sj.dist_bwd = (int16_t)dist_bwd;
        sj.b_slice = 1;
        const SearchOut o = sym_search(p, pl, i_ref, cur, i_cur, sj);
        undef |= o.status != 0;
        cost = o.cost;
        mv_x = o.mv_x;
        mv_y = o.mv_y;
        sym_mvcost = o.mvcost;
    }
    if ((threadIdx.x & 63) == 0) {
        xme_bi_result r;
        r.sym_cost = cost;
        r.sym_mv_x = (int16_t)mv_x;
        r.sym_mv_y = (int16_t)mv_y;
        r.sym_mvcost = sym_mvcost;
        r.bid_cost = XME_MAX_DISTORTION;
        r.fwd_x = (int16_t)fwd_x;
        r.fwd_y = (int16_t)fwd_y;
        r.bwd_x = j.bwd_x;
        r.bwd_y = j.bwd_y;
        r.bid_mvcost = -1;
        r.status = undef ? 2 : 0;
        out[ji] = r;
    }
}

__global__ void __launch_bounds__(256) bi_bid_kernel(SymParams p, GateParams g, SymPlanes pl, int64_t i_ref,
                                                     const uint8_t *cur, int64_t i_cur,
                                                     const xme_bi_job *__restrict__ jobs, int n, int dist_fwd,
                                                     int dist_bwd, xme_bi_result *__restrict__ out)
{
    const int ji = job_index();
    if (ji >= n)
        return;
    const xme_bi_job j = jobs[ji];
    const xme_bi_result first = out[ji];
    bool undef = first.status != 0;
    const GateBox box = gate_box(p, g, j.x, j.y, j.w, j.h);
    const int mv_x = first.sym_mv_x, mv_y = first.sym_mv_y;
    int cost = first.sym_cost;
    int fwd_x = first.fwd_x, fwd_y = first.fwd_y;
    int cost_bid = XME_MAX_DISTORTION, bid_mvcost = -1;
    // md_inter.c:1092-1102
    const bool ok = box.holds(fwd_x, fwd_y) && box.holds(j.bwd_x, j.bwd_y) && gate_mvd(g, j.mvp1_x, j.mvp1_y) &&
                    gate_mvd(g, j.mvp2_x, j.mvp2_y);
    if (ok && !undef) {
        xme_bid_job bj;
        bj.x = j.x;
        bj.y = j.y;
        bj.w = j.w;
        bj.h = j.h;
        bj.fwd_x = (int16_t)fwd_x;
        bj.fwd_y = (int16_t)fwd_y;
        bj.bwd_x = j.bwd_x;
        bj.bwd_y = j.bwd_y;
        bj.mvp1_x = j.mvp1_x;
        bj.mvp1_y = j.mvp1_y;
        bj.mvp2_x = j.mvp2_x;
        bj.mvp2_y = j.mvp2_y;
        const SearchOut o = bid_search(p, pl, i_ref, cur, i_cur, bj);
        undef |= o.status != 0;
        cost_bid = o.cost;
        fwd_x = o.mv_x;
        fwd_y = o.mv_y;
        bid_mvcost = o.mvcost;
    }
    // the stored MVs' own checks (md_inter.c:1107-1122)
    if (!(box.holds(fwd_x, fwd_y) && gate_mvd(g, fwd_x - j.mvp1_x, fwd_y - j.mvp1_y)))
        cost_bid = XME_MAX_DISTORTION;
    if (!(box.holds(j.bwd_x, j.bwd_y) && gate_mvd(g, j.bwd_x - j.mvp2_x, j.bwd_y - j.mvp2_y)))
        cost_bid = XME_MAX_DISTORTION;
    if (!(gate_range_sym(box, mv_x, mv_y, dist_fwd, dist_bwd, true, undef) &&
          gate_mvd(g, mv_x - j.mvp1_x, mv_y - j.mvp1_y)))
        cost = XME_MAX_DISTORTION;
    if ((threadIdx.x & 63) == 0) {
        xme_bi_result r = first;
        r.sym_cost = cost;
        r.bid_cost = cost_bid;
        r.fwd_x = (int16_t)fwd_x;
        r.fwd_y = (int16_t)fwd_y;
        r.bid_mvcost = bid_mvcost;
        r.status = undef ? 2 : 0;
        out[ji] = r;
    }
}

// One reference of pred_inter_search_dual's loop (md_inter.c:1160-1218) for
// every PU: launch r searches reference r with the second reference !r and
// updates the running minimum that launch r - 1 left in out[] (launch 0
// starts it), so the launches on one stream keep the reference's order.
__global__ void __launch_bounds__(256) dual_kernel(SymParams p, GateParams g, SymPlanes pl, int64_t i_ref,
                                                   const uint8_t *cur, int64_t i_cur,
                                                   const xme_dual_job *__restrict__ jobs, int n, int r, int d_fwd,
                                                   int d_bwd, xme_dual_result *__restrict__ out)
{
    const int ji = job_index();
    if (ji >= n)
        return;
    const xme_dual_job &jr = jobs[ji];
    const int x = jr.x, y = jr.y, w = jr.w, h = jr.h;
    int best = XME_MAX_DISTORTION, best_fst = -1, best_snd = -1, best_mvcost = -1;
    bool undef;
    if (r == 0) {
        undef = !caller_pu_ok(p, x, y, w, h);
    } else {
        best = out[ji].cost;
        best_fst = out[ji].fst_ref;
        best_snd = out[ji].snd_ref;
        best_mvcost = out[ji].mvcost;
        undef = out[ji].status != 0;
    }
    const GateBox box = gate_box(p, g, x, y, w, h);
    const int s = r == 0 ? 1 : 0;
    const int m1x = jr.mvp1[r][0], m1y = jr.mvp1[r][1], mvp_x = jr.mvp_x, mvp_y = jr.mvp_y;
    // fst_dual.x <<= 2 on the int16 components of mv_t (md_inter.c:1177-1178)
    int fx = (int16_t)(jr.imv[r][0] * 4), fy = (int16_t)(jr.imv[r][1] * 4);
    int cost = XME_MAX_DISTORTION, mvcost = -1;
    bool ok = gate_range_sym(box, fx, fy, d_fwd, d_bwd, false, undef);
    ok &= gate_mvd(g, fx - m1x, fy - m1y);
    ok &= gate_mvd(g, m1x, m1y);
    ok &= gate_mvd(g, mvp_x, mvp_y);
    if (ok && !undef) {
        xme_sym_job sj;
        sj.x = (int16_t)x;
        sj.y = (int16_t)y;
        sj.w = (uint8_t)w;
        sj.h = (uint8_t)h;
        sj.mv_x = (int16_t)fx;
        sj.mv_y = (int16_t)fy;
        sj.mvp_x = (int16_t)mvp_x;
        sj.mvp_y = (int16_t)mvp_y;
        sj.mvp1_x = (int16_t)m1x;
        sj.mvp1_y = (int16_t)m1y;
        sj.dist_fwd = (int16_t)d_fwd;
        sj.dist_bwd = (int16_t)d_bwd;
        sj.b_slice = 0;
        const SearchOut o = sym_search(p, pl, i_ref, cur, i_cur, sj);
        undef |= o.status != 0;
        fx = o.mv_x;
        fy = o.mv_y;
        cost = o.cost;
        mvcost = o.mvcost;
    }
    // snd_dual (md_inter.c:1192-1193)
    const int sx = gate_scale(fx, d_bwd, d_fwd, undef), sy = gate_scale(fy, d_bwd, d_fwd, undef);
    ok &= gate_range_sym(box, fx, fy, d_fwd, d_bwd, false, undef);
    ok &= gate_mvd(g, fx - m1x, fy - m1y);
    if (ok) {
        // REF_COST(ref_idx) = WEIGHTED_COST(lambda, REF_BITS[ref_idx + 1]),
        // REF_BITS {1, 1, 3, 3, 5, ...} (predict.h:47-52): 1, 3, 3, 5 bits
        // for references 0..3
        cost += xme_weighted_cost(p.lambda, r == 0 ? 1 : (r == 3 ? 5 : 3));
        if (cost < best) {
            best = cost;
            best_fst = r;
            best_snd = s;
            best_mvcost = mvcost;
        }
    }
    if ((threadIdx.x & 63) == 0) {
        xme_dual_result &o = out[ji];
        if (r == 0) {
            for (int q = 1; q < 4; q++) {
                o.fst[q][0] = o.fst[q][1] = 0;
                o.snd[q][0] = o.snd[q][1] = 0;
            }
            o.reserved = 0;
        }
        o.cost = best;
        o.fst_ref = (int8_t)best_fst;
        o.snd_ref = (int8_t)best_snd;
        o.mvcost = best_mvcost;
        o.fst[r][0] = (int16_t)fx;
        o.fst[r][1] = (int16_t)fy;
        o.snd[r][0] = (int16_t)sx;
        o.snd[r][1] = (int16_t)sy;
        o.status = undef ? 2 : 0;
    }
}

} // namespace

extern "C" int xme_me_search_sym_batch(const xme_me_params *prm, const xme_sym_job *d_jobs, int n,
                                       const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref1,
                                       const xme_ref_desc *ref2, xme_sym_result *d_results, void *stream)
{
    if (!prm || n < 0 || !ref1 || !ref2 || prm->use_fractional_me < 0 || prm->use_fractional_me > 2 ||
        prm->frame_w <= 0 || prm->frame_h <= 0 || prm->search_range <= 0) {
        xme_set_error("xme_me_search_sym_batch: bad arguments");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    if (!d_jobs || !d_cur || !d_results || !ref1->planes[0] || !ref2->planes[0]) {
        xme_set_error("xme_me_search_sym_batch: NULL buffer");
        return XME_ERR_ARG;
    }
    // the reference addresses both references with the first one's stride
    // (me.c:1218); the current rows are read as aligned dwords
    if (ref1->stride != ref2->stride || ((uintptr_t)d_cur & 3) || (i_cur & 3)) {
        xme_set_error("xme_me_search_sym_batch: reference strides differ or the current plane is not 4-byte aligned");
        return XME_ERR_ARG;
    }
    SymParams p;
    p.frame_w = prm->frame_w;
    p.frame_h = prm->frame_h;
    p.min_mv[0] = prm->min_mv_range[0];
    p.min_mv[1] = prm->min_mv_range[1];
    p.max_mv[0] = prm->max_mv_range[0];
    p.max_mv[1] = prm->max_mv_range[1];
    p.lambda = prm->lambda_factor;
    p.hadamard = prm->use_hadamard ? 1 : 0;
    p.frac = prm->use_fractional_me;
    p.pmvr = prm->enable_pmvr ? 1 : 0;
    p.max_mvd = xme_mvbits_max_mvd(prm->search_range);
    SymPlanes pl;
    for (int k = 0; k < 16; k++) {
        pl.p1[k] = ref1->planes[k];
        pl.p2[k] = ref2->planes[k];
    }
    const int grid = job_grid(n);
    hipLaunchKernelGGL(sym_kernel, dim3(grid), dim3(64 * SYM_WAVES), 0, (hipStream_t)stream, p, pl,
                       (int64_t)ref1->stride, d_cur, (int64_t)i_cur, d_jobs, n, d_results);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_me_search_bid_batch(const xme_me_params *prm, const xme_bid_job *d_jobs, int n,
                                       const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref1,
                                       const xme_ref_desc *ref2, xme_sym_result *d_results, void *stream)
{
    if (!prm || n < 0 || !ref1 || !ref2 || prm->use_fractional_me < 0 || prm->use_fractional_me > 2 ||
        prm->frame_w <= 0 || prm->frame_h <= 0 || prm->search_range <= 0) {
        xme_set_error("xme_me_search_bid_batch: bad arguments");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    if (!d_jobs || !d_cur || !d_results || !ref1->planes[0] || !ref2->planes[0]) {
        xme_set_error("xme_me_search_bid_batch: NULL buffer");
        return XME_ERR_ARG;
    }
    // both references are addressed with the first one's stride (me.c:1315)
    if (ref1->stride != ref2->stride || ((uintptr_t)d_cur & 3) || (i_cur & 3)) {
        xme_set_error("xme_me_search_bid_batch: reference strides differ or the current plane is not 4-byte aligned");
        return XME_ERR_ARG;
    }
    SymParams p;
    p.frame_w = prm->frame_w;
    p.frame_h = prm->frame_h;
    p.min_mv[0] = prm->min_mv_range[0];
    p.min_mv[1] = prm->min_mv_range[1];
    p.max_mv[0] = prm->max_mv_range[0];
    p.max_mv[1] = prm->max_mv_range[1];
    p.lambda = prm->lambda_factor;
    p.hadamard = prm->use_hadamard ? 1 : 0;
    p.frac = prm->use_fractional_me;
    p.pmvr = prm->enable_pmvr ? 1 : 0;
    p.max_mvd = xme_mvbits_max_mvd(prm->search_range);
    SymPlanes pl;
    for (int k = 0; k < 16; k++) {
        pl.p1[k] = ref1->planes[k];
        pl.p2[k] = ref2->planes[k];
    }
    const int grid = job_grid(n);
    hipLaunchKernelGGL(bid_kernel, dim3(grid), dim3(64 * SYM_WAVES), 0, (hipStream_t)stream, p, pl,
                       (int64_t)ref1->stride, d_cur, (int64_t)i_cur, d_jobs, n, d_results);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

namespace {

// SymParams from the ME parameters; false (with the error set) when they are
// out of range.
bool bi_params(const xme_me_params *prm, const char *who, SymParams &p)
{
    if (!prm || prm->use_fractional_me < 0 || prm->use_fractional_me > 2 || prm->frame_w <= 0 ||
        prm->frame_h <= 0 || prm->search_range <= 0) {
        xme_set_error("%s: bad arguments", who);
        return false;
    }
    p.frame_w = prm->frame_w;
    p.frame_h = prm->frame_h;
    p.min_mv[0] = prm->min_mv_range[0];
    p.min_mv[1] = prm->min_mv_range[1];
    p.max_mv[0] = prm->max_mv_range[0];
    p.max_mv[1] = prm->max_mv_range[1];
    p.lambda = prm->lambda_factor;
    p.hadamard = prm->use_hadamard ? 1 : 0;
    p.frac = prm->use_fractional_me;
    p.pmvr = prm->enable_pmvr ? 1 : 0;
    p.max_mvd = xme_mvbits_max_mvd(prm->search_range);
    return true;
}

// The gate parameters and a distance pair of calculate_distance.
bool gate_params(int lcu_level, int frame_threads, const char *who, GateParams &g)
{
    if (lcu_level < 3 || lcu_level > 6 || frame_threads < 1) {
        xme_set_error("%s: lcu_level must be 3..6 and frame_threads >= 1", who);
        return false;
    }
    g.lcu_level = lcu_level;
    g.frame_threads = frame_threads;
    return true;
}

bool dist_ok(int d) { return d != 0 && d <= 255 && d >= -255; }

} // namespace

extern "C" int xme_pred_inter_search_bi_batch(const xme_me_params *prm, int lcu_level, int frame_threads,
                                              const xme_bi_job *d_jobs, int n, int dist_fwd, int dist_bwd,
                                              const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref_fwd,
                                              const xme_ref_desc *ref_bwd, xme_bi_result *d_results, void *stream)
{
    static const char *who = "xme_pred_inter_search_bi_batch";
    SymParams p;
    GateParams g;
    if (!bi_params(prm, who, p) || !gate_params(lcu_level, frame_threads, who, g))
        return XME_ERR_ARG;
    if (n < 0 || !ref_fwd || !ref_bwd || !dist_ok(dist_fwd) || !dist_ok(dist_bwd)) {
        xme_set_error("%s: bad arguments (n >= 0, both references, distances nonzero within +-255)", who);
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    if (!d_jobs || !d_cur || !d_results || !ref_fwd->planes[0] || !ref_bwd->planes[0]) {
        xme_set_error("%s: NULL buffer", who);
        return XME_ERR_ARG;
    }
    if (ref_fwd->stride != ref_bwd->stride || ((uintptr_t)d_cur & 3) || (i_cur & 3)) {
        xme_set_error("%s: reference strides differ or the current plane is not 4-byte aligned", who);
        return XME_ERR_ARG;
    }
    SymPlanes pl;
    for (int k = 0; k < 16; k++) {
        pl.p1[k] = ref_fwd->planes[k];
        pl.p2[k] = ref_bwd->planes[k];
    }
    const int grid = job_grid(n);
    hipStream_t s = (hipStream_t)stream;
    const int64_t i_ref = (int64_t)ref_fwd->stride;
    hipLaunchKernelGGL(bi_sym_kernel, dim3(grid), dim3(64 * SYM_WAVES), 0, s, p, g, pl, i_ref, d_cur,
                       (int64_t)i_cur, d_jobs, n, dist_fwd, dist_bwd, d_results);
    XME_HIP(hipGetLastError());
    hipLaunchKernelGGL(bi_bid_kernel, dim3(grid), dim3(64 * SYM_WAVES), 0, s, p, g, pl, i_ref, d_cur,
                       (int64_t)i_cur, d_jobs, n, dist_fwd, dist_bwd, d_results);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_pred_inter_search_dual_batch(const xme_me_params *prm, int lcu_level, int frame_threads,
                                                const xme_dual_job *d_jobs, int n, int n_ref, const int dist[4],
                                                const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *refs,
                                                xme_dual_result *d_results, void *stream)
{
    static const char *who = "xme_pred_inter_search_dual_batch";
    SymParams p;
    GateParams g;
    if (!bi_params(prm, who, p) || !gate_params(lcu_level, frame_threads, who, g))
        return XME_ERR_ARG;
    if (n < 0 || n_ref < 2 || n_ref > 4 || !refs || !dist) {
        xme_set_error("%s: bad arguments (n >= 0, 2..4 references and their distances)", who);
        return XME_ERR_ARG;
    }
    for (int r = 0; r < n_ref; r++) {
        if (!dist_ok(dist[r])) {
            xme_set_error("%s: distance %d of reference %d is 0 or beyond +-255", who, dist[r], r);
            return XME_ERR_ARG;
        }
    }
    if (n == 0)
        return XME_OK;
    if (!d_jobs || !d_cur || !d_results || ((uintptr_t)d_cur & 3) || (i_cur & 3)) {
        xme_set_error("%s: NULL buffer or the current plane is not 4-byte aligned", who);
        return XME_ERR_ARG;
    }
    for (int r = 0; r < n_ref; r++) {
        if (!refs[r].planes[0] || refs[r].stride != refs[0].stride) {
            xme_set_error("%s: reference %d has no full-pel plane or another stride", who, r);
            return XME_ERR_ARG;
        }
    }
    // one launch per reference, in the reference's order (each updates the
    // minimum the previous one stored)
    const int grid = job_grid(n);
    for (int r = 0; r < n_ref; r++) {
        const int s = r == 0 ? 1 : 0;
        SymPlanes pl;
        for (int k = 0; k < 16; k++) {
            pl.p1[k] = refs[r].planes[k];
            pl.p2[k] = refs[s].planes[k];
        }
        hipLaunchKernelGGL(dual_kernel, dim3(grid), dim3(64 * SYM_WAVES), 0, (hipStream_t)stream, p, g, pl,
                           (int64_t)refs[0].stride, d_cur, (int64_t)i_cur, d_jobs, n, r, dist[r], dist[s],
                           d_results);
        XME_HIP(hipGetLastError());
    }
    return XME_OK;
}
