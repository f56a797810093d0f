// Sub-pel refinement (me_subpel_refine, source/encoder/me.c:479-579) and the
// store of the full-pel result (me.c:1169-1173) for gfx950.
//
// One wave per (PU, ref).  The candidates of a refinement stage (the start
// point and 8 half-pel points around the integer best, then 8 quarter-pel
// points around the half-pel best) are independent, so a stage evaluates all
// of them at once: lanes take (candidate, 4x4 tile) pairs, compute the tile's
// Hadamard sum (or SAD) from the resident interpolated planes and add it to
// the candidate's LDS counter; the stage then applies the reference's
// sequential strict-'<' update in Spiral order.
#include "xme_dpp.h"
#include "xme_satd.h"
#include "xme_internal.h"

namespace {

__constant__ int8_t c_spiral[9][2] = { { 0, 0 }, { 0, -1 }, { 0, 1 }, { -1, -1 }, { 1, -1 },
                                       { -1, 0 }, { 1, 0 }, { -1, 1 }, { 1, 1 } };
__constant__ int8_t c_spiral2[9][2] = { { 0, 0 }, { 0, -1 }, { -1, -1 }, { -1, 0 }, { -1, 1 },
                                        { 0, 1 }, { 1, 1 }, { 1, 0 }, { 1, -1 } };

// Picture and plane bytes in device memory.  The pointers come from the
// frame descriptor as generic pointers; typing them global gives global (not
// flat) loads.
typedef const __attribute__((address_space(1))) uint8_t *gbytes;
typedef const __attribute__((address_space(1))) uint32_t *gdwords;

// Four bytes at any address.  The dword after the aligned one is read even
// when the address is aligned; reference rows have XME_PAD bytes of padding
// on the right, so it stays inside the plane.
__device__ __forceinline__ uint32_t load_unaligned_dw(gbytes p)
{
    const uintptr_t a = (uintptr_t)p;
    gdwords q = (gdwords)(a & ~(uintptr_t)3);
    return __builtin_amdgcn_alignbyte(q[1], q[0], (uint32_t)(a & 3));
}

// Sum of |H4 D H4^T| over one 4x4 tile (before the final halving), or the SAD,
// of current rows c[0..3] against the reference tile at `ref`.
__device__ __forceinline__ int tile_cost(const uint32_t (&cr)[4], gbytes ref, int64_t i_ref, bool hadamard)
{
    uint32_t rr[4];
#pragma unroll
    for (int y = 0; y < 4; y++)
        rr[y] = load_unaligned_dw(ref + y * i_ref);
    return xme_tile_cost_rows(cr, rr, hadamard);
}

struct RefineCtx {
    gbytes cur;                // PU origin in the current picture
    int64_t i_cur;
    const uint8_t *const *planes;
    int64_t i_ref;
    int64_t off;               // PU offset in the reference planes (p_me->i_bias)
    int w, h;
    bool hadamard;
    int lambda, pmx, pmy;
};

__device__ __forceinline__ gbytes cand_ref(const RefineCtx &c, int mx, int my, int ty, int tx)
{
    const gbytes plane = (gbytes)c.planes[((my & 3) << 2) + (mx & 3)];
    return plane + c.off + (int64_t)((my >> 2) + ty) * c.i_ref + (mx >> 2) + tx;
}

__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

// Candidate k of a refinement stage around (bx, by) (quarter-pel units):
// stage 0 is the start point (k < c0) and the half-pel Spiral points
// (me.c:511-538), stage 1 the quarter-pel points with the PMVR adaptation
// (me.c:546-573, pmvr_adapt_mv me.c:412-424).  *skip marks a point the
// reference does not evaluate.
struct Stage {
    int bx, by, c0, step, stage;
    bool pmvr;
    int ctr_x, ctr_y;
    const int8_t (*pat)[2];
};

__device__ __forceinline__ void cand_mv(const Stage &st, int k, int &mx, int &my, bool &skip)
{
    skip = false;
    if (st.stage == 0) {
        if (k < st.c0) {
            mx = st.bx;
            my = st.by;
            return;
        }
        const int pos = 1 + (k - st.c0) * st.step;
        mx = st.bx + 2 * st.pat[pos][0];
        my = st.by + 2 * st.pat[pos][1];
        return;
    }
    const int pos = 1 + k * st.step;
    if (st.pmvr) {
        if (iabs(st.bx - st.ctr_x) > 2 || iabs(st.by - st.ctr_y) > 2) {
            mx = st.bx + st.pat[pos][0] * 2;
            my = st.by + st.pat[pos][1] * 2;
            skip = iabs(mx - st.ctr_x) <= 2 && iabs(my - st.ctr_y) <= 2;
        } else {
            mx = st.bx + st.pat[pos][0];
            my = st.by + st.pat[pos][1];
            skip = iabs(mx - st.ctr_x) > 2 || iabs(my - st.ctr_y) > 2;
        }
        return;
    }
    mx = st.bx + st.pat[pos][0];
    my = st.by + st.pat[pos][1];
}

__device__ __forceinline__ int pack_cost_idx(int cost, int k)
{
    // costs stay below 2^27 (a 64x64 SATD is below 2^22), 4 bits of order
    return cost == 0x7fffffff ? 0x7fffffff : (cost << 4) | k;
}

// ME_COST_QPEL (me.c:344-349) of candidate (mx, my) with distortion sum s
// (SATD before halving, or SAD).
__device__ __forceinline__ int cand_cost(const RefineCtx &c, int s, int mx, int my)
{
    const int dist = c.hadamard ? (s >> 1) : s;
    return dist + xme_weighted_cost(c.lambda, xme_mvbits(mx - c.pmx) + xme_mvbits(my - c.pmy));
}

// Evaluates the stage's n <= 9 candidates and returns the minimum of
// (cost << 4 | k) over the candidates k >= st.c0 that the reference
// evaluates (0x7fffffff when none); *start gets the cost of candidate 0.
// With P = the PU's 4x4 tile count rounded up to a power of two, a PU of at
// most 64 tiles is one wave's: each group of P lanes takes one candidate at a
// time (a lane = a tile) and the group's last lane completes the sum and the
// cost.  A larger PU (LARGE) is the whole workgroup's: a lane = a tile of
// 4 waves x 64, each wave sums its part of every candidate and the parts meet
// in LDS (`part`, 4 x 16 ints).  The current rows of a lane's tile are loaded
// once.
template <bool LARGE>
__device__ int stage_min(const RefineCtx &c, const Stage &st, int n, int *start, int *part)
{
    const int lane = threadIdx.x & 63;
    const int tw = c.w >> 2, ntiles = tw * (c.h >> 2);
    int P = 1;
    while (P < ntiles)
        P <<= 1;
    int m = 0x7fffffff;
    if constexpr (!LARGE) {        // P <= 64 (the PU lists are split on this)
        const int groups = 64 / P;
        const int g = lane / P, tile = lane % P;
        const bool has_tile = tile < ntiles;
        const int ty = has_tile ? (tile / tw) * 4 : 0, tx = has_tile ? (tile % tw) * 4 : 0;
        uint32_t cr[4];
#pragma unroll
        for (int y = 0; y < 4; y++)
            cr[y] = has_tile ? *(gdwords)(c.cur + (ty + y) * c.i_cur + tx) : 0u;
        int start_v = 0x7fffffff;
        for (int k0 = 0; k0 < n; k0 += groups) {
            const int k = k0 + g;
            int mx = 0, my = 0;
            bool skip = true;
            if (k < n)
                cand_mv(st, k, mx, my, skip);
            int v = 0;
            if (has_tile && !skip)
                v = tile_cost(cr, cand_ref(c, mx, my, ty, tx), c.i_ref, c.hadamard);
            v = group_sum(v, P);
            if (tile == P - 1 && !skip) {
                const int cost = cand_cost(c, v, mx, my);
                if (k == 0)
                    start_v = cost;
                if (k >= st.c0)
                    m = min(m, pack_cost_idx(cost, k));
            }
        }
        *start = __builtin_amdgcn_readlane(start_v, P - 1);   // k = 0: group 0 of the first pass
        return wave_min_key(m);
    } else {
        const int wave = threadIdx.x >> 6;
        const int tile = wave * 64 + lane;
        const bool has_tile = tile < ntiles;
        const int ty = has_tile ? (tile / tw) * 4 : 0, tx = has_tile ? (tile % tw) * 4 : 0;
        uint32_t cr[4];
#pragma unroll
        for (int y = 0; y < 4; y++)
            cr[y] = has_tile ? *(gdwords)(c.cur + (ty + y) * c.i_cur + tx) : 0u;
        for (int k = 0; k < n; k++) {
            int mx, my;
            bool skip;
            cand_mv(st, k, mx, my, skip);
            if (skip)                  // uniform: the candidates do not depend on the lane
                continue;
            int v = 0;
            if (has_tile)
                v = tile_cost(cr, cand_ref(c, mx, my, ty, tx), c.i_ref, c.hadamard);
            v = wave_sum(v);
            if (lane == 0)
                part[wave * 16 + k] = v;
        }
        __syncthreads();
        *start = 0x7fffffff;
        for (int k = 0; k < n; k++) {
            int mx, my;
            bool skip;
            cand_mv(st, k, mx, my, skip);
            if (skip)
                continue;
            const int cost = cand_cost(c, part[k] + part[16 + k] + part[32 + k] + part[48 + k], mx, my);
            if (k == 0)
                *start = cost;
            if (k >= st.c0)
                m = min(m, pack_cost_idx(cost, k));
        }
        __syncthreads();              // `part` is written again by the next stage
        return m;
    }
}

constexpr int SP_WAVES = 4;

// One workgroup of 4 waves takes either one instance of a PU of more than 64
// tiles (the first n_large * n_ref workgroups, so the long instances start
// first) or 4 instances of smaller PUs, one per wave.  Instances are walked
// reference-major (neighbouring PUs of one reference together) and each XCD
// takes a contiguous range of workgroups, so the L2 of an XCD sees the plane
// lines of neighbouring PUs back to back.
//
// The reference updates the best point sequentially with a strict '<' over
// the points of a stage in Spiral order; that is the first minimum among the
// points below the stage's starting cost, i.e. the minimum of
// (cost << 4 | k) compared against the starting cost.
template <bool LARGE>
__device__ void refine(const DevParams &p, const FrameDesc *__restrict__ desc, const DevPu *__restrict__ pus,
                       const DevInst *__restrict__ inst, const uint64_t *__restrict__ best, int pu_i, int r, int *part)
{
    const int lane = threadIdx.x & 63;
    const bool leader = LARGE ? threadIdx.x == 0 : lane == 0;
    const int id = pu_i * p.n_ref + r;
    const DevPu pu = pus[pu_i];
    const DevInst in = inst[id];
    const uint64_t key = best[id];
    xme_result res;
    if (in.status || key >= XME_KEY_NONE) {       // uniform over the instance's waves
        if (leader) {
            // _me_error (me.c:919-921, :1181): only the return value is defined
            res.cost = XME_MAX_DISTORTION;
            res.bcost2 = 0;
            res.bmv_x = res.bmv_y = res.bmv2_x = res.bmv2_y = 0;
            res.mvcost = 0;
            res.status = 1;
            desc->results[id] = res;
        }
        return;
    }
    const uint32_t order = (uint32_t)key;
    const int bcost_int = (int)(key >> 32);
    int bmx, bmy;
    if (order < XME_ORDER_RASTER) {
        // a select, not an index: a dynamically indexed copy of `in` would be
        // placed in LDS
        bmx = order == 0 ? in.mvc[0][0] : in.mvc[1][0];
        bmy = order == 0 ? in.mvc[0][1] : in.mvc[1][1];
    } else {
        const int o = (int)order - XME_ORDER_RASTER, R2 = 2 * p.R;
        bmx = in.omx + (o % R2) - p.R;
        bmy = in.omy + (o / R2) - p.R;
    }
    res.bcost2 = bcost_int;
    res.bmv2_x = (int16_t)bmx;
    res.bmv2_y = (int16_t)bmy;
    res.status = (in.flags & 1) ? 2 : 0;
    int bx = 4 * bmx, by = 4 * bmy, bcost = bcost_int;
    if (p.frac) {
        RefineCtx c;
        c.cur = (gbytes)desc->cur + (int64_t)pu.y * desc->i_cur + pu.x;
        c.i_cur = desc->i_cur;
        c.planes = desc->planes[r];
        c.i_ref = desc->stride[r];
        c.off = (int64_t)pu.y * c.i_ref + pu.x;
        c.w = pu.w;
        c.h = pu.h;
        c.hadamard = p.hadamard != 0;
        c.lambda = p.lambda;
        c.pmx = pu.mvp_x;
        c.pmy = pu.mvp_y;
        Stage st;
        st.step = p.fast ? 2 : 1;
        st.pat = p.fast ? c_spiral2 : c_spiral;
        st.pmvr = p.pmvr != 0;
        st.ctr_x = (pu.mvp_x >> 1) << 1;
        st.ctr_y = (pu.mvp_y >> 1) << 1;
        const int npts = p.fast ? 4 : 8;
        // stage 0: the SATD start cost (me.c:511-516) and the half-pel points
        // around the integer best (me.c:522-538)
        st.stage = 0;
        st.bx = bx;
        st.by = by;
        st.c0 = c.hadamard ? 1 : 0;
        int n = st.c0 + npts;
        int start = 0;
        uint32_t m = (uint32_t)stage_min<LARGE>(c, st, n, &start, part);
        if (c.hadamard)
            bcost = start;
        if ((int)(m >> 4) < bcost && m != 0x7fffffffu) {
            int mx, my;
            bool skip;
            cand_mv(st, (int)(m & 15), mx, my, skip);
            bcost = (int)(m >> 4);
            bx = mx;
            by = my;
        }
        if (p.frac > 1) {                        // quarter-pel (me.c:546-573)
            st.stage = 1;
            st.bx = bx;
            st.by = by;
            st.c0 = 0;
            n = npts;
            m = (uint32_t)stage_min<LARGE>(c, st, n, &start, part);
            if (m != 0x7fffffffu && (int)(m >> 4) < bcost) {
                int mx, my;
                bool skip;
                cand_mv(st, (int)(m & 15), mx, my, skip);
                bcost = (int)(m >> 4);
                bx = mx;
                by = my;
            }
        }
    }
    if (leader) {
        res.cost = bcost;
        res.bmv_x = (int16_t)bx;
        res.bmv_y = (int16_t)by;
        res.mvcost = xme_weighted_cost(p.lambda, xme_mvbits(bx - pu.mvp_x) + xme_mvbits(by - pu.mvp_y));
        desc->results[id] = res;
    }
}

__global__ void __launch_bounds__(256) subpel_kernel(DevParams p, const FrameDesc *__restrict__ desc,
                                                     const DevPu *__restrict__ pus, const DevInst *__restrict__ inst,
                                                     const uint64_t *__restrict__ best,
                                                     const int32_t *__restrict__ list, int n_small, int n_large)
{
    __shared__ int s_part[4 * 16];
    // Workgroup b runs on XCD b % 8.  The large instances take the first
    // blocks in dispatch order (spread over the XCDs), padded to a multiple of
    // 8; the small ones then map to contiguous ranges per XCD.
    const int n_large_wg = n_large * p.n_ref;
    const int large_blocks = (n_large_wg + 7) / 8 * 8;
    const int b = (int)blockIdx.x;
    if (b < large_blocks) {
        if (b < n_large_wg) {
            const int r = b / n_large;
            refine<true>(p, desc, pus, inst, best, list[n_small + b - r * n_large], r, s_part);
        }
        return;
    }
    const int bs = b - large_blocks;                     // bs % 8 == b % 8
    const int per_xcd = ((int)gridDim.x - large_blocks) / 8;
    const int lb = (bs % 8) * per_xcd + bs / 8;
    const int logical = __builtin_amdgcn_readfirstlane(lb * SP_WAVES + (int)(threadIdx.x >> 6));
    if (logical >= n_small * p.n_ref)
        return;
    const int r = logical / n_small;
    refine<false>(p, desc, pus, inst, best, list[logical - r * n_small], r, nullptr);
}

} // namespace

bool xme_subpel_large(int w, int h) { return (w >> 2) * (h >> 2) > 64; }

int xme_launch_subpel(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, const DevInst *d_inst,
                      const uint64_t *d_best, const int32_t *d_list, int n_small, int n_large, hipStream_t s)
{
    // n_large * n_ref workgroups for the large PUs (rounded up to a multiple
    // of 8), then SP_WAVES small instances per workgroup (a multiple of 8
    // workgroups: one range per XCD)
    const int large_blocks = (n_large * p.n_ref + 7) / 8 * 8;
    const int small_blocks = ((n_small * p.n_ref + SP_WAVES - 1) / SP_WAVES + 7) / 8 * 8;
    const int grid = large_blocks + small_blocks;
    if (grid <= 0)
        return XME_OK;
    hipLaunchKernelGGL(subpel_kernel, dim3(grid), dim3(256), 0, s, p, d_desc, d_pus, d_inst, d_best, d_list, n_small,
                       n_large);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
