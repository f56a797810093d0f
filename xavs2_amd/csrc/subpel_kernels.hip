// Sub-pel refinement (me_subpel_refine, source/encoder/me.c:479-579) and the
// store of the full-pel result (me.c:1169-1173) for gfx950.
//
// One wave per (PU, ref).  The candidates of a refinement stage (8 half-pel
// points around the integer best, then 8 quarter-pel points around the
// half-pel best) are independent, so a stage evaluates all of them at once:
// lanes take (candidate, 4x4 tile) pairs, compute the tile's Hadamard sum
// (or SAD) from the resident interpolated planes and add it to the
// candidate's LDS counter; the stage then applies the reference's sequential
// strict-'<' update in Spiral order.
#include "xme_internal.h"

namespace {

__constant__ int8_t c_spiral[9][2] = { { 0, 0 }, { 0, -1 }, { 0, 1 }, { -1, -1 }, { 1, -1 },
                                       { -1, 0 }, { 1, 0 }, { -1, 1 }, { 1, 1 } };
__constant__ int8_t c_spiral2[9][2] = { { 0, 0 }, { 0, -1 }, { -1, -1 }, { -1, 0 }, { -1, 1 },
                                        { 0, 1 }, { 1, 1 }, { 1, 0 }, { 1, -1 } };

__device__ __forceinline__ uint32_t load_unaligned_dw(const uint8_t *p)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint32_t *q = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(a & 3);
    // the second dword holds needed bytes only when the address is unaligned
    return sh ? __builtin_amdgcn_alignbyte(q[1], q[0], sh) : q[0];
}

__device__ __forceinline__ int bytei(uint32_t v, int i) { return (int)((v >> (8 * i)) & 0xFFu); }

// Sum of |H4 D H4^T| over one 4x4 tile (before the final halving), or the SAD.
__device__ __forceinline__ int tile_cost(const uint8_t *cur, int64_t i_cur, const uint8_t *ref, int64_t i_ref,
                                         bool hadamard)
{
    int d[4][4];
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const uint32_t c = *reinterpret_cast<const uint32_t *>(cur + y * i_cur);
        const uint32_t r = load_unaligned_dw(ref + y * i_ref);
#pragma unroll
        for (int x = 0; x < 4; x++)
            d[y][x] = bytei(c, x) - bytei(r, x);
    }
    int s = 0;
    if (!hadamard) {
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int x = 0; x < 4; x++)
                s += abs(d[y][x]);
        return s;
    }
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const int a = d[y][0] + d[y][1], b = d[y][0] - d[y][1];
        const int c = d[y][2] + d[y][3], e = d[y][2] - d[y][3];
        d[y][0] = a + c;
        d[y][1] = b + e;
        d[y][2] = a - c;
        d[y][3] = b - e;
    }
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const int a = d[0][x] + d[1][x], b = d[0][x] - d[1][x];
        const int c = d[2][x] + d[3][x], e = d[2][x] - d[3][x];
        s += abs(a + c) + abs(b + e) + abs(a - c) + abs(b - e);
    }
    return s;
}

struct RefineCtx {
    const uint8_t *cur;        // PU origin in the current picture
    int64_t i_cur;
    const uint8_t *const *planes;
    int64_t i_ref;
    int64_t off;               // PU offset in the reference planes (p_me->i_bias)
    int w, h;
    bool hadamard;
    int lambda, pmx, pmy;
};

// Costs of n <= 8 quarter-pel candidates (ME_COST_QPEL, me.c:344-349).
__device__ void eval_costs(const RefineCtx &c, const int (*mv)[2], int n, int *s_sum, int *out)
{
    const int lane = threadIdx.x;
    if (lane < 8)
        s_sum[lane] = 0;
    __syncthreads();
    const int tw = c.w >> 2, ntiles = tw * (c.h >> 2);
    for (int t = lane; t < n * ntiles; t += 64) {
        const int k = t / ntiles, tile = t - k * ntiles;
        const int ty = (tile / tw) * 4, tx = (tile % tw) * 4;
        const int mx = mv[k][0], my = mv[k][1];
        const uint8_t *plane = c.planes[((my & 3) << 2) + (mx & 3)];
        const uint8_t *ref = plane + c.off + (int64_t)((my >> 2) + ty) * c.i_ref + (mx >> 2) + tx;
        const int v = tile_cost(c.cur + ty * c.i_cur + tx, c.i_cur, ref, c.i_ref, c.hadamard);
        atomicAdd(&s_sum[k], v);
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
        const int dist = c.hadamard ? (s_sum[k] >> 1) : s_sum[k];
        out[k] = dist + xme_weighted_cost(c.lambda, xme_mvbits(mv[k][0] - c.pmx) + xme_mvbits(mv[k][1] - c.pmy));
    }
    __syncthreads();
}

__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

__global__ void __launch_bounds__(64) subpel_kernel(DevParams p, const FrameDesc *__restrict__ desc,
                                                    const DevPu *__restrict__ pus, const DevInst *__restrict__ inst,
                                                    const uint64_t *__restrict__ best, int n_inst)
{
    __shared__ int s_sum[8];
    const int id = blockIdx.x;
    if (id >= n_inst)
        return;
    const int lane = threadIdx.x;
    const DevPu pu = pus[id / p.n_ref];
    const int r = id % p.n_ref;
    const DevInst in = inst[id];
    const uint64_t key = best[id];
    xme_result res;
    if (in.status || key >= XME_KEY_NONE) {
        if (lane == 0) {
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
        bmx = in.mvc[order][0];
        bmy = in.mvc[order][1];
    } else {
        const int o = (int)order - XME_ORDER_RASTER, R2 = 2 * p.R;
        bmx = in.omx + (o % R2) - p.R;
        bmy = in.omy + (o / R2) - p.R;
    }
    res.bcost2 = bcost_int;
    res.bmv2_x = (int16_t)bmx;
    res.bmv2_y = (int16_t)bmy;
    res.status = 0;
    int bx = 4 * bmx, by = 4 * bmy, bcost = bcost_int;
    if (p.frac) {
        RefineCtx c;
        c.cur = desc->cur + (int64_t)pu.y * desc->i_cur + pu.x;
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
        const int step = p.fast ? 2 : 1;
        const int8_t(*pat)[2] = p.fast ? c_spiral2 : c_spiral;
        int mv[8][2], cost[8];
        if (c.hadamard) {                        // me.c:511-516
            mv[0][0] = bx;
            mv[0][1] = by;
            eval_costs(c, mv, 1, s_sum, cost);
            bcost = cost[0];
        }
        // half-pel around the integer best (me.c:522-538)
        int n = 0;
        for (int pos = 1; pos < 9; pos += step, n++) {
            mv[n][0] = bx + 2 * pat[pos][0];
            mv[n][1] = by + 2 * pat[pos][1];
        }
        eval_costs(c, mv, n, s_sum, cost);
        int nbx = bx, nby = by;
        for (int k = 0; k < n; k++)
            if (cost[k] < bcost) {
                bcost = cost[k];
                nbx = mv[k][0];
                nby = mv[k][1];
            }
        bx = nbx;
        by = nby;
        if (p.frac > 1) {                        // quarter-pel (me.c:546-573)
            const int ctr_x = (pu.mvp_x >> 1) << 1, ctr_y = (pu.mvp_y >> 1) << 1;
            n = 0;
            for (int pos = 1; pos < 9; pos += step) {
                int mx, my;
                if (p.pmvr) {                    // pmvr_adapt_mv (me.c:412-424)
                    bool skip;
                    if (iabs(bx - ctr_x) > 2 || iabs(by - ctr_y) > 2) {
                        mx = bx + pat[pos][0] * 2;
                        my = by + pat[pos][1] * 2;
                        skip = iabs(mx - ctr_x) <= 2 && iabs(my - ctr_y) <= 2;
                    } else {
                        mx = bx + pat[pos][0];
                        my = by + pat[pos][1];
                        skip = iabs(mx - ctr_x) > 2 || iabs(my - ctr_y) > 2;
                    }
                    if (skip)
                        continue;
                } else {
                    mx = bx + pat[pos][0];
                    my = by + pat[pos][1];
                }
                mv[n][0] = mx;
                mv[n][1] = my;
                n++;
            }
            if (n > 0)
                eval_costs(c, mv, n, s_sum, cost);
            nbx = bx;
            nby = by;
            for (int k = 0; k < n; k++)
                if (cost[k] < bcost) {
                    bcost = cost[k];
                    nbx = mv[k][0];
                    nby = mv[k][1];
                }
            bx = nbx;
            by = nby;
        }
    }
    if (lane == 0) {
        res.cost = bcost;
        res.bmv_x = (int16_t)bx;
        res.bmv_y = (int16_t)by;
        res.mvcost = xme_weighted_cost(p.lambda, xme_mvbits(bx - pu.mvp_x) + xme_mvbits(by - pu.mvp_y));
        desc->results[id] = res;
    }
}

} // namespace

int xme_launch_subpel(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, const DevInst *d_inst,
                      const uint64_t *d_best, int n_inst, hipStream_t s)
{
    if (n_inst <= 0)
        return XME_OK;
    hipLaunchKernelGGL(subpel_kernel, dim3(n_inst), dim3(64), 0, s, p, d_desc, d_pus, d_inst, d_best, n_inst);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
