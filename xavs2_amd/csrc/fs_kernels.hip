// Integer-pel full search (xavs2_me_search, FS branch, source/encoder/me.c:879-1173)
// for gfx950.
//
// Design:
//  * a candidate-list pass evaluates the <= 2 MV candidates of every (PU, ref)
//    in order and fixes the search centre (me.c:913-921);
//  * the search kernel takes tasks from a work queue; a task stages one
//    reference window and the current pixels of a group of PUs in LDS, and its
//    waves pull items (a band of candidate rows of one PU) from an LDS counter;
//  * a lane owns 4 adjacent candidate columns x K candidate rows.  For every
//    current row it loads one new reference row into a K-deep register ring
//    and issues v_qsad_pk_u16_u8 (4 candidates x 4 pixels per instruction)
//    against each ring row, so a reference byte fetched from LDS feeds 4K
//    absolute differences; 16-bit partial sums are folded into 32-bit ones
//    before they can overflow;
//  * the (cost, order) key of the best legal candidate of a wave goes to the
//    per-(PU, ref) slot with a 64-bit atomic min, which reproduces the
//    first-strict-minimum rule of the reference exactly.
#include "xme_dpp.h"
#include "xme_internal.h"
#include "tune.h"

#include <mutex>
#include <utility>

namespace {

// ---------------------------------------------------------------------------
// Candidate-list pass: a group of 16 lanes per (PU, reference), four groups a
// wave, 16 a workgroup; workgroup b runs on XCD b % 8, and the instances are
// walked reference-major in contiguous ranges per XCD so that neighbouring
// PUs of one reference share an L2.  A PU of more than 16 pixel quads loops
// over them (the groups of a wave run their loops independently; the sums
// follow once the wave has reconverged).
__global__ void __launch_bounds__(256) mvc_kernel(DevParams p, const FrameDesc *__restrict__ desc,
                                                  const DevPu *__restrict__ pus, DevInst *__restrict__ inst,
                                                  uint64_t *__restrict__ best, int n_inst)
{
    constexpr int S = 16;
    const int per_xcd = gridDim.x / 8;
    const int lb = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const int lane = threadIdx.x & 63, l = lane & (S - 1);
    const int logical = lb * 16 + (int)(threadIdx.x >> 4);
    if (__builtin_amdgcn_readfirstlane(lb * 16 + (int)((threadIdx.x >> 6) << 2)) >= n_inst)
        return;   // the whole wave is past the end
    const bool live = logical < n_inst;
    const int li = live ? logical : 0;
    const int r = li / p.n_pu, pu_i = li - r * p.n_pu;
    const int id = pu_i * p.n_ref + r;
    const DevPu pu = pus[pu_i];
    const MvWindow win = xme_mv_window(p, pu.x, pu.y, pu.w, pu.h);
    // add_one_mv_candidate (md_inter.c:840-865) for the MVP and (0,0)
    int cand[2][2], n = 0;
    const int q[2][2] = { { pu.mvp_x, pu.mvp_y }, { 0, 0 } };
    for (int c = 0; c < 2; c++) {
        int x = xme_clip3(win.min_fpel[0], win.max_fpel[0], (q[c][0] + 2) >> 2);
        int y = xme_clip3(win.min_fpel[1], win.max_fpel[1], (q[c][1] + 2) >> 2);
        bool dup = false;
        for (int k = 0; k < n; k++)
            dup |= cand[k][0] == x && cand[k][1] == y;
        if (!dup) {
            cand[n][0] = x;
            cand[n][1] = y;
            n++;
        }
    }
    // 4 pixels per lane step: the current rows are dword aligned (PU x and
    // the stride are multiples of 4); reference dwords are realigned, the
    // extra dword read stays inside the plane's padding.  Both candidates are
    // summed in one pass (their loads in flight together), then the costs are
    // compared in list order.
    typedef const __attribute__((address_space(1))) uint8_t *gbytes;
    typedef const __attribute__((address_space(1))) uint32_t *gdwords;
    const gbytes cur = (gbytes)desc->cur + (int64_t)pu.y * desc->i_cur + pu.x;
    const gbytes ref = (gbytes)desc->planes[r][0] + (int64_t)pu.y * desc->stride[r] + pu.x;
    const int64_t i_ref = desc->stride[r];
    bool ok[2];
    gbytes rp[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        ok[c] = live && c < n && xme_mv_legal(win, cand[c][0], cand[c][1]);
        rp[c] = ref + (int64_t)(ok[c] ? cand[c][1] : 0) * i_ref + (ok[c] ? cand[c][0] : 0);
    }
    uint32_t sad[2] = { 0, 0 };
    const int gw = pu.w >> 2;
    for (int t = l; live && t < gw * pu.h; t += S) {
        const int yy = t / gw, xx = (t - yy * gw) * 4;
        const uint32_t c4 = *(gdwords)(cur + (int64_t)yy * desc->i_cur + xx);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            if (ok[c]) {
                const uintptr_t a = (uintptr_t)(rp[c] + (int64_t)yy * i_ref + xx);
                const gdwords q4 = (gdwords)(a & ~(uintptr_t)3);
                sad[c] = __builtin_amdgcn_sad_u8(c4, __builtin_amdgcn_alignbyte(q4[1], q4[0], (uint32_t)(a & 3)), sad[c]);
            }
        }
    }
    // group sums in the group's last lane
    int tot[2];
#pragma unroll
    for (int c = 0; c < 2; c++)
        tot[c] = group_sum((int)sad[c], S);
    if (!live || l != S - 1)
        return;
    uint64_t key = XME_KEY_NONE;
    int bx = 0, by = 0;
#pragma unroll
    for (int c = 0; c < 2; c++) {
        if (!ok[c])
            continue;
        const int mx = cand[c][0], my = cand[c][1];
        const int cost = tot[c] + xme_weighted_cost(p.lambda, xme_mvbits(4 * mx - pu.mvp_x) + xme_mvbits(4 * my - pu.mvp_y));
        const uint64_t k = xme_key((uint32_t)cost, (uint32_t)c);
        if (k < key) {
            key = k;
            bx = mx;
            by = my;
        }
    }
    DevInst d;
    d.omx = (int16_t)bx;
    d.omy = (int16_t)by;
    for (int c = 0; c < 2; c++) {
        d.mvc[c][0] = (int16_t)(c < n ? cand[c][0] : 0);
        d.mvc[c][1] = (int16_t)(c < n ? cand[c][1] : 0);
    }
    d.n_mvc = (int8_t)n;
    d.status = key == XME_KEY_NONE ? 1 : 0;
    d.flags = 0;
    inst[id] = d;
    best[id] = key;
}

// ---------------------------------------------------------------------------
// Search kernel.

// Byte-aligned dword from two aligned dwords (little endian).
__device__ __forceinline__ uint32_t align_dw(uint32_t hi, uint32_t lo, uint32_t shift)
{
    return __builtin_amdgcn_alignbyte(hi, lo, shift);
}

// Accumulate one strip of SW current columns over all H current rows.
//   win  : LDS window, `pitch` bytes per row; lane reference row r starts at
//          byte offset base + r * pitch (base may be unaligned: `shift`)
//   cur  : LDS current pixels of this strip, `cpitch` bytes per row (4-aligned)
// Window rows live in a ring of K rows kept as the qsad's 8-byte operands
// (pairs of realigned dwords: a ring of 32-bit words read in overlapping pairs
// is merged by the compiler into misaligned 64-bit accesses and ends up in
// scratch).  At current row c, window row c + K - 1 enters the slot freed by
// row c - 1; it is first read by the last qsad of the first column group.
// Current rows are double-buffered: row c + 1 is loaded while row c is
// summed (H is a multiple of K, which is even, so c and cc share parity; the
// load past the last row rereads row H - 1).
template <int K, int SW>
__device__ __forceinline__ void fs_strip(const uint8_t *__restrict__ win, int pitch, int base, const uint8_t *__restrict__ cur,
                                         int cpitch, int H, uint32_t (&acc)[K][4])
{
    constexpr int ND = SW / 4;           // current dwords per row
    constexpr int FLUSH = 256 / SW;      // rows before a 16-bit lane could overflow (SW*rows*255 < 65536)
    const uint32_t shift = (uint32_t)base & 3u;
    const uint32_t *w32 = reinterpret_cast<const uint32_t *>(win + (base & ~3));
    const int pd = pitch >> 2;
    uint64_t ring[K][ND];
    uint64_t a16[K];
#pragma unroll
    for (int k = 0; k < K; k++)
        a16[k] = 0;
    auto load_row = [&](int r, uint64_t (&dst)[ND]) {
        uint32_t raw[ND + 2];
        const uint32_t *src = w32 + r * pd;
#pragma unroll
        for (int t = 0; t < ND + 2; t++)
            raw[t] = src[t];
        uint32_t al[ND + 1];
#pragma unroll
        for (int t = 0; t < ND + 1; t++)
            al[t] = align_dw(raw[t + 1], raw[t], shift);
#pragma unroll
        for (int t = 0; t < ND; t++)
            dst[t] = ((uint64_t)al[t + 1] << 32) | al[t];
    };
#pragma unroll
    for (int k = 0; k < K - 1; k++)
        load_row(k, ring[k]);
    uint32_t cwb[2][ND];
#pragma unroll
    for (int t = 0; t < ND; t++)
        cwb[0][t] = reinterpret_cast<const uint32_t *>(cur)[t];
    for (int c0 = 0; c0 < H; c0 += K) {
        const bool last = c0 + K >= H;
#pragma unroll
        for (int cc = 0; cc < K; cc++) {
            const int c = c0 + cc;
            load_row(c + K - 1, ring[(cc + K - 1) % K]);
            {
                const uint32_t *cp = reinterpret_cast<const uint32_t *>(cur + min(c + 1, H - 1) * cpitch);
#pragma unroll
                for (int t = 0; t < ND; t++)
                    cwb[(cc + 1) & 1][t] = cp[t];
            }
            // k innermost: consecutive qsads update different accumulators; the
            // scheduling barriers keep the compiler from pulling a dependent
            // qsad next to its producer
#pragma unroll
            for (int t = 0; t < ND; t++) {
#pragma unroll
                for (int k = 0; k < K; k++)
                    a16[k] = __builtin_amdgcn_qsad_pk_u16_u8(ring[(cc + k) % K][t], cwb[cc & 1][t], a16[k]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (((c0 + K) % FLUSH) == 0 || last) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                acc[k][0] += (uint32_t)a16[k] & 0xFFFFu;
                acc[k][1] += ((uint32_t)a16[k]) >> 16;
                acc[k][2] += (uint32_t)(a16[k] >> 32) & 0xFFFFu;
                acc[k][3] += (uint32_t)(a16[k] >> 48);
                a16[k] = 0;
            }
        }
    }
}

// The same for a dword-aligned lane base: the 8 reference bytes of every qsad
// are loaded straight into a register pair (no realignment).
template <int K, int SW>
__device__ __forceinline__ void fs_strip_aligned(const uint8_t *__restrict__ win, int pitch, int base,
                                                 const uint8_t *__restrict__ cur, int cpitch, int H,
                                                 uint32_t (&acc)[K][4])
{
    constexpr int ND = SW / 4;
    constexpr int FLUSH = 256 / SW;
    const uint32_t *w32 = reinterpret_cast<const uint32_t *>(win + base);
    const int pd = pitch >> 2;
    uint64_t ring[K][ND];
    uint64_t a16[K];
#pragma unroll
    for (int k = 0; k < K; k++)
        a16[k] = 0;
    auto load_row = [&](int r, uint64_t (&dst)[ND]) {
        const uint32_t *src = w32 + r * pd;
#pragma unroll
        for (int t = 0; t < ND; t++)
            dst[t] = ((uint64_t)src[t + 1] << 32) | src[t];
    };
#pragma unroll
    for (int k = 0; k < K - 1; k++)
        load_row(k, ring[k]);
    uint32_t cwb[2][ND];
#pragma unroll
    for (int t = 0; t < ND; t++)
        cwb[0][t] = reinterpret_cast<const uint32_t *>(cur)[t];
    for (int c0 = 0; c0 < H; c0 += K) {
        const bool last = c0 + K >= H;
#pragma unroll
        for (int cc = 0; cc < K; cc++) {
            const int c = c0 + cc;
            load_row(c + K - 1, ring[(cc + K - 1) % K]);
            {
                const uint32_t *cp = reinterpret_cast<const uint32_t *>(cur + min(c + 1, H - 1) * cpitch);
#pragma unroll
                for (int t = 0; t < ND; t++)
                    cwb[(cc + 1) & 1][t] = cp[t];
            }
#pragma unroll
            for (int t = 0; t < ND; t++) {
#pragma unroll
                for (int k = 0; k < K; k++)
                    a16[k] = __builtin_amdgcn_qsad_pk_u16_u8(ring[(cc + k) % K][t], cwb[cc & 1][t], a16[k]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (((c0 + K) % FLUSH) == 0 || last) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                acc[k][0] += (uint32_t)a16[k] & 0xFFFFu;
                acc[k][1] += ((uint32_t)a16[k]) >> 16;
                acc[k][2] += (uint32_t)(a16[k] >> 32) & 0xFFFFu;
                acc[k][3] += (uint32_t)(a16[k] >> 48);
                a16[k] = 0;
            }
        }
    }
}

// A lane's first candidate column and row in an item: 4 columns x K rows at
// its place in the band, pulled back inside the item when the band is cut
// short (the last band of 2R, or a window split into row ranges).  Such a lane
// then repeats candidates of another lane, which gives the same keys, and its
// reads stay inside the item's window footprint (item_footprint, host side).
__device__ __forceinline__ int lane_col(const FsItem &it, int lc)
{
    return min(it.col0 + 4 * lc, it.col0 + max((int)it.ncols, 4) - 4);
}

__device__ __forceinline__ int lane_row(const FsItem &it, int rg, int K)
{
    return min(it.row0 + rg * K, it.row0 + max((int)it.nrows, K) - K);
}

// Minimum over the 64 lanes, returned to every lane: DPP inside each row of
// 16 lanes (quad swaps, then row rotations by 4 and 8), then the four rows.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, false));  // row_ror:4
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false));  // row_ror:8
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    return min(min(a, b), min(c, d));
}

// Process one item with K rows per lane.
// Per-wave cache of the MV-cost terms of the lane's candidates, valid for one
// item (band), one search centre and one MVP.
struct TermState {
    bool ok;
    int cx, cy;     // search centre the cached MV costs were made for
    int px, py;     // and the MVP
    __device__ bool fresh(const DevInst &in, const DevPu &pu) const
    {
        return ok && cx == in.omx && cy == in.omy && px == pu.mvp_x && py == pu.mvp_y;
    }
    __device__ void set(const DevInst &in, const DevPu &pu)
    {
        ok = true;
        cx = in.omx;
        cy = in.omy;
        px = pu.mvp_x;
        py = pu.mvp_y;
    }
};

template <int K>
__device__ __forceinline__ void fs_item(const DevParams &p, const DevPu &pu, const DevInst &in, const FsItem &it, int inst_id,
                        const uint8_t *__restrict__ win, int pitch, int wx0, int wy0, const uint8_t *__restrict__ cur,
                        int cpitch, int cx0, int cy0, uint64_t *__restrict__ best, int32_t *__restrict__ cost_map,
                        uint32_t *__restrict__ term, TermState &ts)
{
    const int lane = threadIdx.x & 63;
    const int LC = it.LC;
    const int lc = lane % LC, rg = lane / LC;
    const int R = p.R;
    // LDS position of candidate (col 0, row 0): top-left of the block at
    // MV (omx - R, omy - R)
    const int bx = pu.x + in.omx - R - wx0;
    const int by = pu.y + in.omy - R - wy0;
    const int col = lane_col(it, lc);
    const int row = lane_row(it, rg, K);
    const int lane_base = (by + row) * pitch + bx + col;

    uint32_t acc[K][4];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int d = 0; d < 4; d++)
            acc[k][d] = 0;

    const uint8_t *cblk = cur + (pu.y - cy0) * cpitch + (pu.x - cx0);
    int x = 0;
    if ((lane_base & 3) == 0) {      // uniform: col and the pitch are multiples of 4
        for (; x + 16 <= pu.w; x += 16)
            fs_strip_aligned<K, 16>(win, pitch, lane_base + x, cblk + x, cpitch, pu.h, acc);
        if (x + 8 <= pu.w) {
            fs_strip_aligned<K, 8>(win, pitch, lane_base + x, cblk + x, cpitch, pu.h, acc);
            x += 8;
        }
        if (x + 4 <= pu.w)
            fs_strip_aligned<K, 4>(win, pitch, lane_base + x, cblk + x, cpitch, pu.h, acc);
    } else {
        for (; x + 16 <= pu.w; x += 16)
            fs_strip<K, 16>(win, pitch, lane_base + x, cblk + x, cpitch, pu.h, acc);
        if (x + 8 <= pu.w) {
            fs_strip<K, 8>(win, pitch, lane_base + x, cblk + x, cpitch, pu.h, acc);
            x += 8;
        }
        if (x + 4 <= pu.w)
            fs_strip<K, 4>(win, pitch, lane_base + x, cblk + x, cpitch, pu.h, acc);
    }

    // MV cost, legality and the first strict minimum (ME_COST_IPEL, me.c:187-192;
    // COPY3_IF_LT, me.h:70-75).  A lane visits its candidates in evaluation
    // order (rows j, then columns i), so a strict '<' on the cost alone keeps
    // the earliest of equal costs; the order joins the key only at the end.
    const MvWindow mw = xme_mv_window(p, pu.x, pu.y, pu.w, pu.h);
    const int R2 = 2 * R;
    const int col_end = it.col0 + it.ncols, row_end = it.row0 + it.nrows;
    int bits_x[4];
#pragma unroll
    for (int d = 0; d < 4; d++)
        bits_x[d] = xme_mvbits(4 * (in.omx + col + d - R) - pu.mvp_x);
    // uniform fast path: the band is complete and lies inside the legal MV box
    // (for |MV| < 8192 the packed CHECK_MV_RANGE equals the box test)
    const int bmx0 = in.omx + it.col0 - R, bmy0 = in.omy + it.row0 - R;
    const bool full = it.ncols == 4 * it.LC && it.nrows == (64 / it.LC) * K;
    const bool fast = full && bmx0 >= mw.min_fpel[0] && bmx0 + it.ncols - 1 <= mw.max_fpel[0] &&
                      bmy0 >= mw.min_fpel[1] && bmy0 + it.nrows - 1 <= mw.max_fpel[1];
    const uint32_t lam = (uint32_t)p.lambda;   // < 2^24 (checked when the plan is made)
    uint32_t bcost = 0xFFFFFFFFu;
    int bidx = 0;
    if (fast) {
        // term = (MV cost << 5) | candidate index, computed once per item and
        // centre; then key = (SAD << 5) + term and one unsigned min per
        // candidate (SAD < 2^21 and MV cost < 2^15, so no overflow)
        if (!ts.fresh(in, pu)) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int by_bits = xme_mvbits(4 * (in.omy + row + k - R) - pu.mvp_y);
#pragma unroll
                for (int d = 0; d < 4; d++)
                    term[(k * 4 + d) * 64 + lane] =
                        ((__umul24(lam, (uint32_t)(bits_x[d] + by_bits)) >> 16) << 5) | (uint32_t)(k * 4 + d);
            }
            ts.set(in, pu);
        }
        uint32_t m = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int d = 0; d < 4; d++)
                m = min(m, (acc[k][d] << 5) + term[(k * 4 + d) * 64 + lane]);
        bcost = m >> 5;
        bidx = (int)(m & 31u);
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = row + k;
            const int my = in.omy + j - R;
            const int by_bits = xme_mvbits(4 * my - pu.mvp_y);
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int i = col + d;
                const bool legal = i < col_end && j < row_end && xme_mv_legal(mw, in.omx + i - R, my);
                const uint32_t cost = acc[k][d] + (__umul24(lam, (uint32_t)(bits_x[d] + by_bits)) >> 16);
                if (legal && cost < bcost) {
                    bcost = cost;
                    bidx = k * 4 + d;
                }
            }
        }
    }
    // wave minimum of (cost << 11 | order inside the band): the band's
    // candidates in raster order are (local row) * 4 LC + local column < 256 K
    // <= 2^11, and a cost is below 2^21 (SAD <= 64 * 64 * 255, MV cost < 2^14)
    const int lrow = row - it.row0 + (bidx >> 2), lcol = col - it.col0 + (bidx & 3);
    uint32_t lkey = bcost == 0xFFFFFFFFu ? 0xFFFFFFFFu : (bcost << 11) | (uint32_t)(lrow * 4 * LC + lcol);
    lkey = wave_min_u32(lkey);
    if (lkey != 0xFFFFFFFFu && lane == 0) {
        const int lo = (int)(lkey & 2047u);
        const int j = it.row0 + lo / (4 * LC), i = it.col0 + lo % (4 * LC);
        const uint64_t kbest = xme_key(lkey >> 11, (uint32_t)(XME_ORDER_RASTER + j * R2 + i));
        atomicMin(reinterpret_cast<unsigned long long *>(best + inst_id), (unsigned long long)kbest);
    }
    if (cost_map != nullptr) {       // diagnostic output: every raster cost, -1 where illegal
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = row + k;
            const int my = in.omy + j - R;
            const int by_bits = xme_mvbits(4 * my - pu.mvp_y);
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int i = col + d;
                if (i < col_end && j < row_end) {
                    const bool legal = xme_mv_legal(mw, in.omx + i - R, my);
                    const int cost = (int)(acc[k][d] + (__umul24(lam, (uint32_t)(bits_x[d] + by_bits)) >> 16));
                    cost_map[(int64_t)inst_id * R2 * R2 + j * R2 + i] = legal ? cost : -1;
                }
            }
        }
    }
}

// Short PUs (at most 256 pixels) keep their H x 4NQ current pixels in
// registers, NQ = width / 4.  Each of the H + K - 1 window
// rows of the lane is loaded once, realigned to NQ + 1 dwords, and meets every
// current row c = r - k with k in the lane's K candidate rows (a ring of K
// window rows would need K x 2NQ registers).  Window rows are loaded PF rows
// ahead: the first and last rows feed only a few candidate rows, too few
// qsads to cover the LDS latency of the next row.
template <int NQ, bool ALIGNED>
struct ShortRows {
    static constexpr int NW = ALIGNED ? NQ + 1 : NQ + 2;   // dwords loaded per window row
    static constexpr int PF = NQ <= 2 ? 4 : NQ <= 4 ? 2 : 1;   // rows in flight
};

template <int H, int NQ, bool ALIGNED, int r>
__device__ __forceinline__ void short_row(const uint32_t *__restrict__ w32, int pd, uint32_t shift,
                                          const uint32_t (&cw)[H][NQ],
                                          uint32_t (&raw)[ShortRows<NQ, ALIGNED>::PF][ShortRows<NQ, ALIGNED>::NW],
                                          uint64_t (&a16)[XME_NARROW_K])
{
    constexpr int K = XME_NARROW_K;
    constexpr int NW = ShortRows<NQ, ALIGNED>::NW;
    constexpr int PF = ShortRows<NQ, ALIGNED>::PF;
    constexpr int k_lo = r - (H - 1) > 0 ? r - (H - 1) : 0;
    constexpr int k_hi = r < K - 1 ? r : K - 1;
    uint32_t d[NQ + 1];
#pragma unroll
    for (int q = 0; q <= NQ; q++)
        d[q] = ALIGNED ? raw[r % PF][q] : __builtin_amdgcn_alignbyte(raw[r % PF][q + 1], raw[r % PF][q], shift);
    if constexpr (r + PF < H + K - 1) {
        const uint32_t *src = w32 + (r + PF) * pd;
#pragma unroll
        for (int q = 0; q < NW; q++)
            raw[r % PF][q] = src[q];
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const uint64_t w = ((uint64_t)d[q + 1] << 32) | d[q];
#pragma unroll
        for (int k = k_lo; k <= k_hi; k++)
            a16[k] = __builtin_amdgcn_qsad_pk_u16_u8(w, cw[r - k][q], a16[k]);
    }
    __builtin_amdgcn_sched_barrier(0);
}

template <int H, int NQ, bool ALIGNED, int... Rs>
__device__ __forceinline__ void short_rows(std::integer_sequence<int, Rs...>, const uint32_t *__restrict__ w32, int pd,
                                           uint32_t shift, const uint32_t (&cw)[H][NQ],
                                           uint32_t (&raw)[ShortRows<NQ, ALIGNED>::PF][ShortRows<NQ, ALIGNED>::NW],
                                           uint64_t (&a16)[XME_NARROW_K])
{
    (short_row<H, NQ, ALIGNED, Rs>(w32, pd, shift, cw, raw, a16), ...);
}

template <int H, int NQ, bool ALIGNED>
__device__ __forceinline__ void short_strip(const uint8_t *__restrict__ win, int pitch, int base,
                                            const uint8_t *__restrict__ cur, int cpitch,
                                            uint64_t (&a16)[XME_NARROW_K])
{
    constexpr int NW = ShortRows<NQ, ALIGNED>::NW;
    constexpr int PF = ShortRows<NQ, ALIGNED>::PF;
    static_assert(PF <= H + XME_NARROW_K - 1, "prefetch depth");
    const uint32_t shift = (uint32_t)base & 3u;
    const uint32_t *w32 = reinterpret_cast<const uint32_t *>(win + (base & ~3));
    const int pd = pitch >> 2;
    uint32_t cw[H][NQ];
    uint32_t raw[PF][NW];
    // issue order: what the first rows need comes first
#pragma unroll
    for (int q = 0; q < NQ; q++)
        cw[0][q] = reinterpret_cast<const uint32_t *>(cur)[q];
#pragma unroll
    for (int i = 0; i < PF; i++)
#pragma unroll
        for (int q = 0; q < NW; q++)
            raw[i][q] = w32[i * pd + q];
#pragma unroll
    for (int c = 1; c < H; c++) {
        const uint32_t *cp = reinterpret_cast<const uint32_t *>(cur + c * cpitch);
#pragma unroll
        for (int q = 0; q < NQ; q++)
            cw[c][q] = cp[q];
    }
    short_rows<H, NQ, ALIGNED>(std::make_integer_sequence<int, H + XME_NARROW_K - 1>{}, w32, pd, shift, cw, raw, a16);
}

// The packed MV costs of a short-PU lane's 16 x 4 candidates into the wave's
// LDS term area: init[k * 64 + lane] holds columns d = 0..3 of candidate row
// k as u16.  ME_COST_IPEL (me.c:187-192) is (lambda * (bits(4 dx) +
// bits(4 dy))) >> 16; lambda < 2^24 and the bit counts are small, so
// lambda * bx + lambda * by is the same sum (below 2^31) and the products
// come out of the 64-candidate loop.
__device__ __forceinline__ void short_mv_costs(const DevParams &p, const DevPu &pu, const DevInst &in, int col, int row,
                                               uint64_t *__restrict__ init, int lane)
{
    const uint32_t lam = (uint32_t)p.lambda;
    const int R = p.R;
    uint32_t lx[4];
#pragma unroll
    for (int d = 0; d < 4; d++)
        lx[d] = __umul24(lam, (uint32_t)xme_mvbits(4 * (in.omx + col + d - R) - pu.mvp_x));
#pragma unroll
    for (int k = 0; k < XME_NARROW_K; k++) {
        const uint32_t ly = __umul24(lam, (uint32_t)xme_mvbits(4 * (in.omy + row + k - R) - pu.mvp_y));
        const uint32_t lo = ((lx[1] + ly) & 0xFFFF0000u) | ((lx[0] + ly) >> 16);
        const uint32_t hi = ((lx[3] + ly) & 0xFFFF0000u) | ((lx[2] + ly) >> 16);
        init[k * 64 + lane] = ((uint64_t)hi << 32) | lo;
    }
}

// Narrow PUs (8 pixels wide, at most 16 rows) carry little work per
// candidate, so per-call costs dominate the generic path.  Here a lane takes
// K = 16 candidate rows; the SAD and the MV cost of each candidate are summed
// in the qsad's u16 lanes (8 * 16 * 255 + MV cost < 2^16), so the strip needs
// no flush and the epilogue reads the u16 sums directly.  The packed MV costs
// depend only on the lane, the item and the search centre; they are kept in
// the wave's LDS term area (16 x 64 u64 = 8 KB, the size of the generic path's
// term table) across the item's instances.

template <int H>
__device__ __forceinline__ void fs_item_narrow(const DevParams &p, const DevPu &pu, const DevInst &in, const FsItem &it, int inst_id,
                               const uint8_t *__restrict__ win, int pitch, int wx0, int wy0,
                               const uint8_t *__restrict__ cur, int cpitch, int cx0, int cy0,
                               uint64_t *__restrict__ best, int32_t *__restrict__ cost_map,
                               uint32_t *__restrict__ term, TermState &ts)
{
    constexpr int K = XME_NARROW_K;
    const int lane = threadIdx.x & 63;
    const int LC = it.LC;
    const int lc = lane % LC, rg = lane / LC;
    const int R = p.R, R2 = 2 * R;
    const int bx = pu.x + in.omx - R - wx0;
    const int by = pu.y + in.omy - R - wy0;
    const int col = lane_col(it, lc);
    const int row = lane_row(it, rg, K);
    const int lane_base = (by + row) * pitch + bx + col;
    uint64_t *init = reinterpret_cast<uint64_t *>(term);    // [k][lane]
    if (!ts.fresh(in, pu)) {
        short_mv_costs(p, pu, in, col, row, init, lane);
        ts.set(in, pu);
    }
    uint64_t a16[K];
#pragma unroll
    for (int k = 0; k < K; k++)
        a16[k] = init[k * 64 + lane];
    const uint8_t *cblk = cur + (pu.y - cy0) * cpitch + (pu.x - cx0);
    if ((lane_base & 3) == 0)      // uniform: col and the pitch are multiples of 4
        short_strip<H, 2, true>(win, pitch, lane_base, cblk, cpitch, a16);
    else
        short_strip<H, 2, false>(win, pitch, lane_base, cblk, cpitch, a16);

    // first strict minimum in raster order (COPY3_IF_LT, me.h:70-75): the lane's
    // best as (cost << 6 | k * 4 + d), then the wave's as (cost << 12 | order in
    // the band); band order < 4 LC * (64 / LC) * 16 = 4096
    const MvWindow mw = xme_mv_window(p, pu.x, pu.y, pu.w, pu.h);
    const int col_end = it.col0 + it.ncols, row_end = it.row0 + it.nrows;
    const int bmx0 = in.omx + it.col0 - R, bmy0 = in.omy + it.row0 - R;
    const bool full = it.ncols == 4 * LC && it.nrows == (64 / LC) * K;
    const bool fast = full && bmx0 >= mw.min_fpel[0] && bmx0 + it.ncols - 1 <= mw.max_fpel[0] &&
                      bmy0 >= mw.min_fpel[1] && bmy0 + it.nrows - 1 <= mw.max_fpel[1];
    // lane key (cost << 16 | k * 4 + d): the u16 halves become keys with one
    // shift-or or and-or each, and min3 takes two keys at a time
    uint32_t m = 0xFFFFFFFFu;
    if (fast) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t lo = (uint32_t)a16[k], hi = (uint32_t)(a16[k] >> 32);
            const uint32_t k0 = (lo << 16) | (uint32_t)(k * 4 + 0);
            const uint32_t k1 = (lo & 0xFFFF0000u) | (uint32_t)(k * 4 + 1);
            const uint32_t k2 = (hi << 16) | (uint32_t)(k * 4 + 2);
            const uint32_t k3 = (hi & 0xFFFF0000u) | (uint32_t)(k * 4 + 3);
            m = min(m, min(k0, k1));
            m = min(m, min(k2, k3));
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = row + k;
            const int my = in.omy + j - R;
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int i = col + d;
                const bool legal = i < col_end && j < row_end && xme_mv_legal(mw, in.omx + i - R, my);
                const uint32_t v = (uint32_t)(a16[k] >> (16 * d)) & 0xFFFFu;
                if (legal)
                    m = min(m, (v << 16) | (uint32_t)(k * 4 + d));
            }
        }
    }
    uint32_t lkey = 0xFFFFFFFFu;
    if (m != 0xFFFFFFFFu) {
        const int idx = (int)(m & 63u);
        const int lrow = row - it.row0 + (idx >> 2), lcol = col - it.col0 + (idx & 3);
        lkey = ((m >> 16) << 12) | (uint32_t)(lrow * 4 * LC + lcol);
    }
    lkey = wave_min_u32(lkey);
    if (lkey != 0xFFFFFFFFu && lane == 0) {
        const int lo = (int)(lkey & 4095u);
        const int j = it.row0 + lo / (4 * LC), i = it.col0 + lo % (4 * LC);
        const uint64_t kbest = xme_key(lkey >> 12, (uint32_t)(XME_ORDER_RASTER + j * R2 + i));
        atomicMin(reinterpret_cast<unsigned long long *>(best + inst_id), (unsigned long long)kbest);
    }
    if (cost_map != nullptr) {       // diagnostic output: every raster cost, -1 where illegal
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = row + k;
            const int my = in.omy + j - R;
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int i = col + d;
                if (i < col_end && j < row_end) {
                    const bool legal = xme_mv_legal(mw, in.omx + i - R, my);
                    const int v = (int)((a16[k] >> (16 * d)) & 0xFFFFu);
                    cost_map[(int64_t)inst_id * R2 * R2 + j * R2 + i] = legal ? v : -1;
                }
            }
        }
    }
}

// The other PUs of at most 256 pixels (4, 12, 16 or 32 wide; 8 x 32) use the
// same 16 candidate rows per lane, with the packed u16 MV costs cached in the
// wave's LDS term area like the narrow path's.  PUs of at most 128 pixels
// (16 x 8, 16 x 4, 4 x 16, 4 x 8, 4 x 4) start their qsad accumulators at the
// MV cost, as the narrow path does (128 * 255 + MV cost < 2^16); the SAD of
// the larger ones alone may fill the u16 lanes (256 * 255 < 2^16), so their
// MV costs are added in 32 bits in the epilogue.
template <int H, int NQ>
__device__ __forceinline__ void fs_item_wide(const DevParams &p, const DevPu &pu, const DevInst &in,
                                               const FsItem &it, int inst_id, const uint8_t *__restrict__ win,
                                               int pitch, int wx0, int wy0, const uint8_t *__restrict__ cur,
                                               int cpitch, int cx0, int cy0, uint64_t *__restrict__ best,
                                               int32_t *__restrict__ cost_map, uint32_t *__restrict__ term,
                                               TermState &ts)
{
    constexpr int K = XME_NARROW_K;
    constexpr bool PRE = H * 4 * NQ <= 128;    // the MV cost rides in the u16 accumulators
    constexpr int KS = PRE ? 16 : 6;           // the lane key's index bits
    const int lane = threadIdx.x & 63;
    const int LC = it.LC;
    const int lc = lane % LC, rg = lane / LC;
    const int R = p.R, R2 = 2 * R;
    const int bx = pu.x + in.omx - R - wx0;
    const int by = pu.y + in.omy - R - wy0;
    const int col = lane_col(it, lc);
    const int row = lane_row(it, rg, K);
    const int lane_base = (by + row) * pitch + bx + col;
    uint64_t *init = reinterpret_cast<uint64_t *>(term);    // [k][lane]: packed u16 MV costs
    if (!ts.fresh(in, pu)) {
        short_mv_costs(p, pu, in, col, row, init, lane);
        ts.set(in, pu);
    }
    uint64_t a16[K], mvc[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        // read now: the loads land during the strip
        a16[k] = PRE ? init[k * 64 + lane] : 0;
        mvc[k] = PRE ? 0 : init[k * 64 + lane];
    }
    const uint8_t *cblk = cur + (pu.y - cy0) * cpitch + (pu.x - cx0);
    if ((lane_base & 3) == 0)      // uniform: col and the pitch are multiples of 4
        short_strip<H, NQ, true>(win, pitch, lane_base, cblk, cpitch, a16);
    else
        short_strip<H, NQ, false>(win, pitch, lane_base, cblk, cpitch, a16);

    const MvWindow mw = xme_mv_window(p, pu.x, pu.y, pu.w, pu.h);
    const int col_end = it.col0 + it.ncols, row_end = it.row0 + it.nrows;
    const int bmx0 = in.omx + it.col0 - R, bmy0 = in.omy + it.row0 - R;
    const bool full = it.ncols == 4 * LC && it.nrows == (64 / LC) * K;
    const bool fast = full && bmx0 >= mw.min_fpel[0] && bmx0 + it.ncols - 1 <= mw.max_fpel[0] &&
                      bmy0 >= mw.min_fpel[1] && bmy0 + it.nrows - 1 <= mw.max_fpel[1];
    // cost = SAD + MV cost < 2^17: the lane key (cost << 6 | k * 4 + d) fits;
    // with the cost in the u16 lanes, (cost << 16 | k * 4 + d) is one shift-or
    // or and-or per candidate, as in the narrow path
    uint32_t m = 0xFFFFFFFFu;
    if (fast && PRE) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t lo = (uint32_t)a16[k], hi = (uint32_t)(a16[k] >> 32);
            const uint32_t k0 = (lo << 16) | (uint32_t)(k * 4 + 0);
            const uint32_t k1 = (lo & 0xFFFF0000u) | (uint32_t)(k * 4 + 1);
            const uint32_t k2 = (hi << 16) | (uint32_t)(k * 4 + 2);
            const uint32_t k3 = (hi & 0xFFFF0000u) | (uint32_t)(k * 4 + 3);
            m = min(m, min(k0, k1));
            m = min(m, min(k2, k3));
        }
    } else if (fast) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint64_t a = a16[k], mv = mvc[k];
            uint32_t key[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const uint32_t v = ((uint32_t)(a >> (16 * d)) & 0xFFFFu) + ((uint32_t)(mv >> (16 * d)) & 0xFFFFu);
                key[d] = (v << 6) | (uint32_t)(k * 4 + d);
            }
            m = min(m, min(key[0], key[1]));
            m = min(m, min(key[2], key[3]));
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint64_t a = a16[k], mv = mvc[k];
            const int j = row + k;
            const int my = in.omy + j - R;
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int i = col + d;
                const bool legal = i < col_end && j < row_end && xme_mv_legal(mw, in.omx + i - R, my);
                const uint32_t v = ((uint32_t)(a >> (16 * d)) & 0xFFFFu) + ((uint32_t)(mv >> (16 * d)) & 0xFFFFu);
                if (legal)
                    m = min(m, (v << KS) | (uint32_t)(k * 4 + d));
            }
        }
    }
    uint32_t lkey = 0xFFFFFFFFu;
    if (m != 0xFFFFFFFFu) {
        const int idx = (int)(m & 63u);
        const int lrow = row - it.row0 + (idx >> 2), lcol = col - it.col0 + (idx & 3);
        lkey = ((m >> KS) << 12) | (uint32_t)(lrow * 4 * LC + lcol);
    }
    lkey = wave_min_u32(lkey);
    if (lkey != 0xFFFFFFFFu && lane == 0) {
        const int lo = (int)(lkey & 4095u);
        const int j = it.row0 + lo / (4 * LC), i = it.col0 + lo % (4 * LC);
        const uint64_t kbest = xme_key(lkey >> 12, (uint32_t)(XME_ORDER_RASTER + j * R2 + i));
        atomicMin(reinterpret_cast<unsigned long long *>(best + inst_id), (unsigned long long)kbest);
    }
    if (cost_map != nullptr) {       // diagnostic output: every raster cost, -1 where illegal
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint64_t a = a16[k], mv = mvc[k];
            const int j = row + k;
            const int my = in.omy + j - R;
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int i = col + d;
                if (i < col_end && j < row_end) {
                    const bool legal = xme_mv_legal(mw, in.omx + i - R, my);
                    const int v = (int)(((a >> (16 * d)) & 0xFFFFu) + ((mv >> (16 * d)) & 0xFFFFu));
                    cost_map[(int64_t)inst_id * R2 * R2 + j * R2 + i] = legal ? v : -1;
                }
            }
        }
    }
}

// Stage a task's reference window and current pixels in LDS.
__device__ __forceinline__ void stage_task(const FsTask &t, const FrameDesc *__restrict__ desc, const DevParams &p, uint8_t *lds,
                           int pitch, int cpitch)
{
    const uint8_t *ref = desc->planes[t.ref][0];
    const int64_t stride = desc->stride[t.ref];
    // rows/columns outside the padded picture are clamped to it; only illegal
    // candidates (never selected) can read them
    const int xlo = -XME_PAD, xhi = p.frame_w + XME_PAD - 4;
    const int ylo = -XME_PAD, yhi = p.frame_h + XME_PAD - 1;
    const int ndw = t.ww >> 2;
    uint32_t *w32 = reinterpret_cast<uint32_t *>(lds);
    const int total = ndw * t.wh;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int r = e / ndw, q = e - r * ndw;
        const int y = xme_clip3(ylo, yhi, t.wy0 + r);
        const int x = t.wx0 + 4 * q;
        const uint8_t *row = ref + (int64_t)y * stride;
        uint32_t v;
        if (x >= xlo && x <= xhi) {
            v = *reinterpret_cast<const uint32_t *>(row + x);
        } else {
            const uint32_t b = row[xme_clip3(xlo, xhi + 3, x)];
            v = b * 0x01010101u;
        }
        w32[r * (pitch >> 2) + q] = v;
    }
    const uint8_t *cur = desc->cur;
    const int64_t i_cur = desc->i_cur;
    uint32_t *c32 = reinterpret_cast<uint32_t *>(lds + pitch * t.wh);
    const int cdw = t.cw >> 2;
    const int ctotal = cdw * t.ch;
    for (int e = threadIdx.x; e < ctotal; e += blockDim.x) {
        const int r = e / cdw, q = e - r * cdw;
        c32[r * (cpitch >> 2) + q] =
            *reinterpret_cast<const uint32_t *>(cur + (int64_t)(t.cy0 + r) * i_cur + t.cx0 + 4 * q);
    }
}

__device__ __forceinline__ int window_pitch(int ww)
{
    // dword-multiple pitch; an odd dword count spreads rows over LDS banks
    int dw = ww >> 2;
    return (dw | 1) << 2;
}

// Byte offset of the per-wave MV-cost term tables (16-byte aligned, after the
// window and the current pixels); each of the 4 waves owns 8 * 4 * 64 dwords.
__host__ __device__ inline int fs_term_offset(int staged_bytes) { return (staged_bytes + 15) & ~15; }

// KMAX selects the item paths the kernel is compiled with: 4 (K = 4 items;
// fits 128 VGPRs, 4 waves per SIMD), 8 (K = 8 and K = 4 items, 2 waves per
// SIMD) or XME_NARROW_K (the narrow-PU items only).  The narrow path lives in
// its own variant because the register allocation of one kernel holding both
// it and the K = 8 path spills.
template <bool PERSISTENT, int KMAX>
__global__ void __launch_bounds__(256, KMAX == 4 ? 4 : 2) fs_kernel(DevParams p, const FrameDesc *__restrict__ desc,
                                                 const DevPu *__restrict__ pus, const DevInst *__restrict__ inst,
                                                 const FsTask *__restrict__ tasks, int n_tasks,
                                                 const FsItem *__restrict__ items, const int32_t *__restrict__ item_insts,
                                                 uint64_t *__restrict__ best, uint32_t *__restrict__ counters)
{
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ int s_task;
    __shared__ int s_item;
    const int32_t *cost_map = desc->cost_map;
    int task = PERSISTENT ? -1 : (int)blockIdx.x;
    for (;;) {
        if (PERSISTENT) {
            if (threadIdx.x == 0)
                s_task = (int)atomicAdd(&counters[0], 1u);
            __syncthreads();
            task = s_task;
        }
        if (task >= n_tasks)
            break;
        const FsTask t = tasks[task];
        const int pitch = window_pitch(t.ww);
        const int cpitch = t.cw;
        stage_task(t, desc, p, lds, pitch, cpitch);
        if (threadIdx.x == 0)
            s_item = t.item_begin;
        __syncthreads();
        const uint8_t *cur = lds + pitch * t.wh;
        uint32_t *term = reinterpret_cast<uint32_t *>(lds + fs_term_offset(pitch * t.wh + t.cw * t.ch)) +
                         (threadIdx.x >> 6) * (8 * 4 * 64);
        for (;;) {
            int it_i = 0;
            if ((threadIdx.x & 63) == 0)
                it_i = atomicAdd(&s_item, 1);
            it_i = __builtin_amdgcn_readlane(it_i, 0);
            if (it_i >= t.item_end)
                break;
            const FsItem it = items[it_i];
            TermState ts;
            ts.ok = false;
            ts.cx = ts.cy = ts.px = ts.py = 0;
            // the item's instances (<= XME_MAX_GROUP) are fetched at once, one
            // per lane, and handed to the wave one after the other
            const int lane = threadIdx.x & 63;
            int my_id = 0;
            uint4 my_in = make_uint4(0, 0, 0, 0), my_pu = make_uint4(0, 0, 0, 0);
            if (lane < it.count) {
                my_id = item_insts[it.list + lane];
                my_in = reinterpret_cast<const uint4 *>(inst)[my_id];
                my_pu = reinterpret_cast<const uint4 *>(pus)[my_id / p.n_ref];
            }
            for (int g = 0; g < it.count; g++) {
                const int inst_id = __builtin_amdgcn_readlane(my_id, g);
                const uint4 vin = make_uint4(__builtin_amdgcn_readlane(my_in.x, g), __builtin_amdgcn_readlane(my_in.y, g),
                                             __builtin_amdgcn_readlane(my_in.z, g), __builtin_amdgcn_readlane(my_in.w, g));
                const DevInst in = __builtin_bit_cast(DevInst, vin);
                if (in.status)
                    continue;
                const uint4 vpu = make_uint4(__builtin_amdgcn_readlane(my_pu.x, g), __builtin_amdgcn_readlane(my_pu.y, g),
                                             __builtin_amdgcn_readlane(my_pu.z, g), __builtin_amdgcn_readlane(my_pu.w, g));
                const DevPu pu = __builtin_bit_cast(DevPu, vpu);
                if constexpr (KMAX == XME_NARROW_K) {
#define XME_WIDE(HH, NQQ)                                                                                     \
    fs_item_wide<HH, NQQ>(p, pu, in, it, inst_id, lds, pitch, t.wx0, t.wy0, cur, cpitch, t.cx0, t.cy0, best,   \
                          const_cast<int32_t *>(cost_map), term, ts)
#define XME_NARROW(HH)                                                                                        \
    fs_item_narrow<HH>(p, pu, in, it, inst_id, lds, pitch, t.wx0, t.wy0, cur, cpitch, t.cx0, t.cy0, best,      \
                       const_cast<int32_t *>(cost_map), term, ts)
                    // the shapes of xme_narrow_shape (host side)
                    switch (pu.w * 64 + pu.h) {
                    case 8 * 64 + 16: XME_NARROW(16); break;
                    case 8 * 64 + 8: XME_NARROW(8); break;
                    case 8 * 64 + 4: XME_NARROW(4); break;
                    case 16 * 64 + 16: XME_WIDE(16, 4); break;
                    case 16 * 64 + 12: XME_WIDE(12, 4); break;
                    case 16 * 64 + 8: XME_WIDE(8, 4); break;
                    case 16 * 64 + 4: XME_WIDE(4, 4); break;
                    case 32 * 64 + 8: XME_WIDE(8, 8); break;
                    case 8 * 64 + 32: XME_WIDE(32, 2); break;
                    case 12 * 64 + 16: XME_WIDE(16, 3); break;
                    case 4 * 64 + 16: XME_WIDE(16, 1); break;
                    case 4 * 64 + 8: XME_WIDE(8, 1); break;
                    default: XME_WIDE(4, 1); break;   // 4 x 4
                    }
#undef XME_WIDE
#undef XME_NARROW
                } else if (KMAX == 8 && it.K == 8) {
                    fs_item<8>(p, pu, in, it, inst_id, lds, pitch, t.wx0, t.wy0, cur, cpitch, t.cx0, t.cy0, best,
                               const_cast<int32_t *>(cost_map), term, ts);
                } else {
                    fs_item<4>(p, pu, in, it, inst_id, lds, pitch, t.wx0, t.wy0, cur, cpitch, t.cx0, t.cy0, best,
                               const_cast<int32_t *>(cost_map), term, ts);
                }
            }
        }
        __syncthreads();   // LDS is reused by the next task
        if (!PERSISTENT)
            break;
    }
    if (PERSISTENT && threadIdx.x == 0) {
        // the last workgroup out resets the queue for the next launch / replay
        __threadfence();
        const uint32_t done = atomicAdd(&counters[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&counters[0], 0u);
            atomicExch(&counters[1], 0u);
        }
    }
}

} // namespace

int xme_fs_max_lds(void) { return 160 * 1024 - 64; }

// Allow the search kernels the full 160 KB of LDS of a gfx950 CU (once per process).
int xme_fs_prepare(void)
{
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        const void *ks[6] = { xme_fs_kernel_ptr(true, 4), xme_fs_kernel_ptr(true, 8), xme_fs_kernel_ptr(false, 4),
                              xme_fs_kernel_ptr(false, 8), xme_fs_kernel_ptr(true, XME_NARROW_K),
                              xme_fs_kernel_ptr(false, XME_NARROW_K) };
        for (const void *k : ks)
            if (err == hipSuccess)
                err = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, xme_fs_max_lds());
    });
    if (err != hipSuccess)
        return xme_hip_fail(err, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
    return XME_OK;
}

int xme_launch_mvc(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, DevInst *d_inst,
                   uint64_t *d_best, int n_inst, hipStream_t s)
{
    if (n_inst <= 0)
        return XME_OK;
    // a multiple of 8 workgroups (one range per XCD), 16 instances each
    const int grid = ((n_inst + 15) / 16 + 7) / 8 * 8;
    hipLaunchKernelGGL(mvc_kernel, dim3(grid), dim3(256), 0, s, p, d_desc, d_pus, d_inst, d_best, n_inst);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

int xme_launch_fs(const DevParams &p, const FrameDesc *d_desc, const DevPu *d_pus, const DevInst *d_inst,
                  const FsTask *d_tasks, int n_tasks, const FsItem *d_items, const int32_t *d_item_insts,
                  uint64_t *d_best, uint32_t *d_counters, size_t lds_bytes, int grid, bool persistent, int kmax,
                  hipStream_t s)
{
    if (n_tasks <= 0)
        return XME_OK;
    int rc = xme_fs_prepare();
    if (rc)
        return rc;
    const void *k = xme_fs_kernel_ptr(persistent, kmax);
    void *args[] = { (void *)&p, (void *)&d_desc, (void *)&d_pus, (void *)&d_inst, (void *)&d_tasks,
                     (void *)&n_tasks, (void *)&d_items, (void *)&d_item_insts, (void *)&d_best, (void *)&d_counters };
    XME_HIP(hipLaunchKernel(k, dim3(persistent ? grid : n_tasks), dim3(256), args, lds_bytes, s));
    return XME_OK;
}

const void *xme_fs_kernel_ptr(bool persistent, int kmax)
{
    if (kmax == XME_NARROW_K)
        return persistent ? reinterpret_cast<const void *>(&fs_kernel<true, XME_NARROW_K>)
                          : reinterpret_cast<const void *>(&fs_kernel<false, XME_NARROW_K>);
    if (persistent)
        return kmax == 8 ? reinterpret_cast<const void *>(&fs_kernel<true, 8>)
                         : reinterpret_cast<const void *>(&fs_kernel<true, 4>);
    return kmax == 8 ? reinterpret_cast<const void *>(&fs_kernel<false, 8>)
                     : reinterpret_cast<const void *>(&fs_kernel<false, 4>);
}
