// Device-side motion compensation helpers shared by the MC, reconstruction
// and fused residual kernels: the AVS2 interpolation taps (mc.c:73-91),
// cu_get_mc_pos (mc.h:52-66), a block's source in the interpolated planes or
// the full-pel plane with the block filters (mc_luma, mc.c:559-589), and the
// TU prediction of rdo_get_pred_inter (rdo.c:1616-1633).
//
// Pixel FIRs are v_dot4_i32_i8 products of the packed taps with (pixel - 128)
// bytes plus 128 * 64 (every AVS2 filter's taps sum to 64), so each sum is an
// exact 32-bit integer.
#pragma once

#include "xme_internal.h"

namespace {

__constant__ int8_t c_luma[4][8] = { { 0, 0, 0, 64, 0, 0, 0, 0 },
                                     { -1, 4, -10, 57, 19, -7, 3, -1 },
                                     { -1, 4, -11, 40, 40, -11, 4, -1 },
                                     { -1, 3, -7, 19, 57, -10, 4, -1 } };
__constant__ int8_t c_chroma[8][4] = { { 0, 64, 0, 0 }, { -4, 62, 6, 0 }, { -6, 56, 15, -1 }, { -5, 47, 25, -3 },
                                       { -4, 36, 36, -4 }, { -3, 25, 47, -5 }, { -1, 15, 56, -6 }, { 0, 6, 62, -4 } };

__device__ __forceinline__ int clip_pel(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

__device__ __forceinline__ int pack_taps(const int8_t *t)
{
    return (int)((uint32_t)(uint8_t)t[0] | ((uint32_t)(uint8_t)t[1] << 8) | ((uint32_t)(uint8_t)t[2] << 16) |
                 ((uint32_t)(uint8_t)t[3] << 24));
}

// four pixels s[0], s[step], s[2 step], s[3 step] as (pixel - 128) signed bytes
__device__ __forceinline__ int pack_px(const uint8_t *s, int64_t step)
{
    return (int)(((uint32_t)s[0] | ((uint32_t)s[step] << 8) | ((uint32_t)s[2 * step] << 16) |
                  ((uint32_t)s[3 * step] << 24)) ^ 0x80808080u);
}

// sum_k t_k * s[(k - 3) step], k = 0..7 (FLT_8TAP_HOR / FLT_8TAP_VER, mc.c:128-146)
__device__ __forceinline__ int fir8(const uint8_t *s, int64_t step, int tlo, int thi)
{
    return __builtin_amdgcn_sdot4(thi, pack_px(s + step, step),
                                  __builtin_amdgcn_sdot4(tlo, pack_px(s - 3 * step, step), 8192, false), false);
}

// sum_k t_k * s[(k - 1) step], k = 0..3 (FLT_4TAP_HOR / FLT_4TAP_VER)
__device__ __forceinline__ int fir4(const uint8_t *s, int64_t step, int t)
{
    return __builtin_amdgcn_sdot4(t, pack_px(s - step, step), 8192, false);
}

// cu_get_mc_pos (mc.h:52-66); img_size <= 0 disables the clamp.
__device__ __forceinline__ int mc_pos(int img_size, int blk_size, int blk_pos, int mv)
{
    const int imv = mv >> 2, fmv = mv & 7;
    if (img_size > 0) {
        if (blk_pos + imv < -blk_size - 8)
            return (-blk_size - 8) * 4 + fmv;
        if (blk_pos + imv > img_size + 4)
            return (img_size + 4) * 4 + fmv;
    }
    return blk_pos * 4 + mv;
}

struct Planes16c {
    const uint8_t *p[16];
};

// A luma block's source for mc_luma at absolute qpel position (qx, qy):
// the interpolated plane when it exists, else the full-pel plane and the
// block filter's taps.
struct LumaSrc {
    const uint8_t *plane;
    const uint8_t *p0;
    int64_t i_ref;
    int x0, y0, dx, dy, hlo, hhi, vlo, vhi;
};

__device__ __forceinline__ LumaSrc luma_src(const Planes16c &pl, int64_t i_ref, int qx, int qy)
{
    LumaSrc s;
    s.x0 = qx >> 2;
    s.y0 = qy >> 2;
    s.dx = qx & 3;
    s.dy = qy & 3;
    s.i_ref = i_ref;
    s.plane = pl.p[(s.dy << 2) + s.dx];
    s.p0 = pl.p[0];
    s.hlo = pack_taps(c_luma[s.dx]);
    s.hhi = pack_taps(c_luma[s.dx] + 4);
    s.vlo = pack_taps(c_luma[s.dy]);
    s.vhi = pack_taps(c_luma[s.dy] + 4);
    return s;
}

// Pixel (x, y) of the block.
__device__ __forceinline__ int luma_px(const LumaSrc &b, int x, int y)
{
    if (b.plane != nullptr)
        return b.plane[(int64_t)(b.y0 + y) * b.i_ref + b.x0 + x];
    const uint8_t *s = b.p0 + (int64_t)(b.y0 + y) * b.i_ref + b.x0 + x;
    if (b.dy == 0)
        return clip_pel((fir8(s, 1, b.hlo, b.hhi) + 32) >> 6);
    if (b.dx == 0)
        return clip_pel((fir8(s, b.i_ref, b.vlo, b.vhi) + 32) >> 6);
    // intpl_luma_block_ext_c (mc.c:362-398): the mct_t (int16) temporaries hold
    // the horizontal sums unchanged, which lie in [-19 * 255, 88 * 255] for
    // 8-bit pixels
    int a = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        a += (int)c_luma[b.dy][k] * fir8(s + (int64_t)(k - 3) * b.i_ref, 1, b.hlo, b.hhi);
    return clip_pel((a + 2048) >> 12);
}

__device__ __forceinline__ LumaSrc tu_src(const Planes16c &pl, int64_t i_ref, const int32_t *r, int fw, int fh)
{
    const int qx = mc_pos(fw, r[4], r[2], r[6]) + 4 * (r[0] - r[2]);
    const int qy = mc_pos(fh, r[5], r[3], r[7]) + 4 * (r[1] - r[3]);
    return luma_src(pl, i_ref, qx, qy);
}

// The second prediction of a bi-predicted TU: the PU's second MV mv2[2 b],
// mv2[2 b + 1] into the second reference.
__device__ __forceinline__ LumaSrc tu_src2(const Planes16c &pl, int64_t i_ref, const int32_t *r, const int32_t *mv2,
                                           int fw, int fh)
{
    const int qx = mc_pos(fw, r[4], r[2], mv2[0]) + 4 * (r[0] - r[2]);
    const int qy = mc_pos(fh, r[5], r[3], mv2[1]) + 4 * (r[1] - r[3]);
    return luma_src(pl, i_ref, qx, qy);
}

// The TU's prediction pixel: mc_luma, or for BI the rounded average of the
// two (pixf.avg, rdo.c:1624-1631).
template <bool BI>
__device__ __forceinline__ int pred_px(const LumaSrc &a, const LumaSrc &b, int x, int y)
{
    return BI ? (luma_px(a, x, y) + luma_px(b, x, y) + 1) >> 1 : luma_px(a, x, y);
}

// Four prediction pixels (x .. x + 3, y) of the TU, x % 4 == 0, packed in a
// dword: unaligned dword reads of the planes when they exist (the dword after
// the aligned one stays inside the XME_PAD border), else the block filters.
template <bool BI>
__device__ __forceinline__ uint32_t pred4(const LumaSrc &a, const LumaSrc &b, int x, int y)
{
    auto plane4 = [](const LumaSrc &s, int x, int y) {
        const uintptr_t p = (uintptr_t)(s.plane + (int64_t)(s.y0 + y) * s.i_ref + s.x0 + x);
        const uint32_t *q = (const uint32_t *)(p & ~(uintptr_t)3);
        return __builtin_amdgcn_alignbyte(q[1], q[0], (uint32_t)(p & 3));
    };
    if (a.plane != nullptr && (!BI || b.plane != nullptr)) {
        const uint32_t u = plane4(a, x, y);
        if (!BI)
            return u;
        const uint32_t v = plane4(b, x, y);
        return (u | v) - (((u ^ v) >> 1) & 0x7f7f7f7fu);   // (u + v + 1) >> 1 per byte
    }
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        r |= (uint32_t)pred_px<BI>(a, b, x + k, y) << (8 * k);
    return r;
}

// A TU the kernels reconstruct: inside the picture at a column that is a
// multiple of 4, inside its PU, whose sides are 4 to 64 (the clamped
// prediction then stays inside the XME_PAD border).  Others get SSD -1, zero
// levels and no pixels.
template <int N>
__device__ __forceinline__ bool tu_ok(const int32_t *r, int fw, int fh)
{
    return !(r[0] & 3) && r[0] >= 0 && r[1] >= 0 && r[0] + N <= fw && r[1] + N <= fh && r[4] >= 4 && r[4] <= 64 &&
           r[5] >= 4 && r[5] <= 64 && r[0] >= r[2] && r[1] >= r[3] && r[0] + N <= r[2] + r[4] &&
           r[1] + N <= r[3] + r[5];
}

} // namespace
