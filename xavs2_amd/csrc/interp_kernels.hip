// Reference-plane producer for gfx950: border expansion of a plane
// (plane_expand_border, source/common/frame.c:470-507) and the 15 interpolated
// luma planes of a reference picture (interpolate_lcu_row over every LCU row,
// source/common/mc.c:593-785).
//
// Over the whole frame, the row-window algorithm of the reference computes
// every plane on x in [-4, W+4), y in [-4, H+4) from the border-expanded
// full-pel plane and then replicates that area outward, i.e.
//     plane_k(x, y) = interp_k(clamp(x, -4, W+3), clamp(y, -4, H+3)).
// The core kernel computes the [-4, W+4) x [-4, H+4) area tile by tile
// (full-pel tile and its horizontal sums staged in LDS); the border kernel
// then fills the 76-pixel frame around it.  The horizontal sums of 8-bit
// pictures lie in [-4845, 22440], so the reference's int16 temporaries
// (mct_t, mc.c:466-496) hold them without wrapping and int32 is equivalent.
#include <algorithm>

#include "xme_internal.h"

namespace {

// INTPL_FILTERS, source/common/mc.c:73-78
__constant__ int8_t c_taps[4][8] = { { 0, 0, 0, 64, 0, 0, 0, 0 },
                                     { -1, 4, -10, 57, 19, -7, 3, -1 },
                                     { -1, 4, -11, 40, 40, -11, 4, -1 },
                                     { -1, 3, -7, 19, 57, -10, 4, -1 } };

struct Planes16 {
    uint8_t *p[16];
};

typedef short short2v __attribute__((ext_vector_type(2)));

constexpr int TX = 64;   // output columns per tile
constexpr int TY = 16;   // output rows per tile
constexpr int SPW = (TX + 8) / 4;   // staged dwords per row

__device__ __forceinline__ uint32_t clip_pel(int v) { return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

// Four clipped pixels (each 0..255 in an int) packed into one dword, byte i
// from p[i], with v_perm_b32.  Written as shifts and ors, the clips and the
// packing compile to v_ashr_pk_u8_i32; a build of the diagonal planes in that
// form gave wrong pixels (the cause was not isolated), so every plane packs
// this way.
__device__ __forceinline__ uint32_t pack_pels(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3)
{
    return __builtin_amdgcn_perm(__builtin_amdgcn_perm(p3, p2, 0x0C0C0400u), __builtin_amdgcn_perm(p1, p0, 0x0C0C0400u),
                                 0x05040100u);
}

// Taps of a phase packed as two dwords of signed bytes (taps 0-3, 4-7).
__device__ __forceinline__ void packed_taps(int d, int &lo, int &hi)
{
    lo = (int)((uint32_t)(uint8_t)c_taps[d][0] | ((uint32_t)(uint8_t)c_taps[d][1] << 8) |
               ((uint32_t)(uint8_t)c_taps[d][2] << 16) | ((uint32_t)(uint8_t)c_taps[d][3] << 24));
    hi = (int)((uint32_t)(uint8_t)c_taps[d][4] | ((uint32_t)(uint8_t)c_taps[d][5] << 8) |
               ((uint32_t)(uint8_t)c_taps[d][6] << 16) | ((uint32_t)(uint8_t)c_taps[d][7] << 24));
}

// 8-tap filter of 8 pixels held as two dwords of (pixel - 128) signed bytes:
// sum C_k * p_k = sum C_k * (p_k - 128) + 128 * 64, with v_dot4_i32_i8.
__device__ __forceinline__ int fir8(int tlo, int thi, uint32_t p0, uint32_t p1)
{
    return __builtin_amdgcn_sdot4(thi, (int)p1, __builtin_amdgcn_sdot4(tlo, (int)p0, 8192, false), false);
}

// Column j (bytes j of four row dwords) gathered into one dword, rows in order.
__device__ __forceinline__ void transpose4(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t (&col)[4])
{
    const uint32_t t01l = __builtin_amdgcn_perm(a1, a0, 0x05010400u);
    const uint32_t t01h = __builtin_amdgcn_perm(a1, a0, 0x07030602u);
    const uint32_t t23l = __builtin_amdgcn_perm(a3, a2, 0x05010400u);
    const uint32_t t23h = __builtin_amdgcn_perm(a3, a2, 0x07030602u);
    col[0] = __builtin_amdgcn_perm(t23l, t01l, 0x05040100u);
    col[1] = __builtin_amdgcn_perm(t23l, t01l, 0x07060302u);
    col[2] = __builtin_amdgcn_perm(t23h, t01h, 0x05040100u);
    col[3] = __builtin_amdgcn_perm(t23h, t01h, 0x07060302u);
}

// Interpolated planes on the core columns x in [-4, W+4) of rows [ys, ye)
// (the whole core area is ys = -4, ye = H + 4), in tiles whose columns start
// at xs <= -4 (a 64-byte boundary of the output rows); columns left of -4 are
// written too and left to the border pass.  The staged rows below ye + 4 may
// not be final yet (an LCU row still being reconstructed): they only feed
// outputs at or below ye, which are not stored.  The
// full-pel tile (with its 3 + 4 pixel filter margins) is staged in LDS as
// signed bytes (pixel - 128); every 8-tap sum is two v_dot4_i32_i8 with an
// exact 32-bit result, and the horizontal sums feeding the diagonal planes
// are kept in LDS as int32.
__global__ void __launch_bounds__(256) interp_core_kernel(Planes16 pl, int64_t stride, int W, int H, int frac,
                                                          int xs, int ys, int ye, int gx, int n_tiles)
{
    // workgroup b runs on XCD b % 8: each XCD takes a contiguous raster run of
    // tiles, so that the 128-byte lines shared by side-by-side tiles (their
    // staged source rows and the halves of their output rows) meet in one L2
    const int per_xcd = (int)gridDim.x >> 3;
    const int tile = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= n_tiles)
        return;
    const int tx_i = tile % gx, ty_i = tile / gx;
    __shared__ uint32_t sP[TY + 7][SPW];
    __shared__ int32_t sH[3][TY + 7][TX];
    const int x0 = xs + tx_i * TX, y0 = ys + ty_i * TY;
    const uint8_t *P = pl.p[0];
    const int tid = threadIdx.x;
    // stage: column x0 - 4 + 4q .. +3 of rows y0 - 3 .. y0 + TY + 3 (dword aligned
    // since x0 % 4 == 0 and the plane origin and stride are 4-byte aligned)
    for (int e = tid; e < (TY + 7) * SPW; e += 256) {
        const int r = e / SPW, q = e - r * SPW;
        const int y = xme_clip3(-XME_PAD, H + XME_PAD - 1, y0 - 3 + r);
        const int x = x0 - 4 + 4 * q;
        const uint8_t *row = P + (int64_t)y * stride;
        uint32_t v;
        if (x >= -XME_PAD && x + 3 <= W + XME_PAD - 1)
            v = *reinterpret_cast<const uint32_t *>(row + x);
        else
            v = (uint32_t)row[xme_clip3(-XME_PAD, W + XME_PAD - 1, x)] * 0x01010101u;
        sP[r][q] = v ^ 0x80808080u;
    }
    __syncthreads();
    int tl[4], th[4];
#pragma unroll
    for (int d = 1; d <= 3; d++)
        packed_taps(d, tl[d], th[d]);
    // horizontal sums: output column c uses pixels x0 + c - 3 .. x0 + c + 4,
    // i.e. staged bytes c + 1 .. c + 8; a thread does 4 columns of one row
    for (int e = tid; e < (TY + 7) * (TX / 4); e += 256) {
        const int r = e / (TX / 4), c4 = (e - r * (TX / 4)) * 4;
        const uint32_t d0 = sP[r][c4 / 4], d1 = sP[r][c4 / 4 + 1], d2 = sP[r][c4 / 4 + 2], d3 = sP[r][c4 / 4 + 3];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            // bytes c4 + i + 1 .. c4 + i + 8
            const uint32_t sh = (uint32_t)(i + 1);
            const uint32_t lo = sh == 4 ? d1 : __builtin_amdgcn_alignbyte(d1, d0, sh);
            const uint32_t hi = sh == 4 ? d2 : __builtin_amdgcn_alignbyte(d2, d1, sh);
            (void)d3;
#pragma unroll
            for (int dx = 1; dx <= 3; dx++) {
                if (frac < 2 && dx != 2)
                    continue;
                sH[dx - 1][r][c4 + i] = fir8(tl[dx], th[dx], lo, hi);
            }
        }
    }
    __syncthreads();
    const int ty = tid >> 4, c0 = (tid & 15) * 4;
    const int y = y0 + ty, x = x0 + c0;
    const bool active = y < ye && x < W + 4;
    const int64_t off = (int64_t)y * stride + x;
    // horizontal a, b, c
#pragma unroll
    for (int dx = 1; dx <= 3; dx++) {
        if (!active || (frac < 2 && dx != 2))
            continue;
        uint32_t px[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
            px[i] = clip_pel((sH[dx - 1][ty + 3][c0 + i] + 32) >> 6);
        *reinterpret_cast<uint32_t *>(pl.p[dx] + off) = pack_pels(px[0], px[1], px[2], px[3]);
    }
    // vertical d, h, n: pixels of rows y - 3 .. y + 4 at columns x .. x + 3,
    // i.e. staged dword (c0 + 4) / 4 of rows ty .. ty + 7, transposed to columns
    if (active) {
        uint32_t colA[4], colB[4];
        const int q = (c0 + 4) >> 2;
        transpose4(sP[ty][q], sP[ty + 1][q], sP[ty + 2][q], sP[ty + 3][q], colA);
        transpose4(sP[ty + 4][q], sP[ty + 5][q], sP[ty + 6][q], sP[ty + 7][q], colB);
#pragma unroll
        for (int dy = 1; dy <= 3; dy++) {
            if (frac < 2 && dy != 2)
                continue;
            uint32_t px[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
                px[i] = clip_pel((fir8(tl[dy], th[dy], colA[i], colB[i]) + 32) >> 6);
            *reinterpret_cast<uint32_t *>(pl.p[dy << 2] + off) = pack_pels(px[0], px[1], px[2], px[3]);
        }
    }
    // diagonal e..r from the horizontal sums (MC_SHIFT = 12): a thread takes one
    // column c (its lane) and four rows; the vertical 8-tap sums are four
    // v_dot2_i32_i16 on int16 pairs of the horizontal sums, which lie in
    // [-4845, 22440].  The lanes of a quad hold 4 adjacent columns: they
    // transpose their 4 x 4 bytes with DPP broadcasts, and lane i of the quad
    // stores row i of the 4 columns as one dword.
    {
        const int c = tid & 63, g = tid >> 6, i4 = c & 3;
        const int yy = y0 + 4 * g + i4, xx = x0 + (c & ~3);
        const bool st = yy < ye && xx < W + 4;
        const uint32_t sel = (uint32_t)i4 | (uint32_t)(4 + i4) << 8 | 0x0C0C0000u;   // byte i4 of two dwords
        for (int dx = 1; dx <= 3; dx++) {
            if (frac < 2 && dx != 2)
                continue;
            int hv[11];
#pragma unroll
            for (int r = 0; r < 11; r++)
                hv[r] = sH[dx - 1][4 * g + r][c];
            // E[m] = (hv[2m], hv[2m + 1]), O[m] = (hv[2m + 1], hv[2m + 2])
            short2v E[5], O[5];
#pragma unroll
            for (int m = 0; m < 5; m++) {
                E[m] = short2v{ (short)hv[2 * m], (short)hv[2 * m + 1] };
                O[m] = short2v{ (short)hv[2 * m + 1], (short)hv[2 * m + 2] };
            }
            for (int dy = 1; dy <= 3; dy++) {
                if (frac < 2 && dy != 2)
                    continue;
                short2v T[4];
#pragma unroll
                for (int m = 0; m < 4; m++)
                    T[m] = short2v{ (short)c_taps[dy][2 * m], (short)c_taps[dy][2 * m + 1] };
                uint32_t px[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    // rows j .. j + 7: pairs from E for even j, from O for odd j
                    int s = 2048;
#pragma unroll
                    for (int m = 0; m < 4; m++)
                        s = __builtin_amdgcn_sdot2((j & 1) ? O[(j >> 1) + m] : E[(j >> 1) + m], T[m], s, false);
                    px[j] = clip_pel(s >> 12);
                }
                const uint32_t col = pack_pels(px[0], px[1], px[2], px[3]);
                const uint32_t d0 = (uint32_t)__builtin_amdgcn_mov_dpp((int)col, 0x00, 0xF, 0xF, false);
                const uint32_t d1 = (uint32_t)__builtin_amdgcn_mov_dpp((int)col, 0x55, 0xF, 0xF, false);
                const uint32_t d2 = (uint32_t)__builtin_amdgcn_mov_dpp((int)col, 0xAA, 0xF, 0xF, false);
                const uint32_t d3 = (uint32_t)__builtin_amdgcn_mov_dpp((int)col, 0xFF, 0xF, 0xF, false);
                const uint32_t lo = __builtin_amdgcn_perm(d1, d0, sel), hi = __builtin_amdgcn_perm(d3, d2, sel);
                if (st)
                    *reinterpret_cast<uint32_t *>(pl.p[(dy << 2) + dx] + (int64_t)yy * stride + xx) =
                        __builtin_amdgcn_perm(hi, lo, 0x05040100u);
            }
        }
    }
}

// Fill [-pad, W+pad) x [-pad, H+pad) outside the core rectangle [cx0, cx1) x
// [cy0, cy1) with the nearest core value.  The grid covers only the border's
// dwords (core edges and pad are dword aligned), in four regions laid end to
// end: the bands above and below the core (full width; either may be left
// out, as for an LCU row inside the picture), then the strips left and right
// of the core rows.  blockIdx.y selects the plane.
struct BorderGeom {
    int64_t stride;
    int pad, W, H, cx0, cy0, cx1, cy1;
    int wdw;                   // dwords of a full row
    int ldw, rdw;              // dwords of the left and right strips
    int64_t n_top, n_bot, n_side;
};

__global__ void __launch_bounds__(256) border_kernel(Planes16 pl, int first, BorderGeom g)
{
    uint8_t *base = pl.p[first + blockIdx.y];
    if (base == nullptr)
        return;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int x, y;
    if (t < g.n_top) {
        y = -g.pad + (int)(t / g.wdw);
        x = -g.pad + 4 * (int)(t % g.wdw);
    } else if ((t -= g.n_top) < g.n_bot) {
        y = g.cy1 + (int)(t / g.wdw);
        x = -g.pad + 4 * (int)(t % g.wdw);
    } else if ((t -= g.n_bot) < g.n_side) {
        const int sdw = g.ldw + g.rdw, r = (int)(t / sdw), q = (int)(t % sdw);
        y = g.cy0 + r;
        x = q < g.ldw ? -g.pad + 4 * q : g.cx1 + 4 * (q - g.ldw);
    } else {
        return;
    }
    const int sy = xme_clip3(g.cy0, g.cy1 - 1, y);
    const uint8_t *src = base + (int64_t)sy * g.stride;
    uint32_t v;
    if (x >= g.cx0 && x < g.cx1) {
        v = *reinterpret_cast<const uint32_t *>(src + x);
    } else {
        const int sx = x < g.cx0 ? g.cx0 : g.cx1 - 1;
        v = (uint32_t)src[sx] * 0x01010101u;
    }
    *reinterpret_cast<uint32_t *>(base + (int64_t)y * g.stride + x) = v;
}

// The border of `planes` planes from pl.p[first]; every edge and the pad are
// multiples of 4.  top / bottom: fill the band above cy0 / below cy1.
int launch_border(const Planes16 &pl, int first, int planes, int64_t stride, int W, int H, int pad, int cx0,
                  int cy0, int cx1, int cy1, hipStream_t s, bool top = true, bool bottom = true)
{
    BorderGeom g;
    g.stride = stride;
    g.pad = pad;
    g.W = W;
    g.H = H;
    g.cx0 = cx0;
    g.cy0 = cy0;
    g.cx1 = cx1;
    g.cy1 = cy1;
    g.wdw = (W + 2 * pad) >> 2;
    g.ldw = (cx0 + pad) >> 2;
    g.rdw = (W + pad - cx1) >> 2;
    g.n_top = top ? (int64_t)(cy0 + pad) * g.wdw : 0;
    g.n_bot = bottom ? (int64_t)(H + pad - cy1) * g.wdw : 0;
    g.n_side = (int64_t)(cy1 - cy0) * (g.ldw + g.rdw);
    const int64_t n = g.n_top + g.n_bot + g.n_side;
    if (n <= 0)
        return XME_OK;
    hipLaunchKernelGGL(border_kernel, dim3((unsigned)((n + 255) / 256), planes), dim3(256), 0, s, pl, first, g);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

// Planes 1..15 on the core columns of rows [ys, ye), then their border:
// the side strips of those rows, and the bands above / below when top /
// bottom.
int interp_rows(const Planes16 &pl, int64_t stride, int w, int h, int frac, int ys, int ye, bool top, bool bottom,
                hipStream_t s)
{
    // tile columns start where the output planes' rows reach a 64-byte
    // boundary, at or up to 60 columns left of the core's first column -4, so
    // that every 64-column row segment is one whole 64-byte line; the columns
    // left of -4 are overwritten by the border pass
    const int xs = -4 - (int)((reinterpret_cast<uintptr_t>(pl.p[2]) - 4) & 63);
    if (ye > ys) {
        const int gx = (w + 4 - xs + TX - 1) / TX, gy = (ye - ys + TY - 1) / TY;
        const int n_tiles = gx * gy;
        const int grid = (n_tiles + 7) & ~7;     // a multiple of 8: the same tile count per XCD
        hipLaunchKernelGGL(interp_core_kernel, dim3(grid), dim3(256), 0, s, pl, stride, w, h, frac, xs, ys, ye, gx,
                           n_tiles);
        XME_HIP(hipGetLastError());
    }
    return launch_border(pl, 1, 15, stride, w, h, XME_PAD, -4, ys, w + 4, ye, s, top, bottom);
}

// The planes a caller passes: 16 device origins, of which frac 1 needs 0, 2,
// 8 and 10 (frame.c:343-350); false (with the error set) when one is missing
// or unaligned.
bool pick_planes(pel_t *const d_planes[16], int frac, const char *who, Planes16 &pl)
{
    for (int k = 0; k < 16; k++) {
        const bool need = k == 0 || frac == 2 || k == 2 || k == 8 || k == 10;
        pl.p[k] = need ? d_planes[k] : nullptr;
        if (need && (!pl.p[k] || (reinterpret_cast<uintptr_t>(pl.p[k]) & 3))) {
            xme_set_error("%s: plane %d missing or not 4-byte aligned", who, k);
            return false;
        }
    }
    return true;
}

} // namespace

extern "C" int xme_expand_border(pel_t *d_plane, intptr_t stride, int w, int h, int pad, void *stream)
{
    if (!d_plane || w < 1 || h < 1 || (w & 3) || (pad & 3) || pad < 0 || stride < w + 2 * pad ||
        (reinterpret_cast<uintptr_t>(d_plane) & 3) || (stride & 3)) {
        xme_set_error("xme_expand_border: need 4-aligned plane, stride, width and pad");
        return XME_ERR_ARG;
    }
    if (pad == 0)
        return XME_OK;
    Planes16 pl = {};
    pl.p[0] = d_plane;
    return launch_border(pl, 0, 1, (int64_t)stride, w, h, pad, 0, 0, w, h, (hipStream_t)stream);
}

extern "C" int xme_interpolate_frame(pel_t *const d_planes[16], intptr_t stride, int w, int h, int frac, void *stream)
{
    if (!d_planes || !d_planes[0] || w < 8 || h < 8 || (w & 7) || frac < 0 || frac > 2 || (stride & 3) ||
        stride < w + 2 * XME_PAD) {
        xme_set_error("xme_interpolate_frame: need width multiple of 8, stride >= width + %d, multiple of 4",
                      2 * XME_PAD);
        return XME_ERR_ARG;
    }
    if (frac == 0)
        return XME_OK;
    Planes16 pl;
    if (!pick_planes(d_planes, frac, "xme_interpolate_frame", pl))
        return XME_ERR_ARG;
    return interp_rows(pl, (int64_t)stride, w, h, frac, -4, h + 4, true, true, (hipStream_t)stream);
}

extern "C" int xme_expand_border_lcu_row(pel_t *d_plane, intptr_t stride, int w, int h, int pad, int row_log2,
                                         int lcu_y, void *stream)
{
    if (!d_plane || w < 1 || h < 1 || (w & 3) || (pad & 3) || pad < 0 || stride < w + 2 * pad ||
        (reinterpret_cast<uintptr_t>(d_plane) & 3) || (stride & 3) || row_log2 < 2 || row_log2 > 7) {
        xme_set_error("xme_expand_border_lcu_row: need 4-aligned plane, stride, width and pad, row_log2 2..7");
        return XME_ERR_ARG;
    }
    const int n_rows = (h + (1 << row_log2) - 1) >> row_log2;
    if (lcu_y < 0 || lcu_y >= n_rows) {
        xme_set_error("xme_expand_border_lcu_row: LCU row %d outside 0..%d", lcu_y, n_rows - 1);
        return XME_ERR_ARG;
    }
    // xavs2_frame_expand_border_lcurow (frame.c:537-575), one slice: rows
    // [y0, y1), UP_SHIFT 4 rows up except at the picture's top and bottom
    const bool first = lcu_y == 0, last = lcu_y == n_rows - 1;
    const int y0 = (lcu_y << row_log2) - (first ? 0 : 4);
    const int y1 = std::min(h, ((lcu_y + 1) << row_log2) - (last ? 0 : 4));
    if (pad == 0 || y1 <= y0)
        return XME_OK;
    Planes16 pl = {};
    pl.p[0] = d_plane;
    return launch_border(pl, 0, 1, (int64_t)stride, w, h, pad, 0, y0, w, y1, (hipStream_t)stream, first, last);
}

extern "C" int xme_interpolate_lcu_row(pel_t *const d_planes[16], intptr_t stride, int w, int h, int frac,
                                       int lcu_level, int lcu_y, void *stream)
{
    if (!d_planes || !d_planes[0] || w < 8 || h < 8 || (w & 7) || frac < 0 || frac > 2 || (stride & 3) ||
        stride < w + 2 * XME_PAD || lcu_level < 3 || lcu_level > 6) {
        xme_set_error("xme_interpolate_lcu_row: need width multiple of 8, stride >= width + %d, multiple of 4, "
                      "lcu_level 3..6", 2 * XME_PAD);
        return XME_ERR_ARG;
    }
    const int n_rows = (h + (1 << lcu_level) - 1) >> lcu_level;
    if (lcu_y < 0 || lcu_y >= n_rows) {
        xme_set_error("xme_interpolate_lcu_row: LCU row %d outside 0..%d", lcu_y, n_rows - 1);
        return XME_ERR_ARG;
    }
    if (frac == 0)
        return XME_OK;
    Planes16 pl;
    if (!pick_planes(d_planes, frac, "xme_interpolate_lcu_row", pl))
        return XME_ERR_ARG;
    // interpolate_lcu_row (mc.c:748-785), one slice: rows [y0, y1) with
    // PAD_OFFSET 4 and MC_OFFSET 8 (mc.c:95-96)
    const bool first = lcu_y == 0, last = lcu_y == n_rows - 1;
    const int y0 = (lcu_y << lcu_level) - (first ? 4 : 8);
    const int y1 = last ? h + 4 : ((lcu_y + 1) << lcu_level) - 8;
    return interp_rows(pl, (int64_t)stride, w, h, frac, y0, y1, first, last, (hipStream_t)stream);
}
