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

constexpr int TX = 64;   // output columns per tile
constexpr int TY = 16;   // output rows per tile
constexpr int SPW = (TX + 8) / 4;   // staged dwords per row

__device__ __forceinline__ uint32_t clip_pel(int v) { return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

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

// Interpolated planes on the core area x in [-4, W+4), y in [-4, H+4).  The
// full-pel tile (with its 3 + 4 pixel filter margins) is staged in LDS as
// signed bytes (pixel - 128); every 8-tap sum is two v_dot4_i32_i8 with an
// exact 32-bit result, and the horizontal sums feeding the diagonal planes
// are kept in LDS as int32.
__global__ void __launch_bounds__(256) interp_core_kernel(Planes16 pl, int64_t stride, int W, int H, int frac)
{
    __shared__ uint32_t sP[TY + 7][SPW];
    __shared__ int32_t sH[3][TY + 7][TX];
    __shared__ __align__(4) uint8_t sO[9][TY][TX];
    const int x0 = -4 + (int)blockIdx.x * TX, y0 = -4 + (int)blockIdx.y * TY;
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
    const bool active = y < H + 4 && x < W + 4;
    const int64_t off = (int64_t)y * stride + x;
    // horizontal a, b, c
#pragma unroll
    for (int dx = 1; dx <= 3; dx++) {
        if (!active || (frac < 2 && dx != 2))
            continue;
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 4; i++)
            v |= clip_pel((sH[dx - 1][ty + 3][c0 + i] + 32) >> 6) << (8 * i);
        *reinterpret_cast<uint32_t *>(pl.p[dx] + off) = v;
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
            uint32_t v = 0;
#pragma unroll
            for (int i = 0; i < 4; i++)
                v |= clip_pel((fir8(tl[dy], th[dy], colA[i], colB[i]) + 32) >> 6) << (8 * i);
            *reinterpret_cast<uint32_t *>(pl.p[dy << 2] + off) = v;
        }
    }
    // diagonal e..r from the horizontal sums (MC_SHIFT = 12): a thread takes one
    // column and four rows so that every LDS access is a single dword
    {
        const int c = tid & 63, g = tid >> 6;
        for (int dx = 1; dx <= 3; dx++) {
            if (frac < 2 && dx != 2)
                continue;
            int hv[11];
#pragma unroll
            for (int r = 0; r < 11; r++)
                hv[r] = sH[dx - 1][4 * g + r][c];
            for (int dy = 1; dy <= 3; dy++) {
                if (frac < 2 && dy != 2)
                    continue;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int s = 0;
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        s += (int)c_taps[dy][k] * hv[j + k];
                    sO[(dy - 1) * 3 + (dx - 1)][4 * g + j][c] = (uint8_t)clip_pel((s + 2048) >> 12);
                }
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < 9 * TY * (TX / 4); e += 256) {
        const int pk = e / (TY * (TX / 4)), rem = e - pk * (TY * (TX / 4));
        const int r = rem / (TX / 4), q = rem - r * (TX / 4);
        const int dy = pk / 3 + 1, dx = pk % 3 + 1;
        if (frac < 2 && (dx != 2 || dy != 2))
            continue;
        const int yy = y0 + r, xx = x0 + 4 * q;
        if (yy >= H + 4 || xx >= W + 4)
            continue;
        const uint32_t v = *reinterpret_cast<const uint32_t *>(&sO[pk][r][4 * q]);
        *reinterpret_cast<uint32_t *>(pl.p[(dy << 2) + dx] + (int64_t)yy * stride + xx) = v;
    }
}

// Fill [-pad, W+pad) x [-pad, H+pad) outside the core rectangle with the
// nearest core value.  One thread per dword; core edges are dword aligned.
__global__ void __launch_bounds__(256) border_kernel(Planes16 pl, int first, int64_t stride, int W, int H, int pad,
                                                     int cx0, int cy0, int cx1, int cy1)
{
    uint8_t *base = pl.p[first + blockIdx.z];
    if (base == nullptr)
        return;
    const int wdw = (W + 2 * pad) >> 2;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = -pad + (int)blockIdx.y;
    if (q >= wdw)
        return;
    const int x = -pad + 4 * q;
    const bool row_in = y >= cy0 && y < cy1;
    if (row_in && x >= cx0 && x < cx1)
        return;
    const int sy = xme_clip3(cy0, cy1 - 1, y);
    const uint8_t *src = base + (int64_t)sy * stride;
    uint32_t v;
    if (x >= cx0 && x < cx1) {
        v = *reinterpret_cast<const uint32_t *>(src + x);
    } else {
        const int sx = x < cx0 ? cx0 : cx1 - 1;
        v = (uint32_t)src[sx] * 0x01010101u;
    }
    *reinterpret_cast<uint32_t *>(base + (int64_t)y * stride + x) = v;
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
    dim3 grid((((w + 2 * pad) >> 2) + 255) / 256, h + 2 * pad, 1);
    hipLaunchKernelGGL(border_kernel, grid, dim3(256), 0, (hipStream_t)stream, pl, 0, (int64_t)stride, w, h, pad, 0,
                       0, w, h);
    XME_HIP(hipGetLastError());
    return XME_OK;
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
    for (int k = 0; k < 16; k++) {
        const bool need = k == 0 || frac == 2 || k == 2 || k == 8 || k == 10;
        pl.p[k] = need ? d_planes[k] : nullptr;
        if (need && (!pl.p[k] || (reinterpret_cast<uintptr_t>(pl.p[k]) & 3))) {
            xme_set_error("xme_interpolate_frame: plane %d missing or not 4-byte aligned", k);
            return XME_ERR_ARG;
        }
    }
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((w + 8 + TX - 1) / TX, (h + 8 + TY - 1) / TY);
    hipLaunchKernelGGL(interp_core_kernel, grid, dim3(256), 0, s, pl, (int64_t)stride, w, h, frac);
    XME_HIP(hipGetLastError());
    dim3 bgrid((((w + 2 * XME_PAD) >> 2) + 255) / 256, h + 2 * XME_PAD, 15);
    hipLaunchKernelGGL(border_kernel, bgrid, dim3(256), 0, s, pl, 1, (int64_t)stride, w, h, XME_PAD, -4, -4, w + 4,
                       h + 4);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
