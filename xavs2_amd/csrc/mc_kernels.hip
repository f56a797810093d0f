// Motion compensation for gfx950: mc_luma (source/common/mc.c:559-589) from the
// resident interpolated planes, with the block filters intpl_luma_block_*
// (mc.c:342-398) when a plane is absent, and mc_chroma (mc.c:797-829) with the
// 4-tap intpl_chroma_block_* filters (mc.c:275-337).  Block positions follow
// get_mv_for_mc / cu_get_mc_pos (source/common/mc.h:52-75).
//
// One workgroup per block; a thread computes one output pixel.  Pixel FIRs are
// v_dot4_i32_i8 products of the packed taps with (pixel - 128) bytes plus
// 128 * 64 (every AVS2 filter's taps sum to 64), so each sum is an exact 32-bit
// integer; the compiler then has no 16-bit form to narrow the sums into.
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

__global__ void __launch_bounds__(256) mc_chroma_kernel(const uint8_t *__restrict__ pu_, const uint8_t *__restrict__ pv_,
                                                        int64_t i_src, const int32_t *__restrict__ blocks, int n,
                                                        int fw, int fh, uint8_t *__restrict__ du, uint8_t *__restrict__ dv,
                                                        int64_t i_dst, const int64_t *__restrict__ dst_off)
{
    const int b = blockIdx.x;
    if (b >= n)
        return;
    const int32_t *bl = blocks + 6 * b;
    const int lx = bl[0], ly = bl[1], lw = bl[2], lh = bl[3];
    // get_mv_for_mc on the luma block, then chroma 1/8-pel position (rdo.c:1618-1658)
    const int qx = mc_pos(fw, lw, lx, bl[4]), qy = mc_pos(fh, lh, ly, bl[5]);
    const int w = lw >> 1, h = lh >> 1;
    const int px = qx & 7, py = qy & 7;
    const int th = pack_taps(c_chroma[px]);
    const int64_t off = (int64_t)(qy >> 3) * i_src + (qx >> 3);
    for (int t = threadIdx.x; t < 2 * w * h; t += blockDim.x) {
        const int c = t / (w * h), e = t % (w * h), y = e / w, x = e % w;
        const uint8_t *s = (c ? pv_ : pu_) + off + (int64_t)y * i_src + x;
        int v;
        if (px == 0 && py == 0) {
            v = s[0];
        } else if (py == 0) {
            v = clip_pel((fir4(s, 1, th) + 32) >> 6);
        } else if (px == 0) {
            v = clip_pel((fir4(s, i_src, pack_taps(c_chroma[py])) + 32) >> 6);
        } else {
            // intpl_chroma_block_ext_c (mc.c:313-337): 8-bit first pass unshifted
            int acc = 0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                acc += (int)c_chroma[py][k] * fir4(s + (int64_t)(k - 1) * i_src, 1, th);
            v = clip_pel((acc + 2048) >> 12);
        }
        (c ? dv : du)[dst_off[b] + (int64_t)y * i_dst + x] = (uint8_t)v;
    }
}

__global__ void __launch_bounds__(256) mc_luma_kernel(Planes16c pl, int64_t i_ref, const int32_t *__restrict__ blocks,
                                                      int n, int fw, int fh, uint8_t *__restrict__ dst, int64_t i_dst,
                                                      const int64_t *__restrict__ dst_off)
{
    const int b = blockIdx.x;
    if (b >= n)
        return;
    const int32_t *bl = blocks + 6 * b;
    const int lx = bl[0], ly = bl[1], w = bl[2], h = bl[3];
    const int qx = mc_pos(fw, w, lx, bl[4]), qy = mc_pos(fh, h, ly, bl[5]);
    const int x0 = qx >> 2, y0 = qy >> 2, dx = qx & 3, dy = qy & 3;
    const uint8_t *plane = pl.p[(dy << 2) + dx];
    const int hlo = pack_taps(c_luma[dx]), hhi = pack_taps(c_luma[dx] + 4);
    const int vlo = pack_taps(c_luma[dy]), vhi = pack_taps(c_luma[dy] + 4);
    for (int t = threadIdx.x; t < w * h; t += blockDim.x) {
        const int y = t / w, x = t % w;
        int v;
        if (plane != nullptr) {
            v = plane[(int64_t)(y0 + y) * i_ref + x0 + x];
        } else {
            const uint8_t *s = pl.p[0] + (int64_t)(y0 + y) * i_ref + x0 + x;
            if (dy == 0) {
                v = clip_pel((fir8(s, 1, hlo, hhi) + 32) >> 6);
            } else if (dx == 0) {
                v = clip_pel((fir8(s, i_ref, vlo, vhi) + 32) >> 6);
            } else {
                // intpl_luma_block_ext_c (mc.c:362-398): the mct_t (int16)
                // temporaries hold the horizontal sums unchanged, which lie in
                // [-19 * 255, 88 * 255] for 8-bit pixels
                int a = 0;
#pragma unroll
                for (int k = 0; k < 8; k++)
                    a += (int)c_luma[dy][k] * fir8(s + (int64_t)(k - 3) * i_ref, 1, hlo, hhi);
                v = clip_pel((a + 2048) >> 12);
            }
        }
        dst[dst_off[b] + (int64_t)y * i_dst + x] = (uint8_t)v;
    }
}

} // namespace

extern "C" int xme_mc_chroma_batch(const pel_t *d_u, const pel_t *d_v, intptr_t i_src, const int32_t *d_blocks, int n,
                                   int frame_w_luma, int frame_h_luma, pel_t *d_dst_u, pel_t *d_dst_v, intptr_t i_dst,
                                   const int64_t *d_dst_off, void *stream)
{
    if (n < 0 || (n > 0 && (!d_u || !d_v || !d_blocks || !d_dst_u || !d_dst_v || !d_dst_off))) {
        xme_set_error("xme_mc_chroma_batch: bad arguments");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    hipLaunchKernelGGL(mc_chroma_kernel, dim3(n), dim3(128), 0, (hipStream_t)stream, d_u, d_v, (int64_t)i_src,
                       d_blocks, n, frame_w_luma, frame_h_luma, d_dst_u, d_dst_v, (int64_t)i_dst, d_dst_off);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_mc_luma_batch(const xme_ref_desc *ref, const int32_t *d_blocks, int n, int frame_w, int frame_h,
                                 pel_t *d_dst, intptr_t i_dst, const int64_t *d_dst_off, void *stream)
{
    if (!ref || !ref->planes[0] || n < 0 || (n > 0 && (!d_blocks || !d_dst || !d_dst_off))) {
        xme_set_error("xme_mc_luma_batch: bad arguments");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    Planes16c pl;
    for (int k = 0; k < 16; k++)
        pl.p[k] = ref->planes[k];
    hipLaunchKernelGGL(mc_luma_kernel, dim3(n), dim3(256), 0, (hipStream_t)stream, pl, (int64_t)ref->stride, d_blocks,
                       n, frame_w, frame_h, d_dst, (int64_t)i_dst, d_dst_off);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
