// Motion compensation for gfx950: mc_luma (source/common/mc.c:559-589) from the
// resident interpolated planes, with the block filters intpl_luma_block_*
// (mc.c:342-398) when a plane is absent, and mc_chroma (mc.c:797-829) with the
// 4-tap intpl_chroma_block_* filters (mc.c:275-337).  Block positions follow
// get_mv_for_mc / cu_get_mc_pos (source/common/mc.h:52-75).
//
// One workgroup per block; a thread computes one output pixel.  Every sum is a
// plain 32-bit integer expression of the reference's taps and roundings.
#include "xme_internal.h"

namespace {

__constant__ int8_t c_luma[4][8] = { { 0, 0, 0, 64, 0, 0, 0, 0 },
                                     { -1, 4, -10, 57, 19, -7, 3, -1 },
                                     { -1, 4, -11, 40, 40, -11, 4, -1 },
                                     { -1, 3, -7, 19, 57, -10, 4, -1 } };
__constant__ int8_t c_chroma[8][4] = { { 0, 64, 0, 0 }, { -4, 62, 6, 0 }, { -6, 56, 15, -1 }, { -5, 47, 25, -3 },
                                       { -4, 36, 36, -4 }, { -3, 25, 47, -5 }, { -1, 15, 56, -6 }, { 0, 6, 62, -4 } };

__device__ __forceinline__ int clip_pel(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

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
    const int64_t off = (int64_t)(qy >> 3) * i_src + (qx >> 3);
    for (int t = threadIdx.x; t < 2 * w * h; t += blockDim.x) {
        const int c = t / (w * h), e = t % (w * h), y = e / w, x = e % w;
        const uint8_t *s = (c ? pv_ : pu_) + off + (int64_t)y * i_src + x;
        int v;
        if (px == 0 && py == 0) {
            v = s[0];
        } else if (py == 0) {
            v = clip_pel((c_chroma[px][0] * s[-1] + c_chroma[px][1] * s[0] + c_chroma[px][2] * s[1] +
                          c_chroma[px][3] * s[2] + 32) >> 6);
        } else if (px == 0) {
            v = clip_pel((c_chroma[py][0] * s[-i_src] + c_chroma[py][1] * s[0] + c_chroma[py][2] * s[i_src] +
                          c_chroma[py][3] * s[2 * i_src] + 32) >> 6);
        } else {
            int acc = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint8_t *r = s + (int64_t)(k - 1) * i_src;
                const int hs = c_chroma[px][0] * r[-1] + c_chroma[px][1] * r[0] + c_chroma[px][2] * r[1] +
                               c_chroma[px][3] * r[2];
                acc += c_chroma[py][k] * hs;
            }
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
    for (int t = threadIdx.x; t < w * h; t += blockDim.x) {
        const int y = t / w, x = t % w;
        int v;
        if (plane != nullptr) {
            v = plane[(int64_t)(y0 + y) * i_ref + x0 + x];
        } else {
            const uint8_t *s = pl.p[0] + (int64_t)(y0 + y) * i_ref + x0 + x;
            if (dy == 0) {
                int a = 0;
#pragma unroll
                for (int k = 0; k < 8; k++)
                    a += c_luma[dx][k] * s[k - 3];
                v = clip_pel((a + 32) >> 6);
            } else if (dx == 0) {
                int a = 0;
#pragma unroll
                for (int k = 0; k < 8; k++)
                    a += c_luma[dy][k] * s[(int64_t)(k - 3) * i_ref];
                v = clip_pel((a + 32) >> 6);
            } else {
                int a = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const uint8_t *r = s + (int64_t)(k - 3) * i_ref;
                    int hsum = 0;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        hsum += c_luma[dx][j] * r[j - 3];
                    a += c_luma[dy][k] * (int)(int16_t)hsum;     // mct_t temporaries (mc.c:373-378)
                }
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
