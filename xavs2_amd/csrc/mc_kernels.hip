// Motion compensation for gfx950: mc_luma (source/common/mc.c:559-589) from the
// resident interpolated planes, with the block filters intpl_luma_block_*
// (mc.c:342-398) when a plane is absent, and mc_chroma (mc.c:797-829) with the
// 4-tap intpl_chroma_block_* filters (mc.c:275-337).  Block positions follow
// get_mv_for_mc / cu_get_mc_pos (source/common/mc.h:52-75).
//
// One workgroup per block; a thread computes one output pixel (the FIR
// helpers are in xme_mc.h).
#include "xme_dpp.h"
#include "xme_mc.h"

namespace {

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
    const LumaSrc src = luma_src(pl, i_ref, mc_pos(fw, w, lx, bl[4]), mc_pos(fh, h, ly, bl[5]));
    for (int t = threadIdx.x; t < w * h; t += blockDim.x) {
        const int y = t / w, x = t % w;
        dst[dst_off[b] + (int64_t)y * i_dst + x] = (uint8_t)luma_px(src, x, y);
    }
}

// ---- luma inter reconstruction of N x N transform blocks ------------------
//
// A record is {tu_x, tu_y, pu_x, pu_y, pu_w, pu_h, mv_x, mv_y}: the TU lies in
// the PU, whose prediction is mc_luma at the PU's clamped position
// (rdo_get_pred_inter, rdo.c:1616-1633, get_mv_for_mc with the PU size).  One
// wave per TU; lanes over the TU's pixels.
constexpr int RC_WAVES = 4;

// Lanes per TU (P, a power of two: one per 4 pixels, at most 64) and TUs per
// wave for N x N TUs.
template <int N>
struct RcShape {
    static constexpr int Q = N * N / 4, P = Q < 64 ? Q : 64, TPW = 64 / P;
};

// The prediction error cur - pred (sub_ps, pixel.c:780-792) into N x N int16
// blocks, contiguous per TU.  A lane takes 4 pixels of a row at a time.
template <int N, bool BI>
__global__ void __launch_bounds__(256) recon_sub_kernel(Planes16c pl, Planes16c pl2, int64_t i_ref,
                                                        const uint8_t *__restrict__ cur, int64_t i_cur,
                                                        const int32_t *__restrict__ tus,
                                                        const int32_t *__restrict__ mv2, int ntu, int fw, int fh,
                                                        int16_t *__restrict__ resid)
{
    using S = RcShape<N>;
    const int lane = threadIdx.x & 63;
    const int b = ((int)blockIdx.x * RC_WAVES + (int)(threadIdx.x >> 6)) * S::TPW + lane / S::P;
    if (b >= ntu)
        return;
    const int32_t *r = tus + 8 * b;
    uint32_t *out = reinterpret_cast<uint32_t *>(resid + (int64_t)b * N * N);
    if (!tu_ok<N>(r, fw, fh)) {
        for (int q = lane % S::P; q < S::Q; q += S::P) {
            out[2 * q] = 0u;
            out[2 * q + 1] = 0u;
        }
        return;
    }
    const LumaSrc src = tu_src(pl, i_ref, r, fw, fh);
    const LumaSrc src2 = BI ? tu_src2(pl2, i_ref, r, mv2 + 2 * b, fw, fh) : src;
    const uint8_t *c = cur + (int64_t)r[1] * i_cur + r[0];
    for (int q = lane % S::P; q < S::Q; q += S::P) {
        const int y = q / (N / 4), x = (q % (N / 4)) * 4;
        const uint32_t cv = *reinterpret_cast<const uint32_t *>(c + (int64_t)y * i_cur + x);
        const uint32_t pv = pred4<BI>(src, src2, x, y);
        int d[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            d[k] = (int)((cv >> (8 * k)) & 0xff) - (int)((pv >> (8 * k)) & 0xff);
        out[(y * N + x) / 2] = ((uint32_t)d[0] & 0xffffu) | ((uint32_t)d[1] << 16);
        out[(y * N + x) / 2 + 1] = ((uint32_t)d[2] & 0xffffu) | ((uint32_t)d[3] << 16);
    }
}

// The reconstruction clip(pred + residual) (add_ps, pixel.c:794-806) into the
// picture, and its SSD against cur (pixf.ssd, rdo.c:945), summed over the
// TU's P lanes.
template <int N, bool BI>
__global__ void __launch_bounds__(256) recon_add_kernel(Planes16c pl, Planes16c pl2, int64_t i_ref,
                                                        const uint8_t *__restrict__ cur, int64_t i_cur,
                                                        const int32_t *__restrict__ tus,
                                                        const int32_t *__restrict__ mv2, int ntu, int fw, int fh,
                                                        const int16_t *__restrict__ rres, uint8_t *__restrict__ rec,
                                                        int64_t i_rec, int32_t *__restrict__ ssd)
{
    using S = RcShape<N>;
    const int lane = threadIdx.x & 63;
    const int b = ((int)blockIdx.x * RC_WAVES + (int)(threadIdx.x >> 6)) * S::TPW + lane / S::P;
    int e = 0;
    const bool ok = b < ntu && tu_ok<N>(tus + 8 * b, fw, fh);
    if (ok) {
        const int32_t *r = tus + 8 * b;
        const LumaSrc src = tu_src(pl, i_ref, r, fw, fh);
        const LumaSrc src2 = BI ? tu_src2(pl2, i_ref, r, mv2 + 2 * b, fw, fh) : src;
        const uint8_t *c = cur + (int64_t)r[1] * i_cur + r[0];
        uint8_t *d = rec + (int64_t)r[1] * i_rec + r[0];
        const uint32_t *rr = reinterpret_cast<const uint32_t *>(rres + (int64_t)b * N * N);
        for (int q = lane % S::P; q < S::Q; q += S::P) {
            const int y = q / (N / 4), x = (q % (N / 4)) * 4;
            const uint32_t pv = pred4<BI>(src, src2, x, y);
            const uint32_t cv = *reinterpret_cast<const uint32_t *>(c + (int64_t)y * i_cur + x);
            const uint32_t r0 = rr[(y * N + x) / 2], r1 = rr[(y * N + x) / 2 + 1];
            const int res[4] = { (int)(int16_t)(r0 & 0xffffu), (int)(int16_t)(r0 >> 16), (int)(int16_t)(r1 & 0xffffu),
                                 (int)(int16_t)(r1 >> 16) };
            uint32_t o = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int v = clip_pel((int)((pv >> (8 * k)) & 0xff) + res[k]);
                o |= (uint32_t)v << (8 * k);
                const int diff = (int)((cv >> (8 * k)) & 0xff) - v;
                e += diff * diff;
            }
            *reinterpret_cast<uint32_t *>(d + (int64_t)y * i_rec + x) = o;
        }
    }
    e = group_sum(e, S::P);
    if (b < ntu && lane % S::P == S::P - 1)
        ssd[b] = ok ? e : -1;
}

template <int N, bool BI>
int launch_recon(int qp, int half, const uint8_t *cur, int64_t i_cur, const Planes16c &pl, const Planes16c &pl2,
                 int64_t i_ref, int fw, int fh, const int32_t *tus, const int32_t *mv2, int ntu, int16_t *resid,
                 int16_t *level, int16_t *rres, uint8_t *rec, int64_t i_rec, int32_t *nnz, int32_t *ssd, hipStream_t s)
{
    // one wave per group of TPW TUs
    const int waves = (ntu + RcShape<N>::TPW - 1) / RcShape<N>::TPW;
    const int grid = (waves + RC_WAVES - 1) / RC_WAVES;
    hipLaunchKernelGGL((recon_sub_kernel<N, BI>), dim3(grid), dim3(64 * RC_WAVES), 0, s, pl, pl2, i_ref, cur, i_cur,
                       tus, mv2, ntu, fw, fh, resid);
    XME_HIP(hipGetLastError());
    const int rc = xme_residual_run(N, qp, 0, half, resid, level, rres, nnz, ntu, s);
    if (rc != XME_OK)
        return rc;
    hipLaunchKernelGGL((recon_add_kernel<N, BI>), dim3(grid), dim3(64 * RC_WAVES), 0, s, pl, pl2, i_ref, cur, i_cur,
                       tus, mv2, ntu, fw, fh, rres, rec, i_rec, ssd);
    XME_HIP(hipGetLastError());
    return XME_OK;
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

namespace {

template <bool BI>
int recon_entry(const char *name, int n_arg, int qp, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref,
                const xme_ref_desc *ref2, int frame_w, int frame_h, const int32_t *d_tus, const int32_t *d_mv2,
                int ntu, coeff_t *d_resid, coeff_t *d_level, coeff_t *d_rres, pel_t *d_rec, intptr_t i_rec,
                int32_t *d_nnz, int32_t *d_ssd, hipStream_t s)
{
    const int n = n_arg & ~XME_TU_DCT_HALF, half = (n_arg & XME_TU_DCT_HALF) ? 1 : 0;
    if (!ref || !ref->planes[0] || (BI && (!ref2 || !ref2->planes[0] || ref2->stride != ref->stride)) || ntu < 0 ||
        (n != 4 && n != 8 && n != 16 && n != 32 && n != 64) || (half && n < 32) || qp < 0 || qp > 79 ||
        frame_w <= 0 || frame_h <= 0) {
        xme_set_error("%s: bad arguments (n 4, 8, 16, 32 or 64; XME_TU_DCT_HALF with 32 or 64 only)", name);
        return XME_ERR_ARG;
    }
    if (ntu == 0)
        return XME_OK;
    // with every interpolated plane resident the whole path is one launch;
    // otherwise the block filters predict and the stages run apart, through
    // d_resid and d_rres
    bool planes = true;
    for (int k = 1; k < 16; k++)
        planes = planes && ref->planes[k] != nullptr && (!BI || ref2->planes[k] != nullptr);
    if (!d_cur || !d_tus || (BI && !d_mv2) || !d_level || !d_rec || !d_nnz || !d_ssd ||
        (!planes && (!d_resid || !d_rres))) {
        xme_set_error("%s: NULL buffer", name);
        return XME_ERR_ARG;
    }
    if (n == 64 && !planes) {
        xme_set_error("%s: 64x64 TUs need every interpolated plane of the reference", name);
        return XME_ERR_ARG;
    }
    // four pixels per access: the current and reconstructed pictures are read
    // and written as aligned dwords
    if (((uintptr_t)d_cur & 3) || (i_cur & 3) || ((uintptr_t)d_rec & 3) || (i_rec & 3) || ((uintptr_t)d_resid & 7) ||
        ((uintptr_t)d_rres & 7)) {
        xme_set_error("%s: pictures must be 4-byte and coefficient buffers 8-byte aligned", name);
        return XME_ERR_ARG;
    }
    Planes16c pl, pl2;
    for (int k = 0; k < 16; k++) {
        pl.p[k] = ref->planes[k];
        pl2.p[k] = BI ? ref2->planes[k] : nullptr;
    }
    const int64_t ir = (int64_t)ref->stride;
    if (planes)
        return xme_recon_fused(n, qp, half, BI ? 1 : 0, pl.p, pl2.p, ir, d_cur, i_cur, d_rec, i_rec, d_tus, d_mv2, ntu,
                               frame_w, frame_h, d_resid, d_level, d_rres, d_nnz, d_ssd, s);
    switch (n) {
    case 4:
        return launch_recon<4, BI>(qp, half, d_cur, i_cur, pl, pl2, ir, frame_w, frame_h, d_tus, d_mv2, ntu, d_resid,
                                   d_level, d_rres, d_rec, i_rec, d_nnz, d_ssd, s);
    case 8:
        return launch_recon<8, BI>(qp, half, d_cur, i_cur, pl, pl2, ir, frame_w, frame_h, d_tus, d_mv2, ntu, d_resid,
                                   d_level, d_rres, d_rec, i_rec, d_nnz, d_ssd, s);
    case 16:
        return launch_recon<16, BI>(qp, half, d_cur, i_cur, pl, pl2, ir, frame_w, frame_h, d_tus, d_mv2, ntu, d_resid,
                                    d_level, d_rres, d_rec, i_rec, d_nnz, d_ssd, s);
    default:
        return launch_recon<32, BI>(qp, half, d_cur, i_cur, pl, pl2, ir, frame_w, frame_h, d_tus, d_mv2, ntu, d_resid,
                                    d_level, d_rres, d_rec, i_rec, d_nnz, d_ssd, s);
    }
}

} // namespace

extern "C" int xme_inter_recon_batch(int n, int qp, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref,
                                     int frame_w, int frame_h, const int32_t *d_tus, int ntu, coeff_t *d_resid,
                                     coeff_t *d_level, coeff_t *d_rres, pel_t *d_rec, intptr_t i_rec, int32_t *d_nnz,
                                     int32_t *d_ssd, void *stream)
{
    return recon_entry<false>("xme_inter_recon_batch", n, qp, d_cur, i_cur, ref, nullptr, frame_w, frame_h, d_tus,
                              nullptr, ntu, d_resid, d_level, d_rres, d_rec, i_rec, d_nnz, d_ssd, (hipStream_t)stream);
}

extern "C" int xme_inter_recon_bi_batch(int n, int qp, const pel_t *d_cur, intptr_t i_cur, const xme_ref_desc *ref1,
                                        const xme_ref_desc *ref2, int frame_w, int frame_h, const int32_t *d_tus,
                                        const int32_t *d_mv2, int ntu, coeff_t *d_resid, coeff_t *d_level,
                                        coeff_t *d_rres, pel_t *d_rec, intptr_t i_rec, int32_t *d_nnz,
                                        int32_t *d_ssd, void *stream)
{
    return recon_entry<true>("xme_inter_recon_bi_batch", n, qp, d_cur, i_cur, ref1, ref2, frame_w, frame_h, d_tus,
                             d_mv2, ntu, d_resid, d_level, d_rres, d_rec, i_rec, d_nnz, d_ssd, (hipStream_t)stream);
}
