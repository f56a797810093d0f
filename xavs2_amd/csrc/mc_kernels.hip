// Motion compensation for gfx950: mc_luma (source/common/mc.c:559-589) from the
// resident interpolated planes, with the block filters intpl_luma_block_*
// (mc.c:342-398) when a plane is absent, and mc_chroma (mc.c:797-829) with the
// 4-tap intpl_chroma_block_* filters (mc.c:275-337).  Block positions follow
// get_mv_for_mc / cu_get_mc_pos (source/common/mc.h:52-75).
//
// A wave per block, four pixels per lane (the FIR helpers are in xme_mc.h).
#include "xme_dpp.h"
#include "xme_mc.h"

namespace {

// One chroma pixel of mc_chroma at s: the 4-tap filters of the block's phase
// (intpl_chroma_block_hor / _ver / _ext, mc.c:275-337; the 8-bit first pass
// of _ext unshifted).
__device__ __forceinline__ int chroma_px(const uint8_t *s, int64_t i_src, int px, int py, int th, int tv)
{
    if (px == 0 && py == 0)
        return s[0];
    if (py == 0)
        return clip_pel((fir4(s, 1, th) + 32) >> 6);
    if (px == 0)
        return clip_pel((fir4(s, i_src, tv) + 32) >> 6);
    int acc = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        acc += (int)c_chroma[py][k] * fir4(s + (int64_t)(k - 1) * i_src, 1, th);
    return clip_pel((acc + 2048) >> 12);
}

// Bytes p[0..3] and p[4..7] at any alignment, from the three aligned dwords
// around them (inside the plane's border).
__device__ __forceinline__ void load8u(const uint8_t *p, uint32_t &lo, uint32_t &hi)
{
    const uintptr_t a = (uintptr_t)p;
    const uint32_t *q = (const uint32_t *)(a & ~(uintptr_t)3);
    const uint32_t w0 = q[0], w1 = q[1], w2 = q[2], sh = (uint32_t)(a & 3);
    lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
    hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
}

// The horizontal 4-tap sums (taps th, unshifted) of pixels s[0..3] of a row:
// output j takes s[j - 1 .. j + 2], bytes j..j + 3 of s[-1 .. 6].
__device__ __forceinline__ void hsum4(const uint8_t *s, int th, int (&o)[4])
{
    uint32_t lo, hi;
    load8u(s - 1, lo, hi);
#pragma unroll
    for (int j = 0; j < 4; j++)
        o[j] = __builtin_amdgcn_sdot4(th, (int)(__builtin_amdgcn_alignbyte(hi, lo, (uint32_t)j) ^ 0x80808080u), 8192,
                                      false);
}

// Four pixels (each 0..255) as the bytes of a dword, by v_perm: left to
// shifts and ORs, the compiler packs pairs with v_ashr_pk_u8_i32, whose
// upper half is not zero, and the OR of the other pair then corrupts it.
__device__ __forceinline__ uint32_t pack4_pels(const int (&v)[4])
{
    return __builtin_amdgcn_perm(__builtin_amdgcn_perm((uint32_t)v[3], (uint32_t)v[2], 0x0C0C0400u),
                                 __builtin_amdgcn_perm((uint32_t)v[1], (uint32_t)v[0], 0x0C0C0400u), 0x05040100u);
}

// Pixels s[0..3] of mc_chroma packed in a dword (chroma_px of each).
__device__ __forceinline__ uint32_t chroma_px4(const uint8_t *s, int64_t i_src, int px, int py, int th, int tv)
{
    int o[4];
    if (px == 0 && py == 0) {
        uint32_t lo, hi;
        load8u(s, lo, hi);
        return lo;
    }
    if (py == 0) {
        int h[4];
        hsum4(s, th, h);
#pragma unroll
        for (int j = 0; j < 4; j++)
            o[j] = clip_pel((h[j] + 32) >> 6);
        return pack4_pels(o);
    }
    if (px == 0) {
        uint32_t r[4], hi;
#pragma unroll
        for (int k = 0; k < 4; k++)
            load8u(s + (int64_t)(k - 1) * i_src, r[k], hi);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t col = ((r[0] >> (8 * j)) & 0xffu) | (((r[1] >> (8 * j)) & 0xffu) << 8) |
                                 (((r[2] >> (8 * j)) & 0xffu) << 16) | (((r[3] >> (8 * j)) & 0xffu) << 24);
            const int v = __builtin_amdgcn_sdot4(tv, (int)(col ^ 0x80808080u), 8192, false);
            o[j] = clip_pel((v + 32) >> 6);
        }
        return pack4_pels(o);
    }
    int acc[4] = { 0, 0, 0, 0 };
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int h[4];
        hsum4(s + (int64_t)(k - 1) * i_src, th, h);
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[j] += (int)c_chroma[py][k] * h[j];
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
        o[j] = clip_pel((acc[j] + 2048) >> 12);
    return pack4_pels(o);
}

// One wave per block, four blocks per workgroup.  Workgroup b runs on XCD
// b % 8 and takes blocks of one contiguous range per XCD (the grid is a
// multiple of 8), so that the rows of neighbouring blocks, which share
// destination lines, are written through one L2.  A lane writes four pixels
// of a row as one dword where the block's width and destination allow it.
// BI: the rounded average (a + b + 1) >> 1 of the predictions of the two
// references (rdo_get_pred_inter, rdo.c:1656-1666: pixf.avg or
// pixf.average, pixel.c:844-856), block records {x, y, w, h, mv1, mv2}.
struct ChromaSrc {
    const uint8_t *u, *v;
    int64_t off;
    int px, py, th, tv;
};

__device__ __forceinline__ ChromaSrc chroma_src(const uint8_t *pu_, const uint8_t *pv_, int64_t i_src, int fw, int fh,
                                                int lx, int ly, int lw, int lh, int mvx, int mvy)
{
    // get_mv_for_mc on the luma block, then chroma 1/8-pel position (rdo.c:1641-1655)
    const int qx = mc_pos(fw, lw, lx, mvx), qy = mc_pos(fh, lh, ly, mvy);
    ChromaSrc c;
    c.u = pu_;
    c.v = pv_;
    c.px = qx & 7;
    c.py = qy & 7;
    c.th = pack_taps(c_chroma[c.px]);
    c.tv = pack_taps(c_chroma[c.py]);
    c.off = (int64_t)(qy >> 3) * i_src + (qx >> 3);
    return c;
}

template <bool BI>
__global__ void __launch_bounds__(256) mc_chroma_kernel(const uint8_t *__restrict__ pu_, const uint8_t *__restrict__ pv_,
                                                        const uint8_t *__restrict__ pu2, const uint8_t *__restrict__ pv2,
                                                        int64_t i_src, const int32_t *__restrict__ blocks, int n,
                                                        int fw, int fh, uint8_t *__restrict__ du, uint8_t *__restrict__ dv,
                                                        int64_t i_dst, const int64_t *__restrict__ dst_off)
{
    const int lane = threadIdx.x & 63;
    const int lb = (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
    const int b = lb * 4 + (int)(threadIdx.x >> 6);
    if (b >= n)
        return;
    const int32_t *bl = blocks + (BI ? 8 : 6) * b;
    const int lx = bl[0], ly = bl[1], lw = bl[2], lh = bl[3];
    const ChromaSrc c1 = chroma_src(pu_, pv_, i_src, fw, fh, lx, ly, lw, lh, bl[4], bl[5]);
    const ChromaSrc c2 = BI ? chroma_src(pu2, pv2, i_src, fw, fh, lx, ly, lw, lh, bl[6], bl[7]) : c1;
    const int w = lw >> 1, h = lh >> 1;
    uint8_t *ou = du + dst_off[b], *ov = dv + dst_off[b];
    const bool wide = (w & 3) == 0 && ((((uintptr_t)ou | (uintptr_t)ov) & 3) | (uintptr_t)(i_dst & 3)) == 0;
    if (wide) {
        const int q = w >> 2, per = h * q;
        for (int t = lane; t < 2 * per; t += 64) {
            const int c = t >= per, e = c ? t - per : t, y = e / q, x = 4 * (e - y * q);
            const int64_t so = (int64_t)y * i_src + x;
            uint32_t u = chroma_px4((c ? c1.v : c1.u) + c1.off + so, i_src, c1.px, c1.py, c1.th, c1.tv);
            if constexpr (BI) {
                const uint32_t v2 = chroma_px4((c ? c2.v : c2.u) + c2.off + so, i_src, c2.px, c2.py, c2.th, c2.tv);
                u = (u | v2) - (((u ^ v2) >> 1) & 0x7f7f7f7fu);   // (u + v + 1) >> 1 per byte
            }
            *reinterpret_cast<uint32_t *>((c ? ov : ou) + (int64_t)y * i_dst + x) = u;
        }
    } else {
        for (int t = lane; t < 2 * w * h; t += 64) {
            const int c = t / (w * h), e = t % (w * h), y = e / w, x = e % w;
            const int64_t so = (int64_t)y * i_src + x;
            int v = chroma_px((c ? c1.v : c1.u) + c1.off + so, i_src, c1.px, c1.py, c1.th, c1.tv);
            if constexpr (BI)
                v = (v + chroma_px((c ? c2.v : c2.u) + c2.off + so, i_src, c2.px, c2.py, c2.th, c2.tv) + 1) >> 1;
            (c ? ov : ou)[(int64_t)y * i_dst + x] = (uint8_t)v;
        }
    }
}

// mc_luma: a wave per block, blocks in one contiguous range per XCD as in
// mc_chroma_kernel; a lane writes four pixels of a row as one dword (an
// unaligned dword of the interpolated plane, or four block-filter pixels)
// where the block's width and destination allow it.
template <bool BI>
__global__ void __launch_bounds__(256) mc_luma_kernel(Planes16c pl, Planes16c pl2, int64_t i_ref,
                                                      const int32_t *__restrict__ blocks, int n, int fw, int fh,
                                                      uint8_t *__restrict__ dst, int64_t i_dst,
                                                      const int64_t *__restrict__ dst_off)
{
    const int lane = threadIdx.x & 63;
    const int lb = (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
    const int b = lb * 4 + (int)(threadIdx.x >> 6);
    if (b >= n)
        return;
    const int32_t *bl = blocks + (BI ? 8 : 6) * b;
    const int lx = bl[0], ly = bl[1], w = bl[2], h = bl[3];
    const LumaSrc src = luma_src(pl, i_ref, mc_pos(fw, w, lx, bl[4]), mc_pos(fh, h, ly, bl[5]));
    const LumaSrc src2 = BI ? luma_src(pl2, i_ref, mc_pos(fw, w, lx, bl[6]), mc_pos(fh, h, ly, bl[7])) : src;
    uint8_t *o = dst + dst_off[b];
    if ((w & 3) == 0 && (((uintptr_t)o | (uintptr_t)i_dst) & 3) == 0) {
        const int q = w >> 2;
        for (int t = lane; t < h * q; t += 64) {
            const int y = t / q, x = 4 * (t - y * q);
            uint32_t v;
            if (src.plane != nullptr && (!BI || src2.plane != nullptr)) {
                v = pred4<BI>(src, src2, x, y);
            } else {
                const int px[4] = { pred_px<BI>(src, src2, x, y), pred_px<BI>(src, src2, x + 1, y),
                                    pred_px<BI>(src, src2, x + 2, y), pred_px<BI>(src, src2, x + 3, y) };
                v = pack4_pels(px);
            }
            *reinterpret_cast<uint32_t *>(o + (int64_t)y * i_dst + x) = v;
        }
    } else {
        for (int t = lane; t < w * h; t += 64) {
            const int y = t / w, x = t % w;
            o[(int64_t)y * i_dst + x] = (uint8_t)pred_px<BI>(src, src2, x, y);
        }
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
    const int grid = ((n + 3) / 4 + 7) / 8 * 8;
    hipLaunchKernelGGL(mc_chroma_kernel<false>, dim3(grid), dim3(256), 0, (hipStream_t)stream, d_u, d_v, d_u, d_v,
                       (int64_t)i_src, d_blocks, n, frame_w_luma, frame_h_luma, d_dst_u, d_dst_v, (int64_t)i_dst,
                       d_dst_off);
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
    const int grid = ((n + 3) / 4 + 7) / 8 * 8;
    hipLaunchKernelGGL(mc_luma_kernel<false>, dim3(grid), dim3(256), 0, (hipStream_t)stream, pl, pl,
                       (int64_t)ref->stride, d_blocks, n, frame_w, frame_h, d_dst, (int64_t)i_dst, d_dst_off);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_mc_luma_bi_batch(const xme_ref_desc *ref1, const xme_ref_desc *ref2, const int32_t *d_blocks, int n,
                                    int frame_w, int frame_h, pel_t *d_dst, intptr_t i_dst, const int64_t *d_dst_off,
                                    void *stream)
{
    if (!ref1 || !ref1->planes[0] || !ref2 || !ref2->planes[0] || ref1->stride != ref2->stride || n < 0 ||
        (n > 0 && (!d_blocks || !d_dst || !d_dst_off))) {
        xme_set_error("xme_mc_luma_bi_batch: bad arguments (two references of one stride)");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    Planes16c pl, pl2;
    for (int k = 0; k < 16; k++) {
        pl.p[k] = ref1->planes[k];
        pl2.p[k] = ref2->planes[k];
    }
    const int grid = ((n + 3) / 4 + 7) / 8 * 8;
    hipLaunchKernelGGL(mc_luma_kernel<true>, dim3(grid), dim3(256), 0, (hipStream_t)stream, pl, pl2,
                       (int64_t)ref1->stride, d_blocks, n, frame_w, frame_h, d_dst, (int64_t)i_dst, d_dst_off);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_mc_chroma_bi_batch(const pel_t *d_u1, const pel_t *d_v1, const pel_t *d_u2, const pel_t *d_v2,
                                      intptr_t i_src, const int32_t *d_blocks, int n, int frame_w_luma,
                                      int frame_h_luma, pel_t *d_dst_u, pel_t *d_dst_v, intptr_t i_dst,
                                      const int64_t *d_dst_off, void *stream)
{
    if (n < 0 ||
        (n > 0 && (!d_u1 || !d_v1 || !d_u2 || !d_v2 || !d_blocks || !d_dst_u || !d_dst_v || !d_dst_off))) {
        xme_set_error("xme_mc_chroma_bi_batch: bad arguments");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    const int grid = ((n + 3) / 4 + 7) / 8 * 8;
    hipLaunchKernelGGL(mc_chroma_kernel<true>, dim3(grid), dim3(256), 0, (hipStream_t)stream, d_u1, d_v1, d_u2, d_v2,
                       (int64_t)i_src, d_blocks, n, frame_w_luma, frame_h_luma, d_dst_u, d_dst_v, (int64_t)i_dst,
                       d_dst_off);
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
