// Per-call primitives with the exact g_funcs signatures (pixel_cmp_t,
// pixel_cmp_x3_t, pixel_cmp_x4_t; source/common/pixel.h:124-128), for
// installing into an xavs2 intrinsic_func_t after xavs2_init_all_primitives.
//
// A call copies its blocks (host memory, caller strides) into a pinned staging
// buffer of the calling thread, runs the batched GPU kernel with one call on
// the thread's own stream and returns the GPU's result.  Calls are reentrant:
// every thread owns its stream and buffers.  A HIP failure aborts the process
// with a message (the table signatures have no error channel).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "xme_internal.h"
#include "tune.h"

namespace {

struct CallCtx {
    hipStream_t s = nullptr;
    uint8_t *h_buf = nullptr;       // pinned: blocks, then results
    uint8_t *d_buf = nullptr;
    int64_t *d_off = nullptr;       // {0, W*H, 2*W*H, ...}
    int32_t *d_out = nullptr;
    int last_wh = -1;
    ~CallCtx()
    {
        if (s) {
            (void)hipStreamDestroy(s);
            (void)hipHostFree(h_buf);
            (void)hipFree(d_buf);
            (void)hipFree(d_off);
            (void)hipFree(d_out);
        }
    }
};

thread_local CallCtx g_ctx;
constexpr size_t kBufBytes = 5 * 64 * 64 + 64;

[[noreturn]] void die(const char *what, hipError_t e)
{
    fprintf(stderr, "xavs2_amd primitive call: %s: %s\n", what, hipGetErrorString(e));
    abort();
}

#define CK(call)                         \
    do {                                 \
        hipError_t e_ = (call);          \
        if (e_ != hipSuccess)            \
            die(#call, e_);              \
    } while (0)

CallCtx &ctx(int wh)
{
    CallCtx &c = g_ctx;
    if (!c.s) {
        CK(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));
        CK(hipHostMalloc((void **)&c.h_buf, kBufBytes + 64, hipHostMallocDefault));
        CK(hipMalloc((void **)&c.d_buf, kBufBytes));
        CK(hipMalloc((void **)&c.d_off, 8 * sizeof(int64_t)));
        CK(hipMalloc((void **)&c.d_out, 8 * sizeof(int32_t)));
    }
    if (c.last_wh != wh) {
        int64_t off[8];
        for (int i = 0; i < 8; i++)
            off[i] = (int64_t)i * wh;
        off[0] = 0;
        CK(hipMemcpyAsync(c.d_off, off, sizeof(off), hipMemcpyHostToDevice, c.s));
        CK(hipStreamSynchronize(c.s));
        c.last_wh = wh;
    }
    return c;
}

void pack(uint8_t *dst, int w, int h, const pel_t *src, intptr_t stride)
{
    for (int y = 0; y < h; y++)
        memcpy(dst + y * w, src + y * stride, (size_t)w);
}

int32_t run_cmp(int part, const pel_t *p1, intptr_t s1, const pel_t *p2, intptr_t s2, bool satd)
{
    const int w = xme_pu_w(part), h = xme_pu_h(part);
    CallCtx &c = ctx(w * h);
    pack(c.h_buf, w, h, p1, s1);
    pack(c.h_buf + w * h, w, h, p2, s2);
    CK(hipMemcpyAsync(c.d_buf, c.h_buf, (size_t)2 * w * h, hipMemcpyHostToDevice, c.s));
    // offsets {0} for the first block, {w*h} for the second: d_off[0], d_off[1]
    const int rc = satd ? xme_satd_batch(part, c.d_buf, w, c.d_buf, w, c.d_off, c.d_off + 1, 1, c.d_out, c.s)
                        : xme_sad_batch(part, c.d_buf, w, c.d_buf, w, c.d_off, c.d_off + 1, 1, c.d_out, c.s);
    if (rc)
        die(xme_last_error(), hipErrorUnknown);
    int32_t *res = reinterpret_cast<int32_t *>(c.h_buf + kBufBytes);
    CK(hipMemcpyAsync(res, c.d_out, sizeof(int32_t), hipMemcpyDeviceToHost, c.s));
    CK(hipStreamSynchronize(c.s));
    return res[0];
}

void run_xn(int part, int nr, const pel_t *fenc, const pel_t *const *refs, intptr_t stride, int32_t *scores)
{
    const int w = xme_pu_w(part), h = xme_pu_h(part);
    CallCtx &c = ctx(w * h);
    // fenc keeps its FENC_STRIDE layout; references follow, w bytes per row
    uint8_t *hb = c.h_buf;
    pack(hb, XME_FENC_STRIDE, h, fenc, XME_FENC_STRIDE);
    const int fenc_bytes = XME_FENC_STRIDE * h;
    for (int r = 0; r < nr; r++)
        pack(hb + fenc_bytes + r * w * h, w, h, refs[r], stride);
    CK(hipMemcpyAsync(c.d_buf, hb, (size_t)fenc_bytes + (size_t)nr * w * h, hipMemcpyHostToDevice, c.s));
    // reference offsets r * w * h follow d_off[0..nr) relative to d_buf + fenc_bytes
    const int rc = nr == 3 ? xme_sad_x3_batch(part, c.d_buf, c.d_buf + fenc_bytes, w, c.d_off, c.d_off, 1, c.d_out, c.s)
                           : xme_sad_x4_batch(part, c.d_buf, c.d_buf + fenc_bytes, w, c.d_off, c.d_off, 1, c.d_out, c.s);
    if (rc)
        die(xme_last_error(), hipErrorUnknown);
    int32_t *res = reinterpret_cast<int32_t *>(c.h_buf + kBufBytes);
    CK(hipMemcpyAsync(res, c.d_out, sizeof(int32_t) * nr, hipMemcpyDeviceToHost, c.s));
    CK(hipStreamSynchronize(c.s));
    for (int r = 0; r < nr; r++)
        scores[r] = res[r];
}

template <int P>
int32_t sad_fn(const pel_t *p1, intptr_t s1, const pel_t *p2, intptr_t s2)
{
    return run_cmp(P, p1, s1, p2, s2, false);
}

template <int P>
int32_t satd_fn(const pel_t *p1, intptr_t s1, const pel_t *p2, intptr_t s2)
{
    return run_cmp(P, p1, s1, p2, s2, true);
}

template <int P>
void sad_x3_fn(const pel_t *fenc, const pel_t *a, const pel_t *b, const pel_t *c, intptr_t stride, int32_t scores[3])
{
    const pel_t *refs[3] = { a, b, c };
    run_xn(P, 3, fenc, refs, stride, scores);
}

template <int P>
void sad_x4_fn(const pel_t *fenc, const pel_t *a, const pel_t *b, const pel_t *c, const pel_t *d, intptr_t stride,
               int32_t scores[4])
{
    const pel_t *refs[4] = { a, b, c, d };
    run_xn(P, 4, fenc, refs, stride, scores);
}

template <int... Ps>
struct Tables {
    static constexpr xme_pixel_cmp_t sad[sizeof...(Ps)] = { &sad_fn<Ps>... };
    static constexpr xme_pixel_cmp_t satd[sizeof...(Ps)] = { &satd_fn<Ps>... };
    static constexpr xme_pixel_cmp_x3_t x3[sizeof...(Ps)] = { &sad_x3_fn<Ps>... };
    static constexpr xme_pixel_cmp_x4_t x4[sizeof...(Ps)] = { &sad_x4_fn<Ps>... };
};
using T25 = Tables<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24>;

} // namespace

extern "C" int xme_get_sad_table(xme_pixel_cmp_t table[XME_NUM_PU])
{
    if (!table)
        return XME_ERR_ARG;
    for (int i = 0; i < XME_NUM_PU; i++)
        table[i] = T25::sad[i];
    return XME_OK;
}

extern "C" int xme_get_satd_table(xme_pixel_cmp_t table[XME_NUM_PU])
{
    if (!table)
        return XME_ERR_ARG;
    for (int i = 0; i < XME_NUM_PU; i++)
        table[i] = T25::satd[i];
    return XME_OK;
}

extern "C" int xme_get_sad_x3_table(xme_pixel_cmp_x3_t table[XME_NUM_PU])
{
    if (!table)
        return XME_ERR_ARG;
    for (int i = 0; i < XME_NUM_PU; i++)
        table[i] = T25::x3[i];
    return XME_OK;
}

extern "C" int xme_get_sad_x4_table(xme_pixel_cmp_x4_t table[XME_NUM_PU])
{
    if (!table)
        return XME_ERR_ARG;
    for (int i = 0; i < XME_NUM_PU; i++)
        table[i] = T25::x4[i];
    return XME_OK;
}

// ---------------------------------------------------------------------------
// Per-call transform, quant, interpolation and average faces (dctf.dct /
// idct / dct_half / quant / dequant, intpl_*, pixf.avg).  As above: a call
// stages its operands (with the filter margins) into the thread's pinned
// buffer, runs one launch on the thread's stream and copies the results back
// into the caller's memory with the caller's strides.

#include "xme_percall.h"

namespace {

// The thread's growable staging: pinned host and device memory of one size.
struct Stage {
    uint8_t *h = nullptr, *d = nullptr;
    size_t cap = 0;
    ~Stage()
    {
        if (h)
            (void)hipHostFree(h);
        if (d)
            (void)hipFree(d);
    }
    void need(size_t n)
    {
        if (n <= cap)
            return;
        if (h)
            CK(hipHostFree(h));
        if (d)
            CK(hipFree(d));
        cap = std::max(n, cap * 2);
        CK(hipHostMalloc((void **)&h, cap, hipHostMallocDefault));
        CK(hipMalloc((void **)&d, cap));
    }
};

thread_local Stage g_stage;

size_t al(size_t n) { return (n + 255) & ~(size_t)255; }

hipStream_t pc_stream()
{
    CallCtx &c = ctx(16);   // the thread's stream (and the per-call buffers of the cmp faces)
    static thread_local bool prepared = false;
    if (!prepared) {
        if (xme_pc_prepare())
            die(xme_last_error(), hipErrorUnknown);
        prepared = true;
    }
    return c.s;
}

int ilog2i(int n)
{
    int l = 0;
    while ((1 << l) < n)
        l++;
    return l;
}

bool takes_flag(int w, int h) { return (w == 32 && h == 32) || (w == 32 && h == 8) || (w == 8 && h == 32); }

// One dct_t call: dct_WxH_c / idct_WxH_c / dct_*_half_c (transform.c:1123-1597).
void run_xform(int part, int fwd, int half, const coeff_t *src, coeff_t *dst, int stride)
{
    const int w = xme_pu_w(part), h = xme_pu_h(part);
    const bool wave = w == 64 || h == 64;
    const int tw = wave ? w / 2 : w, th = wave ? h / 2 : h;   // the DCT stage's size
    int flag = 0;
    if (wave) {
        flag = 1;                    // the 64-wide sizes run their DCT stage with the flag
    } else if (takes_flag(w, h)) {
        flag = stride & 1;
        stride &= 0xFE;              // the reference's mask of the flagged strides
    }
    hipStream_t s = pc_stream();
    const size_t in_n = fwd ? (size_t)w * h : (size_t)tw * th, out_n = fwd ? (size_t)tw * th : (size_t)w * h;
    g_stage.need(al(in_n * 2) + out_n * 2);
    coeff_t *hin = reinterpret_cast<coeff_t *>(g_stage.h);
    coeff_t *hout = reinterpret_cast<coeff_t *>(g_stage.h + al(in_n * 2));
    if (fwd && !wave) {
        for (int y = 0; y < h; y++)
            memcpy(hin + y * w, src + (intptr_t)y * stride, sizeof(coeff_t) * (size_t)w);
    } else {
        memcpy(hin, src, sizeof(coeff_t) * in_n);     // contiguous input
    }
    int s1 = ilog2i(tw) - 2, s2 = ilog2i(th) + 5, clip2 = 9 + flag;
    if (fwd) {
        if (tw == 32 && th == 8)
            s2 -= flag;
        else
            s1 += flag;
    } else {
        s2 = 12 - flag;
    }
    CK(hipMemcpyAsync(g_stage.d, hin, in_n * 2, hipMemcpyHostToDevice, s));
    if (xme_pc_xform(reinterpret_cast<const int16_t *>(g_stage.d),
                     reinterpret_cast<int16_t *>(g_stage.d + al(in_n * 2)), w, h, fwd, s1, s2, clip2, half, s))
        die(xme_last_error(), hipErrorUnknown);
    CK(hipMemcpyAsync(hout, g_stage.d + al(in_n * 2), out_n * 2, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    if (!fwd && !wave) {
        for (int y = 0; y < h; y++)
            memcpy(dst + (intptr_t)y * stride, hout + y * w, sizeof(coeff_t) * (size_t)w);
    } else {
        memcpy(dst, hout, sizeof(coeff_t) * out_n);
    }
}

template <int P>
void dct_fn(const coeff_t *src, coeff_t *dst, int i_src)
{
    run_xform(P, 1, 0, src, dst, i_src);
}

template <int P>
void idct_fn(const coeff_t *src, coeff_t *dst, int i_dst)
{
    run_xform(P, 0, 0, src, dst, i_dst);
}

template <int P>
void dct_half_fn(const coeff_t *src, coeff_t *dst, int i_src)
{
    run_xform(P, 1, 1, src, dst, i_src);
}

int quant_fn(coeff_t *coef, const int i_coef, const int scale, const int shift, const int add)
{
    hipStream_t s = pc_stream();
    const size_t n = (size_t)std::max(i_coef, 0);
    g_stage.need(al(n * 2) + 64);
    memcpy(g_stage.h, coef, n * 2);
    CK(hipMemcpyAsync(g_stage.d, g_stage.h, n * 2, hipMemcpyHostToDevice, s));
    int *d_nnz = reinterpret_cast<int *>(g_stage.d + al(n * 2));
    if (xme_pc_quant(reinterpret_cast<int16_t *>(g_stage.d), (int)n, scale, shift, add, d_nnz, 1, s))
        die(xme_last_error(), hipErrorUnknown);
    CK(hipMemcpyAsync(g_stage.h, g_stage.d, al(n * 2) + sizeof(int), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    memcpy(coef, g_stage.h, n * 2);
    return *reinterpret_cast<int *>(g_stage.h + al(n * 2));
}

void dequant_fn(coeff_t *coef, const int i_coef, const int scale, const int shift)
{
    hipStream_t s = pc_stream();
    const size_t n = (size_t)std::max(i_coef, 0);
    g_stage.need(al(n * 2));
    memcpy(g_stage.h, coef, n * 2);
    CK(hipMemcpyAsync(g_stage.d, g_stage.h, n * 2, hipMemcpyHostToDevice, s));
    if (xme_pc_quant(reinterpret_cast<int16_t *>(g_stage.d), (int)n, scale, shift, 0, nullptr, 0, s))
        die(xme_last_error(), hipErrorUnknown);
    CK(hipMemcpyAsync(g_stage.h, g_stage.d, n * 2, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    memcpy(coef, g_stage.h, n * 2);
}

// An interpolation call on the GPU.  The source window is rows [y0, h + y1) x
// columns [x0, w + x1) around the block (the filter's reach), of bytes (esz 1)
// or int16 rows (esz 2); n_out output planes of w x h bytes (taps ch[k],
// cv[k] for plane k), and int16 tmp planes when tmp_out.
struct IntplCall {
    int mode, w, h, x0, x1, y0, y1, esz;
    const void *src;
    intptr_t i_src;              // in elements
    const void *src2;            // kPcAvg: the second source, stride i_src2
    intptr_t i_src2;
    int n_out;
    pel_t *dst[3];
    intptr_t i_dst;
    mct_t *tmp[3];
    intptr_t i_tmp;
    const int8_t *ch[3], *cv[3];
    int ntap;
};

void run_intpl(const IntplCall &c)
{
    if (c.w <= 0 || c.h <= 0)
        return;
    hipStream_t s = pc_stream();
    const int sw = c.w + c.x1 - c.x0, sh = c.h + c.y1 - c.y0;
    const size_t src_b = (size_t)sw * sh * c.esz;
    const size_t src2_b = c.src2 ? (size_t)c.w * c.h : 0;
    const size_t out_b = (size_t)c.w * c.h, tmp_b = (size_t)c.w * c.h * 2;
    const size_t o_src2 = al(src_b), o_out = o_src2 + al(src2_b), o_tmp = o_out + al(out_b) * c.n_out;
    g_stage.need(o_tmp + al(tmp_b) * c.n_out);
    uint8_t *hb = g_stage.h;
    const uint8_t *sp = static_cast<const uint8_t *>(c.src) + ((intptr_t)c.y0 * c.i_src + c.x0) * c.esz;
    for (int y = 0; y < sh; y++)
        memcpy(hb + (size_t)y * sw * c.esz, sp + (intptr_t)y * c.i_src * c.esz, (size_t)sw * c.esz);
    if (c.src2)
        for (int y = 0; y < c.h; y++)
            memcpy(hb + o_src2 + (size_t)y * c.w, static_cast<const uint8_t *>(c.src2) + (intptr_t)y * c.i_src2,
                   (size_t)c.w);
    CK(hipMemcpyAsync(g_stage.d, hb, o_out, hipMemcpyHostToDevice, s));
    for (int k = 0; k < c.n_out; k++) {
        PcIntpl p{};
        p.mode = c.mode;
        p.w = c.w;
        p.h = c.h;
        p.i_src = c.esz == 1 ? sw : 0;
        p.i_tmp = c.esz == 2 ? sw : c.w;
        p.i_dst = c.w;
        for (int t = 0; t < c.ntap; t++) {
            p.ch[t] = c.ch[k] ? c.ch[k][t] : 0;
            p.cv[t] = c.cv[k] ? c.cv[k][t] : 0;
        }
        // the block's first sample inside the staged window
        const size_t org = ((size_t)(-c.y0) * sw + (size_t)(-c.x0)) * c.esz;
        const uint8_t *d_src = c.esz == 1 ? g_stage.d + org : nullptr;
        const int16_t *d_tin = c.esz == 2 ? reinterpret_cast<const int16_t *>(g_stage.d + org) : nullptr;
        int16_t *d_tout = c.tmp[k] ? reinterpret_cast<int16_t *>(g_stage.d + o_tmp + al(tmp_b) * k) : nullptr;
        if (xme_pc_intpl(p, d_src, d_tin, c.src2 ? g_stage.d + o_src2 : nullptr, g_stage.d + o_out + al(out_b) * k,
                         d_tout, s))
            die(xme_last_error(), hipErrorUnknown);
    }
    CK(hipMemcpyAsync(hb + o_out, g_stage.d + o_out, (o_tmp - o_out) + al(tmp_b) * c.n_out, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (int k = 0; k < c.n_out; k++) {
        for (int y = 0; y < c.h; y++)
            memcpy(c.dst[k] + (intptr_t)y * c.i_dst, hb + o_out + al(out_b) * k + (size_t)y * c.w, (size_t)c.w);
        if (c.tmp[k])
            for (int y = 0; y < c.h; y++)
                memcpy(c.tmp[k] + (intptr_t)y * c.i_tmp, hb + o_tmp + al(tmp_b) * k + (size_t)y * c.w * 2,
                       (size_t)c.w * 2);
    }
}

IntplCall intpl_call(int mode, int w, int h, int ntap)
{
    IntplCall c{};
    c.mode = mode;
    c.w = w;
    c.h = h;
    c.esz = 1;
    c.n_out = 1;
    c.ntap = ntap;
    return c;
}

// intpl_luma_hor (mc.c:402-418), _x3 (mc.c:465-496): columns -3 .. w+3.
void luma_hor_n(int n, pel_t *const *dst, int i_dst, mct_t *const *tmp, int i_tmp, pel_t *src, int i_src, int w,
                int h, const int8_t *const *coeff)
{
    IntplCall c = intpl_call(kPcLumaHor, w, h, 8);
    c.x0 = -3;
    c.x1 = 4;
    c.src = src;
    c.i_src = i_src;
    c.n_out = n;
    c.i_dst = i_dst;
    c.i_tmp = i_tmp;
    for (int k = 0; k < n; k++) {
        c.dst[k] = dst[k];
        c.tmp[k] = tmp ? tmp[k] : nullptr;
        c.ch[k] = coeff[k];
    }
    run_intpl(c);
}

// intpl_luma_ver (mc.c:422-435), _x3 (mc.c:440-461): rows -3 .. h+3.
void luma_ver_n(int n, pel_t *const *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *const *coeff)
{
    IntplCall c = intpl_call(kPcLumaVer, w, h, 8);
    c.y0 = -3;
    c.y1 = 4;
    c.src = src;
    c.i_src = i_src;
    c.n_out = n;
    c.i_dst = i_dst;
    for (int k = 0; k < n; k++) {
        c.dst[k] = dst[k];
        c.cv[k] = coeff[k];
    }
    run_intpl(c);
}

// intpl_luma_ext (mc.c:500-515), _x3 (mc.c:519-547): int16 rows -3 .. h+3.
void luma_ext_n(int n, pel_t *const *dst, int i_dst, const mct_t *tmp, int i_tmp, int w, int h,
                const int8_t *const *coeff)
{
    IntplCall c = intpl_call(kPcLumaExt, w, h, 8);
    c.y0 = -3;
    c.y1 = 4;
    c.esz = 2;
    c.src = tmp;
    c.i_src = i_tmp;
    c.n_out = n;
    c.i_dst = i_dst;
    for (int k = 0; k < n; k++) {
        c.dst[k] = dst[k];
        c.cv[k] = coeff[k];
    }
    run_intpl(c);
}

void intpl_luma_hor_fn(pel_t *dst, int i_dst, mct_t *tmp, int i_tmp, pel_t *src, int i_src, int w, int h,
                       const int8_t *coeff)
{
    luma_hor_n(1, &dst, i_dst, &tmp, i_tmp, src, i_src, w, h, &coeff);
}

void intpl_luma_ver_fn(pel_t *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *coeff)
{
    luma_ver_n(1, &dst, i_dst, src, i_src, w, h, &coeff);
}

void intpl_luma_ext_fn(pel_t *dst, int i_dst, mct_t *tmp, int i_tmp, int w, int h, const int8_t *coeff)
{
    luma_ext_n(1, &dst, i_dst, tmp, i_tmp, w, h, &coeff);
}

void intpl_luma_ver_x3_fn(pel_t *const dst[3], int i_dst, pel_t *src, int i_src, int w, int h, const int8_t **coeff)
{
    luma_ver_n(3, dst, i_dst, src, i_src, w, h, coeff);
}

void intpl_luma_hor_x3_fn(pel_t *const dst[3], int i_dst, mct_t *const tmp[3], int i_tmp, pel_t *src, int i_src, int w,
                          int h, const int8_t **coeff)
{
    luma_hor_n(3, dst, i_dst, tmp, i_tmp, src, i_src, w, h, coeff);
}

void intpl_luma_ext_x3_fn(pel_t *const dst[3], int i_dst, mct_t *tmp, int i_tmp, int w, int h, const int8_t **coeff)
{
    luma_ext_n(3, dst, i_dst, tmp, i_tmp, w, h, coeff);
}

void intpl_luma_block_hor_fn(pel_t *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *coeff)
{
    luma_hor_n(1, &dst, i_dst, nullptr, 0, src, i_src, w, h, &coeff);
}

// intpl_luma_block_ext (mc.c:364-398): rows and columns -3 .. +3 beyond the block.
void intpl_luma_block_ext_fn(pel_t *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *coeff_h,
                             const int8_t *coeff_v)
{
    IntplCall c = intpl_call(kPcLumaBlockExt, w, h, 8);
    c.x0 = c.y0 = -3;
    c.x1 = c.y1 = 4;
    c.src = src;
    c.i_src = i_src;
    c.dst[0] = dst;
    c.i_dst = i_dst;
    c.ch[0] = coeff_h;
    c.cv[0] = coeff_v;
    run_intpl(c);
}

// intpl_chroma_block_hor / _ver / _ext (mc.c:275-337): 4 taps at -1 .. +2.
void chroma_block(int mode, pel_t *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *ch,
                  const int8_t *cv)
{
    IntplCall c = intpl_call(mode, w, h, 4);
    if (mode != kPcChromaVer) {
        c.x0 = -1;
        c.x1 = 2;
    }
    if (mode != kPcChromaHor) {
        c.y0 = -1;
        c.y1 = 2;
    }
    c.src = src;
    c.i_src = i_src;
    c.dst[0] = dst;
    c.i_dst = i_dst;
    c.ch[0] = ch;
    c.cv[0] = cv;
    run_intpl(c);
}

void intpl_chroma_block_hor_fn(pel_t *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *coeff)
{
    chroma_block(kPcChromaHor, dst, i_dst, src, i_src, w, h, coeff, nullptr);
}

void intpl_chroma_block_ver_fn(pel_t *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *coeff)
{
    chroma_block(kPcChromaVer, dst, i_dst, src, i_src, w, h, nullptr, coeff);
}

void intpl_chroma_block_ext_fn(pel_t *dst, int i_dst, pel_t *src, int i_src, int w, int h, const int8_t *coeff_h,
                               const int8_t *coeff_v)
{
    chroma_block(kPcChromaExt, dst, i_dst, src, i_src, w, h, coeff_h, coeff_v);
}

// pixf.avg[part] (pixel.c:680-719): (src0 + src1 + 1) >> 1; weight unused.
template <int P>
void avg_fn(pel_t *dst, intptr_t dstride, const pel_t *src0, intptr_t sstride0, const pel_t *src1, intptr_t sstride1,
            int weight)
{
    (void)weight;
    IntplCall c = intpl_call(kPcAvg, xme_pu_w(P), xme_pu_h(P), 0);
    c.src = src0;
    c.i_src = sstride0;
    c.src2 = src1;
    c.i_src2 = sstride1;
    c.dst[0] = dst;
    c.i_dst = dstride;
    run_intpl(c);
}

template <int... Ps>
struct XTables {
    static constexpr xme_dct_t dct[sizeof...(Ps)] = { &dct_fn<Ps>... };
    static constexpr xme_dct_t idct[sizeof...(Ps)] = { &idct_fn<Ps>... };
    static constexpr xme_dct_t half[sizeof...(Ps)] = { &dct_half_fn<Ps>... };
    static constexpr xme_pixel_avg_pp_t avg[sizeof...(Ps)] = { &avg_fn<Ps>... };
};
using X25 = XTables<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24>;

// the sizes xavs2_dct_init installs (transform.c:1605-1625)
bool in_dct_table(int part)
{
    const int w = xme_pu_w(part), h = xme_pu_h(part);
    return (w == h && w >= 4) || (w == 16 && h == 4) || (w == 4 && h == 16) || (w == 32 && h == 8) ||
           (w == 8 && h == 32) || (w == 64 && h == 16) || (w == 16 && h == 64);
}

} // namespace

extern "C" int xme_get_dct_table(xme_dct_t dct[XME_NUM_PU], xme_dct_t idct[XME_NUM_PU], xme_dct_t dct_half[XME_NUM_PU])
{
    if (!dct || !idct || !dct_half)
        return XME_ERR_ARG;
    for (int i = 0; i < XME_NUM_PU; i++) {
        const bool on = in_dct_table(i);
        const int w = xme_pu_w(i), h = xme_pu_h(i);
        dct[i] = on ? X25::dct[i] : nullptr;
        idct[i] = on ? X25::idct[i] : nullptr;
        dct_half[i] = (w == h && (w == 32 || w == 64)) ? X25::half[i] : nullptr;
    }
    return XME_OK;
}

extern "C" int xme_get_quant_funcs(xme_quant_t *quant, xme_dequant_t *dequant)
{
    if (!quant || !dequant)
        return XME_ERR_ARG;
    *quant = &quant_fn;
    *dequant = &dequant_fn;
    return XME_OK;
}

extern "C" int xme_get_avg_table(xme_pixel_avg_pp_t table[XME_NUM_PU])
{
    if (!table)
        return XME_ERR_ARG;
    for (int i = 0; i < XME_NUM_PU; i++)
        table[i] = X25::avg[i];
    return XME_OK;
}

extern "C" int xme_get_intpl_funcs(xme_intpl_funcs *f)
{
    if (!f)
        return XME_ERR_ARG;
    f->intpl_luma_hor = &intpl_luma_hor_fn;
    f->intpl_luma_ver = &intpl_luma_ver_fn;
    f->intpl_luma_ext = &intpl_luma_ext_fn;
    f->intpl_luma_ver_x3 = &intpl_luma_ver_x3_fn;
    f->intpl_luma_hor_x3 = &intpl_luma_hor_x3_fn;
    f->intpl_luma_ext_x3 = &intpl_luma_ext_x3_fn;
    f->intpl_luma_block_hor = &intpl_luma_block_hor_fn;
    f->intpl_luma_block_ver = &intpl_luma_ver_fn;       // #define intpl_luma_block_ver_c intpl_luma_ver_c (mc.c:358)
    f->intpl_luma_block_ext = &intpl_luma_block_ext_fn;
    f->intpl_chroma_block_hor = &intpl_chroma_block_hor_fn;
    f->intpl_chroma_block_ver = &intpl_chroma_block_ver_fn;
    f->intpl_chroma_block_ext = &intpl_chroma_block_ext_fn;
    return XME_OK;
}
