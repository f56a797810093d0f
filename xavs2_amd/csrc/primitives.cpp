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
