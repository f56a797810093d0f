// Batched pixel cost primitives for gfx950: pixf.sad / satd / sad_x3 / sad_x4
// (source/common/pixel.c:93-407).  Call i of a batch evaluates the primitive
// on (base + offset[i]) pointers; one wave handles one call.
#include "xme_internal.h"
#include "tune.h"

#include <algorithm>

namespace {

__device__ __forceinline__ uint32_t load_dw_any(const uint8_t *p)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint32_t *q = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(a & 3);
    // the second dword holds needed bytes only when the address is unaligned
    return sh ? __builtin_amdgcn_alignbyte(q[1], q[0], sh) : q[0];
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

// SAD of one W x H block; lanes take 4-pixel groups.
__device__ __forceinline__ int block_sad(const uint8_t *a, int64_t sa, const uint8_t *b, int64_t sb, int w, int h)
{
    const int lane = threadIdx.x & 63;
    const int gw = w >> 2;
    int s = 0;
    for (int g = lane; g < gw * h; g += 64) {
        const int y = g / gw, x = (g - y * gw) * 4;
        s = (int)__builtin_amdgcn_sad_u8(load_dw_any(a + y * sa + x), load_dw_any(b + y * sb + x), (uint32_t)s);
    }
    return wave_sum(s);
}

// 4x4 Hadamard sum (before halving), see satd_4x4 (pixel.c:297-322).
__device__ __forceinline__ int hadamard4x4(const uint8_t *a, int64_t sa, const uint8_t *b, int64_t sb)
{
    int d[4][4];
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const uint32_t u = load_dw_any(a + y * sa), v = load_dw_any(b + y * sb);
#pragma unroll
        for (int x = 0; x < 4; x++)
            d[y][x] = (int)((u >> (8 * x)) & 255u) - (int)((v >> (8 * x)) & 255u);
    }
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const int p = d[y][0] + d[y][1], m = d[y][0] - d[y][1];
        const int q = d[y][2] + d[y][3], n = d[y][2] - d[y][3];
        d[y][0] = p + q;
        d[y][1] = m + n;
        d[y][2] = p - q;
        d[y][3] = m - n;
    }
    int s = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const int p = d[0][x] + d[1][x], m = d[0][x] - d[1][x];
        const int q = d[2][x] + d[3][x], n = d[2][x] - d[3][x];
        s += abs(p + q) + abs(m + n) + abs(p - q) + abs(m - n);
    }
    return s;
}

__global__ void __launch_bounds__(256) sad_batch_kernel(int w, int h, const uint8_t *fenc, int64_t i_fenc,
                                                        const uint8_t *fref, int64_t i_fref,
                                                        const int64_t *__restrict__ off1,
                                                        const int64_t *__restrict__ off2, int n, int32_t *out, int satd)
{
    const int call = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (call >= n)
        return;
    const uint8_t *a = fenc + off1[call];
    const uint8_t *b = fref + off2[call];
    int v;
    if (!satd) {
        v = block_sad(a, i_fenc, b, i_fref, w, h);
    } else {
        const int lane = threadIdx.x & 63, tw = w >> 2, nt = tw * (h >> 2);
        int s = 0;
        for (int t = lane; t < nt; t += 64) {
            const int ty = (t / tw) * 4, tx = (t % tw) * 4;
            s += hadamard4x4(a + ty * i_fenc + tx, i_fenc, b + ty * i_fref + tx, i_fref);
        }
        v = wave_sum(s) >> 1;
    }
    if ((threadIdx.x & 63) == 0)
        out[call] = v;
}

__global__ void __launch_bounds__(256) sad_xn_kernel(int w, int h, int nr, const uint8_t *fenc, const uint8_t *fref,
                                                     int64_t i_fref, const int64_t *__restrict__ off1,
                                                     const int64_t *__restrict__ off2, int n, int32_t *scores)
{
    const int call = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (call >= n)
        return;
    const uint8_t *a = fenc + off1[call];
    for (int r = 0; r < nr; r++) {
        const int v = block_sad(a, XME_FENC_STRIDE, fref + off2[(int64_t)call * nr + r], i_fref, w, h);
        if ((threadIdx.x & 63) == 0)
            scores[(int64_t)call * nr + r] = v;
    }
}

} // namespace

static int check_part(int part, const char *what)
{
    if (part < 0 || part >= XME_NUM_PU) {
        xme_set_error("%s: part %d is not a LumaPU index", what, part);
        return XME_ERR_ARG;
    }
    return XME_OK;
}

static int launch_sad(int part, const pel_t *fenc, intptr_t i_fenc, const pel_t *fref, intptr_t i_fref,
                      const int64_t *o1, const int64_t *o2, int n, int32_t *out, void *stream, int satd)
{
    int rc = check_part(part, satd ? "xme_satd_batch" : "xme_sad_batch");
    if (rc)
        return rc;
    if (n <= 0)
        return XME_OK;
    if (!fenc || !fref || !o1 || !o2 || !out)
        return XME_ERR_ARG;
    hipLaunchKernelGGL(sad_batch_kernel, dim3((n + 3) / 4), dim3(256), 0, (hipStream_t)stream, xme_pu_w(part),
                       xme_pu_h(part), fenc, (int64_t)i_fenc, fref, (int64_t)i_fref, o1, o2, n, out, satd);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_sad_batch(int part, const pel_t *fenc, intptr_t i_fenc, const pel_t *fref, intptr_t i_fref,
                             const int64_t *o1, const int64_t *o2, int n, int32_t *out, void *stream)
{
    return launch_sad(part, fenc, i_fenc, fref, i_fref, o1, o2, n, out, stream, 0);
}

extern "C" int xme_satd_batch(int part, const pel_t *fenc, intptr_t i_fenc, const pel_t *fref, intptr_t i_fref,
                              const int64_t *o1, const int64_t *o2, int n, int32_t *out, void *stream)
{
    return launch_sad(part, fenc, i_fenc, fref, i_fref, o1, o2, n, out, stream, 1);
}

static int launch_xn(int nr, int part, const pel_t *fenc, const pel_t *fref, intptr_t i_fref, const int64_t *o1,
                     const int64_t *o2, int n, int32_t *scores, void *stream)
{
    int rc = check_part(part, "xme_sad_xN_batch");
    if (rc)
        return rc;
    if (n <= 0)
        return XME_OK;
    if (!fenc || !fref || !o1 || !o2 || !scores)
        return XME_ERR_ARG;
    hipLaunchKernelGGL(sad_xn_kernel, dim3((n + 3) / 4), dim3(256), 0, (hipStream_t)stream, xme_pu_w(part),
                       xme_pu_h(part), nr, fenc, fref, (int64_t)i_fref, o1, o2, n, scores);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_sad_x3_batch(int part, const pel_t *fenc, const pel_t *fref, intptr_t i_fref, const int64_t *o1,
                                const int64_t *o2, int n, int32_t *scores, void *stream)
{
    return launch_xn(3, part, fenc, fref, i_fref, o1, o2, n, scores, stream);
}

extern "C" int xme_sad_x4_batch(int part, const pel_t *fenc, const pel_t *fref, intptr_t i_fref, const int64_t *o1,
                                const int64_t *o2, int n, int32_t *scores, void *stream)
{
    return launch_xn(4, part, fenc, fref, i_fref, o1, o2, n, scores, stream);
}

namespace {
__global__ void __launch_bounds__(256) copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
} // namespace

extern "C" int xme_copy_kernel(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!dst || !src || (bytes & 15) || (reinterpret_cast<uintptr_t>(dst) & 15) ||
        (reinterpret_cast<uintptr_t>(src) & 15)) {
        xme_set_error("xme_copy_kernel: 16-byte aligned buffers and size required");
        return XME_ERR_ARG;
    }
    const size_t n = bytes / 16;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(copy_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<uint4 *>(dst), reinterpret_cast<const uint4 *>(src), n);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
