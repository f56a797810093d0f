// Kernels behind the per-call g_funcs faces that have no batched counterpart:
// one transform of the dct_t table (any of its 11 sizes, forward or inverse,
// with the 64-wide sizes' lifting wavelet), quant / dequant of one
// coefficient array, one interpolation call with its own taps, one pixel
// average.  Each call is one small launch (one workgroup for a transform or
// a coefficient array, a grid over the pixels for interpolation): the face
// exists for drop-in parity, not throughput; the batched entry points are the
// fast path.
#include <cmath>
#include <mutex>

#include "xme_percall.h"

namespace {

typedef short short2v __attribute__((ext_vector_type(2)));

// AVS2 DCT basis T_32 (transform.c:59-135, derived from its cosine closed
// form); T_N[k][n] = T_32[k * 32 / N][n].
__constant__ int8_t g_pc_t32[32 * 32];
std::once_flag g_pc_once;
hipError_t g_pc_err = hipSuccess;

void pc_upload_basis()
{
    int8_t t[32 * 32];
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < 32; k++)
        for (int n = 0; n < 32; n++)
            t[k * 32 + n] = (int8_t)(k == 0 ? 32 : std::lround(32.0 * std::sqrt(2.0) * std::cos((2 * n + 1) * k * pi / 64.0)));
    g_pc_err = hipMemcpyToSymbol(HIP_SYMBOL(g_pc_t32), t, sizeof(t));
}

__device__ __forceinline__ int basis(int n, int k, int m) { return g_pc_t32[k * (32 / n) * 32 + m]; }

__device__ __forceinline__ int clip3i(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }

// One forward lifting pass over n samples of x (stride `step` in LDS),
// writing the n / 2 low-band samples to out (stride ostep): mirror two
// samples a side, odd samples minus the floor mean of their neighbours, then
// the even ones plus (l + r + 2) >> 2 (rows) or doubled plus (l + r + 1) >> 1
// (columns); every stored sample is truncated to int16 as the reference's
// coeff_t array keeps it.  The work array holds int: with int16 storage the
// compiler turns the averages into packed 16-bit arithmetic, which this
// library keeps out of its kernels.
__device__ void lift_fwd(const int16_t *x, int step, int16_t *out, int ostep, int n, bool columns)
{
    int e[68];
    int *v = e + 2;
    for (int i = 0; i < n; i++)
        v[i] = x[i * step];
    v[-1] = v[1];
    v[-2] = v[2];
    v[n] = v[n - 2];
    v[n + 1] = v[n - 3];
    for (int i = -1; i < n; i += 2)
        v[i] = (int16_t)(v[i] - ((v[i - 1] + v[i + 1]) >> 1));
    for (int i = 0; i < n; i += 2)
        v[i] = columns ? (int16_t)((v[i] << 1) + ((v[i - 1] + v[i + 1] + 1) >> 1))
                       : (int16_t)(v[i] + ((v[i - 1] + v[i + 1] + 2) >> 2));
    for (int i = 0; i < n / 2; i++)
        out[i * ostep] = (int16_t)v[2 * i];
}

// One inverse lifting pass: n / 2 low samples to the even places, the last
// mirrored (x[n] = x[n - 2]); columns halve the even samples 0..n; odd
// samples are the floor mean of their neighbours.
__device__ void lift_inv(const int16_t *x, int step, int16_t *out, int ostep, int n, bool columns)
{
    int e[66];      // int16 values held in int (as in lift_fwd)
    for (int i = 0; i < n / 2; i++)
        e[2 * i] = x[i * step];
    e[n] = e[n - 2];
    if (columns)
        for (int i = 0; i <= n; i += 2)
            e[i] = (int16_t)(e[i] >> 1);
    for (int i = 1; i < n; i += 2)
        e[i] = (int16_t)((e[i - 1] + e[i + 1]) >> 1);
    for (int i = 0; i < n; i++)
        out[i * ostep] = (int16_t)e[i];
}

struct PcXform {
    int w, h;           // block size as the table names it (64-wide sizes included)
    int fwd;            // 1: dct, 0: idct
    int s1, s2;         // forward shifts, or the inverse second-pass shift
    int clip2;          // inverse second-pass clip depth
    int half;           // dct_half: clear all but the low 16x16
};

// One transform, one workgroup of 256 lanes, the block in LDS.  Forward: the
// 64-wide sizes first reduce the block to the wavelet's low band (a lane a
// row, then a lane a column); then rows (W'-point, shift s1, int16) and
// columns (H'-point, shift s2, int16).  Inverse: columns (H'-point, shift 5,
// 16-bit clip), rows (W'-point, shift s2, clip2), then the inverse wavelet.
__global__ void __launch_bounds__(256) pc_xform_kernel(const int16_t *__restrict__ src, int16_t *__restrict__ dst, PcXform p)
{
    __shared__ int16_t a[64 * 64];
    __shared__ int16_t b[64 * 64];
    const int t = threadIdx.x;
    const bool wave = p.w == 64 || p.h == 64;
    const int w = wave ? p.w / 2 : p.w, h = wave ? p.h / 2 : p.h;
    if (p.fwd) {
        for (int i = t; i < p.w * p.h; i += 256)
            a[i] = src[i];
        __syncthreads();
        if (wave) {
            if (t < p.h)
                lift_fwd(a + t * p.w, 1, b + t * w, 1, p.w, false);
            __syncthreads();
            if (t < w)
                lift_fwd(b + t, w, a + t, w, p.h, true);
            __syncthreads();
        }
        // rows: b[k][j] (h columns) = T_w[k] . a[j]
        const int add1 = (1 << p.s1) >> 1;
        for (int i = t; i < w * h; i += 256) {
            const int j = i / w, k = i - j * w;
            int s = 0;
            for (int m = 0; m < w; m++)
                s += basis(w, k, m) * a[j * w + m];
            b[k * h + j] = (int16_t)((s + add1) >> p.s1);
        }
        __syncthreads();
        const int add2 = (1 << p.s2) >> 1;
        for (int i = t; i < w * h; i += 256) {
            const int r = i / w, k = i - r * w;      // output row r, column k
            int s = 0;
            for (int m = 0; m < h; m++)
                s += basis(h, r, m) * b[k * h + m];
            int16_t v = (int16_t)((s + add2) >> p.s2);
            if (p.half && (r >= 16 || k >= 16))
                v = 0;
            dst[i] = v;
        }
        return;
    }
    for (int i = t; i < w * h; i += 256)
        a[i] = src[i];
    __syncthreads();
    // columns: b[j][m] (h per row j) = clip16((sum_k T_h[k][m] a[k][j] + 16) >> 5)
    for (int i = t; i < w * h; i += 256) {
        const int j = i / h, m = i - j * h;
        int s = 0;
        for (int k = 0; k < h; k++)
            s += basis(h, k, m) * a[k * w + j];
        b[j * h + m] = (int16_t)clip3i(-32768, 32767, (s + 16) >> 5);
    }
    __syncthreads();
    const int add2 = 1 << (p.s2 - 1), hi = (1 << (p.clip2 - 1)) - 1;
    for (int i = t; i < w * h; i += 256) {
        const int j = i / w, m = i - j * w;          // row j, column m
        int s = 0;
        for (int k = 0; k < w; k++)
            s += basis(w, k, m) * b[k * h + j];
        a[i] = (int16_t)clip3i(-hi - 1, hi, (s + add2) >> p.s2);
    }
    __syncthreads();
    if (!wave) {
        for (int i = t; i < w * h; i += 256)
            dst[i] = a[i];
        return;
    }
    if (t < w)
        lift_inv(a + t, w, b + t, w, p.h, true);
    __syncthreads();
    if (t < p.h)
        lift_inv(b + t * w, 1, a + t * p.w, 1, p.w, false);
    __syncthreads();
    for (int i = t; i < p.w * p.h; i += 256)
        dst[i] = a[i];
}

// quant_c / dequant_c (quant.c:109-184) over one coefficient array.
__global__ void __launch_bounds__(256) pc_quant_kernel(int16_t *c, int n, int scale, int shift, int add, int *nnz,
                                                       int quant)
{
    __shared__ int cnt;
    if (threadIdx.x == 0)
        cnt = 0;
    __syncthreads();
    int local = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const int v = c[i];
        if (quant) {
            const int q = (v < 0 ? -1 : 1) * (((v < 0 ? -v : v) * scale + add) >> shift);
            c[i] = (int16_t)q;
            local += (int16_t)q != 0;
        } else if (v != 0) {
            c[i] = (int16_t)clip3i(-32768, 32767, (v * scale + (1 << (shift - 1))) >> shift);
        }
    }
    if (quant) {
        atomicAdd(&cnt, local);
        __syncthreads();
        if (threadIdx.x == 0)
            *nnz = cnt;
    }
}

} // namespace



namespace {

__global__ void __launch_bounds__(256) pc_intpl_kernel(PcIntpl p, const uint8_t *__restrict__ src,
                                                       const int16_t *__restrict__ tin, const uint8_t *__restrict__ src2,
                                                       uint8_t *__restrict__ dst, int16_t *__restrict__ tout)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= p.w || y >= p.h)
        return;
    const uint8_t *s = src + y * p.i_src + x;    // the source sample of (x, y)
    int v = 0;
    switch (p.mode) {
    case kPcLumaHor: {
        for (int k = 0; k < 8; k++)
            v += p.ch[k] * s[k - 3];
        if (tout)
            tout[y * p.i_tmp + x] = (int16_t)v;
        v = (v + 32) >> 6;
        break;
    }
    case kPcLumaVer:
        for (int k = 0; k < 8; k++)
            v += p.cv[k] * s[(k - 3) * p.i_src];
        v = (v + 32) >> 6;
        break;
    case kPcLumaExt: {
        const int16_t *q = tin + y * p.i_tmp + x;
        for (int k = 0; k < 8; k++)
            v += p.cv[k] * q[(k - 3) * p.i_tmp];
        v = (v + 2048) >> 12;
        break;
    }
    case kPcLumaBlockExt: {
        // the row sums as v_dot2_i32_i16 on (pixel, pixel) and (tap, tap)
        // pairs: an int result the compiler cannot narrow to packed 16-bit
        // arithmetic when it is truncated to the mct_t intermediate
        short2v t[4];
        for (int k = 0; k < 4; k++)
            t[k] = short2v{ (short)p.ch[2 * k], (short)p.ch[2 * k + 1] };
        for (int r = 0; r < 8; r++) {
            const uint8_t *row = s + (r - 3) * p.i_src - 3;
            int u = 0;
            for (int k = 0; k < 4; k++)
                u = __builtin_amdgcn_sdot2(short2v{ (short)row[2 * k], (short)row[2 * k + 1] }, t[k], u, false);
            v += p.cv[r] * (int)(int16_t)u;     // the intermediate is mct_t
        }
        v = (v + 2048) >> 12;
        break;
    }
    case kPcChromaHor:
        for (int k = 0; k < 4; k++)
            v += p.ch[k] * s[k - 1];
        v = (v + 32) >> 6;
        break;
    case kPcChromaVer:
        for (int k = 0; k < 4; k++)
            v += p.cv[k] * s[(k - 1) * p.i_src];
        v = (v + 32) >> 6;
        break;
    case kPcChromaExt:
        for (int r = 0; r < 4; r++) {
            int u = 0;
            for (int k = 0; k < 4; k++)
                u += p.ch[k] * s[(r - 1) * p.i_src + k - 1];
            v += p.cv[r] * u;                   // int32 intermediate
        }
        v = (v + 2048) >> 12;
        break;
    default:                                    // kPcAvg: (a + b + 1) >> 1
        v = (s[0] + src2[y * p.i_src + x] + 1) >> 1;
        break;
    }
    dst[y * p.i_dst + x] = (uint8_t)clip3i(0, 255, v);
}

} // namespace

int xme_pc_prepare()
{
    std::call_once(g_pc_once, pc_upload_basis);
    return g_pc_err == hipSuccess ? XME_OK : xme_hip_fail(g_pc_err, "per-call transform basis");
}

// Launch one transform: w x h names the table size; s1 / s2 / clip2 as the
// table entry and its stride flag set them (see primitives.cpp).
int xme_pc_xform(const int16_t *d_src, int16_t *d_dst, int w, int h, int fwd, int s1, int s2, int clip2, int half,
                 hipStream_t s)
{
    PcXform p{ w, h, fwd, s1, s2, clip2, half };
    hipLaunchKernelGGL(pc_xform_kernel, dim3(1), dim3(256), 0, s, d_src, d_dst, p);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

int xme_pc_quant(int16_t *d_c, int n, int scale, int shift, int add, int *d_nnz, int quant, hipStream_t s)
{
    hipLaunchKernelGGL(pc_quant_kernel, dim3(1), dim3(256), 0, s, d_c, n, scale, shift, add, d_nnz, quant);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

int xme_pc_intpl(const PcIntpl &p, const uint8_t *d_src, const int16_t *d_tin, const uint8_t *d_src2, uint8_t *d_dst,
                 int16_t *d_tout, hipStream_t s)
{
    hipLaunchKernelGGL(pc_intpl_kernel, dim3((p.w + 63) / 64, (p.h + 3) / 4), dim3(256), 0, s, p, d_src, d_tin, d_src2,
                       d_dst, d_tout);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
