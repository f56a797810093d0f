// Residual path for gfx950: integer DCT (dct_NxN_c), flat quantisation
// (quant_c), dequantisation (dequant_c) and integer IDCT (idct_NxN_c) for
// N = 4, 8, 16, 32 (source/common/transform.c:167-1312, source/common/quant.c:109-184,
// parameters of tu_quant_forward / tu_quant_inverse, source/encoder/rdo.c:658-718).
//
// A workgroup stages 4096 coefficients (4096 / N^2 blocks) in LDS.  Every 1-D
// pass of the reference's partial butterflies is an exact integer product of a
// matrix row with a block row, so each pass here reads a contiguous row of
// int16 pairs and issues v_dot2_i32_i16 against a matrix row that is uniform
// across the wave (scalar loads); it writes its outputs transposed, which is
// the order the next pass reads (as in the reference, where pass outputs are
// stored at dst[k * line + j]).  The 32-bit sums are then shifted, wrapped to
// int16 (forward) or clipped (inverse) exactly like the reference.
#include <cmath>
#include <mutex>

#include "xme_internal.h"

namespace {

typedef short short2_t __attribute__((ext_vector_type(2)));

// Forward matrices T_N (row k, pairs of columns) and inverse matrices T_N^T
// (row m holds T[k][m], pairs of k), N = 4, 8, 16, 32 at dword offsets
// 0, 8, 40, 168 (N * N / 2 dwords of int16 pairs per matrix).
__constant__ uint32_t c_fwd[8 + 32 + 128 + 512];
__constant__ uint32_t c_inv[8 + 32 + 128 + 512];

__host__ __device__ constexpr int mat_off(int n) { return n == 4 ? 0 : n == 8 ? 8 : n == 16 ? 40 : 168; }

// quant_c / dequant_c parameters
struct QParams {
    int qscale, qshift, qadd;
    int dscale, dshift;
};

enum { DO_FWD = 1, DO_QUANT = 2, DO_DEQUANT = 4, DO_INV = 8 };

template <int N>
struct Geo {
    static constexpr int NB = 4096 / (N * N);       // blocks per workgroup
    static constexpr int ROWS = NB * N;             // block rows per workgroup
    static constexpr int NP = N / 2;                // int16 pairs per row
    // each thread: RT rows x OT outputs per pass; a wave shares its output range
    static constexpr int OT = N == 32 ? 16 : N == 16 ? 16 : N;
    static constexpr int RT = N == 32 ? 1 : N == 16 ? 1 : N == 8 ? 2 : 4;
    static constexpr int OSPLIT = N / OT;           // thread groups splitting the outputs
};

__device__ __forceinline__ int dot_row(const uint32_t *__restrict__ row, const uint32_t *__restrict__ coef, int np,
                                       int acc)
{
    for (int t = 0; t < np; t++)
        acc = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, row[t]), __builtin_bit_cast(short2_t, coef[t]),
                                     acc, false);
    return acc;
}

// One 1-D pass over every row of every block: in[b][r][0..N) -> out[b][m][r]
// for all m (outputs written transposed).  mode 0: forward (wrap to int16),
// 1: inverse (clip to [lo, hi]).
template <int N, bool INV>
__device__ __forceinline__ void pass(const int16_t *__restrict__ in, int16_t *__restrict__ out, int shift, int lo,
                                     int hi)
{
    using G = Geo<N>;
    const int tid = threadIdx.x;
    const int og = tid / (256 / G::OSPLIT);          // output group: uniform across a wave
    const int lt = tid % (256 / G::OSPLIT);
    const uint32_t *mat = (INV ? c_inv : c_fwd) + mat_off(N);
    const int add = (1 << shift) >> 1;
#pragma unroll
    for (int rr = 0; rr < G::RT; rr++) {
        const int grow = lt * G::RT + rr;            // row index within the workgroup
        if (grow >= G::ROWS)
            break;
        const int b = grow / N, r = grow % N;
        uint32_t row[G::NP];
        const uint32_t *src = reinterpret_cast<const uint32_t *>(in + (b * N + r) * N);
#pragma unroll
        for (int t = 0; t < G::NP; t++)
            row[t] = src[t];
#pragma unroll
        for (int o = 0; o < G::OT; o++) {
            const int m = og * G::OT + o;
            const int s = (dot_row(row, mat + m * G::NP, G::NP, 0) + add) >> shift;
            out[(b * N + m) * N + r] = INV ? (int16_t)xme_clip3(lo, hi, s) : (int16_t)s;
        }
    }
}

template <int N>
__global__ void __launch_bounds__(256) residual_kernel(const int16_t *__restrict__ src, int16_t *__restrict__ level,
                                                       int16_t *__restrict__ recon, int32_t *__restrict__ nnz,
                                                       int nblk, QParams q, int mode, int fshift1)
{
    using G = Geo<N>;
    __shared__ __align__(16) int16_t A[4096];
    __shared__ __align__(16) int16_t B[4096];
    __shared__ __align__(16) int16_t L[4096];
    __shared__ int s_nnz[G::NB];
    const int tid = threadIdx.x;
    const int blk0 = blockIdx.x * G::NB;
    const int nb = min(G::NB, nblk - blk0);
    const int nel = nb * N * N;
    const int64_t base = (int64_t)blk0 * N * N;

    // load (16-byte chunks; 4096 elements = 512 chunks)
    if (mode & DO_FWD) {
        for (int e = tid; e < nel / 8; e += 256)
            reinterpret_cast<uint4 *>(A)[e] = reinterpret_cast<const uint4 *>(src + base)[e];
    } else {
        // coefficients arrive row-major [k2][k]; the passes read them as rows of the transpose
        for (int e = tid; e < nel; e += 256) {
            const int b = e / (N * N), rem = e % (N * N), k2 = rem / N, k = rem % N;
            A[(b * N + k) * N + k2] = src[base + e];
        }
    }
    if (tid < G::NB)
        s_nnz[tid] = 0;
    __syncthreads();
    int16_t *coefT = A;           // coefficients as rows of the transpose: coefT[b][col][row]
    if (mode & DO_FWD) {
        pass<N, false>(A, B, fshift1, 0, 0);                       // rows -> B[b][k][j]
        __syncthreads();
        pass<N, false>(B, A, (N == 4 ? 2 : N == 8 ? 3 : N == 16 ? 4 : 5) + 5, 0, 0);  // -> A[b][m][r]: coef(m, r) at A[b][m][r]
        __syncthreads();
        // A now holds coef[b][m][r] with m the vertical frequency (row) and r
        // the horizontal frequency: the transpose layout the inverse reads is
        // A[b][r][m]; quantise in place while transposing into B
        for (int e = tid; e < nel; e += 256) {
            const int b = e / (N * N), rem = e % (N * N), m = rem / N, r = rem % N;
            int c = A[e];
            if (mode & DO_QUANT) {
                const int a = c < 0 ? -c : c;
                c = (int16_t)((c < 0 ? -1 : 1) * ((a * q.qscale + q.qadd) >> q.qshift));
                if (c != 0)
                    atomicAdd(&s_nnz[b], 1);
            }
            L[e] = (int16_t)c;                                        // row-major [m][r]
            int d = c;
            if ((mode & DO_DEQUANT) && c != 0)
                d = xme_clip3(-32768, 32767, (c * q.dscale + (1 << (q.dshift - 1))) >> q.dshift);
            B[(b * N + r) * N + m] = (int16_t)d;
        }
        coefT = B;
        __syncthreads();
        for (int e = tid; e < nel / 8; e += 256)
            reinterpret_cast<uint4 *>(level + base)[e] = reinterpret_cast<const uint4 *>(L)[e];
        if (nnz && tid < nb)
            nnz[blk0 + tid] = s_nnz[tid];
    }
    if (mode & DO_INV) {
        int16_t *other = coefT == A ? B : A;
        pass<N, true>(coefT, other, 5, -32768, 32767);               // -> other[b][m'][j] = tmp[j][m']
        __syncthreads();
        pass<N, true>(other, L, 12, -256, 255);                      // -> L[b][m2][x]: recon at row x, col m2
        __syncthreads();
        // L holds recon transposed: recon[b][x][m2] = L[b][m2][x]
        for (int e = tid; e < nel; e += 256) {
            const int b = e / (N * N), rem = e % (N * N), x = rem / N, m2 = rem % N;
            recon[base + e] = L[(b * N + m2) * N + x];
        }
    }
}

std::once_flag g_tab_once;
hipError_t g_tab_err = hipSuccess;

// AVS2 DCT basis: 32 for the DC row, else round(32*sqrt(2)*cos((2n+1)k*pi/(2N))),
// the values of g_T4..g_T32 (transform.c:59-135); T_N[k][n] = T_32[k*32/N][n].
void upload_tables()
{
    int16_t T32[32][32];
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < 32; k++)
        for (int n = 0; n < 32; n++)
            T32[k][n] = (int16_t)(k == 0 ? 32 : std::lround(32.0 * std::sqrt(2.0) * std::cos((2 * n + 1) * k * pi / 64.0)));
    uint32_t fwd[8 + 32 + 128 + 512], inv[8 + 32 + 128 + 512];
    for (int n : { 4, 8, 16, 32 }) {
        const int s = 32 / n, off = mat_off(n);
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j += 2) {
                const uint16_t a = (uint16_t)T32[k * s][j], b = (uint16_t)T32[k * s][j + 1];
                fwd[off + k * n / 2 + j / 2] = (uint32_t)a | ((uint32_t)b << 16);
                // inverse row m = k, pair (T[j][m], T[j+1][m])
                const uint16_t c = (uint16_t)T32[j * s][k], d = (uint16_t)T32[(j + 1) * s][k];
                inv[off + k * n / 2 + j / 2] = (uint32_t)c | ((uint32_t)d << 16);
            }
    }
    g_tab_err = hipMemcpyToSymbol(HIP_SYMBOL(c_fwd), fwd, sizeof(fwd));
    if (g_tab_err == hipSuccess)
        g_tab_err = hipMemcpyToSymbol(HIP_SYMBOL(c_inv), inv, sizeof(inv));
}

// Quantisation tables, source/common/quant.c:55-97.
const uint16_t kQ[80] = {
    32768, 29775, 27554, 25268, 23170, 21247, 19369, 17770, 16302, 15024, 13777, 12634, 11626, 10624, 9742, 8958,
    8192, 7512, 6889, 6305, 5793, 5303, 4878, 4467, 4091, 3756, 3444, 3161, 2894, 2654, 2435, 2235,
    2048, 1878, 1722, 1579, 1449, 1329, 1218, 1117, 1024, 939, 861, 790, 724, 664, 609, 558,
    512, 470, 430, 395, 362, 332, 304, 279, 256, 235, 215, 197, 181, 166, 152, 140,
    128, 116, 108, 99, 91, 83, 76, 69, 64, 59, 54, 49, 45, 41, 38, 35 };
const uint16_t kIQ[80] = {
    32768, 36061, 38968, 42495, 46341, 50535, 55437, 60424, 32932, 35734, 38968, 42495, 46177, 50535, 55109, 59933,
    65535, 35734, 38968, 42577, 46341, 50617, 55027, 60097, 32809, 35734, 38968, 42454, 46382, 50576, 55109, 60056,
    65535, 35734, 38968, 42495, 46320, 50515, 55109, 60076, 65535, 35744, 38968, 42495, 46341, 50535, 55099, 60087,
    65535, 35734, 38973, 42500, 46341, 50535, 55109, 60097, 32771, 35734, 38965, 42497, 46341, 50535, 55109, 60099,
    32768, 36061, 38968, 42495, 46341, 50535, 55437, 60424, 32932, 35734, 38968, 42495, 46177, 50535, 55109, 59933 };
const uint8_t kIQS[80] = { 15, 15, 15, 15, 15, 15, 15, 15, 14, 14, 14, 14, 14, 14, 14, 14, 14, 13, 13, 13,
                           13, 13, 13, 13, 12, 12, 12, 12, 12, 12, 12, 12, 12, 11, 11, 11, 11, 11, 11, 11,
                           11, 10, 10, 10, 10, 10, 10, 10, 10, 9, 9, 9, 9, 9, 9, 9, 8, 8, 8, 8,
                           8, 8, 8, 8, 7, 7, 7, 7, 7, 7, 7, 7, 6, 6, 6, 6, 6, 6, 6, 6 };

int log2n(int n) { return n == 4 ? 2 : n == 8 ? 3 : n == 16 ? 4 : 5; }

// tu_quant_forward / tu_quant_inverse (rdo.c:658-718) for 8-bit video.
QParams qparams(int n, int qp, int b_intra)
{
    QParams q;
    const int l = log2n(n);
    q.qshift = 15 + 16 - 9 - l;
    q.qscale = kQ[qp];
    q.qadd = (int)(((int64_t)1 << (q.qshift + b_intra)) * 5 / 31);
    q.dscale = kIQ[qp];
    q.dshift = kIQS[qp] + 9 + l - 16;
    return q;
}

int launch(int n, const int16_t *src, int16_t *level, int16_t *recon, int32_t *nnz, int nblk, QParams q, int mode,
           hipStream_t s)
{
    std::call_once(g_tab_once, upload_tables);
    if (g_tab_err != hipSuccess)
        return xme_hip_fail(g_tab_err, "upload transform tables");
    const int fshift1 = log2n(n) - 2;
    switch (n) {
    case 4:
        hipLaunchKernelGGL(residual_kernel<4>, dim3((nblk + Geo<4>::NB - 1) / Geo<4>::NB), dim3(256), 0, s, src,
                           level, recon, nnz, nblk, q, mode, fshift1);
        break;
    case 8:
        hipLaunchKernelGGL(residual_kernel<8>, dim3((nblk + Geo<8>::NB - 1) / Geo<8>::NB), dim3(256), 0, s, src,
                           level, recon, nnz, nblk, q, mode, fshift1);
        break;
    case 16:
        hipLaunchKernelGGL(residual_kernel<16>, dim3((nblk + Geo<16>::NB - 1) / Geo<16>::NB), dim3(256), 0, s, src,
                           level, recon, nnz, nblk, q, mode, fshift1);
        break;
    default:
        hipLaunchKernelGGL(residual_kernel<32>, dim3((nblk + Geo<32>::NB - 1) / Geo<32>::NB), dim3(256), 0, s, src,
                           level, recon, nnz, nblk, q, mode, fshift1);
        break;
    }
    XME_HIP(hipGetLastError());
    return XME_OK;
}

int check(int n, int nblk, const void *a, const void *b, int qp, const char *what)
{
    if ((n != 4 && n != 8 && n != 16 && n != 32) || nblk < 0 || !a || !b || qp < 0 || qp > 79 ||
        (reinterpret_cast<uintptr_t>(a) & 15) || (reinterpret_cast<uintptr_t>(b) & 15)) {
        xme_set_error("%s: n must be 4/8/16/32, qp 0..79, buffers 16-byte aligned", what);
        return XME_ERR_ARG;
    }
    return XME_OK;
}

} // namespace

extern "C" int xme_residual_batch(int n, int qp, int b_intra, const coeff_t *in, coeff_t *level, coeff_t *recon,
                                  int32_t *nnz, int nblk, void *stream)
{
    int rc = check(n, nblk, in, level, qp, "xme_residual_batch");
    if (rc || nblk == 0)
        return rc;
    if (!recon || (reinterpret_cast<uintptr_t>(recon) & 15))
        return XME_ERR_ARG;
    return launch(n, in, level, recon, nnz, nblk, qparams(n, qp, b_intra ? 1 : 0),
                  DO_FWD | DO_QUANT | DO_DEQUANT | DO_INV, (hipStream_t)stream);
}

extern "C" int xme_dct_batch(int n, const coeff_t *src, coeff_t *dst, int nblk, void *stream)
{
    int rc = check(n, nblk, src, dst, 0, "xme_dct_batch");
    if (rc || nblk == 0)
        return rc;
    return launch(n, src, dst, nullptr, nullptr, nblk, qparams(n, 0, 0), DO_FWD, (hipStream_t)stream);
}

extern "C" int xme_idct_batch(int n, const coeff_t *src, coeff_t *dst, int nblk, void *stream)
{
    int rc = check(n, nblk, src, dst, 0, "xme_idct_batch");
    if (rc || nblk == 0)
        return rc;
    return launch(n, src, nullptr, dst, nullptr, nblk, qparams(n, 0, 0), DO_INV, (hipStream_t)stream);
}

namespace {
__global__ void __launch_bounds__(256) quant_kernel(int16_t *coef, int per_blk, int nblk, QParams q, int quant,
                                                    int32_t *nnz)
{
    const int b = blockIdx.x;
    if (b >= nblk)
        return;
    int16_t *c = coef + (int64_t)b * per_blk;
    int cnt = 0;
    for (int i = threadIdx.x; i < per_blk; i += 256) {
        int v = c[i];
        if (quant) {
            const int a = v < 0 ? -v : v;
            v = (int16_t)((v < 0 ? -1 : 1) * ((a * q.qscale + q.qadd) >> q.qshift));
            cnt += v != 0;
        } else if (v != 0) {
            v = xme_clip3(-32768, 32767, (v * q.dscale + (1 << (q.dshift - 1))) >> q.dshift);
        }
        c[i] = (int16_t)v;
    }
    if (quant && nnz) {
        __shared__ int s;
        if (threadIdx.x == 0)
            s = 0;
        __syncthreads();
        atomicAdd(&s, cnt);
        __syncthreads();
        if (threadIdx.x == 0)
            nnz[b] = s;
    }
}
} // namespace

extern "C" int xme_quant_batch(int n, coeff_t *coef, int nblk, int qp, int b_intra, int32_t *nnz, void *stream)
{
    int rc = check(n, nblk, coef, coef, qp, "xme_quant_batch");
    if (rc || nblk == 0)
        return rc;
    hipLaunchKernelGGL(quant_kernel, dim3(nblk), dim3(256), 0, (hipStream_t)stream, coef, n * n, nblk,
                       qparams(n, qp, b_intra ? 1 : 0), 1, nnz);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_dequant_batch(int n, coeff_t *coef, int nblk, int qp, void *stream)
{
    int rc = check(n, nblk, coef, coef, qp, "xme_dequant_batch");
    if (rc || nblk == 0)
        return rc;
    hipLaunchKernelGGL(quant_kernel, dim3(nblk), dim3(256), 0, (hipStream_t)stream, coef, n * n, nblk,
                       qparams(n, qp, 0), 0, nullptr);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
