// Residual path for gfx950: integer DCT (dct_NxN_c), flat quantisation
// (quant_c), dequantisation (dequant_c) and integer IDCT (idct_NxN_c) for
// N = 4, 8, 16, 32 (source/common/transform.c:167-1312, source/common/quant.c:109-184,
// parameters of tu_quant_forward / tu_quant_inverse, source/encoder/rdo.c:658-718).
//
// Every 1-D pass of the reference's partial butterflies is an exact integer
// matrix product, so each pass runs on the matrix cores:
//
//  * A wave owns a 32x32 tile holding (32/N)^2 blocks.  A block-diagonal tile
//    matrix M (T_N repeated on the diagonal; |T| <= 90 fits int8) transforms
//    every block of the tile at once.
//  * A data operand (int16) is split exactly into int8 halves:
//    v = 256 * hi + lo, hi = v >> 8 (signed), lo = v & 255.  With
//    lo_s = lo - 128 (signed) each pass is
//    D = M.lo_s + 128 * (sums of M) + 256 * M.hi, i.e. two
//    v_mfma_i32_32x32x32_i8 and a shift-add.  The 128 * sums term and the
//    rounding offset start the accumulator.
//  * The MFMA result layout (lane = column, registers = rows
//    (i % 4) + 8 (i / 4) + 4 (lane / 32)) is used directly as the next pass's
//    data operand: the summation index K is permuted the same way in the
//    constant operand, so no pass needs a transpose.
//
// Forward:  D1 = X . M^T           (tmp^T; shift log2N - 2, int16 wrap)
//           D2 = M . D1            (coefficients; shift log2N + 5, int16 wrap)
//           quantise, dequantise in registers
// Inverse:  D3 = C^T . M           (first pass; shift 5, clip to int16)
//           D4 = D3 . M            (residual; shift 12, clip to [-256, 255])
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "xme_mc.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// quant_c / dequant_c parameters
struct QParams {
    int qscale, qshift, qadd;
    int dscale, dshift;
};

enum { DO_FWD = 1, DO_QUANT = 2, DO_DEQUANT = 4, DO_INV = 8, DO_HALF = 16 };

// The fused luma inter reconstruction (residual_kernel with FU = 1 or 2):
// each tile's blocks are TUs of the record list {tu_x, tu_y, pu_x, pu_y, pu_w,
// pu_h, mv_x, mv_y}; the prediction comes from the interpolated planes (and
// for FU = 2 the rounded average with planes2 at mv2), the input is cur - pred,
// and the reconstruction clip(pred + residual) goes into rec with its SSD.
// FU = 3 takes the prediction from a picture (planes[0], stride i_ref) at
// the block's own position, blocks listed as {x, y} in `tus`, the plane
// fw x fh: the chroma reconstruction of inter CUs.  Blocks nb1 .. 2 nb1 - 1
// repeat the list on a second plane (planes[1], cur2, rec2; same strides):
// U and V in one launch.
struct FusedArgs {
    const uint8_t *planes[16];
    const uint8_t *planes2[16];
    int64_t i_ref;
    const uint8_t *cur;
    int64_t i_cur;
    uint8_t *rec;
    int64_t i_rec;
    const int32_t *tus;
    const int32_t *mv2;
    int fw, fh;
    int16_t *resid;        // the prediction error, when not null
    int32_t *ssd;
    const uint8_t *cur2;   // FU = 3, second plane
    uint8_t *rec2;
    int64_t nb1;           // FU = 3: blocks per plane
};

// Constant MFMA operands and accumulator starts of one block size, per lane.
struct TileConsts {
    uint32_t f1b[64][4];    // B of D1 = X . M^T: lane col k, K bytes n = 16 (lane / 32) + t: M[k][n]
    uint32_t f2a[64][4];    // A of D2 = M . D1: lane row k2, K bytes j = rho(lane, t): M[k2][j]
    uint32_t ib[64][4];     // B of D3, D4: lane col n, K bytes m = rho(lane, t): M[m][n]
    int32_t rowsum[64];     // 128 * sum_n M[lane % 32][n]
    int32_t colsum[64];     // 128 * sum_m M[m][lane % 32]
    int32_t rowsum_rho[64][16];  // 128 * sum_j M[rho(lane, i)][j]
};
__device__ TileConsts g_tc[4];   // N = 4, 8, 16, 32

__host__ __device__ constexpr int log2n_of(int n) { return n == 4 ? 2 : n == 8 ? 3 : n == 16 ? 4 : 5; }
// tile row held by result register i of a lane in the upper (1) or lower (0) half-wave
__host__ __device__ constexpr int rho(int half, int i) { return (i & 3) + 8 * (i >> 2) + 4 * half; }

__device__ __forceinline__ v16i splat16(int v)
{
    v16i r;
#pragma unroll
    for (int i = 0; i < 16; i++)
        r[i] = v;
    return r;
}

// D = data . B (data as the A operand) with data = 256 hi + lo_s + 128
__device__ __forceinline__ v16i mm_data_a(v4i hi, v4i lo, v4i b, v16i acc)
{
    v16i d = __builtin_amdgcn_mfma_i32_32x32x32_i8(lo, b, acc, 0, 0, 0);
    v16i h = __builtin_amdgcn_mfma_i32_32x32x32_i8(hi, b, splat16(0), 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; i++)
        d[i] += h[i] << 8;
    return d;
}

// D = A . data (data as the B operand)
__device__ __forceinline__ v16i mm_data_b(v4i a, v4i hi, v4i lo, v16i acc)
{
    v16i d = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, lo, acc, 0, 0, 0);
    v16i h = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, hi, splat16(0), 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; i++)
        d[i] += h[i] << 8;
    return d;
}

// Splits the low 16 bits of 16 values into the hi / lo_s byte operands
// (K byte t = value t).  Taking bytes 0 and 1 is the int16 wrap.
__device__ __forceinline__ void split16(const int (&v)[16], v4i &hi, v4i &lo)
{
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t t01 = __builtin_amdgcn_perm((uint32_t)v[4 * w + 1], (uint32_t)v[4 * w], 0x04000501u);
        const uint32_t t23 = __builtin_amdgcn_perm((uint32_t)v[4 * w + 3], (uint32_t)v[4 * w + 2], 0x04000501u);
        hi[w] = (int)__builtin_amdgcn_perm(t23, t01, 0x05040100u);
        lo[w] = (int)(__builtin_amdgcn_perm(t23, t01, 0x07060302u) ^ 0x80808080u);
    }
}

// Splits 16 int16 values held as 8 dwords (value 2d in the low half of d).
__device__ __forceinline__ void split_pairs(const uint32_t (&d)[8], v4i &hi, v4i &lo)
{
#pragma unroll
    for (int w = 0; w < 4; w++) {
        hi[w] = (int)__builtin_amdgcn_perm(d[2 * w + 1], d[2 * w], 0x07050301u);
        lo[w] = (int)(__builtin_amdgcn_perm(d[2 * w + 1], d[2 * w], 0x06040200u) ^ 0x80808080u);
    }
}

__device__ __forceinline__ v4i load4(const uint32_t (&p)[4])
{
    v4i r;
    r[0] = (int)p[0];
    r[1] = (int)p[1];
    r[2] = (int)p[2];
    r[3] = (int)p[3];
    return r;
}

// Offset (in coefficients) of element (y, x) of tile t, split into its row
// and column parts: t * 1024 + rel_y(y) + rel_x(x) (a tile is NB x NB blocks
// of N x N, each block contiguous, row-major within).  A lane's rows rho(half, i) give
// rel_y(rho(half, i)) = rel_y(rho(0, i)) + half * rel_y(4) for every N (the
// 4-row step never leaves an N-row block for N >= 8, and moves one block row
// for N = 4), so a lane's 16 elements are one base plus compile-time offsets.
template <int N>
__host__ __device__ constexpr int rel_y(int y)
{
    return (y / N) * (32 / N) * (N * N) + (y % N) * N;
}
template <int N>
__host__ __device__ constexpr int rel_x(int x)
{
    return (x / N) * (N * N) + x % N;
}

// Four prediction bytes at p (any alignment): the aligned dword pair around
// them, realigned.  The dword after stays inside the plane's XME_PAD border.
__device__ __forceinline__ uint32_t load4u(const uint8_t *p)
{
    const uintptr_t a = (uintptr_t)p;
    const uint32_t *q = (const uint32_t *)(a & ~(uintptr_t)3);
    return __builtin_amdgcn_alignbyte(q[1], q[0], (uint32_t)(a & 3));
}

// Four prediction bytes: from the interpolated planes at any alignment
// (load4u, inside their border), or from a prediction picture (FU = 3) at a
// column that is a multiple of 4 of a 4-byte aligned picture, so that no
// read passes the picture's last byte.
template <int FU>
__device__ __forceinline__ uint32_t load_pred(const uint8_t *p)
{
    if constexpr (FU == 3)
        return *reinterpret_cast<const uint32_t *>(p);
    else
        return load4u(p);
}

__device__ __forceinline__ uint64_t bperm64(int src_lane, uint64_t v)
{
    const int lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(uint32_t)v);
    const int hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(uint32_t)(v >> 32));
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// The 5/3 lifting of a 64-sample line held in registers, v[2 + i] = sample
// i (wavelet_64x64_c / inv_wavelet_64x64_c, transform.c:586-842).  Forward:
// mirror two samples a side, the odd samples minus the floor mean of their
// neighbours, then the even ones plus (l + r + 2) >> 2 (rows) or doubled
// plus (l + r + 1) >> 1 (columns); the low band is v[2 + 2 i].  Every stored
// sample is truncated to int16, as the reference's coeff_t line keeps it.
template <bool COLS>
__device__ __forceinline__ void lift64_fwd(int (&v)[68])
{
    v[1] = v[3];
    v[0] = v[4];
    v[66] = v[64];
    v[67] = v[63];
#pragma unroll
    for (int i = -1; i < 64; i += 2)
        v[2 + i] = (int16_t)(v[2 + i] - ((v[1 + i] + v[3 + i]) >> 1));
#pragma unroll
    for (int i = 0; i < 64; i += 2)
        v[2 + i] = COLS ? (int16_t)((v[2 + i] << 1) + ((v[1 + i] + v[3 + i] + 1) >> 1))
                        : (int16_t)(v[2 + i] + ((v[1 + i] + v[3 + i] + 2) >> 2));
}

// Inverse: the 32 low samples at the even places e[2 i], the last mirrored
// (e[64] = e[62]); columns halve the even samples 0..64; the odd samples are
// the floor mean of their neighbours.
template <bool COLS>
__device__ __forceinline__ void lift64_inv(int (&e)[65])
{
    e[64] = e[62];
    if constexpr (COLS) {
#pragma unroll
        for (int i = 0; i <= 64; i += 2)
            e[i] = (int16_t)(e[i] >> 1);
    }
#pragma unroll
    for (int i = 1; i < 64; i += 2)
        e[i] = (int16_t)((e[i - 1] + e[i + 1]) >> 1);
}

// FU = 0: the residual path on int16 blocks from src.  FU = 1 / 2: the fused
// inter reconstruction (FusedArgs), uni- or bi-predicted; src is unused and
// recon (the reconstructed residual) may be null.  WL (N = 32, FU != 0):
// 64x64 blocks, one per tile (dct_64x64 / idct_64x64, transform.c:1514-1541):
// the prediction error's wavelet low band takes the 32x32 transform with the
// wavelet flag (first forward shift + 1, last inverse shift - 1 and a 10-bit
// clip), quantised at level 6, and the inverse wavelet gives the 64x64
// residual; the prediction error and the reconstructed residual are not
// written (resid, recon unused).
template <int N, int FU, bool WL = false>
__global__ void __launch_bounds__(256) residual_kernel(const int16_t *__restrict__ src, int16_t *__restrict__ level,
                                                       int16_t *__restrict__ recon, int32_t *__restrict__ nnz,
                                                       int nblk, QParams q, int mode, FusedArgs fa)
{
    constexpr int NB = 32 / N, BPT = NB * NB, L2 = log2n_of(N);
    constexpr int W = N < 16 ? N : 16, NP = 16 / W;    // a lane's 16 columns: NP pieces of W
    static_assert(!WL || (N == 32 && FU != 0), "the 64x64 path is a fused 32x32 tile");
    constexpr int NS = WL ? 64 : N;    // the block's side in the picture
    __shared__ int16_t s_res[4][FU ? 32 * 32 : 1];
    __shared__ int s_ssd[4][FU ? BPT : 1];
    // WL: the 64x64 block in LDS, rows of LWS int16 (a dword-aligned stride
    // of 33 dwords, so that a lane per row meets no bank twice)
    constexpr int LWS = 66;
    __shared__ int16_t s_lw[4][WL ? 64 * LWS : 1];
    const int wv = threadIdx.x >> 6;
    const TileConsts &tc = g_tc[L2 - 2];
    const int lane = threadIdx.x & 63, col = lane & 31, half = lane >> 5;
    const int ntiles = (nblk + BPT - 1) / BPT;
    const v4i f1b = load4(tc.f1b[lane]), f2a = load4(tc.f2a[lane]), ib = load4(tc.ib[lane]);
    const int s1 = L2 - 2 + (WL ? 1 : 0), s2 = L2 + 5;
    const int acc_f1 = tc.rowsum[lane] + ((1 << s1) >> 1);
    constexpr int SI2 = WL ? 11 : 12, RLIM = WL ? 512 : 256;    // last inverse shift, clip of the residual
    const int acc_i1 = tc.colsum[lane] + 16, acc_i2 = tc.colsum[lane] + (1 << (SI2 - 1));
    v16i acc_f2;
#pragma unroll
    for (int i = 0; i < 16; i++)
        acc_f2[i] = tc.rowsum_rho[lane][i] + (1 << (s2 - 1));

    // workgroup b runs on XCD b % 8: each XCD takes a contiguous run of the
    // tiles (the grid is a multiple of 8), so that neighbouring TUs, which
    // share current and reference lines, meet in one L2
    const int lb = (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
    for (int64_t t = lb * 4 + (threadIdx.x >> 6); t < ntiles; t += (int64_t)gridDim.x * 4) {
        v4i hi, lo;
        // this lane's elements (rho(half, i), col): e_lane + rel_y(rho(0, i)),
        // in block b_lane + (rho(0, i) / N) * NB
        const int64_t e_lane = t * 1024 + rel_x<N>(col) + half * rel_y<N>(4);
        const int64_t b_lane = t * BPT + col / N + (N == 4 ? half * NB : 0);
        // fused: this lane's 16 current and prediction bytes (tile row `col`,
        // columns 16 half .. 16 half + 15), kept for the reconstruction
        uint32_t c4[4] = { 0u, 0u, 0u, 0u }, p4[4] = { 0u, 0u, 0u, 0u };
        bool tu_good = false;            // lane < BPT: its TU of the tile is reconstructed
        uint64_t my_rec = 0;             // lane < BPT: the TU's first reconstructed pixel
        uint64_t wl_cur = 0, wl_pred = 0, wl_pred2 = 0, wl_rec = 0;   // WL: the block's, in every lane
        // WL: this lane's current and prediction dwords (rows 4 i + lane / 16,
        // columns 4 (lane % 16) ..), loaded once for both ends of the path
        uint32_t wl_c[WL ? 16 : 1], wl_p[WL ? 16 : 1];
        if constexpr (FU != 0) {
            // lane b < BPT takes TU t * BPT + b: validity and base pointers
            uint64_t my_pred = 0, my_pred2 = 0, my_cur = 0;
            const int64_t bt = t * BPT + lane;
            if constexpr (FU == 3) {
                // the prediction picture: block (x, y) of every plane
                if (lane < BPT && bt < nblk) {
                    const bool second = bt >= fa.nb1;
                    const int64_t bi = second ? bt - fa.nb1 : bt;
                    const int bx = fa.tus[2 * bi], by = fa.tus[2 * bi + 1];
                    tu_good = (bx & 3) == 0 && bx >= 0 && by >= 0 && bx <= fa.fw - NS && by <= fa.fh - NS;
                    if (tu_good) {
                        my_pred = (uint64_t)(uintptr_t)((second ? fa.planes[1] : fa.planes[0]) +
                                                        (int64_t)by * fa.i_ref + bx);
                        my_cur = (uint64_t)(uintptr_t)((second ? fa.cur2 : fa.cur) + (int64_t)by * fa.i_cur + bx);
                        my_rec = (uint64_t)(uintptr_t)((second ? fa.rec2 : fa.rec) + (int64_t)by * fa.i_rec + bx);
                    }
                }
            } else if (lane < BPT && bt < nblk) {
                int r[8];
#pragma unroll
                for (int k = 0; k < 8; k++)
                    r[k] = fa.tus[8 * bt + k];
                tu_good = tu_ok<NS>(r, fa.fw, fa.fh);
                if (tu_good) {
                    const int qx = mc_pos(fa.fw, r[4], r[2], r[6]) + 4 * (r[0] - r[2]);
                    const int qy = mc_pos(fa.fh, r[5], r[3], r[7]) + 4 * (r[1] - r[3]);
                    my_pred = (uint64_t)(uintptr_t)(fa.planes[((qy & 3) << 2) + (qx & 3)] + (int64_t)(qy >> 2) * fa.i_ref +
                                                    (qx >> 2));
                    if constexpr (FU == 2) {
                        const int qx2 = mc_pos(fa.fw, r[4], r[2], fa.mv2[2 * bt]) + 4 * (r[0] - r[2]);
                        const int qy2 = mc_pos(fa.fh, r[5], r[3], fa.mv2[2 * bt + 1]) + 4 * (r[1] - r[3]);
                        my_pred2 = (uint64_t)(uintptr_t)(fa.planes2[((qy2 & 3) << 2) + (qx2 & 3)] +
                                                         (int64_t)(qy2 >> 2) * fa.i_ref + (qx2 >> 2));
                    }
                    my_cur = (uint64_t)(uintptr_t)(fa.cur + (int64_t)r[1] * fa.i_cur + r[0]);
                    my_rec = (uint64_t)(uintptr_t)(fa.rec + (int64_t)r[1] * fa.i_rec + r[0]);
                }
            }
            if (lane < BPT)
                s_ssd[wv][lane] = 0;
            if constexpr (WL) {
                wl_cur = bperm64(0, my_cur);
                wl_pred = bperm64(0, my_pred);
                wl_pred2 = FU == 2 ? bperm64(0, my_pred2) : 0;
                wl_rec = bperm64(0, my_rec);
            }
            const int r_in = col % N, c0 = (16 * half) % N;
#pragma unroll
            for (int piece = 0; piece < (WL ? 0 : NP); piece++) {
                const int bl = (col / N) * NB + (16 * half) / N + piece;
                const uint64_t pc = bperm64(bl, my_cur), pp = bperm64(bl, my_pred);
                const uint64_t pp2 = FU == 2 ? bperm64(bl, my_pred2) : 0;
                if (pc != 0) {
                    const uint8_t *cp = (const uint8_t *)(uintptr_t)pc + (int64_t)r_in * fa.i_cur + c0;
                    const uint8_t *pr = (const uint8_t *)(uintptr_t)pp + (int64_t)r_in * fa.i_ref + c0;
                    const uint8_t *pr2 = (const uint8_t *)(uintptr_t)pp2 + (int64_t)r_in * fa.i_ref + c0;
#pragma unroll
                    for (int k = 0; k < W / 4; k++) {
                        c4[piece * (W / 4) + k] = reinterpret_cast<const uint32_t *>(cp)[k];
                        uint32_t u = load_pred<FU>(pr + 4 * k);
                        if constexpr (FU == 2) {
                            const uint32_t v2 = load4u(pr2 + 4 * k);
                            u = (u | v2) - (((u ^ v2) >> 1) & 0x7f7f7f7fu);   // (u + v + 1) >> 1 per byte
                        }
                        p4[piece * (W / 4) + k] = u;
                    }
                }
            }
        }
        if (mode & DO_FWD) {
            // F1 operand: tile row `col`, columns 16 half .. 16 half + 15
            uint32_t d[8];
            if constexpr (WL) {
                // the 64x64 prediction error into s_lw, 16 lanes a row (4
                // rows per access); then lane r lifts row r into its first
                // 32 places, and lanes 0..31 lift the columns in place: rows
                // 0..31 then hold the 32x32 low band
                int16_t *lw = s_lw[wv];
                const int qr = lane >> 4, qc = 4 * (lane & 15);
#pragma unroll
                for (int gi = 0; gi < 16; gi++) {
                    const int g = 4 * gi;
                    uint32_t cw = 0, u = 0;
                    if (wl_cur != 0) {
                        cw = *reinterpret_cast<const uint32_t *>((const uint8_t *)(uintptr_t)wl_cur +
                                                                 (int64_t)(g + qr) * fa.i_cur + qc);
                        u = load_pred<FU>((const uint8_t *)(uintptr_t)wl_pred + (int64_t)(g + qr) * fa.i_ref + qc);
                        if constexpr (FU == 2) {
                            const uint32_t v2 = load_pred<FU>((const uint8_t *)(uintptr_t)wl_pred2 +
                                                              (int64_t)(g + qr) * fa.i_ref + qc);
                            u = (u | v2) - (((u ^ v2) >> 1) & 0x7f7f7f7fu);
                        }
                    }
                    wl_c[gi] = cw;
                    wl_p[gi] = u;
                    int e4[4];
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        e4[j] = (int)((cw >> (8 * j)) & 0xff) - (int)((u >> (8 * j)) & 0xff);
                    uint32_t *dst = reinterpret_cast<uint32_t *>(lw + (g + qr) * LWS + qc);
                    dst[0] = ((uint32_t)e4[0] & 0xffffu) | ((uint32_t)e4[1] << 16);
                    dst[1] = ((uint32_t)e4[2] & 0xffffu) | ((uint32_t)e4[3] << 16);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                {
                    int v[68];
                    const uint32_t *src = reinterpret_cast<const uint32_t *>(lw + lane * LWS);
#pragma unroll
                    for (int k = 0; k < 32; k++) {
                        const uint32_t w2 = src[k];
                        v[2 + 2 * k] = (int)(int16_t)(w2 & 0xffffu);
                        v[3 + 2 * k] = (int)(int16_t)(w2 >> 16);
                    }
                    lift64_fwd<false>(v);
                    uint32_t *dst = reinterpret_cast<uint32_t *>(lw + lane * LWS);
#pragma unroll
                    for (int k = 0; k < 16; k++)
                        dst[k] = ((uint32_t)v[2 + 4 * k] & 0xffffu) | ((uint32_t)v[4 + 4 * k] << 16);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                {
                    // column lane % 32, low-band half lane / 32: rows -2 .. 33
                    // or 30 .. 65 (mirrored at the ends as the lifting
                    // mirrors them) give low samples 0..15 or 16..31; every
                    // lane reads before any writes, so the halves may share
                    // rows 16..33
                    const int cc = lane & 31, b0 = (lane >> 5) ? 30 : -2;
                    int v[36];
#pragma unroll
                    for (int j = 0; j < 36; j++) {
                        int r = b0 + j;
                        r = r < 0 ? -r : (r > 63 ? 126 - r : r);
                        v[j] = lw[r * LWS + cc];
                    }
#pragma unroll
                    for (int j = 1; j < 35; j += 2)
                        v[j] = (int16_t)(v[j] - ((v[j - 1] + v[j + 1]) >> 1));
#pragma unroll
                    for (int j = 2; j <= 32; j += 2)
                        v[j] = (int16_t)((v[j] << 1) + ((v[j - 1] + v[j + 1] + 1) >> 1));
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                    for (int m = 0; m < 16; m++)
                        lw[((b0 + 2) / 2 + m) * LWS + cc] = (int16_t)v[2 + 2 * m];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const uint32_t *row = reinterpret_cast<const uint32_t *>(lw + col * LWS + 16 * half);
#pragma unroll
                for (int k = 0; k < 8; k++)
                    d[k] = row[k];
            } else if constexpr (FU != 0) {
                // the prediction error cur - pred (sub_ps, pixel.c:780-792)
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const uint32_t cw = c4[k / 2] >> (16 * (k & 1)), pw = p4[k / 2] >> (16 * (k & 1));
                    const int e0 = (int)(cw & 0xff) - (int)(pw & 0xff);
                    const int e1 = (int)((cw >> 8) & 0xff) - (int)((pw >> 8) & 0xff);
                    d[k] = ((uint32_t)e0 & 0xffffu) | ((uint32_t)e1 << 16);
                }
                if (fa.resid) {
                    const int64_t blk0 = t * BPT + (col / N) * NB + (16 * half) / N;
                    int16_t *p0 = fa.resid + t * 1024 + rel_y<N>(col) + rel_x<N>(16 * half);
#pragma unroll
                    for (int piece = 0; piece < NP; piece++)
#pragma unroll
                        for (int k = 0; k < W / 2; k++)
                            if (blk0 + piece < nblk)
                                reinterpret_cast<uint32_t *>(p0 + piece * (N * N))[k] = d[piece * (W / 2) + k];
                }
            } else {
                // pieces of W contiguous coefficients: row `col`, columns
                // 16 half + piece W, i.e. blocks (col / N, 16 half / N + piece)
                const int16_t *p0 = src + t * 1024 + rel_y<N>(col) + rel_x<N>(16 * half);
                const int64_t blk0 = t * BPT + (col / N) * NB + (16 * half) / N;
#pragma unroll
                for (int piece = 0; piece < NP; piece++) {
                    const uint32_t *p = reinterpret_cast<const uint32_t *>(p0 + piece * (N * N));
#pragma unroll
                    for (int k = 0; k < W / 2; k++)
                        d[piece * (W / 2) + k] = blk0 + piece < nblk ? p[k] : 0u;
                }
            }
            split_pairs(d, hi, lo);
            v16i D = mm_data_a(hi, lo, f1b, splat16(acc_f1));
            int v[16];
#pragma unroll
            for (int i = 0; i < 16; i++)
                v[i] = D[i] >> s1;
            split16(v, hi, lo);
            D = mm_data_b(f2a, hi, lo, acc_f2);
            // coefficients: row rho(half, i), column col.  Nonzero counts per
            // block: lane b < BPT adds the nonzeros of its block (row block b / NB,
            // column block b % NB) from each row's ballot as it is made, so no
            // ballot stays live (bits 0-31 hold row rho(0, i), bits 32-63 row
            // rho(1, i), bit = column)
            int c[16];
            const int my_rb = lane / NB, my_cb = lane % NB;
            const uint32_t my_cm = N == 32 ? 0xffffffffu : ((1u << N) - 1u) << (my_cb * N);
            int mine = 0;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                c[i] = (int)(int16_t)(D[i] >> s2);
                // dct_32x32_half_c (transform.c:1276-1286): the low 16x16 only
                if (N == 32 && (mode & DO_HALF) && (rho(half, i) >= 16 || col >= 16))
                    c[i] = 0;
                if (mode & DO_QUANT) {
                    const int a = c[i] < 0 ? -c[i] : c[i];
                    const int m = (int)(((uint32_t)a * (uint32_t)q.qscale + (uint32_t)q.qadd) >> q.qshift);
                    c[i] = (int)(int16_t)(c[i] < 0 ? -m : m);
                    const uint64_t nz = __builtin_amdgcn_ballot_w64(c[i] != 0);
                    if (rho(0, i) / N == my_rb)
                        mine += __builtin_popcount((uint32_t)nz & my_cm);
                    if (rho(1, i) / N == my_rb)
                        mine += __builtin_popcount((uint32_t)(nz >> 32) & my_cm);
                }
            }
            // levels (or plain coefficients) out
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (b_lane + (rho(0, i) / N) * NB < nblk)
                    level[e_lane + rel_y<N>(rho(0, i))] = (int16_t)c[i];
            if ((mode & DO_QUANT) && nnz && lane < BPT && t * BPT + lane < nblk)
                nnz[t * BPT + lane] = mine;
            if (mode & DO_INV) {
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    int d = c[i];
                    if ((mode & DO_DEQUANT) && d != 0)
                        d = xme_clip3(-32768, 32767, (d * q.dscale + (1 << (q.dshift - 1))) >> q.dshift);
                    c[i] = d;
                }
                split16(c, hi, lo);
            }
        } else {
            // inverse only: I1 operand lane row j = col, K bytes m = rho(half, t)
            int v[16];
#pragma unroll
            for (int i = 0; i < 16; i++)
                v[i] = b_lane + (rho(0, i) / N) * NB < nblk ? src[e_lane + rel_y<N>(rho(0, i))] : 0;
            split16(v, hi, lo);
        }
        if (mode & DO_INV) {
            v16i D = mm_data_a(hi, lo, ib, splat16(acc_i1));
            int v[16];
#pragma unroll
            for (int i = 0; i < 16; i++)
                v[i] = xme_clip3(-32768, 32767, D[i] >> 5);
            split16(v, hi, lo);
            D = mm_data_a(hi, lo, ib, splat16(acc_i2));
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int rv = xme_clip3(-RLIM, RLIM - 1, D[i] >> SI2);
                if (!WL && recon && b_lane + (rho(0, i) / N) * NB < nblk)
                    recon[e_lane + rel_y<N>(rho(0, i))] = (int16_t)rv;
                if constexpr (FU != 0)
                    s_res[wv][rho(half, i) * 32 + col] = (int16_t)rv;
            }
        }
        if constexpr (WL) {
            // the inverse wavelet of the 32x32 residual in s_res: lanes 0..31
            // a column each into s_lw (64 rows of 32), lane r the row r (64
            // values back into its row); then 16 lanes a row add the
            // prediction, clip, store and sum the block's SSD
            int16_t *lw = s_lw[wv];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            {
                // column lane % 32, half lane / 32: low samples 0..16 give rows
                // 0..31, low samples 16..31 (the last mirrored) rows 32..63;
                // the even samples halved, the odd ones the floor mean
                const int cc = lane & 31, k0 = (lane >> 5) ? 16 : 0;
                int ev[17];
#pragma unroll
                for (int m = 0; m < 17; m++) {
                    const int k = k0 + m > 31 ? 31 : k0 + m;
                    ev[m] = (int16_t)((int)s_res[wv][k * 32 + cc] >> 1);
                }
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    lw[(2 * (k0 + m)) * LWS + cc] = (int16_t)ev[m];
                    lw[(2 * (k0 + m) + 1) * LWS + cc] = (int16_t)((ev[m] + ev[m + 1]) >> 1);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            {
                int e[65];
                const uint32_t *src = reinterpret_cast<const uint32_t *>(lw + lane * LWS);
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    const uint32_t w2 = src[k];
                    e[4 * k] = (int)(int16_t)(w2 & 0xffffu);
                    e[4 * k + 2] = (int)(int16_t)(w2 >> 16);
                }
                lift64_inv<false>(e);
                uint32_t *dst = reinterpret_cast<uint32_t *>(lw + lane * LWS);
#pragma unroll
                for (int k = 0; k < 32; k++)
                    dst[k] = ((uint32_t)e[2 * k] & 0xffffu) | ((uint32_t)e[2 * k + 1] << 16);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            int e2 = 0;
            if (wl_rec != 0) {
                const int qr = lane >> 4, qc = 4 * (lane & 15);
#pragma unroll
                for (int gi = 0; gi < 16; gi++) {
                    const int r = 4 * gi + qr;
                    const uint32_t cw = wl_c[gi], u = wl_p[gi];
                    const uint32_t *res = reinterpret_cast<const uint32_t *>(lw + r * LWS + qc);
                    const uint32_t r01 = res[0], r23 = res[1];
                    const int rv[4] = { (int)(int16_t)(r01 & 0xffffu), (int)(int16_t)(r01 >> 16),
                                        (int)(int16_t)(r23 & 0xffffu), (int)(int16_t)(r23 >> 16) };
                    uint32_t o = 0;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int v = clip_pel((int)((u >> (8 * j)) & 0xff) + rv[j]);
                        const int dc = (int)((cw >> (8 * j)) & 0xff) - v;
                        o |= (uint32_t)v << (8 * j);
                        e2 += dc * dc;
                    }
                    *reinterpret_cast<uint32_t *>((uint8_t *)(uintptr_t)wl_rec + (int64_t)r * fa.i_rec + qc) = o;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
                e2 += __shfl_xor(e2, off, 64);
            if (lane == 0 && t < nblk)
                fa.ssd[t] = tu_good ? e2 : -1;
            // the next tile rewrites s_lw and s_res only after these reads
            __builtin_amdgcn_wave_barrier();
        } else if constexpr (FU != 0) {
            // the residual, now by rows: this lane's 16 columns of row `col`;
            // reconstruction clip(pred + residual) (add_ps, pixel.c:794-806)
            // and its SSD (pixf.ssd, rdo.c:1497-1499) per TU
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int r_in = col % N, c0 = (16 * half) % N;
#pragma unroll
            for (int piece = 0; piece < NP; piece++) {
                const int bl = (col / N) * NB + (16 * half) / N + piece;
                const uint64_t prc = bperm64(bl, my_rec);
                if (prc == 0)
                    continue;
                uint8_t *rp = (uint8_t *)(uintptr_t)prc + (int64_t)r_in * fa.i_rec + c0;
                int e = 0;
#pragma unroll
                for (int k = 0; k < W / 4; k++) {
                    const int kk = piece * (W / 4) + k;
                    uint32_t o = 0;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int x = 16 * half + 4 * kk + j;
                        const int pv = (int)((p4[kk] >> (8 * j)) & 0xff), cv = (int)((c4[kk] >> (8 * j)) & 0xff);
                        const int v = clip_pel(pv + s_res[wv][col * 32 + x]);
                        o |= (uint32_t)v << (8 * j);
                        e += (cv - v) * (cv - v);
                    }
                    reinterpret_cast<uint32_t *>(rp)[k] = o;
                }
                atomicAdd(&s_ssd[wv][bl], e);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < BPT && t * BPT + lane < nblk)
                fa.ssd[t * BPT + lane] = tu_good ? s_ssd[wv][lane] : -1;
            // the next tile rewrites s_res and s_ssd only after these reads
            __builtin_amdgcn_wave_barrier();
        }
    }
}

std::once_flag g_tab_once;
hipError_t g_tab_err = hipSuccess;

// AVS2 DCT basis: 32 for the DC row, else round(32*sqrt(2)*cos((2n+1)k*pi/(2N))),
// the values of g_T4..g_T32 (transform.c:59-135); T_N[k][n] = T_32[k*32/N][n].
void upload_tables()
{
    int T32[32][32];
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < 32; k++)
        for (int n = 0; n < 32; n++)
            T32[k][n] = k == 0 ? 32 : (int)std::lround(32.0 * std::sqrt(2.0) * std::cos((2 * n + 1) * k * pi / 64.0));
    static TileConsts tc[4];
    for (int l2 = 2; l2 <= 5; l2++) {
        const int n = 1 << l2, s = 32 / n;
        TileConsts &c = tc[l2 - 2];
        // block-diagonal tile matrix
        auto M = [&](int r, int q) { return r / n == q / n ? T32[(r % n) * s][q % n] : 0; };
        for (int lane = 0; lane < 64; lane++) {
            const int col = lane & 31, half = lane >> 5;
            uint8_t f1b[16], f2a[16], ib[16];
            for (int t = 0; t < 16; t++) {
                f1b[t] = (uint8_t)(int8_t)M(col, 16 * half + t);
                f2a[t] = (uint8_t)(int8_t)M(col, rho(half, t));
                ib[t] = (uint8_t)(int8_t)M(rho(half, t), col);
            }
            memcpy(c.f1b[lane], f1b, 16);
            memcpy(c.f2a[lane], f2a, 16);
            memcpy(c.ib[lane], ib, 16);
            int rs = 0, cs = 0;
            for (int k = 0; k < 32; k++) {
                rs += M(col, k);
                cs += M(k, col);
            }
            c.rowsum[lane] = 128 * rs;
            c.colsum[lane] = 128 * cs;
            for (int i = 0; i < 16; i++) {
                int r = 0;
                for (int k = 0; k < 32; k++)
                    r += M(rho(half, i), k);
                c.rowsum_rho[lane][i] = 128 * r;
            }
        }
    }
    g_tab_err = hipMemcpyToSymbol(HIP_SYMBOL(g_tc), tc, sizeof(tc));
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

int log2n(int n) { return n == 4 ? 2 : n == 8 ? 3 : n == 16 ? 4 : n == 32 ? 5 : 6; }

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

template <int FU>
int launch_fu(int n, const int16_t *src, int16_t *level, int16_t *recon, int32_t *nnz, int nblk, QParams q, int mode,
              const FusedArgs &fa, hipStream_t s)
{
    std::call_once(g_tab_once, upload_tables);
    if (g_tab_err != hipSuccess)
        return xme_hip_fail(g_tab_err, "upload transform tables");
    // one wave per 32x32 tile of (32/n)^2 blocks (a 64x64 block's tile is its
    // wavelet low band); waves loop over tiles
    const int bpt = n >= 32 ? 1 : (32 / n) * (32 / n);
    const int64_t tiles = ((int64_t)nblk + bpt - 1) / bpt;
    // a multiple of 8 workgroups, one range of tiles per XCD
    const int grid = (int)((std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 256 * 16)) + 7) / 8 * 8);
    switch (n) {
    case 4:
        hipLaunchKernelGGL((residual_kernel<4, FU>), dim3(grid), dim3(256), 0, s, src, level, recon, nnz, nblk, q, mode,
                           fa);
        break;
    case 8:
        hipLaunchKernelGGL((residual_kernel<8, FU>), dim3(grid), dim3(256), 0, s, src, level, recon, nnz, nblk, q, mode,
                           fa);
        break;
    case 16:
        hipLaunchKernelGGL((residual_kernel<16, FU>), dim3(grid), dim3(256), 0, s, src, level, recon, nnz, nblk, q,
                           mode, fa);
        break;
    case 32:
        hipLaunchKernelGGL((residual_kernel<32, FU>), dim3(grid), dim3(256), 0, s, src, level, recon, nnz, nblk, q,
                           mode, fa);
        break;
    default:
        if constexpr (FU != 0)
            hipLaunchKernelGGL((residual_kernel<32, FU, true>), dim3(grid), dim3(256), 0, s, src, level, nullptr, nnz,
                               nblk, q, mode, fa);
        break;
    }
    XME_HIP(hipGetLastError());
    return XME_OK;
}

// tab_qp_scale_chroma, source/common/cudata.c:98-106.
const uint8_t kQpScaleChroma[64] = { 0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15,
                                     16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31,
                                     32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 42, 43, 43, 44, 44,
                                     45, 45, 46, 46, 47, 47, 48, 48, 48, 49, 49, 49, 50, 50, 50, 51 };

int launch(int n, const int16_t *src, int16_t *level, int16_t *recon, int32_t *nnz, int nblk, QParams q, int mode,
           hipStream_t s)
{
    return launch_fu<0>(n, src, level, recon, nnz, nblk, q, mode, FusedArgs{}, s);
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

int xme_residual_run(int n, int qp, int b_intra, int half, const int16_t *in, int16_t *level, int16_t *recon,
                     int32_t *nnz, int nblk, hipStream_t s)
{
    return launch(n, in, level, recon, nnz, nblk, qparams(n, qp, b_intra ? 1 : 0),
                  DO_FWD | DO_QUANT | DO_DEQUANT | DO_INV | (half ? DO_HALF : 0), s);
}

int xme_recon_fused(int n, int qp, int half, int bi, const uint8_t *const planes[16], const uint8_t *const planes2[16],
                    int64_t i_ref, const uint8_t *cur, int64_t i_cur, uint8_t *rec, int64_t i_rec, const int32_t *tus,
                    const int32_t *mv2, int ntu, int fw, int fh, int16_t *resid, int16_t *level, int16_t *rres,
                    int32_t *nnz, int32_t *ssd, hipStream_t s)
{
    FusedArgs fa{};
    for (int k = 0; k < 16; k++) {
        fa.planes[k] = planes[k];
        fa.planes2[k] = bi ? planes2[k] : nullptr;
    }
    fa.i_ref = i_ref;
    fa.cur = cur;
    fa.i_cur = i_cur;
    fa.rec = rec;
    fa.i_rec = i_rec;
    fa.tus = tus;
    fa.mv2 = mv2;
    fa.fw = fw;
    fa.fh = fh;
    fa.resid = resid;
    fa.ssd = ssd;
    const int mode = DO_FWD | DO_QUANT | DO_DEQUANT | DO_INV | (half ? DO_HALF : 0);
    const QParams q = qparams(n, qp, 0);
    return bi ? launch_fu<2>(n, nullptr, level, rres, nnz, ntu, q, mode, fa, s)
              : launch_fu<1>(n, nullptr, level, rres, nnz, ntu, q, mode, fa, s);
}

extern "C" int xme_chroma_qp(int luma_qp) { return kQpScaleChroma[luma_qp < 0 ? 0 : (luma_qp > 63 ? 63 : luma_qp)]; }

namespace {

// U and V (or one plane when cur_v is null) of nblk blocks in one launch.
int recon_pred(const char *who, int n_arg, int qp, const pel_t *cur_u, const pel_t *cur_v, intptr_t i_cur,
               const pel_t *pred_u, const pel_t *pred_v, intptr_t i_pred, int plane_w, int plane_h,
               const int32_t *blocks, int nblk, coeff_t *level, pel_t *rec_u, pel_t *rec_v, intptr_t i_rec,
               int32_t *nnz, int32_t *ssd, hipStream_t s)
{
    const int n = n_arg & ~XME_TU_DCT_HALF, half = (n_arg & XME_TU_DCT_HALF) ? 1 : 0;
    if ((n != 4 && n != 8 && n != 16 && n != 32 && n != 64) || (half && n < 32) || qp < 0 || qp > 79 || nblk < 0 ||
        plane_w <= 0 || plane_h <= 0) {
        xme_set_error("%s: bad arguments (n 4, 8, 16, 32 or 64; XME_TU_DCT_HALF with 32 or 64 only; qp 0..79)", who);
        return XME_ERR_ARG;
    }
    if (nblk == 0)
        return XME_OK;
    const bool two = cur_v != nullptr;
    if (!cur_u || !pred_u || !blocks || !level || !rec_u || !nnz || !ssd || (two && (!pred_v || !rec_v))) {
        xme_set_error("%s: NULL buffer", who);
        return XME_ERR_ARG;
    }
    // every picture is read or written as aligned dwords
    const uintptr_t al = (uintptr_t)cur_u | (uintptr_t)rec_u | (uintptr_t)pred_u |
                         (two ? ((uintptr_t)cur_v | (uintptr_t)rec_v | (uintptr_t)pred_v) : 0);
    if ((al & 3) || (i_cur & 3) || (i_rec & 3) || (i_pred & 3) || ((uintptr_t)level & 15)) {
        xme_set_error("%s: pictures must be 4-byte and the levels 16-byte aligned", who);
        return XME_ERR_ARG;
    }
    FusedArgs fa{};
    fa.planes[0] = pred_u;
    fa.planes[1] = two ? pred_v : pred_u;
    fa.i_ref = i_pred;
    fa.cur = cur_u;
    fa.cur2 = two ? cur_v : cur_u;
    fa.i_cur = i_cur;
    fa.rec = rec_u;
    fa.rec2 = two ? rec_v : rec_u;
    fa.i_rec = i_rec;
    fa.tus = blocks;
    fa.nb1 = nblk;
    fa.fw = plane_w;
    fa.fh = plane_h;
    fa.ssd = ssd;
    const int mode = DO_FWD | DO_QUANT | DO_DEQUANT | DO_INV | (half ? DO_HALF : 0);
    return launch_fu<3>(n, nullptr, level, nullptr, nnz, two ? 2 * nblk : nblk, qparams(n, qp, 0), mode, fa, s);
}

} // namespace

extern "C" int xme_recon_pred_batch(int n, int qp, const pel_t *d_cur, intptr_t i_cur, const pel_t *d_pred,
                                    intptr_t i_pred, int plane_w, int plane_h, const int32_t *d_blocks, int nblk,
                                    coeff_t *d_level, pel_t *d_rec, intptr_t i_rec, int32_t *d_nnz, int32_t *d_ssd,
                                    void *stream)
{
    return recon_pred("xme_recon_pred_batch", n, qp, d_cur, nullptr, i_cur, d_pred, nullptr, i_pred, plane_w, plane_h,
                      d_blocks, nblk, d_level, d_rec, nullptr, i_rec, d_nnz, d_ssd, (hipStream_t)stream);
}

extern "C" int xme_recon_chroma_batch(int n, int qp, const pel_t *d_cur_u, const pel_t *d_cur_v, intptr_t i_cur,
                                      const pel_t *d_pred_u, const pel_t *d_pred_v, intptr_t i_pred, int plane_w,
                                      int plane_h, const int32_t *d_blocks, int nblk, coeff_t *d_level,
                                      pel_t *d_rec_u, pel_t *d_rec_v, intptr_t i_rec, int32_t *d_nnz, int32_t *d_ssd,
                                      void *stream)
{
    if (nblk > 0 && !d_cur_v) {
        xme_set_error("xme_recon_chroma_batch: NULL buffer");
        return XME_ERR_ARG;
    }
    if ((int64_t)nblk * 2 > INT32_MAX) {
        xme_set_error("xme_recon_chroma_batch: too many blocks");
        return XME_ERR_ARG;
    }
    return recon_pred("xme_recon_chroma_batch", n, qp, d_cur_u, d_cur_v, i_cur, d_pred_u, d_pred_v, i_pred, plane_w,
                      plane_h, d_blocks, nblk, d_level, d_rec_u, d_rec_v, i_rec, d_nnz, d_ssd, (hipStream_t)stream);
}

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
