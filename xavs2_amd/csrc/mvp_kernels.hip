// MV predictors (get_mvp_default, source/encoder/md_inter.c:470-524) for a
// batch of PUs, and the UMH search's encoder state (get_mvp_default_sad,
// md_inter.c:527-608, and fast_me_prepare_info, md_inter.c:640-718) for a
// batch of (PU, reference) jobs, gfx950.  One thread per record: the work is
// a few dozen integer and floating-point operations, so a batch is bound by
// reading its records.
#include "xme_internal.h"

namespace {

struct MvpTables {
    int32_t dpoc[XME_MVP_REFS], multi[XME_MVP_REFS];
};

__device__ __forceinline__ int iabs_(int v) { return v < 0 ? -v : v; }

// get_mvp_type_default (md_inter.c:333-368): 0 median, 1 left, 2 top, 3 top-right.
__device__ __forceinline__ int mvp_type(int ref, int rl, int ru, int rur, int cx, int cy, int cw, int ch)
{
    if (rl != -1 && ru == -1 && rur == -1)
        return 1;
    if (rl == -1 && ru != -1 && rur == -1)
        return 2;
    if (rl == -1 && ru == -1 && rur != -1)
        return 3;
    if (cw < ch)
        return cx == 0 ? (rl == ref ? 1 : 0) : (rur == ref ? 3 : 0);
    if (cw > ch)
        return cy == 0 ? (ru == ref ? 2 : 0) : (rl == ref ? 1 : 0);
    return 0;
}

// scale_mv_default (md_inter.c:405-410); *bad where the product leaves int.
__device__ __forceinline__ int scale_default(int mv, int dist, int scale, bool &bad)
{
    const int64_t prod = (int64_t)iabs_(mv) * dist * scale;
    bad |= prod + 8192 > 0x7fffffffLL || prod < 0;
    const int v = (mv > 0 ? 1 : (mv < 0 ? -1 : 0)) * (int)((prod + 8192) >> 14);
    return min(max(v, -32768), 32767);
}

// derive_median_mv (md_inter.c:370-402): which pair it averages, 0 (a, b),
// 1 (b, c) or 2 (c, a), as its mvp_type return.
__device__ __forceinline__ int median_src(int a, int b, int c)
{
    if ((a < 0 && b > 0 && c > 0) || (a > 0 && b < 0 && c < 0))
        return 1;
    if ((b < 0 && a > 0 && c > 0) || (b > 0 && a < 0 && c < 0))
        return 2;
    if ((c < 0 && a > 0 && b > 0) || (c > 0 && a < 0 && b < 0))
        return 0;
    const int dab = iabs_(a - b), dbc = iabs_(b - c), dca = iabs_(c - a);
    const int m = min(dab, min(dbc, dca));
    return m == dab ? 0 : (m == dbc ? 1 : 2);
}

// derive_median_mv (md_inter.c:372-401); C's division truncates toward zero.
__device__ __forceinline__ int median_mv(int a, int b, int c)
{
    if ((a < 0 && b > 0 && c > 0) || (a > 0 && b < 0 && c < 0))
        return (int16_t)((b + c) / 2);
    if ((b < 0 && a > 0 && c > 0) || (b > 0 && a < 0 && c < 0))
        return (int16_t)((c + a) / 2);
    if ((c < 0 && a > 0 && b > 0) || (c > 0 && a < 0 && b < 0))
        return (int16_t)((a + b) / 2);
    const int dab = iabs_(a - b), dbc = iabs_(b - c), dca = iabs_(c - a);
    const int m = min(dab, min(dbc, dca));
    if (m == dab)
        return (int16_t)((a + b) / 2);
    if (m == dbc)
        return (int16_t)((b + c) / 2);
    return (int16_t)((c + a) / 2);
}

// get_mvp_default for one record: status 0, 2 (an index outside the lists)
// or 3 (a scaling product leaves int); *src is the neighbour the MVP's y
// component came from (0 left, 1 top, 2 top-right or its stand-in), as
// get_mvp_default_sad picks its SAD.
__device__ __forceinline__ int mvp_core(const xme_mvp_job &j, int b_slice, const MvpTables &t, int &px, int &py,
                                        int &src)
{
    px = py = 0;
    src = 0;
    const int rl = j.ref[0], ru = j.ref[1], rur = j.ur_available ? j.ref[3] : j.ref[2];
    const int refs[3] = { rl, ru, rur };
    // a table index outside the reference lists cannot be derived
    bool bad = j.ref_idx < 0 || j.ref_idx >= XME_MVP_REFS;
#pragma unroll
    for (int k = 0; k < 3; k++)
        bad |= refs[k] >= XME_MVP_REFS || refs[k] < -1;
    if (bad)
        return 2;
    const int c_src = j.ur_available ? 3 : 2;
    int mv[3][2] = { { j.mv[0][0], j.mv[0][1] }, { j.mv[1][0], j.mv[1][1] }, { j.mv[c_src][0], j.mv[c_src][1] } };
    const int type = mvp_type(j.ref_idx, rl, ru, rur, j.cb_x, j.cb_y, j.cb_w, j.cb_h);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (refs[k] < 0) {
            mv[k][0] = mv[k][1] = 0;
            continue;
        }
        // B slices: the list's distance (B_BWD = 0, B_FWD = 1, avs2_defs.h:138-139);
        // others: the searched reference's distance and the neighbour's multiplier
        const int dist = b_slice ? t.dpoc[j.bwd_2nd ? 0 : 1] : t.dpoc[j.ref_idx];
        const int scale = b_slice ? t.multi[j.bwd_2nd ? 0 : 1] : t.multi[refs[k]];
        mv[k][1] = scale_default(mv[k][1], dist, scale, bad);
        mv[k][0] = scale_default(mv[k][0], dist, scale, bad);
    }
    if (type == 0) {
        px = (int16_t)median_mv(mv[0][0], mv[1][0], mv[2][0]);
        py = (int16_t)median_mv(mv[0][1], mv[1][1], mv[2][1]);
        src = median_src(mv[0][1], mv[1][1], mv[2][1]);
    } else {
        px = (int16_t)mv[type - 1][0];
        py = (int16_t)mv[type - 1][1];
        src = type - 1;
    }
    return bad ? 3 : 0;
}

__global__ void __launch_bounds__(256) mvp_kernel(const xme_mvp_job *__restrict__ jobs, int n, int b_slice,
                                                  MvpTables t, xme_mvp_result *__restrict__ out)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    const xme_mvp_job j = jobs[i];
    xme_mvp_result r;
    int px, py, src;
    r.status = mvp_core(j, b_slice, t, px, py, src);
    r.mvp_x = (int16_t)px;
    r.mvp_y = (int16_t)py;
    out[i] = r;
}

// tab_umh_alpha_2nd / _3rd (md_inter.c:49-54): float literals held as doubles.
__constant__ float c_alpha2[8] = { 0.0f, 0.01f, 0.01f, 0.01f, 0.02f, 0.03f, 0.03f, 0.04f };
__constant__ float c_alpha3[8] = { 0.0f, 0.06f, 0.07f, 0.07f, 0.08f, 0.12f, 0.11f, 0.15f };

struct UmhTables {
    int32_t dpoc[XME_MVP_REFS], multi[XME_MVP_REFS];
    double bsize[8];
};

// (int16_t)(mv * dpoc_curr / (float)dpoc_src) (md_inter.c:673-674, :690-691):
// the int product in float, an IEEE division, truncation.
__device__ __forceinline__ int scale_float(int mv, int dpoc_curr, int dpoc_src)
{
    const float f = (float)(mv * dpoc_curr) / (float)dpoc_src;
    return (int16_t)(int)f;
}

__global__ void __launch_bounds__(256) umh_state_kernel(const xme_umh_job *__restrict__ jobs, int n, int slice_type,
                                                        UmhTables t, xme_umh_state *__restrict__ out)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    const xme_umh_job j = jobs[i];
    xme_umh_state o = {};
    const int b_slice = slice_type == 2, ref_idx = j.mvp.ref_idx, mode = j.mode;
    MvpTables mt;
#pragma unroll
    for (int k = 0; k < XME_MVP_REFS; k++) {
        mt.dpoc[k] = t.dpoc[k];
        mt.multi[k] = t.multi[k];
    }
    int px, py, src;
    int st = mvp_core(j.mvp, b_slice, mt, px, py, src);
    if (st == 2 || mode < 1 || mode > 7) {
        o.status = 2;
        out[i] = o;
        return;
    }
    o.mvp_x = (int16_t)px;
    o.mvp_y = (int16_t)py;
    o.pred_sad_space = j.nb_sad[src];
    xme_umh_ctx &c = o.ctx;
    // the upper layer (md_inter.c:648-661)
    int up = -1;
    if (mode == 4 || mode == 5)
        up = 0;          // PRED_2NxnU / 2NxnD: PRED_2NxN
    else if (mode == 6 || mode == 7)
        up = 1;          // PRED_nLx2N / nRx2N: PRED_Nx2N
    else if (mode > 1)
        up = 2;          // PRED_2NxN / Nx2N: PRED_2Nx2N
    int sad_up = 0;
    if (up >= 0) {
        c.mvp1_x = j.best_up[up][0];
        c.mvp1_y = j.best_up[up][1];
        sad_up = j.mincost_up[up] / 2;
    }
    c.pred_sad_uplayer = sad_up;
    // the nearest reference (md_inter.c:664-679); mvp2 is unused for ref 0
    int sad_ref = 0;
    if (b_slice && ref_idx == 1) {
        c.mvp2_x = (int16_t)(-j.best_prev[0]);
        c.mvp2_y = (int16_t)(-j.best_prev[1]);
        sad_ref = j.mincost_prev;
    } else if (ref_idx > 0) {
        if (t.dpoc[ref_idx - 1] == 0) {
            // a division by a zero distance: the state stops here
            o.status = 4;
            out[i] = o;
            return;
        }
        c.mvp2_x = scale_float(j.best_prev[0], t.dpoc[ref_idx], t.dpoc[ref_idx - 1]);
        c.mvp2_y = scale_float(j.best_prev[1], t.dpoc[ref_idx], t.dpoc[ref_idx - 1]);
        sad_ref = j.mincost_prev;
    }
    // the co-located MV (md_inter.c:682-698)
    if (j.col_valid && j.col_ref >= 0) {
        if (j.col_dpoc == 0) {
            o.status = 4;
            out[i] = o;
            return;
        }
        c.mvp3_x = scale_float(j.col_mv[0], t.dpoc[ref_idx], j.col_dpoc);
        c.mvp3_y = scale_float(j.col_mv[1], t.dpoc[ref_idx], j.col_dpoc);
    }
    c.ref_idx = ref_idx;
    const int pred_sad = (!b_slice && ref_idx > 0) ? sad_ref : (mode == 1 ? o.pred_sad_space : sad_up);
    c.pred_sad = pred_sad;
    if (pred_sad != 0) {
        // dist_t (int32) square, wrapping, then double (md_inter.c:710-715)
        const int sq = (int)((uint32_t)pred_sad * (uint32_t)pred_sad);
        const double threshold = t.bsize[mode] / (double)sq;
        c.beta2 = threshold - (double)c_alpha2[mode];
        c.beta3 = threshold - (double)c_alpha3[mode];
    }
    o.status = st;
    out[i] = o;
}

} // namespace

extern "C" int xme_umh_prepare_batch(const xme_umh_job *d_jobs, int n, int slice_type,
                                     const int32_t ref_dpoc[XME_MVP_REFS], const int32_t ref_dpoc_multi[XME_MVP_REFS],
                                     const double umh_bsize[8], xme_umh_state *d_out, void *stream)
{
    if (n < 0 || slice_type < 0 || slice_type > 3 || !ref_dpoc || !ref_dpoc_multi || !umh_bsize ||
        (n > 0 && (!d_jobs || !d_out))) {
        xme_set_error("xme_umh_prepare_batch: bad arguments");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    UmhTables t;
    for (int k = 0; k < XME_MVP_REFS; k++) {
        t.dpoc[k] = ref_dpoc[k];
        t.multi[k] = ref_dpoc_multi[k];
    }
    for (int k = 0; k < 8; k++)
        t.bsize[k] = umh_bsize[k];
    hipLaunchKernelGGL(umh_state_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_jobs, n,
                       slice_type, t, d_out);
    XME_HIP(hipGetLastError());
    return XME_OK;
}

extern "C" int xme_umh_bsize(int i_qp, int slice_type, double bsize[8])
{
    static const int quant_coef[6] = { 13107, 11916, 10082, 9362, 8192, 7282 };
    if (!bsize || i_qp < 0 || i_qp > 79) {
        xme_set_error("xme_umh_bsize: bad arguments");
        return XME_ERR_ARG;
    }
    const int per = i_qp / 6, rem = i_qp % 6, q_bits = 15 + per;
    const int qp_const = slice_type == 0 ? (1 << q_bits) / 3 : (1 << q_bits) / 6;
    const int threshold_4x4 = ((1 << q_bits) - qp_const) / quant_coef[rem];
    const double step = (float)threshold_4x4 / (4 * 5.61f);
    for (int m = 0; m < 8; m++)
        bsize[m] = 0;
    bsize[7] = (16 * 16) * step;    // PRED_nRx2N
    bsize[6] = 4 * bsize[7];        // PRED_nLx2N
    bsize[5] = 4 * bsize[7];        // PRED_2NxnD
    bsize[4] = 4 * bsize[5];        // PRED_2NxnU
    bsize[3] = 4 * bsize[4];        // PRED_Nx2N
    bsize[2] = 4 * bsize[4];        // PRED_2NxN
    bsize[1] = 4 * bsize[2];        // PRED_2Nx2N
    return XME_OK;
}

extern "C" int xme_get_mvp_batch(const xme_mvp_job *d_jobs, int n, int b_slice, const int32_t ref_dpoc[XME_MVP_REFS],
                                 const int32_t ref_dpoc_multi[XME_MVP_REFS], xme_mvp_result *d_out, void *stream)
{
    if (n < 0 || !ref_dpoc || !ref_dpoc_multi || (n > 0 && (!d_jobs || !d_out))) {
        xme_set_error("xme_get_mvp_batch: bad arguments");
        return XME_ERR_ARG;
    }
    if (n == 0)
        return XME_OK;
    MvpTables t;
    for (int k = 0; k < XME_MVP_REFS; k++) {
        t.dpoc[k] = ref_dpoc[k];
        t.multi[k] = ref_dpoc_multi[k];
    }
    hipLaunchKernelGGL(mvp_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_jobs, n, b_slice, t,
                       d_out);
    XME_HIP(hipGetLastError());
    return XME_OK;
}
