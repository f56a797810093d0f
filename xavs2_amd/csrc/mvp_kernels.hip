// MV predictors (get_mvp_default, source/encoder/md_inter.c:470-524) for a
// batch of PUs, gfx950.  One thread per PU: the work is a few dozen integer
// operations on the PU's four neighbours, so the batch is bound by reading
// its 28-byte records.
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

__global__ void __launch_bounds__(256) mvp_kernel(const xme_mvp_job *__restrict__ jobs, int n, int b_slice,
                                                  MvpTables t, xme_mvp_result *__restrict__ out)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    const xme_mvp_job j = jobs[i];
    xme_mvp_result r;
    r.mvp_x = r.mvp_y = 0;
    r.status = 0;
    const int rl = j.ref[0], ru = j.ref[1], rur = j.ur_available ? j.ref[3] : j.ref[2];
    const int refs[3] = { rl, ru, rur };
    // a table index outside the reference lists cannot be derived
    bool bad = j.ref_idx < 0 || j.ref_idx >= XME_MVP_REFS;
#pragma unroll
    for (int k = 0; k < 3; k++)
        bad |= refs[k] >= XME_MVP_REFS || refs[k] < -1;
    if (bad) {
        r.status = 2;
        out[i] = r;
        return;
    }
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
        r.mvp_x = (int16_t)median_mv(mv[0][0], mv[1][0], mv[2][0]);
        r.mvp_y = (int16_t)median_mv(mv[0][1], mv[1][1], mv[2][1]);
    } else {
        r.mvp_x = (int16_t)mv[type - 1][0];
        r.mvp_y = (int16_t)mv[type - 1][1];
    }
    r.status = bad ? 3 : 0;
    out[i] = r;
}

} // namespace

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
