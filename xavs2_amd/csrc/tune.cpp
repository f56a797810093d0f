// Tile-shape autotuner for the integer search.  A choice (K rows per lane,
// LC lanes per candidate row) is timed on the device for a PU shape and
// search range; winners are kept for the process and used by later plans.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "xme_internal.h"
#include "tune.h"

namespace {
std::mutex g_mu;
std::map<std::pair<int, int>, TuneChoice> g_table;
bool g_forced = false;
TuneChoice g_force[XME_NUM_PU];

int pow2_at_least(int v)
{
    int p = 1;
    while (p < v)
        p <<= 1;
    return p;
}
} // namespace

namespace {
int g_kmax = 0;
}

int xme_fs_kmax(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_kmax == 0) {
        // K = 8 (2 waves per SIMD) measured faster than K = 4 (4 waves per
        // SIMD) on the 2160p mixed-PU job: per-qsad overhead dominates
        const char *e = getenv("XME_FS_K");
        g_kmax = (e && atoi(e) == 4) ? 4 : 8;
    }
    return g_kmax;
}

int xme_fs_wg_per_cu(void)
{
    static const int v = [] {
        const char *e = getenv("XME_FS_WG_PER_CU");
        const int n = e ? atoi(e) : 2;
        return n >= 1 && n <= 4 ? n : 2;
    }();
    return v;
}

void xme_fs_set_kmax(int kmax)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_kmax = kmax == 8 ? 8 : 4;
}

TuneChoice xme_tune_default(int part, int R)
{
    TuneChoice c;
    // enough 4-column lanes to span the 2R candidate columns (at most a wave)
    c.LC = std::min(64, std::max(8, pow2_at_least((2 * R + 3) / 4)));
    c.K = xme_narrow_shape(part) ? 16 : (xme_pu_h(part) % 8 == 0) ? 8 : 4;
    return c;
}

TuneChoice xme_tune_lookup(int part, int R)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_forced)
        return g_force[part];
    auto it = g_table.find({ part, R });
    return it != g_table.end() ? it->second : xme_tune_default(part, R);
}

void xme_tune_store(int part, int R, TuneChoice c)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[{ part, R }] = c;
}

void xme_tune_force(const TuneChoice *c)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (c) {
        for (int i = 0; i < XME_NUM_PU; i++)
            g_force[i] = *c;
        g_forced = true;
    } else {
        g_forced = false;
    }
}

extern "C" int xme_tune_get(int part, int R, int *k_rows, int *lanes_per_row)
{
    if (part < 0 || part >= XME_NUM_PU)
        return XME_ERR_ARG;
    TuneChoice c = xme_tune_lookup(part, R);
    if (k_rows)
        *k_rows = c.K;
    if (lanes_per_row)
        *lanes_per_row = c.LC;
    return XME_OK;
}

// Times integer-only full search over a synthetic 256x256-pixel region tiled
// with PUs of each requested shape, for every candidate tile shape.
extern "C" int xme_tune(const xme_me_params *prm, const int *parts, int n_parts, void *stream)
{
    if (!prm || !parts || n_parts <= 0)
        return XME_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int W = 384, H = 320, AW = 256, AH = 256;
    const int stride = W + 2 * XME_PAD;
    const size_t plane_bytes = (size_t)stride * (H + 2 * XME_PAD + 1) + 64;
    uint8_t *d_ref = nullptr, *d_cur = nullptr;
    xme_result *d_res = nullptr;
    XME_HIP(hipMalloc(&d_ref, plane_bytes));
    XME_HIP(hipMalloc(&d_cur, (size_t)W * H));
    std::vector<uint8_t> h(plane_bytes);
    uint32_t x = 12345;
    for (auto &b : h) {
        x = x * 1664525u + 1013904223u;
        b = (uint8_t)(x >> 24);
    }
    XME_HIP(hipMemcpy(d_ref, h.data(), plane_bytes, hipMemcpyHostToDevice));
    XME_HIP(hipMemcpy(d_cur, h.data() + 777, (size_t)W * H, hipMemcpyHostToDevice));
    XME_HIP(hipMalloc(&d_res, sizeof(xme_result) * 8192));

    xme_me_params p = *prm;
    p.frame_w = W;
    p.frame_h = H;
    p.use_fractional_me = 0;
    p.me_method = XME_ME_FS;       // the tiles are those of the full search
    xme_ref_desc ref;
    memset(&ref, 0, sizeof(ref));
    ref.planes[0] = d_ref + (size_t)XME_PAD * stride + XME_PAD;
    ref.stride = stride;
    int tuned = 0;
    for (int n = 0; n < n_parts; n++) {
        const int part = parts[n];
        if (part < 0 || part >= XME_NUM_PU)
            continue;
        const int pw = xme_pu_w(part), ph = xme_pu_h(part);
        std::vector<xme_pu> pus;
        for (int y = 32; y + ph <= 32 + AH; y += ph)
            for (int xx = 64; xx + pw <= 64 + AW; xx += pw) {
                xme_pu q;
                memset(&q, 0, sizeof(q));
                q.x = (int16_t)xx;
                q.y = (int16_t)y;
                q.w = (uint8_t)pw;
                q.h = (uint8_t)ph;
                pus.push_back(q);
            }
        if ((int)pus.size() > 8192)
            pus.resize(8192);
        std::vector<TuneChoice> cands;
        for (int K : { 4, 8, 16 }) {
            if (K == 16 && !xme_narrow_shape(part))
                continue;
            if (K != 16 && ph % K)
                continue;
            for (int LC : { 8, 16, 32, 64 }) {
                if (4 * LC > 2 * 2 * p.search_range + 3 && LC > 8)
                    continue;   // more than half the lanes would be idle
                cands.push_back(TuneChoice{ K, LC });
            }
        }
        double best_t = 1e30;
        TuneChoice best_c = xme_tune_default(part, p.search_range);
        for (const TuneChoice &c : cands) {
            xme_tune_force(&c);
            xme_plan *pl = nullptr;
            int rc = xme_plan_create(&p, pus.data(), (int)pus.size(), 1, XME_PLAN_BATCHED, &pl);
            xme_tune_force(nullptr);
            if (rc)
                continue;
            double t_best = 1e30;
            for (int rep = 0; rep < 4 && rc == 0; rep++) {
                hipStreamSynchronize(s);
                auto t0 = std::chrono::steady_clock::now();
                rc = xme_me_run(pl, d_cur, W, &ref, d_res, nullptr, s);
                hipStreamSynchronize(s);
                double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (rep > 0)
                    t_best = std::min(t_best, t);
            }
            xme_plan_destroy(pl);
            if (rc == 0 && t_best < best_t) {
                best_t = t_best;
                best_c = c;
            }
        }
        xme_tune_store(part, p.search_range, best_c);
        tuned++;
    }
    hipFree(d_ref);
    hipFree(d_cur);
    hipFree(d_res);
    return tuned;
}

extern "C" int xme_tune_set(int part, int R, int k_rows, int lanes_per_row)
{
    const bool lc_ok = lanes_per_row == 8 || lanes_per_row == 16 || lanes_per_row == 32 || lanes_per_row == 64;
    const bool k_ok = k_rows == 4 || k_rows == 8 || (k_rows == 16 && part >= 0 && part < XME_NUM_PU &&
                                                        xme_narrow_shape(part));
    if (part < 0 || part >= XME_NUM_PU || R < 1 || R > 128 || !k_ok || !lc_ok) {
        xme_set_error("xme_tune_set: part 0..24, R 1..128, K 4 or 8 (16 for 8- or 16-wide shapes of <= 16 rows), "
                      "LC 8/16/32/64");
        return XME_ERR_ARG;
    }
    xme_tune_store(part, R, TuneChoice{ k_rows, lanes_per_row });
    return XME_OK;
}
