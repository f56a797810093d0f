// Continuous-batching frame-ME server: many concurrent encodes share one GPU.
//
// The reference encodes one video per process: its frame threads call
// xavs2_me_search per PU (source/encoder/me.c:879-1173) against references
// whose 15 interpolated planes live in host memory (source/common/frame.c:335-367,
// built by interpolate_lcu_row, source/common/mc.c:748-785).  Here a stream
// (one encode) keeps its last n_ref references and their planes resident in
// device memory as a ring of n_ref + 1 slots; a submitted frame is copied
// into the slot of the oldest reference, searched against the previous
// min(frame, n_ref) frames with the stream's plan, and then border-expanded
// and interpolated in place so that it becomes the newest reference.
//
// Scheduling: whenever one of the max_inflight slots (one HIP stream each) is
// free, the scheduler admits, among the queued frames that are next in their
// own stream, the one with the earliest deadline (submit time plus twice its
// search work at a nominal device rate, so a small frame does not wait behind
// a large one that arrived just before it), if its working memory fits the
// budget beside everything resident.  Frames admitted into different slots
// overlap on the device; a frame's search waits (by an event, on the device)
// only for the previous frame of its own stream to have become a reference.
// Results are copied back into pinned host memory and a host callback stamps
// the completion time, frees the slot and wakes the scheduler; the frame's
// reference build follows on the slot's stream, and a second callback marks
// it done, which is what closing the stream waits for.  A failed reference
// build marks the stream failed: its later frames and submissions return
// that error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "xme_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxSlots = 16;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Blocks of one size are reused; nothing is returned to HIP until the budget
// needs the room or the server ends.  The pool is bookkeeping only: blocks are
// allocated and freed by the caller outside the server lock, because a HIP
// allocation or free may wait for the device, and the device's host callbacks
// take that lock.
struct BlockPool {
    std::multimap<size_t, void *> free_blocks;
    size_t cached = 0;

    void *take(size_t n)
    {
        auto it = free_blocks.find(n);
        if (it == free_blocks.end())
            return nullptr;
        void *p = it->second;
        free_blocks.erase(it);
        cached -= n;
        return p;
    }
    void put(void *p, size_t n)
    {
        if (!p)
            return;
        free_blocks.emplace(n, p);
        cached += n;
    }
    // removes cached blocks until at most `keep` bytes stay cached
    std::vector<void *> trim(size_t keep)
    {
        std::vector<void *> out;
        while (cached > keep && !free_blocks.empty()) {
            auto it = free_blocks.begin();
            out.push_back(it->second);
            cached -= it->first;
            free_blocks.erase(it);
        }
        return out;
    }
};

void free_device_blocks(const std::vector<void *> &v)
{
    for (void *p : v)
        (void)hipFree(p);
}

void free_host_blocks(const std::vector<void *> &v)
{
    for (void *p : v)
        (void)hipHostFree(p);
}

struct Stream;

struct Job {
    int64_t id = 0;
    Stream *st = nullptr;
    int64_t frame = 0;
    int n_refs = 0;
    int n_results = 0;
    size_t dev_bytes = 0;      // counted against the budget while admitted
    uint8_t *h_pic = nullptr;  // pinned copy of the picture, cur_stride x h
    size_t h_pic_bytes = 0;
    xme_result *h_res = nullptr;
    size_t h_res_bytes = 0;
    uint8_t *d_cur = nullptr;
    size_t d_cur_bytes = 0;
    xme_result *d_res = nullptr;
    size_t d_res_bytes = 0;
    Clock::time_point t_submit, t_done;
    double deadline = 0;       // admission order: seconds since the server's start
    int slot = -1;
    int rc = XME_OK;
    bool done = false;
    xme_server *srv = nullptr;
};

struct Stream {
    int id = 0;
    int w = 0, h = 0;
    int n_pu = 0;
    intptr_t stride = 0;         // reference planes
    intptr_t cur_stride = 0;     // staged current picture
    size_t plane_bytes = 0;
    std::vector<uint8_t *> ring; // n_ref + 1 slots of 16 planes
    std::vector<xme_plan *> plans;   // plans[r - 1] searches r references
    hipEvent_t ref_ready[2] = { nullptr, nullptr };
    int64_t next_frame = 0;
    int64_t next_admit = 0;      // the frame the scheduler may admit next
    int64_t resident = 0;
    int64_t work = 0;            // byte absdiffs of a frame's search with n_ref references
    int pending = 0;             // submitted and not yet done
    int building = 0;            // admitted frames whose reference build has not finished
    int failed = XME_OK;         // a reference build failed: the stream cannot go on
    xme_server *srv = nullptr;
    bool closing = false;
    std::mutex submit_mu;        // one submission of the stream at a time: frame order

    uint8_t *plane(int slot, int k) const
    {
        return ring[slot] + (size_t)k * plane_bytes + (size_t)XME_PAD * stride + XME_PAD;
    }
};

} // namespace

struct xme_server {
    xme_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_sched;   // scheduler: queue / slot / memory changes
    std::condition_variable cv_done;    // waiters: a job finished
    std::deque<Job *> queue;
    std::unordered_map<int64_t, Job *> jobs;
    std::map<int, Stream *> streams;
    hipStream_t slots[kMaxSlots] = {};
    bool slot_busy[kMaxSlots] = {};
    int64_t next_job = 1;
    int next_stream = 1;
    int64_t resident = 0, inflight = 0, peak = 0;
    BlockPool dev_pool, host_pool;
    std::thread sched;
    bool stop = false;
    int running = 0;                    // admitted, not done
    // statistics
    int64_t submitted = 0, completed = 0, rejected = 0;
    std::vector<double> lat_ms;
    bool have_first = false;
    Clock::time_point t_first, t_last;
    Clock::time_point t_start = Clock::now();

    // The queued frame to admit next: the earliest deadline among frames that
    // are next in their stream (ties by arrival); null when none may go.
    std::deque<Job *>::iterator pick_locked()
    {
        auto best = queue.end();
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Job *j = *it;
            if (j->frame != j->st->next_admit)
                continue;
            if (best == queue.end() || j->deadline < (*best)->deadline)
                best = it;
        }
        return best;
    }

    bool can_admit_locked()
    {
        if (queue.empty())
            return false;
        bool slot = false;
        for (int i = 0; i < cfg.max_inflight; i++)
            slot = slot || !slot_busy[i];
        if (!slot)
            return false;
        auto it = pick_locked();
        return it != queue.end() && resident + inflight + (int64_t)(*it)->dev_bytes <= cfg.memory_budget;
    }
};

namespace {

void finish_locked(xme_server *S, Job *j)
{
    j->done = true;
    j->t_done = Clock::now();
    S->slot_busy[j->slot] = false;
    S->inflight -= (int64_t)j->dev_bytes;
    S->running--;
    S->dev_pool.put(j->d_cur, j->d_cur_bytes);
    S->dev_pool.put(j->d_res, j->d_res_bytes);
    j->d_cur = nullptr;
    j->d_res = nullptr;
    S->host_pool.put(j->h_pic, j->h_pic_bytes);
    j->h_pic = nullptr;
    j->st->pending--;
    S->completed++;
    if (j->rc == XME_OK && j->n_refs > 0) {
        S->lat_ms.push_back(std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count());
        S->t_last = j->t_done;
    }
    S->cv_sched.notify_all();
    S->cv_done.notify_all();
}

// Runs on a HIP runtime thread once the job's results are in host memory; it
// only does bookkeeping (no HIP calls are allowed here).  After it the job may
// be gone (a waiter deletes it).
void on_results(void *arg)
{
    Job *j = static_cast<Job *>(arg);
    xme_server *S = j->srv;
    std::lock_guard<std::mutex> lk(S->mu);
    finish_locked(S, j);
}

// Runs once a frame's reference build (and its event) has finished.
void on_built(void *arg)
{
    Stream *st = static_cast<Stream *>(arg);
    xme_server *S = st->srv;
    std::lock_guard<std::mutex> lk(S->mu);
    st->building--;
    S->cv_done.notify_all();
}

// Device memory for a job: a cached block of the size, else a fresh one,
// after trimming cached blocks that would push the device past the budget.
void *dev_get(xme_server *S, size_t n)
{
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (void *p = S->dev_pool.take(n))
            return p;
        // the job's bytes are already counted in `inflight`
        const int64_t room = S->cfg.memory_budget - S->resident - S->inflight;
        drop = S->dev_pool.trim(room > 0 ? (size_t)room : 0);
    }
    free_device_blocks(drop);
    void *p = nullptr;
    return hipMalloc(&p, n) == hipSuccess ? p : nullptr;
}

void *host_get(xme_server *S, size_t n)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (void *p = S->host_pool.take(n))
            return p;
    }
    void *p = nullptr;
    return hipHostMalloc(&p, n, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}

// Queues a frame's work on its slot's stream.  *results_queued tells whether
// the results callback was queued (after which the job must not be touched).
int launch_job(xme_server *S, Job *j, bool *results_queued)
{
    *results_queued = false;
    Stream *st = j->st;
    const int n_slots = (int)st->ring.size();
    const int64_t frame = j->frame;      // j may be gone once on_results is queued
    const int cur_slot = (int)(frame % n_slots);
    hipStream_t s = S->slots[j->slot];
    if (st->failed != XME_OK) {
        xme_set_error("xme_server: stream %d cannot go on (an earlier reference build failed)", st->id);
        return st->failed;
    }
    j->d_cur = static_cast<uint8_t *>(dev_get(S, j->d_cur_bytes));
    if (j->d_res_bytes)
        j->d_res = static_cast<xme_result *>(dev_get(S, j->d_res_bytes));
    if (!j->d_cur || (j->d_res_bytes && !j->d_res)) {
        xme_set_error("xme_server: device allocation failed");
        return XME_ERR_NOMEM;
    }
    // the upload overlaps whatever the stream's previous frame still runs
    XME_HIP(hipMemcpyAsync(j->d_cur, j->h_pic, (size_t)st->cur_stride * st->h, hipMemcpyHostToDevice, s));
    if (j->frame > 0)
        XME_HIP(hipStreamWaitEvent(s, st->ref_ready[(j->frame - 1) & 1], 0));
    // the frame takes the ring slot of its stream's oldest reference
    uint8_t *cur = st->plane(cur_slot, 0);
    XME_HIP(hipMemcpy2DAsync(cur, st->stride, j->d_cur, st->cur_stride, st->w, st->h, hipMemcpyDeviceToDevice, s));
    if (j->n_refs > 0) {
        xme_ref_desc refs[XME_MAX_REFS];
        for (int r = 0; r < j->n_refs; r++) {
            const int rs = (int)((j->frame - 1 - r) % n_slots);
            for (int k = 0; k < 16; k++) {
                const bool need = k == 0 || S->cfg.use_fractional_me == 2 ||
                                  (S->cfg.use_fractional_me == 1 && (k == 2 || k == 8 || k == 10));
                refs[r].planes[k] = need ? st->plane(rs, k) : nullptr;
            }
            refs[r].stride = st->stride;
        }
        int rc = xme_me_run_graph(st->plans[j->n_refs - 1], cur, st->stride, refs, j->d_res, s);
        if (rc)
            return rc;
        XME_HIP(hipMemcpyAsync(j->h_res, j->d_res, j->d_res_bytes, hipMemcpyDeviceToHost, s));
    }
    XME_HIP(hipLaunchHostFunc(s, on_results, j));
    *results_queued = true;
    // the frame becomes its stream's newest reference
    int rc = xme_expand_border(cur, st->stride, st->w, st->h, XME_PAD, s);
    if (rc)
        return rc;
    if (S->cfg.use_fractional_me > 0) {
        pel_t *pl[16];
        for (int k = 0; k < 16; k++)
            pl[k] = st->plane(cur_slot, k);
        rc = xme_interpolate_frame(pl, st->stride, st->w, st->h, S->cfg.use_fractional_me, s);
        if (rc)
            return rc;
    }
    XME_HIP(hipEventRecord(st->ref_ready[frame & 1], s));
    XME_HIP(hipLaunchHostFunc(s, on_built, st));
    return XME_OK;
}

void sched_loop(xme_server *S)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_sched.wait(lk, [S] { return S->can_admit_locked() || (S->stop && S->queue.empty()); });
        if (!S->can_admit_locked())
            break;                       // stopping with an empty queue
        auto it = S->pick_locked();
        Job *j = *it;
        S->queue.erase(it);
        Stream *st = j->st;
        st->next_admit++;
        st->building++;
        int slot = 0;
        while (S->slot_busy[slot])
            slot++;
        S->slot_busy[slot] = true;
        j->slot = slot;
        S->inflight += (int64_t)j->dev_bytes;
        S->peak = std::max(S->peak, S->resident + S->inflight);
        S->running++;
        lk.unlock();
        bool results_queued = false;
        const int rc = launch_job(S, j, &results_queued);
        if (rc != XME_OK) {
            // nothing after the failing call was queued: wait for what was
            (void)hipStreamSynchronize(S->slots[slot]);
        }
        lk.lock();
        if (rc != XME_OK) {
            if (!results_queued) {
                j->rc = rc;
                finish_locked(S, j);     // the job's results never come
            }
            // the frame's reference was not built: the stream cannot go on
            if (st->failed == XME_OK)
                st->failed = rc;
            st->building--;
            S->cv_done.notify_all();
        }
    }
}

// Frees a stream's device memory; called without the server lock.
void free_stream(Stream *st)
{
    for (xme_plan *p : st->plans)
        xme_plan_destroy(p);
    for (uint8_t *r : st->ring)
        (void)hipFree(r);
    for (hipEvent_t e : st->ref_ready)
        if (e)
            (void)hipEventDestroy(e);
    delete st;
}

} // namespace

extern "C" int xme_server_create(const xme_server_config *cfg, xme_server **out)
{
    if (!cfg || !out || cfg->memory_budget <= 0 || cfg->max_inflight < 1 || cfg->max_inflight > kMaxSlots ||
        cfg->search_range < 1 || cfg->search_range > 128 || cfg->n_ref < 1 || cfg->n_ref > XME_MAX_REFS ||
        cfg->lambda_factor < 0 || cfg->use_fractional_me < 0 || cfg->use_fractional_me > 2 ||
        cfg->workgroup_share_pct < 0 || cfg->workgroup_share_pct > 100 || cfg->full_share_below_absdiff < 0 ||
        (cfg->me_method != XME_ME_FS && cfg->me_method != XME_ME_DIA && cfg->me_method != XME_ME_HEX &&
         cfg->me_method != XME_ME_TZ)) {
        xme_set_error("xme_server_create: bad configuration");
        return XME_ERR_ARG;
    }
    auto *S = new xme_server();
    S->cfg = *cfg;
    if (hipGetDevice(&S->device) != hipSuccess)
        S->device = 0;
    for (int i = 0; i < cfg->max_inflight; i++) {
        hipError_t e = hipStreamCreateWithFlags(&S->slots[i], hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (int k = 0; k < i; k++)
                (void)hipStreamDestroy(S->slots[k]);
            delete S;
            return xme_hip_fail(e, "xme_server_create");
        }
    }
    S->sched = std::thread(sched_loop, S);
    *out = S;
    return XME_OK;
}

extern "C" int xme_server_destroy(xme_server *S)
{
    if (!S)
        return XME_ERR_ARG;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_sched.notify_all();
    }
    S->sched.join();
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [S] {
            if (S->running != 0)
                return false;
            for (auto &kv : S->streams)
                if (kv.second->building != 0)
                    return false;
            return true;
        });
    }
    for (int i = 0; i < S->cfg.max_inflight; i++) {
        (void)hipStreamSynchronize(S->slots[i]);
        (void)hipStreamDestroy(S->slots[i]);
    }
    for (auto &kv : S->jobs) {
        S->host_pool.put(kv.second->h_res, kv.second->h_res_bytes);
        delete kv.second;
    }
    for (auto &kv : S->streams)
        free_stream(kv.second);
    free_device_blocks(S->dev_pool.trim(0));
    free_host_blocks(S->host_pool.trim(0));
    delete S;
    return XME_OK;
}

extern "C" int xme_server_open_stream(xme_server *S, int w, int h, const xme_pu *pus, int n_pu, int *stream_id)
{
    if (!S || !pus || n_pu <= 0 || !stream_id || w < 16 || h < 16 || w > 8192 || h > 8192 || (w & 7) || (h & 7)) {
        xme_set_error("xme_server_open_stream: sizes must be multiples of 8 in [16, 8192] and PUs given");
        return XME_ERR_ARG;
    }
    auto st = std::make_unique<Stream>();
    st->srv = S;
    // search work of a frame per reference, for the admission deadlines: PU
    // area times the candidates (the full window, or about 64 pattern points)
    {
        int64_t area = 0;
        for (int i = 0; i < n_pu; i++)
            area += (int64_t)pus[i].w * pus[i].h;
        const int64_t r2 = 2 * (int64_t)S->cfg.search_range;
        st->work = area * (S->cfg.me_method == XME_ME_FS ? r2 * r2 : 64);
    }
    st->w = w;
    st->h = h;
    st->n_pu = n_pu;
    st->stride = (intptr_t)align_up((size_t)w + 2 * XME_PAD, 256);
    st->cur_stride = (intptr_t)align_up((size_t)w, 64);
    st->plane_bytes = align_up((size_t)(h + 2 * XME_PAD + 1) * st->stride + 256, 256);
    const int n_slots = S->cfg.n_ref + 1;
    const int n_planes = S->cfg.use_fractional_me > 0 ? 16 : 1;
    const size_t ring_bytes = (size_t)n_planes * st->plane_bytes;

    xme_me_params prm{};
    prm.frame_w = w;
    prm.frame_h = h;
    prm.search_range = S->cfg.search_range;
    prm.lambda_factor = S->cfg.lambda_factor;
    // horizontal limit of encoder_decide_mv_range (encoder.c:849-892) and a
    // vertical limit that leaves every in-picture candidate of [-R, R) legal
    prm.min_mv_range[0] = -8192;
    prm.max_mv_range[0] = 8191;
    prm.min_mv_range[1] = -2048;
    prm.max_mv_range[1] = 2047;
    prm.use_hadamard = S->cfg.use_hadamard;
    prm.use_fractional_me = S->cfg.use_fractional_me;
    prm.use_fast_sub_me = S->cfg.use_fast_sub_me;
    prm.enable_pmvr = S->cfg.enable_pmvr;
    prm.me_method = S->cfg.me_method;
    // the plans first (they validate the PUs); their memory counts as resident
    for (int r = 1; r <= S->cfg.n_ref; r++) {
        xme_plan *p = nullptr;
        int rc = xme_plan_create(&prm, pus, n_pu, r, XME_PLAN_BATCHED, &p);
        if (rc) {
            for (xme_plan *q : st->plans)
                xme_plan_destroy(q);
            return rc;
        }
        st->plans.push_back(p);
        st->resident += (int64_t)xme_plan_device_bytes(p);
        const int64_t split = S->cfg.launch_split_absdiff;
        xme_plan_set_launch_split(p, split == 0 ? (int64_t)8000000000LL : split < 0 ? 0 : split);
        // the share caps the streams of large frames; a small frame's search
        // takes every workgroup that frees, so it finishes before the large
        // frames around it instead of sharing the device with them
        const bool small = S->cfg.full_share_below_absdiff > 0 && st->work < S->cfg.full_share_below_absdiff;
        const int share = small ? 100 : S->cfg.workgroup_share_pct;
        if (share > 0 && share < 100) {
            int g = 0, full = 0;
            xme_plan_workgroups(p, &g, &full);
            xme_plan_set_max_workgroups(p, std::max(1, (full * share + 99) / 100));
        }
    }
    st->resident += (int64_t)(ring_bytes * n_slots);
    std::vector<void *> drop;
    int64_t room = 0;
    bool fits = false;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        room = S->cfg.memory_budget - S->resident - S->inflight;
        fits = st->resident <= room;
        if (fits) {
            drop = S->dev_pool.trim((size_t)(room - st->resident));
            S->resident += st->resident;   // reserved before the allocations below
            S->peak = std::max(S->peak, S->resident + S->inflight);
        }
    }
    free_device_blocks(drop);
    if (!fits) {
        for (xme_plan *q : st->plans)
            xme_plan_destroy(q);
        xme_set_error("xme_server_open_stream: %lld resident bytes exceed the %lld bytes the memory budget has free",
                      (long long)st->resident, (long long)room);
        return XME_ERR_BUDGET;
    }
    hipError_t e = hipSuccess;
    for (int i = 0; i < n_slots && e == hipSuccess; i++) {
        void *p = nullptr;
        e = hipMalloc(&p, ring_bytes);
        if (e == hipSuccess) {
            st->ring.push_back(static_cast<uint8_t *>(p));
            // a frame-0 search never reads it, but keep the planes defined
            e = hipMemset(p, 0, ring_bytes);
        }
    }
    for (auto &ev : st->ref_ready)
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        int rc = xme_hip_fail(e, "xme_server_open_stream");
        const int64_t r = st->resident;
        free_stream(st.release());
        std::lock_guard<std::mutex> lk(S->mu);
        S->resident -= r;
        return rc;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    Stream *sp = st.release();
    sp->id = S->next_stream++;
    *stream_id = sp->id;
    S->streams[sp->id] = sp;
    return XME_OK;
}

extern "C" int xme_server_close_stream(xme_server *S, int stream_id)
{
    if (!S)
        return XME_ERR_ARG;
    Stream *st = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->streams.find(stream_id);
        if (it == S->streams.end() || it->second->closing) {
            xme_set_error("xme_server_close_stream: unknown stream %d", stream_id);
            return XME_ERR_ARG;
        }
        st = it->second;
        st->closing = true;
        // every admitted frame's reference build has finished once building is 0
        S->cv_done.wait(lk, [st] { return st->pending == 0 && st->building == 0; });
        for (auto &kv : S->jobs)
            if (kv.second->st == st)
                kv.second->st = nullptr;
        S->streams.erase(it);
    }
    const int64_t r = st->resident;
    free_stream(st);
    std::lock_guard<std::mutex> lk(S->mu);
    S->resident -= r;
    S->cv_sched.notify_all();
    return XME_OK;
}

extern "C" int xme_server_submit(xme_server *S, int stream_id, const pel_t *luma, intptr_t stride, int64_t *job_id)
{
    if (!S || !luma || !job_id) {
        xme_set_error("xme_server_submit: bad arguments");
        return XME_ERR_ARG;
    }
    Stream *st = nullptr;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->streams.find(stream_id);
        if (it == S->streams.end() || it->second->closing) {
            xme_set_error("xme_server_submit: unknown stream %d", stream_id);
            return XME_ERR_ARG;
        }
        st = it->second;
        if (stride < st->w) {
            xme_set_error("xme_server_submit: stride %lld < width %d", (long long)stride, st->w);
            return XME_ERR_ARG;
        }
        if (st->failed != XME_OK) {
            xme_set_error("xme_server_submit: stream %d cannot go on (a reference build failed)", stream_id);
            return st->failed;
        }
        st->pending++;           // holds the stream open until this submission ends
    }
    auto release = [S, st] {
        std::lock_guard<std::mutex> lk(S->mu);
        st->pending--;
        S->cv_done.notify_all();
    };
    std::lock_guard<std::mutex> order(st->submit_mu);
    auto j = std::make_unique<Job>();
    j->st = st;
    j->srv = S;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        j->frame = st->next_frame;
        j->n_refs = (int)std::min<int64_t>(j->frame, S->cfg.n_ref);
        j->n_results = st->n_pu * j->n_refs;
        j->d_cur_bytes = align_up((size_t)st->cur_stride * st->h, 256);
        j->d_res_bytes = sizeof(xme_result) * (size_t)j->n_results;
        j->dev_bytes = j->d_cur_bytes + align_up(j->d_res_bytes, 256);
        if (S->resident + (int64_t)j->dev_bytes > S->cfg.memory_budget) {
            S->rejected++;
            xme_set_error("xme_server_submit: a frame of stream %d needs %zu bytes beside the resident %lld; "
                          "budget %lld",
                          stream_id, j->dev_bytes, (long long)S->resident, (long long)S->cfg.memory_budget);
            st->pending--;
            S->cv_done.notify_all();
            return XME_ERR_BUDGET;
        }
    }
    j->h_pic_bytes = (size_t)st->cur_stride * st->h;
    j->h_res_bytes = std::max<size_t>(j->d_res_bytes, 64);
    j->h_pic = static_cast<uint8_t *>(host_get(S, j->h_pic_bytes));
    j->h_res = static_cast<xme_result *>(host_get(S, j->h_res_bytes));
    if (!j->h_pic || !j->h_res) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->host_pool.put(j->h_pic, j->h_pic_bytes);
            S->host_pool.put(j->h_res, j->h_res_bytes);
        }
        release();
        xme_set_error("xme_server_submit: pinned host allocation failed");
        return XME_ERR_NOMEM;
    }
    for (int y = 0; y < st->h; y++)
        memcpy(j->h_pic + (size_t)y * st->cur_stride, luma + (size_t)y * stride, (size_t)st->w);
    std::lock_guard<std::mutex> lk(S->mu);
    Job *jp = j.release();
    st->next_frame++;            // frame order of the stream = submission order
    jp->t_submit = Clock::now();
    // twice the frame's search work at a nominal 100 T absdiff/s
    jp->deadline = std::chrono::duration<double>(jp->t_submit - S->t_start).count() +
                   2.0 * (double)st->work * jp->n_refs / 1e14;
    jp->id = S->next_job++;
    if (!S->have_first) {
        S->have_first = true;
        S->t_first = jp->t_submit;
    }
    S->submitted++;
    S->jobs[jp->id] = jp;
    S->queue.push_back(jp);      // pending stays counted until the job is done
    *job_id = jp->id;
    S->cv_sched.notify_all();
    return XME_OK;
}

extern "C" int xme_server_job_size(xme_server *S, int64_t job_id, int *n_results)
{
    if (!S || !n_results)
        return XME_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) {
        xme_set_error("xme_server_job_size: unknown job %lld", (long long)job_id);
        return XME_ERR_ARG;
    }
    *n_results = it->second->n_results;
    return XME_OK;
}

extern "C" int xme_server_wait(xme_server *S, int64_t job_id, xme_result *results, double *latency_ms)
{
    if (!S)
        return XME_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) {
        xme_set_error("xme_server_wait: unknown job %lld", (long long)job_id);
        return XME_ERR_ARG;
    }
    Job *j = it->second;
    S->cv_done.wait(lk, [j] { return j->done; });
    const int rc = j->rc;
    if (rc == XME_OK && results && j->n_results)
        memcpy(results, j->h_res, sizeof(xme_result) * (size_t)j->n_results);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
    S->host_pool.put(j->h_res, j->h_res_bytes);
    S->jobs.erase(it);
    delete j;
    if (rc != XME_OK)
        xme_set_error("xme_server_wait: job %lld failed", (long long)job_id);
    return rc;
}

extern "C" int xme_server_stats_get(xme_server *S, xme_server_stats *out)
{
    if (!S || !out)
        return XME_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    memset(out, 0, sizeof(*out));
    out->submitted = S->submitted;
    out->completed = S->completed;
    out->rejected = S->rejected;
    out->resident_bytes = S->resident;
    out->inflight_bytes = S->inflight;
    out->peak_bytes = S->peak;
    if (!S->lat_ms.empty()) {
        std::vector<double> v = S->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&v](double q) {
            // nearest rank
            size_t k = (size_t)std::ceil(q * (double)v.size());
            return v[std::min(v.size() - 1, k > 0 ? k - 1 : 0)];
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v)
            sum += x;
        out->mean_ms = sum / (double)v.size();
        out->max_ms = v.back();
        const double span = std::chrono::duration<double>(S->t_last - S->t_first).count();
        out->throughput_fps = span > 0 ? (double)v.size() / span : 0.0;
    }
    return XME_OK;
}

extern "C" int xme_server_stats_reset(xme_server *S)
{
    if (!S)
        return XME_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    S->submitted = S->completed = S->rejected = 0;
    S->lat_ms.clear();
    S->have_first = false;
    S->peak = S->resident + S->inflight;
    return XME_OK;
}
