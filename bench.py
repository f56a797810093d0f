#!/usr/bin/env python3
"""Benchmark of the MI355X motion-estimation cost path.

Workload (one step = one frame-level ME job): batched multi-reference ME at
3840x2160 8-bit -- 4 reference pictures (seeded smooth + noise fields), a
current picture built from per-region mixtures of the references displaced by
MVs in [-48, 48]^2, a seeded mixed PU tiling (8x8 ... 64x64 incl. 64x16,
16x64, 32x8, 8x32), full search over [-64, 64)^2 with cost SAD + lambda * MV
bits (lambda factor 4 * 65536, MVP (0,0)), then half- and quarter-pel SATD
refinement.  The whole frame is one persistent launch replayed from a HIP
graph.

Printed: one JSON line with candidate evaluations per second (batched path),
the roofline of the integer-search kernel, the plain one-launch-per-PU-shape
GPU path on the same job, per-frame latency p50/p99, and the CPU oracle timed
on a bounded sample of the same job.  Side legs: the DIA, HEX and UMH
pattern searches on the same frame, the residual transform, and the serving
loop.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ME cand-evals/s + roofline frac, batched vs plain; p50/p99 frame-ME latency"
VALU_PEAK_TABSDIFF = 256 * 4 * 16 * 2.4e9 * 4 / 1e12     # v_sad_u8 / v_qsad: 157.3 T absdiff/s
HBM_PEAK_GBS = 8000.0
# The copy of known size that scales the HBM traffic counters, run beside the
# search step in each counter pass.
TRAFFIC_COPY_MIB = 256


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def build_inputs(w, h, n_ref, seed):
    from xavs2_amd import synth
    refs = [synth.smooth_field(w, h, seed=seed + 101 * r) for r in range(n_ref)]
    cur = synth.multi_ref_current(refs, 32, 48, seed=seed + 7)
    pus = synth.mixed_pus(w, h, seed=seed + 3)
    return refs, cur, pus


def legal_counts(w, h, x, y, pw, ph, R, mvr_x=(-8192, 8191), mvr_y=(-2048, 2047)):
    """Legal full-pel displacements per axis of a PU searched around (0,0)
    (MV limits of md_inter.c:910-926, raster [-R, R))."""
    def axis(pos, size, frame, lim):
        lo = min(max((-64 - pos) * 4, lim[0]), lim[1])
        hi = min(max((frame + 64 - pos - size) * 4, lim[0]), lim[1])
        lo, hi = (lo >> 2) + 6, (hi >> 2) - 6
        return max(0, min(R - 1, hi) - max(-R, lo) + 1)
    return axis(x, pw, w, mvr_x), axis(y, ph, h, mvr_y)


def host_cpus():
    """The host's CPUs as this process sees them: nproc (os.cpu_count), the
    CPUs it may run on (affinity), the cgroup CPU quota, the CPU model, and the
    worker count used for CPU work: the CPUs this job is given, at most 16
    (the host's share per GPU)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = affinity if quota is None else max(1, min(affinity, int(quota)))
    return {"nproc": nproc, "affinity": affinity, "cgroup_quota_cpus": quota, "model": model,
            "threads": max(1, min(16, avail))}


def cpu_baseline(prm_args, refs, cur, pus, gpu_res, n_ref, budget_s, threads, host=None):
    """The oracle (CPU restatement) timed on a bounded sample of the same job:
    all PUs of the first CTU rows against reference 0, grown until the sample
    takes about budget_s seconds, on `threads` workers; then the first CTU row
    on one thread.  Also checks the GPU results of the sample."""
    from oracle import oracle as O
    w, h = refs[0].shape[1], refs[0].shape[0]
    p = O.Plane(w, h)
    p.set_picture(refs[0])
    p.expand_border()
    planes = O.interp_planes(p, 2)
    cp = O.Plane(w, h)
    cp.set_picture(cur)
    prm = O.me_param(w, h, *prm_args)
    rows = sorted({pu[1] // 64 for pu in pus})
    sel_rows = 1
    elapsed = 0.0
    while True:
        idx = [i for i, pu in enumerate(pus) if pu[1] // 64 < sel_rows]
        arr = np.array([(pus[i][0], pus[i][1], pus[i][2], pus[i][3], 0, 0) for i in idx], dtype=O.PU_DTYPE)
        t0 = time.perf_counter()
        res = O.me_search_batch(prm, arr, cp, planes, nthreads=threads)
        elapsed = time.perf_counter() - t0
        if elapsed * 2 > budget_s or sel_rows >= len(rows):
            break
        sel_rows = min(len(rows), max(sel_rows + 1, int(sel_rows * budget_s / max(elapsed, 1e-3) * 0.8)))
    # legal integer candidates of the sample (MVP (0,0): centre (0,0), window clipped by the MV range)
    R = prm.search_range
    n_cand = 0
    for i in idx:
        x, y, pw, ph = pus[i]
        nx, ny = legal_counts(w, h, x, y, pw, ph, R)
        n_cand += nx * ny
    ok = True
    for k, i in enumerate(idx):
        g = gpu_res[i * n_ref + 0]
        for f in ("cost", "bcost2", "bmv_x", "bmv_y", "bmv2_x", "bmv2_y", "mvcost", "status"):
            if int(g[f]) != int(res[k][f]):
                ok = False
    # one thread: the PUs of the first CTU row that fit about a quarter of the budget
    row0 = [i for i in idx if pus[i][1] // 64 == 0]
    n1 = len(row0)
    while True:
        sub = row0[:n1]
        arr1 = np.array([(pus[i][0], pus[i][1], pus[i][2], pus[i][3], 0, 0) for i in sub], dtype=O.PU_DTYPE)
        t0 = time.perf_counter()
        O.me_search_batch(prm, arr1, cp, planes, nthreads=1)
        t1 = time.perf_counter() - t0
        if t1 < budget_s / 2 or n1 <= 8:
            break
        n1 = max(8, int(n1 * budget_s / 4 / t1))
    c1 = sum(legal_counts(w, h, pus[i][0], pus[i][1], pus[i][2], pus[i][3], R)[0] *
             legal_counts(w, h, pus[i][0], pus[i][1], pus[i][2], pus[i][3], R)[1] for i in sub)
    host = host or host_cpus()
    return {
        "value": n_cand / elapsed,
        "unit": "cand-evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(idx)} PUs ({sel_rows} CTU rows of 3840x2160) x 1 ref, full search [-64,64)^2 + "
                  f"half/quarter-pel SATD, {elapsed:.2f} s on {threads} threads",
        "single_thread": {"value": c1 / t1, "unit": "cand-evals/s",
                          "sample": f"{len(sub)} PUs of the first CTU row, {t1:.2f} s on 1 thread"},
        "host": host,
        "parity_with_gpu_on_sample": ok,
    }


def umh_contexts(X, n_pu, n_ref, seed):
    """Seeded synthetic encoder state for the UMH search, one record per (PU,
    reference): predictor MVs within +-24 pixels (a third of mvp1 and mvp3
    absent), ref_idx = the reference index, predicted SADs up to 6000 and
    betas in [-1, 0.6] / [-1, 0.2], so that a share of the searches stops
    early at each termination point."""
    rng = np.random.default_rng(seed)
    n = n_pu * n_ref
    c = np.zeros(n, dtype=X.UMH_CTX_DTYPE)
    for f in ("mvp1", "mvp2", "mvp3"):
        for a in ("_x", "_y"):
            c[f + a] = rng.integers(-96, 97, n)
    for f in ("mvp1", "mvp3"):
        z = rng.random(n) < 0.3
        c[f + "_x"][z] = 0
        c[f + "_y"][z] = 0
    c["ref_idx"] = np.arange(n) % n_ref
    c["pred_sad"] = rng.integers(0, 6000, n)
    c["pred_sad_uplayer"] = rng.integers(0, 6000, n)
    c["beta2"] = rng.uniform(-1.0, 0.6, n)
    c["beta3"] = rng.uniform(-1.0, 0.2, n)
    return c


def pattern_leg(X, w, h, R, lam, pus, n_ref, dcur, d_refs, refs, cur, stream, steps, cpu):
    """The DIA, HEX and UMH integer searches (+ the same sub-pel refinement)
    on the same frame and PUs: job time per frame (graph replays),
    search-kernel time (HIP events), integer points evaluated, and parity with
    the oracle for every PU against reference 0, which is also the CPU time
    reported.  UMH runs at big-hexagon level 2 on seeded synthetic encoder
    state (umh_contexts)."""
    from oracle import oracle as O
    pu_arr = X.make_pus([(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in pus])
    planes0, cp = None, None
    if cpu:
        p = O.Plane(w, h)
        p.set_picture(refs[0])
        p.expand_border()
        planes0 = O.interp_planes(p, 2)
        cp = O.Plane(w, h)
        cp.set_picture(cur)
    threads = host_cpus()["threads"]
    out = {}
    ctx = umh_contexts(X, len(pus), n_ref, seed=2027)
    for name, method in (("dia", X.ME_DIA), ("hex", X.ME_HEX), ("umh", X.ME_UMH)):
        level = 2 if method == X.ME_UMH else 0
        plan = X.Plan(X.me_params(w, h, R, lam, me_method=method, umh_big_hex_level=level), pu_arr, n_ref,
                      X.PLAN_BATCHED)
        if method == X.ME_UMH:
            plan.set_umh_context(ctx)
        for _ in range(2):
            plan.run_graph(dcur, d_refs, stream.h)
        stream.sync()
        res = plan.fetch()
        X.lib().xme_device_sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.run_graph(dcur, d_refs, stream.h)
        stream.sync()
        X.lib().xme_device_sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        points, absdiffs = plan.search_count()
        search_ms, job_ms, _ = plan.time(dcur, d_refs, stream.h, 5)
        stable = bool(np.array_equal(plan.fetch(), res))
        n_inst = len(pus) * n_ref
        leg = {
            "ms_per_frame": ms,
            "pu_refs_per_s": n_inst / (ms * 1e-3),
            "search_kernel_ms": search_ms,
            "job_ms_events": job_ms,
            "points_per_frame": points,
            "points_per_pu_ref": points / n_inst,
            "absdiffs_per_frame": absdiffs,
            "search_tabsdiff_per_s": absdiffs / (search_ms * 1e-3) / 1e12,
            "results_stable_across_replays": stable,
        }
        if cpu:
            arr = np.array([(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in pus], dtype=O.PU_DTYPE)
            prm_o = O.me_param(w, h, R, lam, me_method=method, umh_big_hex_level=level)
            c0 = np.ascontiguousarray(ctx[0::n_ref]).view(O.UMH_CTX_DTYPE) if method == X.ME_UMH else None
            t0 = time.perf_counter()
            exp = O.me_search_batch(prm_o, arr, cp, planes0, nthreads=threads, ctxs=c0)
            el = time.perf_counter() - t0
            got = res[0::n_ref]
            leg["parity_with_oracle_ref0"] = bool(all(np.array_equal(got[f], exp[f]) for f in exp.dtype.names))
            leg["cpu_oracle"] = {"pu_refs_per_s": len(pus) / el, "threads": threads, "seconds": el,
                                 "sample": f"all {len(pus)} PUs x reference 0"}
        log(f"{name}: {ms:.3f} ms/frame, search {search_ms:.3f} ms, {points / n_inst:.1f} points per PU-ref")
        out[name] = leg
        del plan
    return out


def bipred_leg(X, w, h, R, lam, pus, gpu_res, n_ref, dcur, d_refs, refs, cur, reps, cpu):
    """The bi-prediction refinements of every PU between references 0 and 1,
    as the encoder runs them: symmetric (xme_me_search_sym_batch; the
    dual-hypothesis search, md_inter.c:1165-1186, start = 4 x the PU's
    full-pel best against reference 0, F-frame distances 2 and 4, so the second
    MV is twice the first; 5 half- and 4 quarter-pel points) and
    bi-directional (xme_me_search_bid_batch; md_inter.c:1089-1094, forward
    start and fixed backward MV = 4 x the full-pel bests against references 0
    and 1; 9 half- and 8 quarter-pel points).  MVPs (0,0), SATD.  Kernel time
    from back-to-back launches on one stream; parity and the CPU time on
    every 4th PU."""
    import ctypes as C
    from oracle import oracle as O
    r0, r1 = gpu_res[0::n_ref], gpu_res[1::n_ref]
    arr = np.array(pus, dtype=np.int64)
    n = len(pus)
    sym = np.zeros(n, dtype=X.SYM_JOB_DTYPE)
    bid = np.zeros(n, dtype=X.BID_JOB_DTYPE)
    for jobs in (sym, bid):
        jobs["x"], jobs["y"], jobs["w"], jobs["h"] = arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3]
    sym["mv_x"], sym["mv_y"] = 4 * r0["bmv2_x"], 4 * r0["bmv2_y"]
    sym["dist_fwd"], sym["dist_bwd"], sym["b_slice"] = 2, 4, 0
    bid["fwd_x"], bid["fwd_y"] = 4 * r0["bmv2_x"], 4 * r0["bmv2_y"]
    bid["bwd_x"], bid["bwd_y"] = 4 * r1["bmv2_x"], 4 * r1["bmv2_y"]
    prm = X.me_params(w, h, R, lam)
    d1, d2 = d_refs[0].desc(), d_refs[1].desc()
    stream = X.Stream()
    area = int((arr[:, 2] * arr[:, 3]).sum())
    planes, cp = None, None
    if cpu:
        planes = []
        for r in refs[:2]:
            p = O.Plane(w, h)
            p.set_picture(r)
            p.expand_border()
            planes.append(O.interp_planes(p, 2))
        cp = O.Plane(w, h)
        cp.set_picture(cur)
    prm_o = O.me_param(w, h, R, lam)
    out = {}
    for name, jobs, entry, points in (("sym", sym, "xme_me_search_sym_batch", 9),
                                      ("bid", bid, "xme_me_search_bid_batch", 17)):
        d_jobs = X.upload(jobs)
        d_res = X.DeviceBuffer(n * X.SYM_RESULT_DTYPE.itemsize)

        def launch():
            X.check(getattr(X.lib(), entry)(C.byref(prm), C.c_void_p(d_jobs.ptr), n, C.c_void_p(dcur.origin),
                                            dcur.stride, C.byref(d1), C.byref(d2), C.c_void_p(d_res.ptr),
                                            stream.h), entry)

        for _ in range(3):
            launch()
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            launch()
        stream.sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        res = d_res.download(n, X.SYM_RESULT_DTYPE)
        leg = {
            "ms_per_batch": ms,
            "pus_per_s": n / (ms * 1e-3),
            "candidate_pixels_per_s": area * points / (ms * 1e-3),
            "points_per_pu": points,
            "status_ok": bool(np.all(res["status"] == 0)),
        }
        if cpu:
            idx = range(0, n, 4)
            t0 = time.perf_counter()
            if name == "sym":
                exp = [O.me_search_sym(prm_o, dict(pix_x=int(jobs["x"][i]), pix_y=int(jobs["y"][i]),
                                                   w=int(jobs["w"][i]), h=int(jobs["h"][i]),
                                                   mv_x=int(jobs["mv_x"][i]), mv_y=int(jobs["mv_y"][i]), mvp_x=0,
                                                   mvp_y=0, mvp1_x=0, mvp1_y=0, dist_fwd=2, dist_bwd=4, b_slice=0),
                                       cp, planes[0], planes[1]) for i in idx]
            else:
                exp = [O.me_search_bid(prm_o, dict(pix_x=int(jobs["x"][i]), pix_y=int(jobs["y"][i]),
                                                   w=int(jobs["w"][i]), h=int(jobs["h"][i]),
                                                   fwd_x=int(jobs["fwd_x"][i]), fwd_y=int(jobs["fwd_y"][i]),
                                                   bwd_x=int(jobs["bwd_x"][i]), bwd_y=int(jobs["bwd_y"][i]),
                                                   mvp1_x=0, mvp1_y=0, mvp2_x=0, mvp2_y=0),
                                       cp, planes[0], planes[1]) for i in idx]
            el = time.perf_counter() - t0
            got = [(int(res["cost"][i]), int(res["mv_x"][i]), int(res["mv_y"][i]), int(res["mvcost"][i]))
                   for i in idx]
            leg["parity_with_oracle_every_4th"] = got == [tuple(int(v) for v in e) for e in exp]
            leg["cpu_oracle"] = {"pus_per_s": len(exp) / el, "threads": 1, "seconds": el,
                                 "sample": f"every 4th PU ({len(exp)}), one thread, Python loop over the C oracle"}
        log(f"{name}: {ms:.3f} ms per batch of {n} PUs")
        out[name] = leg
    out["workload"] = (f"{n} PUs of the frame, references 0 and 1; sym: start = 4 x full-pel best of reference 0, "
                       "distances 2 / 4 (F frame); bid: forward / backward = 4 x full-pel bests of references "
                       "0 / 1; MVP (0,0), SATD, quarter-pel")
    out["callers"] = callers_leg(X, w, h, R, lam, arr, gpu_res, n_ref, dcur, d_refs, refs, cp, planes, reps, cpu,
                                 stream)
    return out


def callers_leg(X, w, h, R, lam, arr, gpu_res, n_ref, dcur, d_refs, refs, cp, planes, reps, cpu, stream):
    """The callers of the refinements over the frame's PUs, LCU level 6, one
    frame thread: pred_inter_search_bi (xme_pred_inter_search_bi_batch, a B
    frame between references 0 and 1 at distances 2 / 2: the gates, the
    symmetric search from 4 x the full-pel best of reference 0 and the
    bi-directional one with the backward MV 4 x the full-pel best of
    reference 1) and pred_inter_search_dual (xme_pred_inter_search_dual_batch,
    an F frame over every reference at distances 1, 2, ...: per reference the
    gates and the symmetric search from its full-pel best, plus REF_COST).
    MVPs (0,0).  Parity with the oracle and the CPU time on every 8th PU."""
    import ctypes as C
    from oracle import oracle as O
    n = len(arr)
    bi = np.zeros(n, dtype=X.BI_JOB_DTYPE)
    dual = np.zeros(n, dtype=X.DUAL_JOB_DTYPE)
    for jobs in (bi, dual):
        jobs["x"], jobs["y"], jobs["w"], jobs["h"] = arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3]
    r0, r1 = gpu_res[0::n_ref], gpu_res[1::n_ref]
    bi["fwd_imv_x"], bi["fwd_imv_y"] = r0["bmv2_x"], r0["bmv2_y"]
    bi["bwd_x"], bi["bwd_y"] = 4 * r1["bmv2_x"], 4 * r1["bmv2_y"]
    nd = min(n_ref, 4)
    dist = list(range(1, nd + 1))
    for r in range(nd):
        rr = gpu_res[r::n_ref]
        dual["imv"][:, r, 0], dual["imv"][:, r, 1] = rr["bmv2_x"], rr["bmv2_y"]
    prm = X.me_params(w, h, R, lam)
    L = X.lib()
    d_bi, d_dual = X.upload(bi), X.upload(dual)
    d_bres = X.DeviceBuffer(n * X.BI_RESULT_DTYPE.itemsize)
    d_dres = X.DeviceBuffer(n * X.DUAL_RESULT_DTYPE.itemsize)
    r01 = (d_refs[0].desc(), d_refs[1].desc())
    descs = (X.RefDesc * 4)(*[r.desc() for r in d_refs[:nd]])
    dists = (C.c_int * 4)(*(dist + [0] * (4 - nd)))

    def run_bi():
        X.check(L.xme_pred_inter_search_bi_batch(C.byref(prm), 6, 1, C.c_void_p(d_bi.ptr), n, 2, 2,
                                                 C.c_void_p(dcur.origin), dcur.stride, C.byref(r01[0]),
                                                 C.byref(r01[1]), C.c_void_p(d_bres.ptr), stream.h), "bi caller")

    def run_dual():
        X.check(L.xme_pred_inter_search_dual_batch(C.byref(prm), 6, 1, C.c_void_p(d_dual.ptr), n, nd, dists,
                                                   C.c_void_p(dcur.origin), dcur.stride, descs,
                                                   C.c_void_p(d_dres.ptr), stream.h), "dual caller")

    out = {}
    for name, run, d_res, dtype, points in (("bi", run_bi, d_bres, X.BI_RESULT_DTYPE, 26),
                                            ("dual", run_dual, d_dres, X.DUAL_RESULT_DTYPE, 9 * nd)):
        for _ in range(3):
            run()
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        stream.sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        res = d_res.download(n, dtype)
        leg = {"ms_per_batch": ms, "pus_per_s": n / (ms * 1e-3), "searched_points_per_pu_at_most": points,
               "status_ok": bool(np.all(res["status"] == 0))}
        if name == "bi":
            leg["sym_ran"] = int(np.sum(res["sym_mvcost"] >= 0))
            leg["bid_ran"] = int(np.sum(res["bid_mvcost"] >= 0))
        else:
            leg["references"] = nd
            leg["winners"] = {str(r): int(np.sum(res["fst_ref"] == r)) for r in range(-1, nd)}
        if cpu:
            idx = range(0, n, 8)
            prm_o = O.me_param(w, h, R, lam)
            t0 = time.perf_counter()
            if name == "bi":
                exp = [O.pred_inter_search_bi(prm_o, {"pix_x": int(bi["x"][i]), "pix_y": int(bi["y"][i]),
                                                      "w": int(bi["w"][i]), "h": int(bi["h"][i]),
                                                      "fwd_imv_x": int(bi["fwd_imv_x"][i]),
                                                      "fwd_imv_y": int(bi["fwd_imv_y"][i]),
                                                      "bwd_x": int(bi["bwd_x"][i]), "bwd_y": int(bi["bwd_y"][i]),
                                                      "mvp_x": 0, "mvp_y": 0, "mvp1_x": 0, "mvp1_y": 0,
                                                      "mvp2_x": 0, "mvp2_y": 0}, 2, 2, cp, planes[0], planes[1])
                       for i in idx]
                got = [tuple(int(res[f][i]) for f in ("sym_cost", "sym_mv_x", "sym_mv_y", "sym_mvcost",
                                                      "bid_cost", "fwd_x", "fwd_y", "bwd_x", "bwd_y",
                                                      "bid_mvcost")) for i in idx]
                exp = [tuple(e[f] for f in ("sym_cost", "sym_mv_x", "sym_mv_y", "sym_mvcost", "bid_cost", "fwd_x",
                                            "fwd_y", "bwd_x", "bwd_y", "bid_mvcost")) for e in exp]
            else:
                for r in refs[len(planes):nd]:
                    p = O.Plane(w, h)
                    p.set_picture(r)
                    p.expand_border()
                    planes.append(O.interp_planes(p, 2))
                exp = [O.pred_inter_search_dual(prm_o, {"pix_x": int(dual["x"][i]), "pix_y": int(dual["y"][i]),
                                                        "w": int(dual["w"][i]), "h": int(dual["h"][i]),
                                                        "mvp_x": 0, "mvp_y": 0, "mvp1": [(0, 0)] * nd,
                                                        "imv": [tuple(int(v) for v in dual["imv"][i, r])
                                                                for r in range(nd)]},
                                                dist, cp, planes[:nd]) for i in idx]
                got = [(int(res["cost"][i]), int(res["fst_ref"][i]), int(res["snd_ref"][i]), int(res["mvcost"][i]),
                        [tuple(int(v) for v in res["fst"][i, r]) for r in range(nd)],
                        [tuple(int(v) for v in res["snd"][i, r]) for r in range(nd)]) for i in idx]
                exp = [tuple(e) for e in exp]
            el = time.perf_counter() - t0
            leg["parity_with_oracle_every_8th"] = got == exp
            leg["cpu_oracle"] = {"pus_per_s": len(exp) / el, "threads": 1, "seconds": el,
                                 "sample": f"every 8th PU ({len(exp)}), one thread, Python loop over the C oracle"}
        log(f"{name} caller: {ms:.3f} ms per batch of {n} PUs")
        out[name] = leg
    return out


def recon_leg(X, w, h, pus, gpu_res, n_ref, dcur, d_refs, refs, cur, reps, cpu, qp=32):
    """Luma inter reconstruction of the frame (xme_inter_recon_batch): every
    PU split into the largest square TUs (4..64) that tile it, a 64x64 PU
    one 64x64 TU (wavelet), predicted from reference 0 at 4 x its full-pel
    best, at QP 32; one call per TU size, each one launch (the planes are
    resident; no prediction-error or residual buffers).  Algorithmic HBM
    bytes: 3 per pixel (current and prediction read, reconstruction written)
    and 2 per level (n x n levels, 32 x 32 for a 64x64 TU).  Parity with the
    oracle on the TUs of every 4th PU."""
    import ctypes as C
    from oracle import oracle as O
    r0 = gpu_res[0::n_ref]
    by_n = {4: [], 8: [], 16: [], 32: [], 64: []}
    sample = {4: [], 8: [], 16: [], 32: [], 64: []}
    for i, (x, y, pw, ph) in enumerate(pus):
        n = 64
        while pw % n or ph % n:
            n >>= 1
        mv = (4 * int(r0["bmv2_x"][i]), 4 * int(r0["bmv2_y"][i]))
        for ty in range(y, y + ph, n):
            for tx in range(x, x + pw, n):
                by_n[n].append((tx, ty, x, y, pw, ph, mv[0], mv[1]))
                sample[n].append(i % 4 == 0)
    rec = X.DevicePlane(w, h, pad=0, stride=dcur.stride)
    rd = d_refs[0].desc()
    stream = X.Stream()
    bufs = []
    for n, rows in by_n.items():
        if not rows:
            continue
        t = np.array(rows, np.int32)
        ntu = len(t)
        nl = min(n, 32)
        bufs.append((n, ntu, X.upload(t), None, X.DeviceBuffer(2 * ntu * nl * nl), None, X.DeviceBuffer(4 * ntu),
                     X.DeviceBuffer(4 * ntu)))

    def frame():
        for n, ntu, d_t, _resid, level, _rres, nnz, ssd in bufs:
            X.check(X.lib().xme_inter_recon_batch(n, qp, C.c_void_p(dcur.origin), dcur.stride, C.byref(rd), w, h,
                                                  d_t.at(), ntu, None, level.at(), None,
                                                  C.c_void_p(rec.origin), rec.stride, nnz.at(), ssd.at(), stream.h),
                    "xme_inter_recon_batch")

    for _ in range(3):
        frame()
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        frame()
    stream.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    pixels = sum(ntu * n * n for n, ntu, *_ in bufs)
    alg_bytes = sum(ntu * (3 * n * n + 2 * min(n, 32) ** 2) for n, ntu, *_ in bufs)
    leg = {
        "workload": f"{w}x{h} luma, {sum(b[1] for b in bufs)} TUs ("
                    + ", ".join(f"{b[1]} of {b[0]}x{b[0]}" for b in bufs) + f"), reference 0, QP {qp}",
        "ms_per_frame": ms,
        "pixels_per_s": pixels / (ms * 1e-3),
        "algorithmic_bytes_per_frame": alg_bytes,
        "algorithmic_gb_per_s": alg_bytes / (ms * 1e-3) / 1e9,
        "hbm_frac": alg_bytes / (ms * 1e-3) / 8e12,
    }
    if cpu:
        p = O.Plane(w, h)
        p.set_picture(refs[0])
        p.expand_border()
        planes = O.interp_planes(p, 2)
        cp = O.Plane(w, h)
        cp.set_picture(cur)
        full = rec.buf.download(rec.rows * rec.stride, np.uint8).reshape(rec.rows, rec.stride)[:h, :w]
        ok, el, n_tu = True, 0.0, 0
        for n, ntu, d_t, resid, level, rres, nnz, ssd in bufs:
            sel = np.array(sample[n])
            t = np.array(by_n[n], np.int32)[sel]
            t0 = time.perf_counter()
            el_level, el_rec, el_nnz, el_ssd = O.inter_recon_batch(n, qp, cp, planes, w, h, t)
            el += time.perf_counter() - t0
            n_tu += len(t)
            nl = min(n, 32)
            g_level = level.download(ntu * nl * nl, np.int16).reshape(ntu, nl, nl)[sel]
            ok = ok and np.array_equal(g_level, el_level) and np.array_equal(nnz.download(ntu, np.int32)[sel], el_nnz)
            ok = ok and np.array_equal(ssd.download(ntu, np.int32)[sel], el_ssd)
            for (tx, ty, *_r) in t:
                ok = ok and np.array_equal(full[ty:ty + n, tx:tx + n], el_rec[ty:ty + n, tx:tx + n])
        leg["parity_with_oracle_every_4th_pu"] = bool(ok)
        leg["cpu_oracle"] = {"pixels_per_s": sum(len(np.array(by_n[n])[np.array(sample[n])]) * n * n
                                                 for n in by_n if by_n[n]) / el,
                             "threads": 1, "seconds": el, "sample": f"the {n_tu} TUs of every 4th PU"}
    # the luma MC of every PU into a prediction picture (xme_mc_luma_batch),
    # the input of xme_recon_pred_batch for CUs of several PUs
    mvs = [(4 * int(r0["bmv2_x"][i]), 4 * int(r0["bmv2_y"][i])) for i in range(len(pus))]
    pblocks = np.array([(x, y, pw, ph) + mvs[i] for i, (x, y, pw, ph) in enumerate(pus)], np.int32)
    pred = X.DevicePlane(w, h, pad=0, stride=dcur.stride)
    poff = np.array([int(b[1]) * pred.stride + int(b[0]) for b in pblocks], np.int64)
    d_pb, d_poff = X.upload(pblocks.ravel()), X.upload(poff)

    def mc():
        X.check(X.lib().xme_mc_luma_batch(C.byref(rd), d_pb.at(), len(pblocks), w, h, C.c_void_p(pred.origin),
                                          pred.stride, d_poff.at(), stream.h), "xme_mc_luma_batch")

    for _ in range(3):
        mc()
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        mc()
    stream.sync()
    leg["mc_luma_ms_per_frame"] = (time.perf_counter() - t0) * 1e3 / reps
    if cpu:
        gp = pred.buf.download(pred.rows * pred.stride, np.uint8).reshape(pred.rows, pred.stride)
        ok = True
        for i in range(0, len(pblocks), 8):
            x, y, pw, ph, mx, my = (int(v) for v in pblocks[i])
            qx = O.lib().xo_get_mc_pos(w, pw, x, mx)
            qy = O.lib().xo_get_mc_pos(h, ph, y, my)
            e = np.zeros((ph, pw), np.uint8)
            O.lib().xo_mc_luma(O.ptr(e), pw, qx, qy, pw, ph, O.plane_ptrs(planes), planes[0].stride)
            ok = ok and np.array_equal(gp[y:y + ph, x:x + pw], e)
        leg["mc_luma_parity_with_oracle_every_8th_pu"] = bool(ok)
    log(f"recon: {ms:.3f} ms per frame; luma MC of every PU {leg['mc_luma_ms_per_frame']:.3f} ms")
    return leg


def chroma_cus(pus):
    """The CU of each PU of the mixed tiling: a square PU is its own CU; the
    four 64x16 / 16x64 PUs of a 64x64 CU and the four 32x8 / 8x32 of a 32x32
    share theirs.  Returns {(x, y, size)} in first-seen order."""
    seen = {}
    for (x, y, pw, ph) in pus:
        s = max(pw, ph)
        cu = (x & ~(s - 1), y & ~(s - 1), s)
        seen.setdefault(cu, None)
    return list(seen)


def chroma_recon_leg(X, w, h, pus, gpu_res, n_ref, refs, cur, reps, cpu, luma_qp=32):
    """Chroma reconstruction of the frame's inter CUs (cu_recon_chroma): each
    PU's U and V prediction by mc_chroma (xme_mc_chroma_batch) into two
    prediction pictures, at reference 0 and 4 x the PU's full-pel best, then
    each CU's (size / 2)^2 block of both planes through xme_recon_chroma_batch
    at the chroma QP of luma QP 32; one MC launch and one launch per block
    size for both planes.  4:2:0 planes: U = the 2x2 mean of the luma pictures,
    V = 255 - U.  Algorithmic HBM bytes of the reconstruction: 5 per chroma
    pixel (current and prediction read, levels (int16) and reconstruction
    written).  Parity with the oracle on every 4th CU."""
    import ctypes as C
    from oracle import oracle as O
    cw, ch = w // 2, h // 2

    def planes420(img):
        u = img[: 2 * ch, : 2 * cw].reshape(ch, 2, cw, 2).astype(np.int32).sum(axis=(1, 3))
        u = ((u + 2) >> 2).astype(np.uint8)
        return u, (255 - u).astype(np.uint8)

    ref_uv, cur_uv = planes420(refs[0]), planes420(cur)
    d_ref = []
    for img in ref_uv:
        d = X.DevicePlane(cw, ch)
        d.upload_picture(img)
        d.expand_border()
        d_ref.append(d)
    d_cur = []
    for img in cur_uv:
        d = X.DevicePlane(cw, ch, pad=0, stride=((cw + 63) // 64) * 64)
        d.upload_picture(img)
        d_cur.append(d)
    pred = [X.DevicePlane(cw, ch, pad=0, stride=d_cur[0].stride) for _ in range(2)]
    rec = [X.DevicePlane(cw, ch, pad=0, stride=d_cur[0].stride) for _ in range(2)]
    r0 = gpu_res[0::n_ref]
    mvs = np.stack([4 * r0["bmv2_x"].astype(np.int32), 4 * r0["bmv2_y"].astype(np.int32)], axis=1)
    blocks = np.array([(x, y, pw, ph, int(mvs[i, 0]), int(mvs[i, 1])) for i, (x, y, pw, ph) in enumerate(pus)],
                      np.int32)
    off = np.array([(int(b[1]) // 2) * pred[0].stride + int(b[0]) // 2 for b in blocks], np.int64)
    d_pu, d_off = X.upload(blocks.ravel()), X.upload(off)
    cus = chroma_cus(pus)
    qp_c = X.lib().xme_chroma_qp(luma_qp)
    by_n = {}
    for k, (x, y, s) in enumerate(cus):
        by_n.setdefault(s // 2, []).append((x // 2, y // 2, k % 4 == 0))
    stream = X.Stream()
    bufs = []
    for n in sorted(by_n):
        b = np.array([(x, y) for x, y, _ in by_n[n]], np.int32)
        nb = len(b)
        bufs.append((n, nb, b, X.upload(b.ravel()),
                     (X.DeviceBuffer(2 * 2 * nb * n * n), X.DeviceBuffer(2 * 4 * nb), X.DeviceBuffer(2 * 4 * nb))))

    def mc():
        X.check(X.lib().xme_mc_chroma_batch(C.c_void_p(d_ref[0].origin), C.c_void_p(d_ref[1].origin),
                                            d_ref[0].stride, d_pu.at(), len(blocks), w, h,
                                            C.c_void_p(pred[0].origin), C.c_void_p(pred[1].origin), pred[0].stride,
                                            d_off.at(), stream.h), "xme_mc_chroma_batch")

    def recon():
        for n, nb, _b, d_b, (level, nnz, ssd) in bufs:
            X.check(X.lib().xme_recon_chroma_batch(n, qp_c, C.c_void_p(d_cur[0].origin), C.c_void_p(d_cur[1].origin),
                                                   d_cur[0].stride, C.c_void_p(pred[0].origin),
                                                   C.c_void_p(pred[1].origin), pred[0].stride, cw, ch, d_b.at(), nb,
                                                   level.at(), C.c_void_p(rec[0].origin), C.c_void_p(rec[1].origin),
                                                   rec[0].stride, nnz.at(), ssd.at(), stream.h),
                    "xme_recon_chroma_batch")

    def timed(fn):
        for _ in range(3):
            fn()
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        stream.sync()
        return (time.perf_counter() - t0) * 1e3 / reps

    mc_ms = timed(mc)
    recon_ms = timed(recon)
    pixels = 2 * sum(nb * n * n for n, nb, *_ in bufs)
    leg = {
        "workload": f"{cw}x{ch} U and V of the {w}x{h} frame's {len(cus)} CUs ("
                    + ", ".join(f"{nb} of {n}x{n}" for n, nb, *_ in bufs)
                    + f" per plane), {len(blocks)} PUs predicted from reference 0, chroma QP {qp_c} (luma {luma_qp})",
        "mc_ms_per_frame": mc_ms,
        "recon_ms_per_frame": recon_ms,
        "recon_pixels_per_s": pixels / (recon_ms * 1e-3),
        "recon_algorithmic_gb_per_s": 5.0 * pixels / (recon_ms * 1e-3) / 1e9,
        "recon_hbm_frac": 5.0 * pixels / (recon_ms * 1e-3) / 8e12,
    }
    if cpu:
        ok, el, n_blk = True, 0.0, 0
        o_ref = []
        for img in ref_uv:
            op = O.Plane(cw, ch, pad=40)
            op.set_picture(img)
            op.expand_border()
            o_ref.append(op)
        g_pred = [pr.buf.download(pr.rows * pr.stride, np.uint8).reshape(pr.rows, pr.stride)[:ch, :cw]
                  for pr in pred]
        g_rec = [rc.buf.download(rc.rows * rc.stride, np.uint8).reshape(rc.rows, rc.stride)[:ch, :cw]
                 for rc in rec]
        e_pred = [np.zeros((ch, cw), np.uint8) for _ in range(2)]
        sampled = {(x, y, s) for k, (x, y, s) in enumerate(cus) if k % 4 == 0}
        for (x, y, pw, ph, mx, my) in blocks:
            s = max(int(pw), int(ph))
            if (int(x) & ~(s - 1), int(y) & ~(s - 1), s) not in sampled:
                continue
            qx = O.lib().xo_get_mc_pos(w, int(pw), int(x), int(mx))
            qy = O.lib().xo_get_mc_pos(h, int(ph), int(y), int(my))
            bu, bv = O.mc_chroma(qx, qy, int(pw) // 2, int(ph) // 2, o_ref[0], o_ref[1])
            e_pred[0][y // 2:y // 2 + ph // 2, x // 2:x // 2 + pw // 2] = bu
            e_pred[1][y // 2:y // 2 + ph // 2, x // 2:x // 2 + pw // 2] = bv
        for n, nb, b, _d_b, (level, nnz, ssd) in bufs:
            sel = np.array([t[2] for t in by_n[n]])
            bs = b[sel]
            g_level = level.download(2 * nb * n * n, np.int16).reshape(2, nb, n, n)
            g_nnz, g_ssd = nnz.download(2 * nb, np.int32).reshape(2, nb), ssd.download(2 * nb, np.int32).reshape(2, nb)
            for p in range(2):
                t0 = time.perf_counter()
                el_level, el_rec, el_nnz, el_ssd = O.recon_pred_batch(n, qp_c, cur_uv[p], e_pred[p], bs)
                el += time.perf_counter() - t0
                n_blk += len(bs)
                ok = ok and np.array_equal(g_level[p][sel], el_level)
                ok = ok and np.array_equal(g_nnz[p][sel], el_nnz) and np.array_equal(g_ssd[p][sel], el_ssd)
                for (x, y) in bs:
                    ok = ok and np.array_equal(g_pred[p][y:y + n, x:x + n], e_pred[p][y:y + n, x:x + n])
                    ok = ok and np.array_equal(g_rec[p][y:y + n, x:x + n], el_rec[y:y + n, x:x + n])
        leg["parity_with_oracle_every_4th_cu"] = bool(ok)
        leg["cpu_oracle"] = {"recon_pixels_per_s": 2 * sum(int(np.sum([t[2] for t in by_n[n]])) * n * n
                                                           for n in by_n) / el,
                             "threads": 1, "seconds": el,
                             "sample": f"the U and V blocks of every 4th CU ({n_blk} blocks)"}
    log(f"chroma: MC {mc_ms:.3f} ms, recon {recon_ms:.3f} ms per frame")
    return leg


def transform_leg(X, blocks_total, reps, seed=91):
    """DCT -> flat quant -> dequant -> IDCT over blocks_total int16 blocks, a
    quarter each of 4x4, 8x8, 16x16 and 32x32 (Laplacian b=6 residuals clipped
    to +-255, 30 % all-zero blocks), at QP 22, 27, 32, 37, 42.  Algorithmic HBM
    bytes: 6 per coefficient (int16 residual in, level and reconstruction out)."""
    from oracle import oracle as O
    qps = (22, 27, 32, 37, 42)
    rng = np.random.default_rng(seed)
    per = blocks_total // 4
    bufs = []
    n_elem = 0
    sample = {}
    for n in (4, 8, 16, 32):
        host = np.empty((per, n, n), np.int16)
        step = max(1, (1 << 24) // (n * n))
        for b0 in range(0, per, step):
            b1 = min(per, b0 + step)
            v = np.rint(rng.laplace(0.0, 6.0, (b1 - b0, n, n))).clip(-255, 255)
            v[rng.random(b1 - b0) < 0.3] = 0
            host[b0:b1] = v
        sample[n] = host[:256].copy()
        src = X.upload(host)
        bufs.append((n, src, X.DeviceBuffer(host.nbytes), X.DeviceBuffer(host.nbytes), X.DeviceBuffer(4 * per)))
        n_elem += host.size
        del host
    stream = X.Stream()

    def one_pass(qp):
        for n, src, lv, rc, nz in bufs:
            X.check(X.lib().xme_residual_batch(n, qp, 0, src.at(), lv.at(), rc.at(), nz.at(), per, stream.h),
                    "xme_residual_batch")

    one_pass(qps[0])
    stream.sync()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for qp in qps:
            one_pass(qp)
        stream.sync()
        times.append((time.perf_counter() - t0) / len(qps))
    ms = min(times) * 1e3
    gbs = 6.0 * n_elem / (ms * 1e-3) / 1e9
    # the last pass (QP 42) against the oracle on the first 256 blocks of each size
    ok = True
    for n, src, lv, rc, nz in bufs:
        el, er, en = O.residual_batch(n, qps[-1], 0, sample[n])
        ok = ok and np.array_equal(lv.download(el.shape, np.int16), el) and \
            np.array_equal(rc.download(er.shape, np.int16), er) and np.array_equal(nz.download(256, np.int32), en)
    return {
        "workload": f"{blocks_total} int16 blocks (1/4 each 4x4, 8x8, 16x16, 32x32; Laplacian b=6, +-255, 30 % zero), "
                    "DCT -> flat quant -> dequant -> IDCT, inter rounding, QP 22/27/32/37/42",
        "coefficients": int(n_elem),
        "ms_per_qp_pass": ms,
        "coef_per_s": n_elem / (ms * 1e-3),
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "bytes_per_coefficient": 6},
        "parity_with_oracle_on_sample": bool(ok),
    }


SERVE_STREAMS = [(1280, 720)] * 6 + [(1920, 1080)] * 6 + [(3840, 2160)] * 4


def serving_leg(X, seconds, loads, slots, n_cycle, share=0, split=0, seed=77, small_work=0):
    """16 concurrent encodes (6 x 720p, 6 x 1080p, 4 x 2160p), each a moving
    smooth+noise sequence (pan <= 2 px/frame, local block MVs in [-32, 32]^2),
    frames arriving as Poisson processes at 30-60 fps per stream (scaled by the
    load factor); every frame is a 2-reference ME job over a mixed PU tiling,
    R = 64, through the server under a 64 GB budget."""
    import queue
    import threading

    from xavs2_amd import synth
    rng = np.random.default_rng(seed)
    fps = rng.uniform(30.0, 60.0, len(SERVE_STREAMS))
    log(f"serving: generating {len(SERVE_STREAMS)} sequences of {n_cycle} frames")
    seqs = [synth.moving_sequence(w, h, n_cycle, seed=seed + 10 * i) for i, (w, h) in enumerate(SERVE_STREAMS)]
    srv = X.Server(64 << 30, max_inflight=slots, search_range=64, n_ref=2, workgroup_share_pct=share,
                   launch_split_absdiff=split, full_share_below_absdiff=small_work)
    sids = []
    for i, (w, h) in enumerate(SERVE_STREAMS):
        pus = [(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in synth.mixed_pus(w, h, seed + 10 * i + 1)]
        sids.append(srv.open_stream(w, h, X.make_pus(pus)))
    frame_no = [0] * len(SERVE_STREAMS)

    def submit(i):
        j = srv.submit(sids[i], seqs[i][frame_no[i] % n_cycle])
        frame_no[i] += 1
        return j

    # warm-up: every stream reaches two references (and captures its plans' graphs)
    warm = [submit(i) for _ in range(3) for i in range(len(SERVE_STREAMS))]
    for j in warm:
        srv.wait(j)
    out = []
    for load in loads:
        sched = []
        for i in range(len(SERVE_STREAMS)):
            t = float(rng.exponential(1.0 / (fps[i] * load)))
            while t < seconds:
                sched.append((t, i))
                t += float(rng.exponential(1.0 / (fps[i] * load)))
        sched.sort()
        srv.reset_stats()
        q = queue.Queue()
        lat = [[] for _ in SERVE_STREAMS]

        def waiter():
            while True:
                item = q.get()
                if item is None:
                    return
                i, j = item
                lat[i].append(srv.wait(j)[1])

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        late = 0.0
        for t, i in sched:
            dt = t0 + t - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            else:
                late = max(late, -dt)
            q.put((i, submit(i)))
        q.put(None)
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        per_class = {}
        for (w, h) in sorted(set(SERVE_STREAMS)):
            v = np.concatenate([np.array(lat[i]) for i in range(len(SERVE_STREAMS)) if SERVE_STREAMS[i] == (w, h)])
            per_class[f"{w}x{h}"] = {"frames": int(len(v)), "p50_ms": float(np.percentile(v, 50)),
                                     "p99_ms": float(np.percentile(v, 99)), "max_ms": float(v.max())}
        offered = len(sched) / seconds
        rec = {
            "load": load,
            "offered_fps": offered,
            "achieved_fps": st["throughput_fps"],
            "frames": int(st["completed"]),
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"], "max_ms": st["max_ms"],
            "p99_over_p50": st["p99_ms"] / st["p50_ms"] if st["p50_ms"] else None,
            "per_resolution": per_class,
            "peak_device_bytes": int(st["peak_bytes"]),
            "submit_late_max_ms": late * 1e3,
            "wall_s": wall,
        }
        log(f"serving load {load}: offered {offered:.0f} fps, achieved {rec['achieved_fps']:.0f} fps, "
            f"p50 {rec['p50_ms']:.2f} ms p99 {rec['p99_ms']:.2f} ms")
        out.append(rec)
    resident = srv.stats()["resident_bytes"]
    srv.close()
    rated = None
    for rec in out:
        keeps_up = rec["achieved_fps"] >= 0.97 * rec["offered_fps"]
        tails = all(c["p99_ms"] <= 2.0 * c["p50_ms"] for c in rec["per_resolution"].values())
        if keeps_up and tails:
            rated = rec["load"]
    return {
        "workload": "16 concurrent encodes (6x1280x720, 6x1920x1080, 4x3840x2160), Poisson arrivals at 30-60 fps "
                    "per stream x load, 2-ref ME over mixed PUs, R=64, 64 GB budget, "
                    f"{slots} device slots on {os.environ.get('GPU_MAX_HW_QUEUES', '4')} hardware queues, "
                    f"search launches capped at {share or 100} % of the device"
                    + (f" for streams of at least {small_work:.0e} absdiffs per reference (smaller ones uncapped)"
                       if small_work > 0 else "") + " and "
                    f"split at {'8e9' if split == 0 else ('none' if split < 0 else f'{split:.0e}')} absdiffs, "
                    f"{seconds:.0f} s per load",
        "stream_fps": [round(float(f), 2) for f in fps],
        "resident_bytes": int(resident),
        "loads": out,
        "rated_load": rated,
        "rated_load_rule": "largest load with achieved >= 97 % of offered fps and p99 <= 2 x p50 within every "
                           "resolution class",
    }


def copy_probe(X, mib):
    """A plain device copy of known size: FETCH_SIZE / WRITE_SIZE of its
    dispatches against mib MiB give the scale of the HBM traffic counters."""
    n = mib << 20
    src, dst = X.DeviceBuffer(n), X.DeviceBuffer(n)
    src.zero()
    for _ in range(20):
        X.check(X.lib().xme_copy_kernel(dst.at(), src.at(), n, None), "xme_copy_kernel")
    X.lib().xme_device_sync()
    print(json.dumps({"copy_probe_bytes": n, "dispatches": 20}))


def traffic_probe(X, args):
    """Child of measure_traffic, run under rocprofv3 --pmc: the copy of
    TRAFFIC_COPY_MIB MiB 3 times, then the bench's search step (the plan's
    graph) 3 times after one warm-up."""
    n = TRAFFIC_COPY_MIB << 20
    src, dst = X.DeviceBuffer(n), X.DeviceBuffer(n)
    src.zero()
    for _ in range(3):
        X.check(X.lib().xme_copy_kernel(dst.at(), src.at(), n, None), "xme_copy_kernel")
    X.lib().xme_device_sync()
    del src, dst
    s = setup_search(X, args)
    for _ in range(4):
        s["plan"].run_graph(s["dcur"], s["d_refs"], s["stream"].h)
    s["stream"].sync()
    print(json.dumps({"traffic_probe": "done"}), flush=True)


def measure_traffic(args, timeout_s=240):
    """roofline.traffic: HBM bytes per search step (its two persistent
    dispatches), from two rocprofv3 --pmc children of this script in
    --traffic-probe mode, FETCH_SIZE then WRITE_SIZE (KiB per dispatch), each
    scaled by what the same counter reads for the copy of TRAFFIC_COPY_MIB MiB
    in the same run.  Returns (bytes or None, details)."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, {"error": "rocprofv3 not found"}
    workload = ["--width", str(args.width), "--height", str(args.height), "--refs", str(args.refs), "--range",
                str(args.range)] + (["--no-tune"] if args.no_tune else []) + (["--retune"] if args.retune else []) + \
               (["--pu-shape", args.pu_shape] if args.pu_shape else [])
    env = dict(os.environ, TMPDIR="/tmp")
    details, total = {}, 0.0
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="xme_traffic_", dir="/tmp")
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "t", "--", sys.executable,
               os.path.abspath(__file__), "--serve-hw-queues", "0", "--traffic-probe"] + workload
        log(f"traffic: {counter} pass")
        proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                                start_new_session=True)
        try:
            out, _ = proc.communicate(timeout=timeout_s)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.communicate()
            shutil.rmtree(d, ignore_errors=True)
            return None, {"error": f"{counter} pass timed out"}
        path = os.path.join(d, "t_counter_collection.csv")
        # the profiler may end with an error of its own after the program has
        # finished; its results are whole once the probe printed its last line
        if "traffic_probe" not in out or not os.path.exists(path):
            shutil.rmtree(d, ignore_errors=True)
            return None, {"error": f"{counter} pass failed (exit {proc.returncode})"}
        copy_kib, search_kib = [], {}
        with open(path) as f:
            for row in csv.DictReader(f):
                name, v = row["Kernel_Name"], float(row["Counter_Value"])
                if "copy_kernel" in name:
                    copy_kib.append(v)
                elif "fs_kernel<true" in name:
                    key = name[name.index("fs_kernel<"):]
                    search_kib.setdefault(key[:key.index(">") + 1], []).append(v)
        shutil.rmtree(d, ignore_errors=True)
        if not copy_kib or not search_kib:
            return None, {"error": f"{counter} pass: no copy or search dispatch in the counters"}
        scale = (TRAFFIC_COPY_MIB << 20) / (float(np.median(copy_kib)) * 1024.0)
        per_step = sum(float(np.median(v)) for v in search_kib.values()) * 1024.0 * scale
        details[counter] = {"kib_per_dispatch": {k: float(np.median(v)) for k, v in search_kib.items()},
                            "copy_kib": float(np.median(copy_kib)), "scale": scale, "bytes_per_step": per_step}
        total += per_step
    return total, details


def setup_search(X, args):
    """The bench's search job: inputs, resident reference planes, tile
    shapes and the batched plan."""
    w, h, n_ref, R = args.width, args.height, args.refs, args.range
    lam = 4 * 65536
    log(f"generating {w}x{h} inputs, {n_ref} refs")
    refs, cur, pus = build_inputs(w, h, n_ref, seed=2026)
    if args.pu_shape:
        pw, ph = (int(v) for v in args.pu_shape.lower().split("x"))
        pus = [(x, y, pw, ph) for y in range(0, h - ph + 1, ph) for x in range(0, w - pw + 1, pw)]
        log(f"one-shape tiling: {len(pus)} PUs of {pw}x{ph}")
    stream = X.Stream()
    d_refs = []
    for r in refs:
        d = X.DeviceRef(w, h, frac=2)
        d.load(r, stream.h)
        d_refs.append(d)
    dcur = X.DevicePlane(w, h, pad=0, stride=((w + 255) // 256) * 256)
    dcur.upload_picture(cur)
    stream.sync()
    prm = X.me_params(w, h, R, lam, use_hadamard=1, use_fractional_me=2, use_fast_sub_me=0, enable_pmvr=0)
    part_of = {(X.PU_W[i], X.PU_H[i]): i for i in range(25)}
    parts = sorted({part_of[(pw, ph)] for (_, _, pw, ph) in pus})
    tiles, tuned_now = {}, []
    if not args.no_tune:
        from xavs2_amd import tiles as T
        tiles, tuned_now = T.install(prm, parts, stream.h, retune=args.retune)
        log(f"tile shapes for {len(parts)} PU shapes at R={R}: {len(parts) - len(tuned_now)} from the shipped "
            f"table, {len(tuned_now)} tuned now")
    pu_arr = X.make_pus([(x, y, pw, ph, 0, 0) for (x, y, pw, ph) in pus])
    log(f"planning {len(pus)} PUs x {n_ref} refs")
    plan = X.Plan(prm, pu_arr, n_ref, X.PLAN_BATCHED)
    return dict(w=w, h=h, n_ref=n_ref, R=R, lam=lam, refs=refs, cur=cur, pus=pus, stream=stream, d_refs=d_refs,
                dcur=dcur, prm=prm, tiles=tiles, tuned_now=tuned_now, pu_arr=pu_arr, plan=plan)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-reps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default tile shapes, no table, no tuning")
    ap.add_argument("--retune", action="store_true", help="ignore the shipped tile table and tune every shape")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-seconds", type=float, default=2.5)
    ap.add_argument("--serve-loads", default="0.1,0.2,0.4,0.7,1.0")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (0: leave the environment's): the serving leg's "
                         "slots are HIP streams, and streams beyond the hardware queues share them")
    ap.add_argument("--serve-frames", type=int, default=4, help="distinct frames per stream, cycled")
    ap.add_argument("--serve-split", type=float, default=0,
                    help="search work per launch in the serving leg, absdiffs (0: server default, <0: no split)")
    ap.add_argument("--serve-share", type=int, default=50,
                    help="%% of a full-device search launch per frame in the serving leg (0 = 100); a cap "
                         "leaves CUs for frames that arrive while a large one searches")
    ap.add_argument("--serve-small-work", type=float, default=5e10,
                    help="streams whose search work per reference is below this many absdiffs search on the "
                         "whole device (the 720p and 1080p ones at R = 64); --serve-share caps the others")
    ap.add_argument("--no-transform", action="store_true", help="skip the residual-transform leg")
    ap.add_argument("--no-pattern", action="store_true", help="skip the DIA / HEX / UMH pattern-search leg")
    ap.add_argument("--no-bipred", action="store_true", help="skip the bi-prediction (sym / bid) leg")
    ap.add_argument("--no-recon", action="store_true", help="skip the inter-reconstruction leg")
    ap.add_argument("--transform-blocks", type=int, default=1 << 20)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the 16-byte copy kernel over MIB MiB 20 times (HBM counter calibration)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure roofline.traffic")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="child mode of the traffic measurement: a copy and the search step, nothing printed "
                         "but a marker line")
    ap.add_argument("--width", type=int, default=3840)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--refs", type=int, default=4)
    ap.add_argument("--range", type=int, default=64)
    ap.add_argument("--pu-shape", default="", metavar="WxH",
                    help="tile the picture with PUs of this one shape instead of the mixed tiling (profiling; "
                         "the metric is quoted on the mixed tiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the ME path runs on a single MI355X; --gpus {args.gpus} is not "
              "supported (several GPUs would each run whole streams with their resident references).",
              file=sys.stderr)
        sys.exit(2)

    # read by the HIP runtime when it starts (the first HIP call below)
    if args.serve_hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.serve_hw_queues)
    from xavs2_amd import lib as X

    if args.copy_probe:
        copy_probe(X, args.copy_probe)
        return
    if args.traffic_probe:
        traffic_probe(X, args)
        return

    # the serving leg runs first: its device slots are the process's first HIP
    # streams, so each gets a hardware queue of its own (up to GPU_MAX_HW_QUEUES)
    serving = None
    if not args.no_serve:
        if not args.no_tune:
            from xavs2_amd import tiles as T
            T.install(X.me_params(1920, 1080, 64, 4 * 65536), list(range(25)), retune=args.retune)
        loads = [float(v) for v in args.serve_loads.split(",") if v]
        serving = serving_leg(X, args.serve_seconds, loads, args.serve_slots, args.serve_frames, args.serve_share,
                              int(args.serve_split), small_work=int(args.serve_small_work))

    su = setup_search(X, args)
    w, h, n_ref, R, lam = su["w"], su["h"], su["n_ref"], su["R"], su["lam"]
    refs, cur, pus, stream, d_refs, dcur = su["refs"], su["cur"], su["pus"], su["stream"], su["d_refs"], su["dcur"]
    prm, tiles, tuned_now, pu_arr, plan = su["prm"], su["tiles"], su["tuned_now"], su["pu_arr"], su["plan"]
    n_inst, n_cand, n_absdiff = plan.info()

    # warmup (also captures the graph)
    for _ in range(max(1, args.warmup)):
        plan.run_graph(dcur, d_refs, stream.h)
    stream.sync()
    gpu_res = plan.fetch()

    # timed region: K graph replays bracketed by synchronizations
    log(f"timing {args.steps} steps")
    X.lib().xme_device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run_graph(dcur, d_refs, stream.h)
    stream.sync()
    X.lib().xme_device_sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = n_cand * args.steps / elapsed
    again = plan.fetch()
    stable = bool(np.array_equal(again, gpu_res))

    # per-frame latency: each replay synchronized on its own
    lat = []
    for _ in range(args.steps):
        t = time.perf_counter()
        plan.run_graph(dcur, d_refs, stream.h)
        stream.sync()
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat)

    # roofline of the integer-search kernel: HIP events on its stream
    search_ms, job_ms, _ = plan.time(dcur, d_refs, stream.h, 5)
    achieved = n_absdiff / (search_ms * 1e-3) / 1e12
    log(f"search kernel {search_ms:.3f} ms, {achieved:.1f} T absdiff/s")

    # plain path: one launch per PU shape, no persistent queue, no graph
    log("timing the plain path")
    plain = X.Plan(prm, pu_arr, n_ref, X.PLAN_PLAIN)
    _, plain_ms, _ = plain.time(dcur, d_refs, stream.h, max(1, args.plain_reps))
    plain_res = plain.fetch(stream.h)
    plain_equal = bool(np.array_equal(plain_res, gpu_res))
    plain_value = n_cand / (plain_ms * 1e-3)

    del plain
    pattern = None
    if not args.no_pattern:
        log("pattern-search leg (DIA, HEX, UMH)")
        pattern = pattern_leg(X, w, h, R, lam, pus, n_ref, dcur, d_refs, refs, cur, stream, args.steps,
                              not args.no_cpu)
    bipred = None
    if not args.no_bipred and n_ref >= 2:
        log("bi-prediction leg (sym, bid)")
        bipred = bipred_leg(X, w, h, R, lam, pus, gpu_res, n_ref, dcur, d_refs, refs, cur, max(20, args.steps),
                            not args.no_cpu)
    recon = None
    if not args.no_recon:
        log("inter-reconstruction leg")
        recon = recon_leg(X, w, h, pus, gpu_res, n_ref, dcur, d_refs, refs, cur, max(20, args.steps),
                          not args.no_cpu)
        log("chroma reconstruction leg")
        recon["chroma"] = chroma_recon_leg(X, w, h, pus, gpu_res, n_ref, refs, cur, max(20, args.steps),
                                           not args.no_cpu)
    transform = None
    if not args.no_transform:
        log(f"residual transform leg: {args.transform_blocks} blocks")
        transform = transform_leg(X, args.transform_blocks, reps=5)
    cpu = None
    if not args.no_cpu:
        host = host_cpus()
        threads = host["threads"]
        log(f"timing the CPU oracle on a sample ({threads} threads; host {host})")
        cpu = cpu_baseline((R, lam, 1, 2, 0, 0), refs, cur, pus, gpu_res, n_ref, args.cpu_seconds, threads, host)
    traffic, traffic_detail = None, {"skipped": "--no-traffic"}
    if not args.no_traffic:
        traffic, traffic_detail = measure_traffic(args)
        log(f"traffic: {traffic} B per search step")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cand-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"batched multi-ref ME {w}x{h} 8-bit, {n_ref} refs, {len(pus)} "
                        + (f"PUs of {args.pu_shape}" if args.pu_shape else
                           "mixed PUs (8x8..64x64 incl. 64x16/16x64/32x8/8x32)")
                        + f", full search [-{R},{R})^2 + half/quarter-pel SATD refine, one persistent launch "
                        "replayed from a HIP graph",
            "pu_refs": int(n_inst),
            "integer_candidates_per_step": int(n_cand),
            "absdiffs_per_step": int(n_absdiff),
            "lambda_factor": lam,
            "tiles_K_LC": {str(k): v for k, v in tiles.items()},
            "tiles_tuned_in_run": tuned_now,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_TABSDIFF,
            "unit": "Tabsdiff/s",
            "frac": achieved / VALU_PEAK_TABSDIFF,
            "traffic": traffic,
            "traffic_unit": "HBM bytes per search step, measured in this run: rocprofv3 --pmc FETCH_SIZE and "
                            "WRITE_SIZE on a child running the same step, each scaled by its reading for a "
                            f"{TRAFFIC_COPY_MIB} MiB copy in the same pass",
            "traffic_detail": traffic_detail,
            "kernel": "integer full search: fs_kernel<true,8> (generic PUs) + fs_kernel<true,16> (PUs of <= 256 px), both persistent",
            "kernel_ms": search_ms,
        },
        "plain": {"value": plain_value, "ms_per_step": plain_ms, "results_equal": plain_equal},
        "speedup_vs_plain": value / plain_value,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean())},
        "results_stable_across_replays": stable,
        "cpu_baseline": cpu,
        "pattern_search": pattern,
        "bipred": bipred,
        "inter_recon": recon,
        "transform": transform,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
