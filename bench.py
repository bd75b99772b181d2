#!/usr/bin/env python3
"""Imputation throughput of the MI355X hot path: batched vs plain, and serving
latency (one JSON line on stdout; progress on stderr).

Sections (all on synthetic data from the seeded generators of
swiftimpute_amd.synth; every input is resident in HBM before a timed region):

  batch  (the headline `value`) -- 512 uneven jobs against one shared
         20,000-haplotype x 200,000-marker panel: regions log-uniform in
         2,000-80,000 markers, cohorts log-uniform in 1-2,000 samples, L
         uniform in {4, 8, 16}, typed density uniform in {1/10, 1/20, 1/50},
         Ne = 10,000.  One step = every job imputed once (forward-backward,
         posterior dosage, Philox sampling).  Batched form: one persistent
         work-queue launch per pass and state class, the classes side by side
         on forked streams, replayed from a captured hipGraph.  Plain form: si_fb_run once per job, one after another on
         one stream.  Both write the same outputs; the bench checks that they
         are equal bit for bit.
  serve  -- a resident 40,000 x 500,000 panel; Poisson request arrivals;
         each request a cohort (log-uniform 1-5,000 samples) over a region
         (log-uniform 1,000-100,000 markers), L = 8, Ne = 10,000, typed every
         20th marker.  Reports arrival-to-result p50/p99 latency and served
         genotypes/s, beside the same requests submitted all at once
         (batched-offline).
  plain  -- 20,000 x 100,000 panel, 10,000 targets typed every 20th marker,
         L = 16, Ne = 20,000, one job per launch: the per-kernel roofline
         (chain and block passes timed with HIP events on their stream), the
         log-space step's transcendental rate, and the PBWT index build.
  cfg1   -- 4,432 x 1,644 panel, 444 targets typed every 4th marker, L = 8,
         FP32 log-space: end to end through the C++ API (VCF in, VCF out,
         timed per stage) beside the reference's published 24.8 s, the
         kernel alone, and the CPU baseline on the whole cohort.
  precision -- 8,000 x 100,000 panel with heavy-tailed spacing: linear and
         log-space FP32 against the FP64 oracle for Ne in {1e3, 1e4, 1e5}.

Imputed genotype = one (target sample, untyped marker) pair.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (= FP32 matrix) peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
# v_exp_f32 / v_log_f32 issue at a quarter of the FP32 lane rate:
# 256 CUs x 64 lanes x 2.4 GHz / 4 per second
PEAK_TRANSCENDENTAL_TPS = 256 * 64 * 2.4e9 / 4 / 1e12


def host_threads():
    """Worker threads for the CPU baselines: the CPUs this process may run on
    (its affinity mask), at most 16 -- a one-GPU share of the GPU box."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info():
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        aff = None
    return {"os_cpu_count": os.cpu_count(), "affinity_cpus": aff}


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def load_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc
    summary (profiles/traffic.json, plain 20,000 x 100,000 workload), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"].get(kernel_key, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError, KeyError):
        return None


def sync():
    import torch
    torch.cuda.synchronize()


# --------------------------------------------------------------------------
# batch: config-4 job set, batched vs plain
# --------------------------------------------------------------------------
def bench_batch(args):
    import numpy as np
    import torch

    from swiftimpute_amd import _native, device, synth

    H, M, pool_B, ne = 20000, 200000, 2000, 10000
    n_jobs = args.jobs
    log(f"[batch] generating panel H={H} M={M}")
    pan = synth.make_panel(H, M, seed=12345)
    log(f"[batch] generating a pool of {pool_B} targets typed at every marker")
    pool = synth.make_targets(pan, pool_B, 1, seed=12346, with_truth=False)
    specs = synth.uneven_jobs(n_jobs, M, pool_B)
    log("[batch] building the PBWT index and state tables for L in {4, 8, 16}")
    idx16 = device.DeviceIndex(pan.cols, H, 16, keep_prefix=True, with_div=False)
    idx = {16: idx16, 8: idx16.derive(8), 4: idx16.derive(4)}
    idx16.drop_prefix()
    torch.cuda.empty_cache()
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    cm = pan.genetic_cM()

    log(f"[batch] uploading {n_jobs} jobs")
    jobs = []
    units = untyped = 0
    for sp in specs:
        pitch = (sp.M + 15) // 16 * 16
        c = synth.region_codes(pool.codes, sp, pitch)
        untyped += int((c[:, : sp.M] == synth.CODE_MISSING).sum())
        units += sp.B * sp.M
        tr = device.Transitions.build(pan.pos[sp.m0:sp.m0 + sp.M],
                                      None if cm is None else cm[sp.m0:sp.m0 + sp.M], ne)
        jobs.append(device.RegionJob(idx[sp.L], sp.m0, sp.M, torch.from_numpy(c).cuda(), tr,
                                     sp.seed, tables))
    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        log(f"[batch] CPU baseline: reference CPU path on a sample of the jobs, {threads} threads")
        v, dt, n, parts = cpu_baseline_jobs(pan, pool.codes, specs, ne, args.cpu_jobs,
                                            args.cpu_samples, args.cpu_markers, threads)
        cpu = {"value": v, "unit": "genotypes/s", "cores": threads, "kind": "port",
               "sample": f"the first {args.cpu_jobs} jobs of the same job set, each cut to at most "
                         f"{args.cpu_samples} samples x {args.cpu_markers} markers "
                         f"({', '.join(parts)}; {n} imputed genotypes in {dt:.1f} s): PBWT sweep "
                         "+ per-(sample, marker) std::sort selection + L^2 FP32 forward-backward "
                         "with sqrt(M) replay + sampling, threaded over samples",
               "host": host_info()}
        log(f"[batch] CPU baseline {v:.4g} genotypes/s")
        # the reference runs single-threaded: the same path on one thread, on
        # the first job cut to 4 samples
        v1, dt1, n1, parts1 = cpu_baseline_jobs(pan, pool.codes, specs, ne, 1, 4,
                                                args.cpu_markers, 1)
        cpu["single_thread"] = {"value": v1, "cores": 1,
                                "sample": f"{parts1[0]} ({n1} imputed genotypes in {dt1:.1f} s)"}
        log(f"[batch] CPU baseline, one thread: {v1:.4g} genotypes/s")
        # the stronger CPU form: selection once per marker (it is
        # sample-independent) and the O(M L) forward-backward
        vh, dth, nh, _ = cpu_baseline_jobs(pan, pool.codes, specs, ne, args.cpu_jobs,
                                           args.cpu_samples, args.cpu_markers, threads,
                                           hoisted=True)
        cpu["hoisted"] = {"value": vh, "cores": threads,
                          "sample": f"the same sample ({nh} imputed genotypes in {dth:.2f} s): "
                                    "PBWT sweep + one state table per marker + rank-1 FP64 "
                                    "forward-backward + sampling"}
        log(f"[batch] CPU baseline, hoisted: {vh:.4g} genotypes/s")
    del pool
    cjobs = [j.c_job() for j in jobs]
    n_log = sum(j.logspace for j in jobs)
    sync()
    stream = torch.cuda.Stream()   # graph capture needs a created (non-null) stream
    sh = ctypes.c_void_p(stream.cuda_stream)
    lib = _native.hip()

    # plain: one job per launch
    arr1 = [(_native.FbJob * 1)(cj) for cj in cjobs]
    ws_plain = max(int(lib.si_fb_workspace_bytes(a, 1)) for a in arr1)
    ws = torch.empty(ws_plain, dtype=torch.uint8, device="cuda")

    def plain_step():
        for a in arr1:
            _native.check(lib.si_fb_run(a, 1, ctypes.c_void_p(ws.data_ptr()), ws_plain, sh),
                          "si_fb_run")

    plain_steps = max(1, min(args.steps, args.plain_steps))
    log(f"[batch] plain: {min(args.warmup, 1)} warmup + {plain_steps} timed steps")
    for _ in range(min(args.warmup, 1)):
        plain_step()
    sync()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        plain_step()
    sync()
    plain_ms = (time.perf_counter() - t0) * 1e3 / plain_steps
    ref = [{k: v.clone() for k, v in j.out.items()} for j in jobs]
    del ws
    torch.cuda.empty_cache()
    log(f"[batch] plain {plain_ms:.1f} ms/step")

    # batched: persistent work-queue launches replayed from a hipGraph, lane
    # shapes and checkpoint block picked per class by the autotuner
    batch = device.Batch(cjobs)
    sync()
    t0 = time.perf_counter()
    batch.tune(stream=stream)
    sync()
    tune_s = time.perf_counter() - t0
    classes = batch.classes()
    for lp, lg, var, _ in classes:
        log(f"[batch] class Lp={lp}{' log' if lg else ''}: chain {var[0]} lanes/sample, "
            f"block {var[1]} lanes/sample, checkpoint block {var[2]}")
    log(f"[batch] autotuning took {tune_s:.2f} s")

    def time_graph(steps):
        for _ in range(args.warmup):
            batch.run(use_graph=True, stream=stream)
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.run(use_graph=True, stream=stream)
        sync()
        return (time.perf_counter() - t0) * 1e3 / steps

    batch.run(use_graph=False, stream=stream)   # eager once (also a correctness reference)
    sync()
    log(f"[batch] batched: {args.warmup} warmup + {args.steps} timed graph replays "
        f"(workspace {batch.ws_bytes / 2**30:.2f} GiB)")
    batch_ms = time_graph(args.steps)
    log(f"[batch] batched {batch_ms:.1f} ms/step")
    # the graph launch timed with HIP events on its own stream
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    evs[0].record(stream)
    for _ in range(args.steps):
        batch.run(use_graph=True, stream=stream)
    evs[1].record(stream)
    sync()
    graph_ms = evs[0].elapsed_time(evs[1]) / args.steps
    flops = sum(12.0 * sp.L * sp.B * sp.M for sp in specs)
    per_class = {}
    for sp in specs:
        per_class.setdefault(device.padded_states(sp.L), 0.0)
        per_class[device.padded_states(sp.L)] += 12.0 * sp.L * sp.B * sp.M
    step_roofline = {
        "bound": "valu", "kernel": "batched step: one graph launch holding every state class's "
                                   "chain and block kernels",
        "achieved": flops / (graph_ms * 1e-3) / 1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": flops / (graph_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
        "graph_ms": graph_ms, "flops_per_step": flops,
        "flops_per_class": {str(k): v for k, v in sorted(per_class.items())},
        "algorithmic": "sum over jobs of 12*L*B*M FLOP (forward 4L, backward 4L, posterior and "
                       "dosage 4L per (sample, marker))",
    }
    log(f"[batch] graph launch {graph_ms:.1f} ms: {step_roofline['achieved']:.2f} TFLOP/s "
        f"algorithmic ({100 * step_roofline['frac']:.1f} % of FP32 peak)")
    equal = all(torch.equal(j.out[k], r[k]) for j, r in zip(jobs, ref) for k in r)
    if not equal:
        raise SystemExit("batched outputs differ from the plain path")
    # the same batch with every class at the default variant
    for k, (lp, _, _, _) in enumerate(classes):
        batch.set_variant(k, device.default_variant(lp))
    default_ms = time_graph(max(1, min(args.steps, 2)))
    log(f"[batch] batched with default variants {default_ms:.1f} ms/step")
    equal = all(torch.equal(j.out[k], r[k]) for j, r in zip(jobs, ref) for k in r)
    if not equal:
        raise SystemExit("batched outputs at the default variants differ from the plain path")
    batch.close()
    autotune = {
        "tune_s": tune_s,
        "classes": [{"padded_states": lp, "logspace": lg, "chain_lanes_per_sample": v[0],
                     "block_lanes_per_sample": v[1], "checkpoint_block": v[2]}
                    for lp, lg, v, _ in classes],
        "tuned_ms_per_step": batch_ms,
        "default_variant_ms_per_step": default_ms,
        "candidates": "lanes per sample Lp/8, Lp/4, Lp/2 (chain), Lp/4, Lp/2 (block); "
                      "checkpoint block 16 or 32; each candidate warmed on 16-marker cuts, "
                      "then timed twice on each class's whole jobs; beside other classes the "
                      "chain keeps Lp/4 lanes and another block shape needs a 10% gain",
    }
    return {
        "cpu_baseline": cpu,
        "units": units, "untyped": untyped, "batch_ms": batch_ms, "plain_ms": plain_ms,
        "plain_steps": plain_steps, "n_jobs": n_jobs, "n_logspace_jobs": n_log,
        "autotune": autotune, "step_roofline": step_roofline,
        "L_counts": {str(L): sum(sp.L == L for sp in specs) for L in (4, 8, 16)},
    }


# --------------------------------------------------------------------------
# serve: config-5 request stream against a resident panel
# --------------------------------------------------------------------------
def _pcie_rates(nbytes=1 << 30, reps=3):
    """Page-locked host <-> device copy rates (GB/s): each direction alone,
    and both at once on two streams (the link is full duplex)."""
    import torch

    dev = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dev2 = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    host2 = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    host.fill_(1)
    host2.fill_(2)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()

    def timed(fn):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    def d2h():
        with torch.cuda.stream(s1):
            host.copy_(dev, non_blocking=True)

    def h2d():
        with torch.cuda.stream(s2):
            dev2.copy_(host2, non_blocking=True)

    t_d2h, t_h2d = timed(d2h), timed(h2d)
    t_both = timed(lambda: (d2h(), h2d()))
    out = {"d2h_GBps": nbytes / t_d2h / 1e9, "h2d_GBps": nbytes / t_h2d / 1e9,
           "both_GBps_each": nbytes / t_both / 1e9, "copy_bytes": nbytes,
           "method": "1 GiB page-locked <-> device copies, best of 3, each direction alone "
                     "and both at once on two streams"}
    del dev, dev2, host, host2
    return out


def bench_serve(args):
    import numpy as np
    import torch

    from swiftimpute_amd import serving, synth

    H, M, pool_B, L, ne, te = args.serve_H, args.serve_M, 5000, 8, 10000, 20
    log(f"[serve] generating panel H={H} M={M}")
    pan = synth.make_panel(H, M, seed=12345)
    log(f"[serve] generating a pool of {pool_B} targets typed at every marker")
    pool = synth.make_targets(pan, pool_B, 1, seed=12346, with_truth=False)

    def request_set(rate, seconds, seed):
        reqs = synth.poisson_requests(rate, seconds, M, pool_B, L=L, typed_every=te, seed=seed)
        codes = [synth.region_codes(pool.codes, sp) for _, sp in reqs]
        return reqs, codes

    def pinned_outputs(reqs):
        # the client's result arrays, page-locked (the server copies into them
        # from the device directly), allocated and touched before the clock
        outs, keep = [], []
        for _, sp in reqs:
            ht = torch.empty((sp.B, 2, sp.M), dtype=torch.uint8, pin_memory=True)
            dt = torch.empty((sp.B, sp.M), dtype=torch.float32, pin_memory=True)
            h, d = ht.numpy(), dt.numpy()
            h.fill(0)
            d.fill(0)
            keep.append((ht, dt))
            outs.append((h, d))
        return outs, keep

    def online(srv, reqs, codes, outs):
        lat = [None] * len(reqs)
        ids = [None] * len(reqs)
        done = threading.Event()

        def waiter():
            for i in range(len(reqs)):
                while ids[i] is None:
                    time.sleep(0.0005)
                lat[i] = srv.wait(ids[i]).latency_s
            done.set()

        th = threading.Thread(target=waiter, daemon=True)
        th.start()
        t_start = time.perf_counter()
        for i, ((t_arr, sp), c) in enumerate(zip(reqs, codes)):
            now = time.perf_counter() - t_start
            if t_arr > now:
                time.sleep(t_arr - now)
            ids[i] = srv.submit_into(sp.m0, sp.m0 + sp.M, c, outs[i][0], outs[i][1], seed=sp.seed)
        done.wait()
        t_total = time.perf_counter() - t_start
        lat_ms = np.array(lat) * 1e3
        untyped = sum(int((c == synth.CODE_MISSING).sum()) for c in codes)
        busy_s = t_total - reqs[0][0] if reqs else t_total
        return {"requests": len(reqs), "p50_ms": float(np.percentile(lat_ms, 50)),
                "p99_ms": float(np.percentile(lat_ms, 99)), "mean_ms": float(lat_ms.mean()),
                "genotypes_per_s": untyped / busy_s}

    def offline(srv, reqs, codes, outs):
        before = srv.stats()
        t1 = time.perf_counter()
        ids = [srv.submit_into(sp.m0, sp.m0 + sp.M, c, o[0], o[1], seed=sp.seed)
               for (_, sp), c, o in zip(reqs, codes, outs)]
        t_sub = time.perf_counter() - t1
        for i in ids:
            srv.wait(i)
        wall = time.perf_counter() - t1
        after = srv.stats()
        d = {k: after[k] - before[k] for k in after
             if k not in ("max_batch_requests", "peak_bytes")}
        d["submit_s"] = t_sub
        return wall, d

    def make_server(plain=False):
        return serving.Server(pan.cols, H, pan.pos, pan.genetic_cM(), L=L, ne=ne, max_streams=4,
                              max_batch=64, hbm_budget_gb=args.serve_budget_gb,
                              max_queue_delay_us=args.queue_delay_us, plain=plain)

    pcie = _pcie_rates()
    log(f"[serve] PCIe: D2H {pcie['d2h_GBps']:.1f} GB/s, H2D {pcie['h2d_GBps']:.1f} GB/s, "
        f"both at once {pcie['both_GBps_each']:.1f} GB/s each")

    main_reqs, main_codes = request_set(args.rate, args.serve_seconds, 12345)
    untyped_main = sum(int((c == synth.CODE_MISSING).sum()) for c in main_codes)
    log(f"[serve] {len(main_reqs)} requests at {args.rate} req/s over {args.serve_seconds} s")
    log("[serve] starting the server (PBWT index build on the device)")
    t0 = time.perf_counter()
    srv = make_server()
    startup_s = time.perf_counter() - t0
    log(f"[serve] server ready in {startup_s:.1f} s")
    srv.wait(srv.submit(0, 2000, np.full((4, 2000), 3, np.uint8)))  # kernels warmed once

    # arrival-rate sweep (online latency and throughput per rate)
    sweep = {}
    for k, rate in enumerate(args.serve_rates):
        reqs, codes = request_set(rate, args.sweep_seconds, 777 + k)
        if not reqs:
            continue
        outs, keep = pinned_outputs(reqs)
        r = online(srv, reqs, codes, outs)
        sweep[str(rate)] = r
        log(f"[serve] {rate} req/s: {r['requests']} requests, p50 {r['p50_ms']:.1f} ms, "
            f"p99 {r['p99_ms']:.1f} ms, {r['genotypes_per_s']:.3g} genotypes/s")
        del outs, keep

    outs, keep = pinned_outputs(main_reqs)
    main_online = online(srv, main_reqs, main_codes, outs)
    log(f"[serve] online at {args.rate} req/s: p50 {main_online['p50_ms']:.1f} ms, "
        f"p99 {main_online['p99_ms']:.1f} ms")
    t_off, st_off = offline(srv, main_reqs, main_codes, outs)
    peak = srv.stats()["peak_bytes"]
    srv.close()
    log(f"[serve] offline: {t_off:.2f} s for {len(main_reqs)} requests")
    # the plain form on the same requests: each request its own launches
    srv = make_server(plain=True)
    srv.wait(srv.submit(0, 2000, np.full((4, 2000), 3, np.uint8)))
    t_plain, st_plain = offline(srv, main_reqs, main_codes, outs)
    srv.close()
    log(f"[serve] offline, plain form: {t_plain:.2f} s")
    del outs, keep

    out_bytes = st_off["d2h_bytes"]
    in_bytes = st_off["h2d_bytes"]
    # the link's copy rate with both directions busy is its rate for their sum
    pcie_s = max(out_bytes / (pcie["d2h_GBps"] * 1e9), in_bytes / (pcie["h2d_GBps"] * 1e9),
                 (in_bytes + out_bytes) / (2.0 * pcie["both_GBps_each"] * 1e9))

    def split(st, wall):
        return {"wall_s": wall, "submit_s": st["submit_s"], "kernel_busy_s": st["kernel_s"],
                "h2d_busy_s": st["h2d_s"],
                "d2h_busy_s": st["d2h_s"], "launch_host_s": st["launch_host_s"],
                "copy_host_s": st["copy_host_s"], "batches": st["batches"],
                "h2d_bytes": st["h2d_bytes"], "d2h_bytes": st["d2h_bytes"]}

    return {
        "workload": f"{H} x {M} resident panel; Poisson arrivals, cohorts log-uniform 1-5000, "
                    f"regions log-uniform 1000-100000 markers; L={L}, Ne={ne}, typed every "
                    f"{te}th marker; 4 streams, up to 64 requests per batch",
        "rate_sweep": {"seconds_per_rate": args.sweep_seconds, "per_rate": sweep},
        "requests": len(main_reqs),
        "rate_req_per_s": args.rate,
        "seconds": args.serve_seconds,
        "p50_ms": main_online["p50_ms"],
        "p99_ms": main_online["p99_ms"],
        "mean_ms": main_online["mean_ms"],
        "genotypes_per_s": main_online["genotypes_per_s"],
        "untyped_genotypes": untyped_main,
        "offline_genotypes_per_s": untyped_main / t_off,
        "offline_plain_genotypes_per_s": untyped_main / t_plain,
        "offline_batched_vs_plain": t_plain / t_off,
        "offline_split": split(st_off, t_off),
        "offline_plain_split": split(st_plain, t_plain),
        "pcie": pcie,
        "pcie_bound": {
            "s": pcie_s,
            "genotypes_per_s": untyped_main / pcie_s,
            "frac": pcie_s / t_off,
            "in_bytes": in_bytes,
            "out_bytes": out_bytes,
            "note": "every request's results (2 haplotype bytes and a 4-byte dosage per "
                    "(sample, marker)) cross to the host and its codes (1 byte per (sample, "
                    "marker), rows padded to 16) cross to the device; bound = the slowest of "
                    "results alone at the D2H rate, codes alone at the H2D rate, and both at "
                    "the rate the link gives both directions at once (their sum)",
        },
        "peak_admitted_bytes": peak,
        "server_startup_s": startup_s,
        "note": "online genotypes/s is bounded by the arrival rate; offline = every request "
                "queued at once on the same server; busy times are sums over batches on 4 "
                "streams, so they can exceed the wall time",
    }


# --------------------------------------------------------------------------
# plain: config-2 cohort, per-kernel roofline, PBWT build, CPU baseline
# --------------------------------------------------------------------------
def cpu_baseline(pan, tg, L, ne, n_samples, n_markers, threads):
    """The reference CPU path (oracle.cpu_reference_impute: PBWT sweep,
    per-(sample, marker) std::sort selection, L^2 FP32 forward-backward with
    sqrt(M) checkpoint replay, inverse-CDF sampling) on a bounded sample."""
    import numpy as np

    import oracle
    from swiftimpute_amd import synth

    W = pan.W
    b = pan.cols[:n_markers].view(np.uint8).reshape(n_markers, W * 8)
    ref = np.unpackbits(b, axis=1, bitorder="little")[:, : pan.H].copy()
    gl = synth.HARDCALL_PALETTE[tg.codes[:n_samples, :n_markers]]
    g = oracle.genetic_map(pan.pos[:n_markers])
    Td, To = oracle.transitions(g, ne)
    t0 = time.perf_counter()
    oracle.cpu_reference_impute(ref, gl, Td, To, L, 12345, nthreads=threads)
    dt = time.perf_counter() - t0
    untyped = int((tg.codes[:n_samples, :n_markers] == synth.CODE_MISSING).sum())
    return untyped / dt, dt, untyped


def cpu_baseline_jobs(pan, pool_codes, specs, ne, n_jobs, max_samples, max_markers, threads,
                      hoisted=False):
    """The reference CPU path on a bounded sample of a job set: the first
    n_jobs jobs, each cut to its first max_samples samples and first
    max_markers markers of its region.  hoisted: the stronger CPU form
    (PBWT sweep, one state table per marker, rank-1 FP64 forward-backward,
    sampling) on the same sample."""
    import numpy as np

    import oracle
    from swiftimpute_amd import synth

    cm = pan.genetic_cM()
    total_dt, untyped, parts = 0.0, 0, []
    for sp in specs[:n_jobs]:
        Mc, Bc = min(sp.M, max_markers), min(sp.B, max_samples)
        b = pan.cols[sp.m0:sp.m0 + Mc].view(np.uint8).reshape(Mc, pan.W * 8)
        ref = np.unpackbits(b, axis=1, bitorder="little")[:, : pan.H].copy()
        cut = synth.JobSpec(sp.m0, Mc, sp.s0, Bc, sp.L, sp.typed_every, sp.seed)
        c = synth.region_codes(pool_codes, cut)
        gl = synth.HARDCALL_PALETTE[c]
        g = oracle.genetic_map(pan.pos[sp.m0:sp.m0 + Mc], None if cm is None else cm[sp.m0:sp.m0 + Mc])
        Td, To = oracle.transitions(g, ne)
        t0 = time.perf_counter()
        if hoisted:
            p, _ = oracle.pbwt_build(ref)
            _, _, a0, a1 = oracle.state_table(ref, p, sp.L)
            o = oracle.fb_f64_fast(a0, a1, gl, Td, To, want_prob=True, nthreads=threads)
            oracle.sample_replay(o["prob"].astype(np.float32), a0, a1, sp.seed)
        else:
            oracle.cpu_reference_impute(ref, gl, Td, To, sp.L, sp.seed, nthreads=threads)
        total_dt += time.perf_counter() - t0
        untyped += int((c == synth.CODE_MISSING).sum())
        parts.append(f"{Bc}x{Mc} L={sp.L}")
    return untyped / total_dt, total_dt, untyped, parts


def bench_plain(args):
    import torch

    from swiftimpute_amd import _native, device, synth

    w = synth.WORKLOADS["cfg2"]
    H, M, B, L, ne, te = w["H"], w["M"], w["B"], w["L"], w["ne"], w["typed_every"]
    stream = torch.cuda.current_stream()
    log(f"[plain] generating panel H={H} M={M} and {B} targets typed every {te}")
    pan = synth.make_panel(H, M, seed=12345)
    tg = synth.make_targets(pan, B, te, seed=12346, with_truth=False)
    untyped = int((tg.codes[:, :M] == synth.CODE_MISSING).sum())

    log("[plain] building the PBWT index on the GPU")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    sync()
    ev[0].record(stream)
    idx = device.DeviceIndex(pan.cols, H, L, stream=stream)
    ev[1].record(stream)
    sync()
    index_ms = ev[0].elapsed_time(ev[1])
    pbwt_bytes = M * (pan.W * 8 + 4 * H + 4 * H)
    log(f"[plain] index built in {index_ms:.1f} ms (panel upload, PBWT, state table)")
    # the PBWT sweep alone (three launches), on resident inputs and outputs
    lib = _native.hip()
    wsb = int(lib.si_pbwt_workspace_bytes(H, M))
    pws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    sh = ctypes.c_void_p(stream.cuda_stream)
    reps = 3
    for k in range(reps + 1):
        if k == 1:
            ev[0].record(stream)
        _native.check(lib.si_pbwt_sweep(device._ptr(idx.cols), H, M, device._ptr(idx.prefix),
                                        device._ptr(idx.div), device._ptr(pws), wsb, sh),
                      "si_pbwt_sweep")
    ev[1].record(stream)
    sync()
    sweep_ms = ev[0].elapsed_time(ev[1]) / reps
    log(f"[plain] PBWT sweep {sweep_ms:.2f} ms ({pbwt_bytes / (sweep_ms * 1e-3) / 1e9:.0f} GB/s)")
    del pws
    idx.drop_prefix()
    torch.cuda.empty_cache()

    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), ne)
    codes = torch.from_numpy(tg.codes).cuda()

    def time_form(logspace):
        """Tune, warm up and time the plain two-pass step of one numeric form;
        returns (ms/step, chain ms, block ms, variant)."""
        name = "log-space" if logspace else "linear"
        job = device.Job(idx, codes, synth.HARDCALL_PALETTE, tr, seed=12345, logspace=logspace)
        job.alloc_outputs(ds=True, ap=False, hap=True, prob=False, loglik=False)
        arr = (_native.FbJob * 1)(job.c_job())
        ws_bytes = int(lib.si_fb_workspace_bytes(arr, 1))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        variant = device.tune_jobs([arr[0]], stream=stream, workspace=ws)
        var = ctypes.byref(_native.FbVariant(*variant))
        log(f"[plain] {name}: tuned variant: chain {variant[0]} lanes/sample, block {variant[1]} "
            f"lanes/sample, checkpoint block {variant[2]}")

        def step(evs=None):
            if evs:
                evs[0].record(stream)
            _native.check(lib.si_fb_run_variant(arr, 1, var, ctypes.c_void_p(ws.data_ptr()),
                                                ws_bytes, _native.SI_FB_PASS_CHAIN, sh),
                          "chain pass")
            if evs:
                evs[1].record(stream)
            _native.check(lib.si_fb_run_variant(arr, 1, var, ctypes.c_void_p(ws.data_ptr()),
                                                ws_bytes, _native.SI_FB_PASS_BLOCK, sh),
                          "block pass")
            if evs:
                evs[2].record(stream)

        for _ in range(args.warmup):
            step()
        sync()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
        log(f"[plain] {name}: timing {args.steps} steps")
        t0 = time.perf_counter()
        for k in range(args.steps):
            step(evs[k])
        sync()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        chain_ms = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
        block_ms = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps
        log(f"[plain] {name}: {ms:.3f} ms/step (chain {chain_ms:.3f} ms, block {block_ms:.3f} ms)")
        ds = job.out.ds
        assert torch.isfinite(ds).all().item() and float(ds.min()) >= -1e-6 and float(ds.max()) <= 2 + 1e-5
        assert int(job.out.hap.max()) <= 1
        del ws, job
        torch.cuda.empty_cache()
        return ms, chain_ms, block_ms, variant

    # the default form (linear when every switch probability is >= 1e-30), then
    # the log-space form the workload also names
    ms, chain_ms, block_ms, variant = time_form(False)
    lms, lchain_ms, lblock_ms, lvariant = time_form(True)

    units = B * M
    # The HMM's algorithmic count is 12*L FLOP per (sample, marker): forward
    # 4L and backward 4L (the chain pass) and posterior + dosage 4L (the block
    # pass, whose forward replay and backward re-walk are not counted).
    chain_flops = 8.0 * L * units
    block_flops = 4.0 * L * units
    # the block pass is one fb_block_kernel launch; the chain pass of these fast
    # jobs is fb_tmask_kernel + fb_tlist_kernel + fb_chain_fast_kernel
    dom, dom_ms, dom_flops = (("fb_block_kernel", block_ms, block_flops) if block_ms >= chain_ms
                              else ("fb_chain_fast_kernel", chain_ms, chain_flops))
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    # log-space: one exp per state per pass, 3L transcendentals per unit
    trans_tps = 3.0 * L * units / (lms * 1e-3) / 1e12
    out = {
        "workload": f"{H} haplotypes x {M} markers panel, {B} targets typed every {te}th marker, "
                    f"L={L}, Ne={ne}, plain one-job-per-launch",
        "genotypes_per_s": untyped / (ms * 1e-3),
        "all_marker_genotypes_per_s": units / (ms * 1e-3),
        "ms_per_step": ms,
        "passes_ms": {"chain_pass": chain_ms, "block_pass": block_ms},
        "pass_kernels": {"chain_pass": "fb_tmask_kernel + fb_tlist_kernel + fb_chain_fast_kernel",
                         "block_pass": "fb_block_kernel"},
        "variant": {"chain_lanes_per_sample": variant[0], "block_lanes_per_sample": variant[1],
                    "checkpoint_block": variant[2]},
        "form": "linear FP32 (per-marker sum-normalised forward, FP64 backward scale)",
        "log_space": {
            "form": "log-space FP32 (max-normalised, per-state log-sum-exp)",
            "genotypes_per_s": untyped / (lms * 1e-3),
            "ms_per_step": lms,
            "passes_ms": {"chain_pass": lchain_ms, "block_pass": lblock_ms},
            "variant": {"chain_lanes_per_sample": lvariant[0],
                        "block_lanes_per_sample": lvariant[1], "checkpoint_block": lvariant[2]},
            "transcendental": {
                "achieved": trans_tps, "peak": PEAK_TRANSCENDENTAL_TPS, "unit": "T/s",
                "frac": trans_tps / PEAK_TRANSCENDENTAL_TPS,
                "algorithmic": "3*L exp per (sample, marker) over the whole step; peak = 256 CUs "
                               "x 64 lanes x 2.4 GHz / 4 (quarter-rate v_exp_f32)"},
        },
        "pass_rooflines": {
            "chain_pass": {"flops": chain_flops, "ms": chain_ms,
                                "achieved": chain_flops / (chain_ms * 1e-3) / 1e12,
                                "frac": chain_flops / (chain_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                                "algorithmic": "8*L FLOP per unit (forward 4L, backward 4L)"},
            "block_pass": {"flops": block_flops, "ms": block_ms,
                                "achieved": block_flops / (block_ms * 1e-3) / 1e12,
                                "frac": block_flops / (block_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                                "algorithmic": "4*L FLOP per unit (posterior and dosage)"},
            "whole_step": {"flops": 12.0 * L * units, "ms": ms,
                           "achieved": 12.0 * L * units / (ms * 1e-3) / 1e12,
                           "frac": 12.0 * L * units / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS},
        },
        "roofline": {
            "bound": "valu",
            "kernel": dom + " (plain 20,000 x 100,000 workload)",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": load_traffic(dom),
            "algorithmic": f"{4 if dom == 'fb_block_kernel' else 8}*L FLOP per (sample, marker)"
                           f" x {units} units per launch",
        },
        "pbwt": {
            "workload": f"{H} haplotypes x {M} markers, prefix and divergence rows",
            "sweep_ms": sweep_ms,
            "bytes": pbwt_bytes,
            "algorithmic": "M * (8*ceil(H/64) + 4H + 4H) bytes: column read, prefix and "
                           "divergence rows written",
            "achieved_GBps": pbwt_bytes / (sweep_ms * 1e-3) / 1e9,
            "peak_GBps": PEAK_HBM_GBS,
            "frac": pbwt_bytes / (sweep_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "index_build_ms_incl_upload_and_state_table": index_ms,
        },
    }
    if not args.no_cpu_baseline and "batch" not in args.sections.split(","):
        threads = host_threads()
        ns, nm = min(args.cpu_samples, B), min(args.cpu_markers, M)
        log(f"[plain] CPU baseline: reference CPU path on {ns} samples x {nm} markers, "
            f"{threads} threads")
        v, dt, n = cpu_baseline(pan, tg, L, ne, ns, nm, threads)
        out["cpu_baseline"] = {
            "value": v, "unit": "genotypes/s", "cores": threads, "kind": "port",
            "sample": f"first {ns} targets x first {nm} markers of the 20,000 x 100,000 panel "
                      f"and cohort ({n} imputed genotypes in {dt:.1f} s); PBWT sweep + "
                      "per-(sample, marker) std::sort selection + L^2 FP32 forward-backward "
                      "with sqrt(M) replay + sampling",
        }
        log(f"[plain] CPU baseline {v:.4g} genotypes/s")
    return out


def bench_precision(args):
    """Config-3 precision stress: 8,000 x 100,000 panel with heavy-tailed
    spacing (log-normal, median 20 kb, sigma 1.5) and 50x hotspots on 1% of
    intervals, 2,000 targets at 2% genotype error typed every 20th marker,
    L = 8, per-marker transitions, Ne in {1e3, 1e4, 1e5}.  Both FP32 kernels
    run on the whole cohort; the FP64 oracle on the first `precision_samples`
    samples gives the max dosage error and the genotype (argmax GP)
    discordance; discordance against the generator's truth is over the whole
    cohort."""
    import numpy as np
    import torch

    import oracle
    from swiftimpute_amd import device, synth

    H, M, B, L, te = 8000, 100000, 2000, 8, 20
    log(f"[precision] generating panel H={H} M={M} (heavy-tailed spacing, hotspots)")
    pan = synth.make_panel(H, M, seed=12345, spacing_mode=1)
    tg = synth.make_targets(pan, B, te, seed=12346, geno_error=0.02, with_truth=True)
    idx = device.DeviceIndex(pan.cols, H, L, keep_prefix=False, with_div=False)
    a0, a1 = idx.state_alleles()
    codes = torch.from_numpy(tg.codes).cuda()
    untyped = tg.codes[:, :M] == synth.CODE_MISSING
    truth_g = tg.truth[:, 0, :].astype(np.int32) + tg.truth[:, 1, :]
    ns = min(args.precision_samples, B)
    gl_sub = synth.HARDCALL_PALETTE[tg.codes[:ns, :M]]

    def gp(ap0, ap1):
        ap0 = ap0.astype(np.float64)
        ap1 = ap1.astype(np.float64)
        return np.stack([(1 - ap0) * (1 - ap1), ap0 * (1 - ap1) + (1 - ap0) * ap1, ap0 * ap1])

    def gp_call(ap0, ap1):
        return np.argmax(gp(ap0, ap1), axis=0)

    out = {"workload": f"{H} x {M} panel, heavy-tailed spacing + hotspots, {B} targets at 2% "
                       f"genotype error typed every {te}th marker, L={L}, per-marker "
                       f"transitions; FP64 oracle on the first {ns} samples",
           "per_ne": {}}
    for ne in (1000, 10000, 100000):
        tr = device.Transitions.build(pan.pos, pan.genetic_cM(), ne, per_marker=True)
        o = oracle.fb_f64_fast(a0, a1, gl_sub, tr.Td, tr.To, per_marker=True, want_prob=False)
        ref_gp = np.sort(gp(o["ap"][..., 0], o["ap"][..., 1]), axis=0)
        ref_call = gp_call(o["ap"][..., 0], o["ap"][..., 1])
        # calls the FP64 posterior decides by more than 1e-5 (the model often
        # gives exactly tied genotype probabilities, where any rounding flips
        # the argmax)
        decisive = (ref_gp[2] - ref_gp[1]) > 1e-5
        row = {"min_switch_probability": tr.min_switch()}
        for name, logspace in (("linear_f32", False), ("log_f32", True)):
            if not logspace and tr.min_switch() < 1e-30:
                row[name] = None
                continue
            job = device.Job(idx, codes, synth.HARDCALL_PALETTE, tr, seed=12345, logspace=logspace)
            job.alloc_outputs(ds=True, ap=True, hap=False, prob=False, loglik=False)
            device.run_jobs([job.c_job()])
            sync()
            ds = job.out.ds.cpu().numpy()
            ap = job.out.ap.cpu().numpy()
            call = gp_call(ap[..., 0], ap[..., 1])
            u = untyped[:ns]
            row[name] = {
                "max_abs_ds_err_vs_f64": float(np.abs(ds[:ns] - o["ds"]).max()),
                "discordance_vs_f64": float((call[:ns][u] != ref_call[u]).mean()),
                "discordance_vs_f64_decisive_calls": float(
                    (call[:ns][u & decisive] != ref_call[u & decisive]).mean()),
                "discordance_vs_truth": float((call[untyped] != truth_g[untyped]).mean()),
            }
            del job
        row["f64_tied_call_fraction"] = float(1.0 - decisive[untyped[:ns]].mean())
        row["f64_discordance_vs_truth_subset"] = float(
            (ref_call[untyped[:ns]] != truth_g[:ns][untyped[:ns]]).mean())
        out["per_ne"][str(ne)] = row
        log(f"[precision] Ne={ne}: {row}")
    return out


# --------------------------------------------------------------------------
# cfg1: the reference's published run's shape, end to end and kernel-only
# --------------------------------------------------------------------------
REF_PUBLISHED_S = 24.789          # TEST_RESULTS.md:32 (NVIDIA GB10, load -> write)
REF_PUBLISHED_GPS = 14700.0       # README.md:165 (genotypes/s)


def bench_cfg1(args):
    """4,432 haplotypes x 1,644 markers, 444 targets typed every 4th marker,
    L = 8, Ne = 1e4, FP32 log-space, Philox seed 12345.  End to end: the
    synthetic panel and cohort written as VCF text (untimed), then the C++
    API's impute(ref_vcf, tgt_vcf, out_vcf, config) path timed per stage
    (load, index, impute, write), as the reference's executable reports it.
    Kernel-only: the same job on device-resident inputs, timed with HIP events.
    CPU: the reference loop structure on the whole cohort."""
    import tempfile

    import numpy as np
    import torch

    from swiftimpute_amd import _native, device, synth

    w = synth.WORKLOADS["cfg1"]
    H, M, B, L, ne, te = w["H"], w["M"], w["B"], w["L"], w["ne"], w["typed_every"]
    log(f"[cfg1] generating panel H={H} M={M} and {B} targets typed every {te}")
    pan = synth.make_panel(H, M, seed=12345)
    tg = synth.make_targets(pan, B, te, seed=12346, with_truth=False)
    untyped = int((tg.codes[:, :M] == synth.CODE_MISSING).sum())
    out = {"workload": f"{H} haplotypes x {M} markers, {B} targets typed every {te}th marker, "
                       f"L={L}, Ne={ne}, FP32 log-space, Philox seed 12345",
           "untyped_genotypes": untyped}
    lib = _native.api()
    with tempfile.TemporaryDirectory(prefix="si_cfg1_") as d:
        rv, tv, ov = (os.path.join(d, n) for n in ("panel.vcf", "targets.vcf", "out.vcf"))
        synth.write_panel_vcf(rv, pan)
        synth.write_targets_vcf(tv, pan, tg)
        opts = _native.ImputeOpts()
        lib.si_impute_defaults(ctypes.byref(opts))
        opts.num_states, opts.ne, opts.precision = L, ne, 2
        runs = []
        for k in range(2):
            t = _native.ImputeTiming()
            rc = lib.si_impute_vcf(rv.encode(), tv.encode(), ov.encode(), ctypes.byref(opts),
                                   ctypes.byref(t))
            if rc != 0:
                raise SystemExit(f"si_impute_vcf failed ({rc}): {lib.si_impute_last_error().decode()}")
            runs.append({n: getattr(t, n) for n, _ in _native.ImputeTiming._fields_})
            log(f"[cfg1] end to end ({'first' if k == 0 else 'second'} call): "
                f"{t.total_s * 1e3:.1f} ms (load {1e3 * (t.reference_load_s + t.target_load_s):.1f}, "
                f"index {1e3 * t.index_s:.1f}, impute {1e3 * t.impute_s:.1f}, write {1e3 * t.write_s:.1f})")
        assert runs[0]["untyped_genotypes"] == untyped
        out_lines = sum(1 for l in open(ov) if not l.startswith("#"))
        assert out_lines == M
    e2e = runs[0]
    out["end_to_end"] = {
        "stages_s_first_call": {k: e2e[k] for k in ("reference_load_s", "target_load_s", "index_s",
                                                   "impute_s", "write_s", "total_s")},
        "stages_s_second_call": {k: runs[1][k] for k in ("reference_load_s", "target_load_s",
                                                        "index_s", "impute_s", "write_s", "total_s")},
        "genotypes_per_s": untyped / e2e["total_s"],
        "reference_published": {"total_s": REF_PUBLISHED_S, "genotypes_per_s": REF_PUBLISHED_GPS,
                                "imputation_stage_s": 24.133,
                                "note": "NVIDIA GB10, Phylos panel of the same shape with 50% of "
                                        "genotypes masked at random (this cohort masks 75%); "
                                        "TEST_RESULTS.md:26-40"},
        "speedup_vs_published_total": REF_PUBLISHED_S / e2e["total_s"],
        "speedup_vs_published_imputation_stage": 24.133 / e2e["impute_s"],
        "what": "the C++ API's impute(ref_vcf, tgt_vcf, out_vcf, config) via si_impute_vcf: VCF "
                "parsing, device index build, impute() with its uploads and downloads, VCF "
                "writing (GT, DS, GP, AP, INFO)",
    }
    # kernel-only: the plain job on resident inputs
    idx = device.DeviceIndex(pan.cols, H, L, keep_prefix=False, with_div=False)
    tr = device.Transitions.build(pan.pos, None, ne)
    job = device.Job(idx, torch.from_numpy(tg.codes).cuda(), synth.HARDCALL_PALETTE, tr,
                     seed=12345, logspace=True)
    job.alloc_outputs(ds=True, ap=True, hap=True, prob=False, loglik=False)
    arr = (_native.FbJob * 1)(job.c_job())
    wsb = int(_native.hip().si_fb_workspace_bytes(arr, 1))
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sh = ctypes.c_void_p(stream.cuda_stream)
    hip = _native.hip()

    def step():
        _native.check(hip.si_fb_run(arr, 1, ctypes.c_void_p(ws.data_ptr()), wsb, sh), "si_fb_run")

    for _ in range(3):
        step()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = 20
    ev[0].record(stream)
    for _ in range(reps):
        step()
    ev[1].record(stream)
    sync()
    kms = ev[0].elapsed_time(ev[1]) / reps
    out["kernel_only"] = {"ms": kms, "genotypes_per_s": untyped / (kms * 1e-3),
                          "what": "si_fb_run (log-space FP32, both passes) on device-resident "
                                  "inputs and outputs"}
    log(f"[cfg1] kernel-only {kms:.3f} ms")
    if not args.no_cpu_baseline:
        threads = host_threads()
        log(f"[cfg1] CPU baseline: reference loop structure on the whole cohort, {threads} threads")
        v, dt, n = cpu_baseline(pan, tg, L, ne, B, M, threads)
        v1, dt1, n1 = cpu_baseline(pan, tg, L, ne, 16, M, 1)
        out["cpu_baseline"] = {
            "value": v, "unit": "genotypes/s", "cores": threads, "kind": "port",
            "sample": f"the whole cohort ({n} imputed genotypes in {dt:.1f} s)",
            "single_thread": {"value": v1, "cores": 1,
                              "sample": f"first 16 targets, all markers ({n1} imputed genotypes "
                                        f"in {dt1:.1f} s)"},
            "host": host_info()}
        log(f"[cfg1] CPU baseline {v:.4g} genotypes/s ({threads} threads), {v1:.4g} (1 thread)")
    return out


def copy_calibration():
    """A plain device copy of 1 GiB (three launches), the reference for
    scaling the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc."""
    import torch
    n = 1 << 28
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    sync()
    print(json.dumps({"copy_bytes_per_launch": n * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sections", default="batch,serve,plain,cfg1,precision")
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="timed steps of the one-job-per-launch comparison")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    ap.add_argument("--serve-rates", type=lambda x: [float(v) for v in x.split(",")],
                    default=[2.0, 5.0, 10.0, 20.0, 50.0])
    ap.add_argument("--sweep-seconds", type=float, default=4.0)
    ap.add_argument("--serve-H", type=int, default=40000)
    ap.add_argument("--serve-M", type=int, default=500000)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--queue-delay-us", type=int, default=2000)
    ap.add_argument("--cpu-jobs", type=int, default=8)
    ap.add_argument("--cpu-samples", type=int, default=16)
    ap.add_argument("--cpu-markers", type=int, default=4000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision-samples", type=int, default=32)
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the multi-GPU "
              "path is not part of this build.", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration()
        return
    sections = [s for s in args.sections.split(",") if s]
    res = {}
    for s in sections:
        res[s] = {"batch": bench_batch, "serve": bench_serve, "plain": bench_plain,
                  "cfg1": bench_cfg1, "precision": bench_precision}[s](args)
        torch.cuda.empty_cache()

    out = {
        "metric": "imputed genotypes/s, batched vs plain kernel on 1 GPU; p50/p99 request latency",
        "unit": "genotypes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
    }
    if "batch" in res:
        b = res["batch"]
        out.update({
            "value": b["untyped"] / (b["batch_ms"] * 1e-3),
            "ms_per_step": b["batch_ms"],
            "config": {
                "workload": f"{b['n_jobs']} uneven jobs on one 20000 x 200000 panel: regions "
                            "log-uniform 2000-80000 markers, cohorts log-uniform 1-2000 samples, "
                            "L in {4,8,16}, typed 1/10, 1/20 or 1/50, Ne=10000; batched "
                            "persistent work-queue launches (state classes side by side) "
                            "replayed from a hipGraph",
                "jobs": b["n_jobs"], "haplotypes": 20000, "markers": 200000,
                "sample_marker_units": b["units"], "untyped_genotypes": b["untyped"],
                "jobs_per_L": b["L_counts"], "logspace_jobs": b["n_logspace_jobs"],
                "form": "batched",
            },
            "batched_vs_plain": {
                "plain_genotypes_per_s": b["untyped"] / (b["plain_ms"] * 1e-3),
                "plain_ms_per_step": b["plain_ms"],
                "plain_timed_steps": b["plain_steps"],
                "speedup": b["plain_ms"] / b["batch_ms"],
                "outputs_bit_equal": True,
            },
            "autotune": b["autotune"],
            "roofline_batched_step": b["step_roofline"],
        })
    elif "plain" in res:
        p = res["plain"]
        out.update({"value": p["genotypes_per_s"], "ms_per_step": p["ms_per_step"],
                    "config": {"workload": p["workload"], "form": "plain"}})
    if "precision" in res:
        out["precision"] = res["precision"]
    if "cfg1" in res:
        out["config1_4432x1644"] = res["cfg1"]
    if "batch" in res and res["batch"]["cpu_baseline"]:
        out["cpu_baseline"] = res["batch"]["cpu_baseline"]
    if "serve" in res:
        out["serving"] = res["serve"]
    if "plain" in res:
        p = dict(res["plain"])
        out["roofline"] = p.pop("roofline")
        out["pbwt"] = p.pop("pbwt")
        if "cpu_baseline" in p:
            out["cpu_baseline"] = p.pop("cpu_baseline")
        out["plain_20000x100000"] = p
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
