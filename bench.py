#!/usr/bin/env python3
"""Imputation throughput of the MI355X hot path (one JSON line on stdout).

Workload (default "cfg2"): a synthetic 20,000-haplotype x 100,000-marker panel
(mosaics of 40 founders, 50 kb spacing on a 1 cM/Mb map, mutation 1e-3) and
10,000 diploid targets typed at 5,000 evenly spaced markers (every 20th),
imputing the other 95,000; L = 16 states, Ne = 20,000, FP32 forward-backward.

One step = one plain (one job per launch) pass of the hot path over the whole
cohort with every input already resident in HBM: the chain pass and the block
pass of si_fb_run (forward-backward, posterior dosage, Philox sampling).  The
PBWT index is built once at set-up (timed separately, reported under "pbwt").

Progress goes to stderr; the JSON result is the only stdout line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (= FP32 matrix) peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def load_traffic(kernel_key):
    """HBM bytes per launch of a kernel from a committed rocprofv3 --pmc summary
    (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_key, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(pan, tg, L, ne, n_samples, n_markers, threads):
    """The reference CPU path (oracle.cpu_reference_impute: PBWT sweep,
    per-(sample, marker) std::sort selection, L^2 FP32 forward-backward with
    sqrt(M) checkpoint replay, inverse-CDF sampling) on a bounded sample."""
    import numpy as np

    import oracle
    from swiftimpute_amd import synth

    ref = pan.alleles()[:n_markers] if pan.M * pan.H <= 4e8 else None
    if ref is None:  # unpack only the rows we need
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cfg2", choices=["cfg1", "cfg2", "cfg3"])
    ap.add_argument("--cpu-samples", type=int, default=64)
    ap.add_argument("--cpu-markers", type=int, default=5000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is not part of this build (see DESIGN.md, sharding).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from swiftimpute_amd import _native, device, synth

    w = synth.WORKLOADS[args.workload]
    H, M, B, L, ne, te = w["H"], w["M"], w["B"], w["L"], w["ne"], w["typed_every"]
    threads = min(16, os.cpu_count() or 1)
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()

    log(f"generating panel H={H} M={M}")
    pan = synth.make_panel(H, M, seed=12345, spacing_mode=w["spacing_mode"])
    log(f"generating {B} targets typed every {te}")
    tg = synth.make_targets(pan, B, te, seed=12346, with_truth=False)
    untyped = int((tg.codes[:, :M] == synth.CODE_MISSING).sum())

    # ---- index build (PBWT sweep + state table), timed once
    log("building the PBWT index on the GPU")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    idx = device.DeviceIndex(pan.cols, H, L, stream=stream)
    ev[1].record(stream)
    torch.cuda.synchronize()
    index_ms = ev[0].elapsed_time(ev[1])
    W = (H + 63) // 64
    pbwt_bytes = M * (W * 8 + 4 * H + 4 * H)
    log(f"index built in {index_ms:.1f} ms")
    idx.prefix = None
    idx.div = None
    torch.cuda.empty_cache()

    per_marker = w["spacing_mode"] != 0
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), ne, per_marker=per_marker)
    codes = torch.from_numpy(tg.codes).cuda()
    job = device.Job(idx, codes, synth.HARDCALL_PALETTE, tr, seed=12345)
    job.alloc_outputs(ds=True, ap=False, hap=True, prob=False, loglik=False)
    cj = job.c_job()
    import ctypes
    arr = (_native.FbJob * 1)(cj)
    lib = _native.hip()
    ws_bytes = int(lib.si_fb_workspace_bytes(arr, 1))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    sh = ctypes.c_void_p(stream.cuda_stream)

    def step(evs=None):
        if evs:
            evs[0].record(stream)
        _native.check(lib.si_fb_run_passes(arr, 1, ctypes.c_void_p(ws.data_ptr()), ws_bytes,
                                           _native.SI_FB_PASS_CHAIN, sh), "chain pass")
        if evs:
            evs[1].record(stream)
        _native.check(lib.si_fb_run_passes(arr, 1, ctypes.c_void_p(ws.data_ptr()), ws_bytes,
                                           _native.SI_FB_PASS_BLOCK, sh), "block pass")
        if evs:
            evs[2].record(stream)

    log(f"warmup {args.warmup} steps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / args.steps
    chain_ms = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
    block_ms = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps
    log(f"{ms:.3f} ms/step (chain {chain_ms:.3f} ms, block {block_ms:.3f} ms)")

    # sanity of the outputs of the last step (size-independent properties)
    ds = job.out.ds
    assert torch.isfinite(ds).all().item() and float(ds.min()) >= -1e-6 and float(ds.max()) <= 2 + 1e-5
    hap = job.out.hap
    assert int(hap.max()) <= 1

    units = B * M
    block_flops = 12.0 * L * units   # forward replay 4L + backward 4L + posterior/dosage 4L
    chain_flops = 8.0 * L * units    # forward 4L + backward 4L
    dom, dom_ms, dom_flops = (("fb_block_kernel", block_ms, block_flops) if block_ms >= chain_ms
                              else ("fb_chain_kernel", chain_ms, chain_flops))
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    result = {
        "metric": "imputed genotypes/s",
        "value": untyped / (ms * 1e-3),
        "unit": "genotypes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {H} haplotypes x {M} markers panel, {B} targets typed "
                        f"every {te}th marker, L={L}, Ne={ne}, plain one-job-per-launch",
            "haplotypes": H, "markers": M, "samples": B, "states": L, "ne": ne,
            "typed_every": te, "transitions": "per-marker" if per_marker else "first-interval",
            "form": "plain",
        },
        "all_marker_genotypes_per_s": units / (ms * 1e-3),
        "passes_ms": {"fb_chain_kernel": chain_ms, "fb_block_kernel": block_ms},
        "roofline": {
            "bound": "valu",
            "kernel": dom,
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": load_traffic(dom),
            "algorithmic": f"{12 if dom == 'fb_block_kernel' else 8}*L FLOP per (sample, marker)"
                           f" x {units} units per launch",
        },
        "pbwt": {
            "index_build_ms": index_ms,
            "bytes": pbwt_bytes,
            "achieved_GBps_incl_state_table": pbwt_bytes / (index_ms * 1e-3) / 1e9,
            "peak_GBps": PEAK_HBM_GBS,
        },
    }
    del ws
    if not args.no_cpu_baseline:
        ns, nm = min(args.cpu_samples, B), min(args.cpu_markers, M)
        log(f"CPU baseline: reference CPU path on {ns} samples x {nm} markers, {threads} threads")
        v, dt, n = cpu_baseline(pan, tg, L, ne, ns, nm, threads)
        result["cpu_baseline"] = {
            "value": v, "unit": "genotypes/s", "cores": threads, "kind": "port",
            "sample": f"first {ns} targets x first {nm} markers of the same panel/cohort "
                      f"({n} imputed genotypes in {dt:.1f} s); PBWT sweep + per-(sample, marker) "
                      "std::sort selection + L^2 FP32 forward-backward with sqrt(M) replay + "
                      "sampling; the replay cost per marker grows with sqrt(M), so the full "
                      "100k-marker run would be slower per genotype",
        }
        log(f"CPU baseline {v:.4g} genotypes/s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
