"""Windowed compatibility mode: the reference's windowed forward-backward
(run_windowed_forward_backward, src/api/imputer.cpp:1044-1174, with
forward_pass_windowed_kernel / backward_pass_windowed_kernel,
src/kernels/forward_backward.cu:406-626), which its default configuration
runs when M > window_size.

The HIP path is checked against the oracle's FP64 restatement of that mode
(oracle.fb_f64_windowed, itself pinned by path enumeration in test_oracle.py)
to the same bars as the exact model: dosage and allele probabilities within
1e-5, sampled haplotypes bit-exact against the host replay.  Windows whose
starts fall inside a checkpoint block, short last windows, both numeric forms,
every kernel variant and the batched form are covered."""
import numpy as np
import pytest
import torch

import oracle
from swiftimpute_amd import device, synth

pytestmark = pytest.mark.gpu

DS_TOL = 1e-5


def _codes(rng, B, M, typed_every=4):
    codes = np.full((B, M), 3, np.uint8)
    codes[:, ::typed_every] = rng.integers(0, 3, (B, (M + typed_every - 1) // typed_every))
    return codes


def _job(idx, codes, tr, window, logspace, seed=4242):
    B, M = codes.shape
    pitch = (M + 15) // 16 * 16
    c = np.full((B, pitch), 3, np.uint8)
    c[:, :M] = codes
    job = device.Job(idx, torch.from_numpy(c).cuda(), synth.HARDCALL_PALETTE, tr, seed=seed,
                     logspace=logspace, window=window)
    job.alloc_outputs(ds=True, ap=True, hap=True, prob=True, loglik=True)
    return job


def _result(job):
    torch.cuda.synchronize()
    o = job.out
    return {k: getattr(o, k).cpu().numpy() for k in ("ds", "ap", "hap", "prob", "loglik")}


def _run(idx, codes, tr, window, logspace, variant=None):
    job = _job(idx, codes, tr, window, logspace)
    device.run_jobs([job.c_job()], variant=variant)
    return _result(job)


def _world(L, M, H=None, seed=0, per_marker=True):
    pan = synth.make_panel(H or max(6 * L, 80), M, seed=seed + L, spacing_mode=1)
    idx = device.DeviceIndex(pan.cols, pan.H, L, keep_prefix=False)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000, per_marker=per_marker)
    if not per_marker and tr.min_switch() < 1e-30:
        raise AssertionError("test panel needs the log-space kernel")
    return pan, idx, tr


def _check(idx, codes, tr, window, res, per_marker, seed=4242):
    a0, a1 = idx.state_alleles()
    gl = synth.HARDCALL_PALETTE[codes]
    ref = oracle.fb_f64_windowed(a0, a1, gl, tr.Td, tr.To, window, per_marker=per_marker)
    err = np.abs(res["ds"] - ref["ds"]).max()
    assert err <= DS_TOL, f"max |DS - DS_fp64| = {err:.3g}"
    assert np.abs(res["ap"] - ref["ap"]).max() <= DS_TOL
    assert np.abs(res["prob"] - ref["prob"]).max() <= DS_TOL
    ll = np.abs(res["loglik"] - ref["loglik"]) / np.maximum(1.0, np.abs(ref["loglik"]))
    assert ll.max() < 1e-5
    hap = oracle.sample_replay(res["prob"], a0, a1, seed)
    assert np.array_equal(res["hap"], hap), "sampled haplotypes differ from the host replay"
    return ref


@pytest.mark.parametrize("L,M,W,per_marker,logspace", [
    (8, 700, 100, True, False),    # window starts inside checkpoint blocks
    (8, 700, 96, False, False),    # window starts on checkpoint blocks
    (16, 1000, 250, True, True),
    (5, 333, 40, True, False),     # many short windows, a short last one
    (4, 650, 64, False, True),
    (32, 500, 33, True, False),
    (64, 300, 37, False, False),
    (1, 200, 50, True, False),
])
def test_windowed_matches_oracle(L, M, W, per_marker, logspace):
    print(f"L={L} M={M} W={W} per_marker={per_marker} logspace={logspace}", flush=True)
    rng = np.random.default_rng(L * 31 + W)
    pan, idx, tr = _world(L, M, per_marker=per_marker)
    codes = _codes(rng, 7, M)
    res = _run(idx, codes, tr, W, logspace)
    ref = _check(idx, codes, tr, W, res, per_marker)
    if L > 1:  # the windowed model is a different answer from the exact one
        exact = oracle.fb_f64(*idx.state_alleles(), synth.HARDCALL_PALETTE[codes], tr.Td, tr.To,
                              per_marker=per_marker)
        assert np.abs(ref["ds"] - exact["ds"]).max() > 1e-3


def test_reference_default_window():
    # window_size = 10000 (src/api/imputer.hpp:53) over 21000 markers: three windows
    rng = np.random.default_rng(5)
    L, M = 8, 21000
    pan, idx, tr = _world(L, M, H=200, seed=3, per_marker=False)
    codes = _codes(rng, 3, M, typed_every=8)
    res = _run(idx, codes, tr, 10000, False)
    _check(idx, codes, tr, 10000, res, False)


def test_window_at_least_m_is_the_exact_model():
    rng = np.random.default_rng(9)
    L, M = 8, 300
    pan, idx, tr = _world(L, M)
    codes = _codes(rng, 5, M)
    base = _run(idx, codes, tr, 0, False)
    for W in (M, M + 1, 100000):
        got = _run(idx, codes, tr, W, False)
        for k in base:
            assert np.array_equal(got[k], base[k]), (W, k)


@pytest.mark.parametrize("L,W,logspace", [(16, 100, False), (8, 72, True), (64, 45, False)])
def test_every_variant_gives_identical_windowed_outputs(L, W, logspace):
    rng = np.random.default_rng(L + W)
    M = 517
    pan, idx, tr = _world(L, M)
    codes = _codes(rng, 9, M)
    ref = _run(idx, codes, tr, W, logspace)
    for v in device.variants(L):
        print(f"L={L} W={W} variant={v}", flush=True)
        got = _run(idx, codes, tr, W, logspace, variant=v)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (v, k)


def test_batched_windowed_jobs_equal_plain():
    # windowed and exact jobs of several lengths in one batch, replayed from a
    # graph twice: each equals its plain run
    rng = np.random.default_rng(17)
    pan, idx, tr = _world(8, 900)
    cm = pan.genetic_cM()
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    jobs, want, keep = [], [], []
    for M, W in [(900, 100), (640, 0), (500, 64), (333, 40), (900, 900), (700, 32)]:
        codes = np.ascontiguousarray(_codes(rng, int(rng.integers(2, 9)), M))
        sub = device.DeviceIndex(np.ascontiguousarray(pan.cols[:M]), pan.H, 8, keep_prefix=False)
        trs = device.Transitions.build(pan.pos[:M], None if cm is None else cm[:M], 10000,
                                       per_marker=True)
        want.append(_run(sub, codes, trs, W, False))
        j = _job(sub, codes, trs, W, False)
        cj = tables.apply(j.c_job())
        jobs.append(j)
        keep.append((sub, trs, cj))
    b = device.Batch([cj for _, _, cj in keep])
    stream = torch.cuda.Stream()
    for rep in range(2):
        for j in jobs:
            for v in (j.out.ds, j.out.ap, j.out.prob, j.out.loglik, j.out.hap):
                v.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            b.run(use_graph=True, stream=stream)
        stream.synchronize()
        for j, w in zip(jobs, want):
            got = _result(j)
            for k in w:
                assert np.array_equal(got[k], w[k]), (rep, k)
    b.close()
