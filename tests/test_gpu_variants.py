"""Kernel variants and the autotuner.

Every sum over states is the same binary tree over the padded states whatever
the lane shape, so every variant (lanes per sample in each pass) must give
bit-identical outputs; the default
variant is itself checked against the oracle in test_gpu_parity.py.  The
tuners must pick a built variant, leave the jobs' outputs untouched while
timing, and a tuned batch must still equal the plain path bit for bit."""
import ctypes

import numpy as np
import pytest
import torch

from swiftimpute_amd import _native, device, synth
from test_gpu_batch import RegionJob, make_jobs, plain_results

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    pan = synth.make_panel(700, 430, seed=41, spacing_mode=1)
    tg = synth.make_targets(pan, 37, 6, seed=42)
    return pan, tg, device.SharedTables(synth.HARDCALL_PALETTE)


def _job(world, L, logspace, per_marker, M=430, m0=0, seed=77):
    pan, tg, tables = world
    idx = device.DeviceIndex(pan.cols, pan.H, L, keep_prefix=False)
    return RegionJob(idx, pan, m0, m0 + M, tg.codes, 10000, tables, seed=seed, logspace=logspace,
                     per_marker=per_marker)


def _run(rj, variant):
    cj = rj.cjob()
    device.run_jobs([cj], variant=variant)
    return rj.result()


@pytest.mark.parametrize("L,logspace,per_marker,M", [
    (4, False, True, 430), (5, True, True, 430), (8, False, False, 417), (16, False, True, 430),
    (16, True, False, 33), (33, False, True, 430), (64, True, True, 200), (64, False, False, 17),
    (100, False, True, 200), (128, True, False, 33), (256, False, True, 64),
])
def test_every_variant_gives_identical_outputs(world, L, logspace, per_marker, M):
    rj = _job(world, L, logspace, per_marker, M=M)
    ref = _run(rj, None)
    vs = device.variants(L)
    assert device.default_variant(L) in vs and len(vs) >= (2 if L > 64 else 4)
    for v in vs:
        print(f"L={L} log={logspace} pm={per_marker} M={M} variant={v}", flush=True)
        got = _run(rj, v)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (v, k)


def test_unbuilt_variant_is_refused(world):
    rj = _job(world, 16, False, True, M=64)
    cj = rj.cjob()
    arr = (_native.FbJob * 1)(cj)
    for bad in [(3, 4, 16), (4, 4, 24), (4, 4, 32), (4, 1, 16), (16, 4, 16)]:
        rc = _native.hip().si_fb_run_variant(arr, 1, ctypes.byref(_native.FbVariant(*bad)), None, 0,
                                             3, None)
        assert rc == _native.SI_ERR_UNSUPPORTED, bad


def test_tune_picks_a_built_variant_and_leaves_outputs(world):
    rj = _job(world, 16, False, True)
    cj = rj.cjob()
    v = device.tune_jobs([cj], max_markers=256)
    assert v in device.variants(16)
    r = rj.result()  # still the fill values of RegionJob.cjob
    assert (r["ds"] == -1.0).all() and (r["hap"] == 7).all() and (r["loglik"] == 0).all()


def test_batch_tunes_every_class_and_equals_plain():
    pan, jobs = make_jobs(seed=13, n_jobs=12)
    ref = plain_results(jobs)
    b = device.Batch([rj.cjob() for rj in jobs])
    before = b.classes()
    assert before and not any(t for *_, t in before)
    b.run(use_graph=False)
    after = b.classes()
    for lp, log, var, tuned in after:
        assert tuned and var in device.variants(lp), (lp, log, var)
    for rj, r in zip(jobs, ref):
        got = rj.result()
        for k in r:
            assert np.array_equal(got[k], r[k]), (k, rj.B, rj.M, rj.idx.L)
    # a fixed variant per class, other than the tuned one, replayed from a graph
    for k, (lp, log, var, _) in enumerate(after):
        alt = [v for v in device.variants(lp) if v != var]
        b.set_variant(k, alt[k % len(alt)])
    assert all(not t for *_, t in b.classes())
    for rj in jobs:
        for v in rj.out.values():
            v.fill_(0)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        b.run(use_graph=True, stream=stream)
    stream.synchronize()
    for rj, r in zip(jobs, ref):
        got = rj.result()
        for k in r:
            assert np.array_equal(got[k], r[k]), (k, rj.B, rj.M, rj.idx.L)
    b.close()


def test_second_batch_of_the_same_kind_reuses_the_tuning():
    pan, jobs = make_jobs(seed=13, n_jobs=12)
    b1 = device.Batch([rj.cjob() for rj in jobs])
    b1.tune()
    b2 = device.Batch([rj.cjob() for rj in jobs])
    b2.tune()
    assert [c[2] for c in b1.classes()] == [c[2] for c in b2.classes()]
    b1.close()
    b2.close()
