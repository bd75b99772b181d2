"""si_fb_run_noalloc: the forward-backward pass with its checkpoint workspace
in the library's static per-device scratch (no allocation in the call).

Its outputs must equal si_fb_run's (caller workspace) bit for bit, on calls
made one after another on the same scratch, the second smaller than the
first (so stale checkpoints from the larger call lie beyond what the second
reads and writes), and jobs above the scratch size are refused."""
import ctypes

import numpy as np
import pytest
import torch

from swiftimpute_amd import _native, device, synth

pytestmark = pytest.mark.gpu


def _job(B, M, L, seed, logspace, per_marker):
    pan = synth.make_panel(max(4 * L, 256), M, seed=seed, spacing_mode=1)
    tg = synth.make_targets(pan, B, 5, seed=seed + 1)
    idx = device.DeviceIndex(pan.cols, pan.H, L, keep_prefix=False, with_div=False)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000, per_marker=per_marker)
    job = device.Job(idx, torch.from_numpy(tg.codes).cuda(), synth.HARDCALL_PALETTE, tr,
                     seed=seed, logspace=logspace)
    job.alloc_outputs(ds=True, ap=True, hap=True, prob=True, loglik=True)
    return job


def _outputs(job):
    o = job.out
    return {k: getattr(o, k).clone() for k in ("ds", "ap", "hap", "prob", "loglik")}


def _clear(job):
    o = job.out
    for k in ("ds", "ap", "prob", "loglik"):
        getattr(o, k).fill_(-7.0)
    o.hap.fill_(9)


@pytest.mark.parametrize("variant", [None, "other"])
def test_noalloc_equals_caller_workspace_on_consecutive_calls(variant):
    lib = _native.hip()
    stream = torch.cuda.current_stream()
    sh = ctypes.c_void_p(stream.cuda_stream)
    # a larger call, then a smaller one with other shapes and forms
    shapes = [(48, 3000, 16, 11, True, True), (7, 517, 8, 12, False, False)]
    for B, M, L, seed, logspace, per_marker in shapes:
        job = _job(B, M, L, seed, logspace, per_marker)
        cj = job.c_job()
        arr = (_native.FbJob * 1)(cj)
        need = int(lib.si_fb_workspace_bytes(arr, 1))
        assert 0 < need <= _native.SI_FB_NOALLOC_SCRATCH_BYTES
        var = None
        if variant is not None:
            lanes = device.lane_candidates(L, _native.SI_FB_PASS_CHAIN)
            var = ctypes.byref(_native.FbVariant(lanes[-1], device.default_variant(L)[1], 16))
        ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        _native.check(lib.si_fb_run_variant(arr, 1, var, ctypes.c_void_p(ws.data_ptr()), need, 3,
                                            sh), "si_fb_run_variant")
        torch.cuda.synchronize()
        want = _outputs(job)
        _clear(job)
        torch.cuda.synchronize()
        _native.check(lib.si_fb_run_noalloc(arr, 1, var, sh), "si_fb_run_noalloc")
        torch.cuda.synchronize()
        got = _outputs(job)
        for k in want:
            assert torch.equal(got[k], want[k]), (B, M, L, k)


def test_noalloc_refuses_jobs_above_the_scratch():
    lib = _native.hip()
    j = _native.FbJob()
    j.n_samples, j.n_markers, j.n_states = 10000, 100000, 16
    arr = (_native.FbJob * 1)(j)
    assert int(lib.si_fb_workspace_bytes(arr, 1)) > _native.SI_FB_NOALLOC_SCRATCH_BYTES
    rc = lib.si_fb_run_noalloc(arr, 1, None, None)
    assert rc == _native.SI_ERR_WORKSPACE
    assert b"static scratch" in lib.si_last_error()
