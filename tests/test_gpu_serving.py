"""Serving form: a resident panel answering (cohort, region) requests.

A request's result must not depend on what it was batched with: each request
served alone (one request per batch, one stream) equals the same request
served in a mix (many per batch, several streams), and both equal the plain
one-job-per-launch path over the same region (itself checked against the
oracle in test_gpu_parity.py)."""
import numpy as np
import pytest
import torch

from swiftimpute_amd import _native, device, serving, synth

pytestmark = pytest.mark.gpu

L = 8
NE = 10000


def _requests(pan, tg, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        M = int(np.exp(rng.uniform(0, np.log(pan.M))))
        m0 = int(rng.integers(0, pan.M - M + 1))
        B = int(rng.integers(1, tg.B + 1))
        s0 = int(rng.integers(0, tg.B - B + 1))
        out.append((m0, m0 + M, np.ascontiguousarray(tg.codes[s0:s0 + B, m0:m0 + M]), 700 + i))
    return out


@pytest.fixture(scope="module")
def world():
    pan = synth.make_panel(1200, 1800, seed=31)
    tg = synth.make_targets(pan, 48, 10, seed=32)
    return pan, tg


def _plain(pan, idx, tables, m0, m1, codes, seed, idx_m0=0, window=0, pos=None):
    """The plain path over [m0, m1); idx holds the markers from idx_m0 on
    (0: a panel-wide index); window: the windowed model; pos: positions in
    place of the panel's."""
    base = m0 - idx_m0
    M = m1 - m0
    B = codes.shape[0]
    pitch = (M + 15) // 16 * 16
    c = np.full((B, pitch), 3, np.uint8)
    c[:, :M] = codes
    cm = pan.genetic_cM()
    p = pan.pos if pos is None else pos
    tr = device.Transitions.build(p[m0:m1], None if cm is None else cm[m0:m1], NE)
    cd = torch.from_numpy(c).cuda()
    ds = torch.full((B, M), -1.0, device="cuda")
    hap = torch.full((B, 2, M), 7, dtype=torch.uint8, device="cuda")
    j = _native.FbJob()
    j.n_samples, j.n_markers, j.n_states = B, M, L
    j.flags = (_native.SI_FB_LOGSPACE | _native.SI_FB_BANDED) if tr.min_switch() < 1e-30 else 0
    j.window = window
    j.code_pitch = pitch
    j.seed = seed
    s = tr.step
    j.step = _native.StepCoef(float(s[0]), float(s[1]), float(s[2]), float(s[3]))
    j.steps = device._ptr(tr.steps_dev)
    import ctypes
    j.sa = ctypes.c_void_p(idx.sa.data_ptr() + base * idx.Lp)
    j.mmask = ctypes.c_void_p(idx.mmask.data_ptr() + base)
    j.codes = device._ptr(cd)
    tables.apply(j)
    j.ds, j.hap = device._ptr(ds), device._ptr(hap)
    device.run_jobs([j])
    torch.cuda.synchronize()
    return hap.cpu().numpy(), ds.cpu().numpy()


def test_alone_equals_mix_equals_plain(world):
    pan, tg = world
    reqs = _requests(pan, tg, 14, seed=4)
    cm = pan.genetic_cM()
    with serving.Server(pan.cols, pan.H, pan.pos, cm, L=L, ne=NE, max_streams=1,
                        max_batch=1) as srv:
        alone = [srv.wait(srv.submit(*r)) for r in reqs]
        st1 = srv.stats()
    assert st1["batches"] == len(reqs) and st1["max_batch_requests"] == 1
    with serving.Server(pan.cols, pan.H, pan.pos, cm, L=L, ne=NE, max_streams=3, max_batch=6,
                        max_queue_delay_us=200000) as srv:
        ids = [srv.submit(*r) for r in reqs]
        mix = [srv.wait(i) for i in ids]
        st2 = srv.stats()
    assert st2["requests"] == len(reqs)
    assert st2["max_batch_requests"] > 1, st2
    assert st2["batches"] < len(reqs), st2
    for r, a, b in zip(reqs, alone, mix):
        assert np.array_equal(a.hap, b.hap), r[:2]
        assert np.array_equal(a.ds, b.ds), r[:2]
        assert a.latency_s > 0 and b.latency_s > 0
    # results written straight into caller arrays: the same values; every
    # other request's arrays are page-locked, which the device copy fills
    # directly (no staging)
    keep = []

    def out_arrays(k, shape_h, shape_d):
        if k % 2 == 0:
            return np.full(shape_h, 9, np.uint8), np.full(shape_d, -1, np.float32)
        ht = torch.full(shape_h, 9, dtype=torch.uint8, pin_memory=True)
        dt = torch.full(shape_d, -1.0, dtype=torch.float32, pin_memory=True)
        keep.append((ht, dt))
        return ht.numpy(), dt.numpy()

    with serving.Server(pan.cols, pan.H, pan.pos, cm, L=L, ne=NE, max_streams=2, max_batch=4) as srv:
        outs = [out_arrays(k, (r[2].shape[0], 2, r[1] - r[0]), (r[2].shape[0], r[1] - r[0]))
                for k, r in enumerate(reqs)]
        ids = [srv.submit_into(r[0], r[1], r[2], h, d, seed=r[3]) for r, (h, d) in zip(reqs, outs)]
        into = [srv.wait(i) for i in ids]
        st3 = srv.stats()
    # the page-locked requests (odd k) took the direct device copy, the others
    # the staging copy
    assert st3["direct_copies"] == len(keep) == len(reqs) // 2, st3
    for r, a, b, (h, d) in zip(reqs, alone, into, outs):
        assert b.hap is h and b.ds is d
        assert np.array_equal(a.hap, h) and np.array_equal(a.ds, d), r[:2]

    idx = device.DeviceIndex(pan.cols, pan.H, L, keep_prefix=False)
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    for r, a in zip(reqs, alone):
        hap, ds = _plain(pan, idx, tables, *r)
        assert np.array_equal(a.hap, hap), r[:2]
        assert np.array_equal(a.ds, ds), r[:2]


def test_plain_mode_equals_batched_and_splits_time(world):
    # the plain form (one job per launch) on the same requests gives the
    # batched form's results, and so do servers whose page-locked codes pool
    # is too small for some requests or absent (codes in pageable memory,
    # uploaded through slot staging); all report bytes moved and where time
    # went
    pan, tg = world
    reqs = _requests(pan, tg, 10, seed=8)
    cm = pan.genetic_cM()
    got = {}
    for key, plain, staging_gb in (("batched", False, 4.0), ("plain", True, 4.0),
                                   ("small pool", False, 2.0 ** -16), ("no pool", False, 0.0)):
        with serving.Server(pan.cols, pan.H, pan.pos, cm, L=L, ne=NE, max_streams=2, max_batch=4,
                            max_queue_delay_us=100000, plain=plain, staging_gb=staging_gb) as srv:
            ids = [srv.submit(*r) for r in reqs]
            got[key] = [srv.wait(i) for i in ids]
            st = srv.stats()
        assert st["requests"] == len(reqs)
        assert st["h2d_bytes"] == sum(r[2].shape[0] * ((r[1] - r[0] + 15) // 16 * 16) for r in reqs)
        assert st["d2h_bytes"] == sum(r[2].shape[0] * (r[1] - r[0]) * 6 for r in reqs)
        for k in ("launch_host_s", "h2d_s", "kernel_s", "d2h_s", "copy_host_s"):
            assert st[k] >= 0.0, (k, st)
        assert st["kernel_s"] > 0.0 and st["launch_host_s"] > 0.0, st
    for key in ("plain", "small pool", "no pool"):
        for r, a, b in zip(reqs, got["batched"], got[key]):
            assert np.array_equal(a.hap, b.hap), (key, r[:2])
            assert np.array_equal(a.ds, b.ds), (key, r[:2])


def test_windowed_and_coincident_requests_equal_plain(world):
    # The reference's windowed model on long requests (window_size), and
    # requests over coincident markers (switch probability at the -999
    # floor: the banded log form), served in a mix: each equals the plain
    # path with the same model.
    pan, tg = world
    cm = pan.genetic_cM()
    idx = device.DeviceIndex(pan.cols, pan.H, L, keep_prefix=False)
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    reqs = _requests(pan, tg, 8, seed=21)
    reqs.append((100, 1500, np.ascontiguousarray(tg.codes[:9, 100:1500]), 900))
    with serving.Server(pan.cols, pan.H, pan.pos, cm, L=L, ne=NE, max_streams=2, max_batch=4,
                        max_queue_delay_us=100000, window_size=400) as srv:
        got = [srv.wait(srv.submit(*r)) for r in reqs]
    for r, g in zip(reqs, got):
        hap, ds = _plain(pan, idx, tables, *r, window=400)
        assert np.array_equal(g.hap, hap) and np.array_equal(g.ds, ds), r[:2]
    pos = pan.pos.copy()
    pos[300:700] = pos[300]  # a run of coincident markers
    with serving.Server(pan.cols, pan.H, pos, None, L=L, ne=NE, max_streams=2, max_batch=4,
                        max_queue_delay_us=100000) as srv:
        ids = [srv.submit(*r) for r in reqs]
        got = [srv.wait(i) for i in ids]
    for r, g in zip(reqs, got):
        hap, ds = _plain(pan, idx, tables, *r, pos=pos)
        assert np.array_equal(g.hap, hap) and np.array_equal(g.ds, ds), r[:2]


def test_admission_under_tight_budget(world):
    pan, tg = world
    reqs = _requests(pan, tg, 10, seed=8)
    big = max(r[2].shape[0] * (r[1] - r[0]) for r in reqs)
    # room for about two of the largest requests at a time
    budget = 3 * (big * (2 + 4 + 1) + big * L * 4) / (1 << 30) + 1e-4
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE, max_streams=4,
                        max_batch=8) as ref_srv:
        ref = [ref_srv.wait(ref_srv.submit(*r)) for r in reqs]
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE, max_streams=4,
                        max_batch=8, hbm_budget_gb=budget) as srv:
        ids = [srv.submit(*r) for r in reqs]
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["requests"] == len(reqs)
    assert st["peak_bytes"] <= budget * (1 << 30)
    for a, b in zip(ref, got):
        assert np.array_equal(a.hap, b.hap) and np.array_equal(a.ds, b.ds)


def test_request_above_budget_is_refused(world):
    pan, tg = world
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE,
                        hbm_budget_gb=1e-4) as srv:
        with pytest.raises(_native.NativeError, match="budget"):
            srv.submit(0, pan.M, np.ascontiguousarray(tg.codes[:, : pan.M]))


def test_gpu_compat_selection_equals_plain(world):
    # a server built with the reference's GPU-kernel selection serves the same
    # results as the plain path over a state table built that way
    pan, tg = world
    reqs = _requests(pan, tg, 5, seed=12)
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE, max_streams=2,
                        max_batch=4, gpu_compat_selection=True) as srv:
        ids = [srv.submit(*r) for r in reqs]
        got = [srv.wait(i) for i in ids]
    idx = device.DeviceIndex(pan.cols, pan.H, L, keep_prefix=False, gpu_compat=True)
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    for r, a in zip(reqs, got):
        hap, ds = _plain(pan, idx, tables, *r)
        assert np.array_equal(a.hap, hap), r[:2]
        assert np.array_equal(a.ds, ds), r[:2]


def test_region_resweep_equals_a_panel_of_the_region(world):
    # region_resweep: a request's states come from the PBWT of its region alone,
    # swept from the identity order at m0 -- the plain path over an index built
    # from the region's columns
    pan, tg = world
    reqs = _requests(pan, tg, 6, seed=21)
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE, max_streams=2,
                        max_batch=3, region_resweep=True) as srv:
        ids = [srv.submit(*r) for r in reqs]
        got = [srv.wait(i) for i in ids]
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE, max_streams=2,
                        max_batch=3) as srv:
        resident = [srv.wait(srv.submit(*r)) for r in reqs]
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    differ = 0
    for r, a, b in zip(reqs, got, resident):
        m0, m1 = r[0], r[1]
        idx = device.DeviceIndex(np.ascontiguousarray(pan.cols[m0:m1]), pan.H, L, keep_prefix=False)
        hap, ds = _plain(pan, idx, tables, *r, idx_m0=m0)
        assert np.array_equal(a.hap, hap), r[:2]
        assert np.array_equal(a.ds, ds), r[:2]
        differ += m0 > 0 and not np.array_equal(a.ds, b.ds)
    assert differ > 0  # re-sweeping from m0 changes the states of most regions


def test_bad_codes_are_refused(world):
    pan, tg = world
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE) as srv:
        c = np.ascontiguousarray(tg.codes[:3, 100:400])
        c[2, 17] = 4
        with pytest.raises(_native.NativeError, match="outside 0..3"):
            srv.submit(100, 400, c)
        c[2, 17] = 3
        r = srv.wait(srv.submit(100, 400, c))
        assert r.hap.shape == (3, 2, 300)


def test_failed_launch_waits_for_in_flight_copies(world, monkeypatch):
    # The second batch's launch fails on the host (injected); the first batch
    # may still be on the device, copying into page-locked caller arrays.  No
    # wait may return before that copy has landed: the arrays must not change
    # after wait() returns, whatever the request's status.
    pan, tg = world
    monkeypatch.setenv("SI_SERVER_FAIL_AT_BATCH", "2")
    big = np.ascontiguousarray(np.repeat(tg.codes[:, : pan.M], 40, axis=0))  # 1,920 samples
    B, M = big.shape
    ht = torch.full((B, 2, M), 9, dtype=torch.uint8, pin_memory=True)
    dt = torch.full((B, M), -1.0, dtype=torch.float32, pin_memory=True)
    with serving.Server(pan.cols, pan.H, pan.pos, pan.genetic_cM(), L=L, ne=NE, max_streams=2,
                        max_batch=1) as srv:
        first = srv.submit_into(0, M, big, ht.numpy(), dt.numpy(), seed=3)
        second = srv.submit(0, 500, np.ascontiguousarray(tg.codes[:4, :500]), seed=4)
        statuses = []
        for rid in (first, second):
            try:
                srv.wait(rid)
                statuses.append(0)
            except _native.NativeError:
                statuses.append(1)
        snap_h, snap_d = ht.numpy().copy(), dt.numpy().copy()
        import time
        time.sleep(0.3)
        assert np.array_equal(ht.numpy(), snap_h) and np.array_equal(dt.numpy(), snap_d)
        assert statuses[1] == 1  # the injected failure is reported
        with pytest.raises(_native.NativeError, match="injected"):
            srv.submit(0, 100, np.ascontiguousarray(tg.codes[:2, :100]))
