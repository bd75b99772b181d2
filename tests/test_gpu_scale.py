"""GPU parity at the benchmark workloads' full sizes.

Every workload the benchmark quotes runs here at its real shape, checked
against the CPU oracle where the oracle finishes in seconds and through
size-independent properties where it does not:

* 20,000 haplotypes x 100,000 markers, 10,000 targets typed every 20th
  marker, L = 16, Ne = 2e4: every PBWT prefix row and state-table row against
  the oracle's sweep (streamed in marker windows, so the host never holds the
  8 GB arrays), every divergence row by its closed form (row m is m); the
  full-cohort forward-backward in log-space and linear FP32 with a 100-sample
  subset (four runs of 25 spread over the cohort) against the FP64 oracle
  (dosage, allele probabilities, state probabilities within 1e-5) and the
  host Philox replay (haplotypes bit-exact), and the subset rows equal to the
  full-cohort rows bit for bit.
* 512 uneven jobs on one 20,000 x 200,000 panel: the panel's PBWT rows and
  the L = 4, 8, 16 state tables against the oracle (streamed), the batched
  graph-replayed form equal to the plain one-job-per-launch form bit for bit,
  and the first two samples of every job against the FP64 oracle and the
  replay (every L class, the longest regions included).
* 40,000 haplotypes (the partition form of the sweep): the sweep over 8
  default chunks against the oracle; served requests on a 40,000 x 30,000
  panel alone = in a mix = plain, two samples of each against the FP64
  oracle; and served requests on the full 40,000 x 500,000 panel alone = in a
  mix = plain.

Transition constants handed to the oracle come from the oracle's own
restatement (transition.cu:83-121), and the library's are checked equal to
them, so an error in either would show.
"""
import concurrent.futures as cf
import ctypes

import numpy as np
import pytest
import torch

import oracle
from swiftimpute_amd import _native, device, serving, synth
from test_gpu_serving import _plain

pytestmark = pytest.mark.gpu

TOL = 1e-5
WIN = 5000  # markers per streamed oracle window


def unpack(cols, H):
    """[Mw][W] uint64 bit-packed rows -> [Mw][H] uint8 alleles."""
    Mw = cols.shape[0]
    b = np.ascontiguousarray(cols).view(np.uint8).reshape(Mw, -1)
    return np.unpackbits(b, axis=1, bitorder="little")[:, :H]


def oracle_transitions(pos, cM, ne):
    g = oracle.genetic_map(pos, cM)
    return oracle.transitions(g, ne)


def check_index_streamed(pan, prefix, tables, div=None, win=WIN):
    """PBWT prefix rows (device [M][H] int32) and state tables {L: DeviceIndex}
    against the oracle's sweep, window by window; div (if given) by its
    closed form on the device."""
    H, M = pan.H, pan.M
    st = oracle.PbwtStream(H)
    sa_host = {L: ix.state_alleles() for L, ix in tables.items()}
    states = {L: ix.states.cpu().numpy().view(np.uint32) for L, ix in tables.items()}
    for m0 in range(0, M, win):
        m1 = min(M, m0 + win)
        ref = unpack(pan.cols[m0:m1], H)
        p, _ = st.next(ref, with_div=False)
        got = prefix[m0:m1].cpu().numpy().view(np.uint32)
        assert np.array_equal(got, p), f"prefix rows {m0}..{m1}"
        if div is not None:
            want = torch.arange(m0, m1, device=div.device, dtype=torch.int32)[:, None]
            assert bool((div[m0:m1] == want).all()), f"divergence rows {m0}..{m1}"
        for L in tables:
            h0, _, a0, a1 = oracle.state_table(ref, p, L)
            assert np.array_equal(states[L][m0:m1], h0), f"L={L} states {m0}..{m1}"
            assert np.array_equal(sa_host[L][0][m0:m1], a0) and np.array_equal(sa_host[L][1][m0:m1], a1)
    assert st.m == M


def check_against_f64(ds, ap, prob, hap, a0, a1, gl, Td, To, per_marker, seed, sample_base,
                      nthreads=None):
    """One set of samples' outputs against the FP64 oracle and the replay."""
    ref = oracle.fb_f64_fast(a0, a1, gl, Td, To, per_marker=per_marker, nthreads=nthreads)
    err = float(np.abs(ds - ref["ds"]).max())
    assert err <= TOL, f"max |DS - DS_fp64| = {err:.3g}"
    assert float(np.abs(ap - ref["ap"]).max()) <= TOL
    assert float(np.abs(prob - ref["prob"]).max()) <= TOL
    want = oracle.sample_replay(prob, a0, a1, seed, sample_base=sample_base)
    assert np.array_equal(hap, want), "sampled haplotypes differ from the host replay"
    return err


# --------------------------------------------------------------------------
# 20,000 x 100,000, 10,000 targets, L = 16
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg2():
    w = synth.WORKLOADS["cfg2"]
    pan = synth.make_panel(w["H"], w["M"], seed=12345)
    tg = synth.make_targets(pan, w["B"], w["typed_every"], seed=12346, with_truth=False)
    return w, pan, tg


def test_cfg2_index_full_panel(cfg2):
    w, pan, _ = cfg2
    idx = device.DeviceIndex(pan.cols, pan.H, w["L"])
    torch.cuda.synchronize()
    check_index_streamed(pan, idx.prefix, {w["L"]: idx}, div=idx.div)


@pytest.mark.parametrize("logspace", [True, False])
def test_cfg2_fb_full_cohort_with_f64_subset(cfg2, logspace):
    w, pan, tg = cfg2
    H, M, B, L, ne = w["H"], w["M"], w["B"], w["L"], w["ne"]
    idx = device.DeviceIndex(pan.cols, H, L, keep_prefix=False, with_div=False)
    a0, a1 = idx.state_alleles()
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), ne)
    Td, To = oracle_transitions(pan.pos, pan.genetic_cM(), ne)
    assert np.array_equal(tr.Td, Td) and np.array_equal(tr.To, To)
    codes = torch.from_numpy(tg.codes).cuda()
    full = device.Job(idx, codes, synth.HARDCALL_PALETTE, tr, seed=12345, logspace=logspace)
    full.alloc_outputs(ds=True, ap=False, hap=True, prob=False, loglik=False)
    device.run_jobs([full.c_job()])
    torch.cuda.synchronize()
    worst = 0.0
    for s0 in (0, 3333, 6666, B - 25):
        sub = device.Job(idx, codes[s0:s0 + 25].contiguous(), synth.HARDCALL_PALETTE, tr,
                         seed=12345, logspace=logspace, sample_base=s0)
        sub.alloc_outputs(ds=True, ap=True, hap=True, prob=True, loglik=False)
        device.run_jobs([sub.c_job()])
        torch.cuda.synchronize()
        assert torch.equal(sub.out.ds, full.out.ds[s0:s0 + 25]), s0
        assert torch.equal(sub.out.hap, full.out.hap[s0:s0 + 25]), s0
        gl = synth.HARDCALL_PALETTE[tg.codes[s0:s0 + 25, :M]]
        worst = max(worst, check_against_f64(
            sub.out.ds.cpu().numpy(), sub.out.ap.cpu().numpy(), sub.out.prob.cpu().numpy(),
            sub.out.hap.cpu().numpy(), a0, a1, gl, Td, To, False, 12345, s0))
        del sub
    print(f"{'log' if logspace else 'linear'}: max |DS - DS_fp64| over 100 samples = {worst:.3g}")


# --------------------------------------------------------------------------
# 512 uneven jobs on one 20,000 x 200,000 panel
# --------------------------------------------------------------------------
def test_cfg4_full_jobset():
    H, M, pool_B, ne = 20000, 200000, 2000, 10000
    pan = synth.make_panel(H, M, seed=12345)
    pool = synth.make_targets(pan, pool_B, 1, seed=12346, with_truth=False)
    specs = synth.uneven_jobs(512, M, pool_B)
    assert {sp.L for sp in specs} == {4, 8, 16}
    idx16 = device.DeviceIndex(pan.cols, H, 16, keep_prefix=True, with_div=False)
    idx = {16: idx16, 8: idx16.derive(8), 4: idx16.derive(4)}
    torch.cuda.synchronize()
    check_index_streamed(pan, idx16.prefix, idx, win=10000)
    idx16.drop_prefix()
    for ix in idx.values():
        ix.drop_prefix()
    torch.cuda.empty_cache()
    alleles = {L: ix.state_alleles() for L, ix in idx.items()}
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    cm = pan.genetic_cM()
    jobs, trs = [], []
    for sp in specs:
        pitch = (sp.M + 15) // 16 * 16
        c = synth.region_codes(pool.codes, sp, pitch)
        tr = device.Transitions.build(pan.pos[sp.m0:sp.m0 + sp.M],
                                      None if cm is None else cm[sp.m0:sp.m0 + sp.M], ne)
        trs.append(tr)
        jobs.append(device.RegionJob(idx[sp.L], sp.m0, sp.M, torch.from_numpy(c).cuda(), tr,
                                     sp.seed, tables))
    cjobs = [j.c_job() for j in jobs]
    # plain: one job per launch
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ws = None
        for cj in cjobs:
            ws = device.run_jobs([cj], workspace=ws, stream=stream)
    torch.cuda.synchronize()
    plain = [{k: v.clone() for k, v in j.out.items()} for j in jobs]
    del ws
    for j in jobs:
        for v in j.out.values():
            v.fill_(0)
    # batched: persistent work-queue launches, captured into a hipGraph on the
    # first run and replayed on the second
    batch = device.Batch(cjobs)
    batch.run(use_graph=True, stream=stream)
    torch.cuda.synchronize()
    for j in jobs:
        for v in j.out.values():
            v.fill_(0)
    batch.run(use_graph=True, stream=stream)
    torch.cuda.synchronize()
    for n, (j, p) in enumerate(zip(jobs, plain)):
        for k in p:
            assert torch.equal(j.out[k], p[k]), (n, k)
    batch.close()
    del plain
    torch.cuda.empty_cache()

    # the first two samples of every job against the FP64 oracle and the replay
    subs = []
    for sp, tr, j in zip(specs, trs, jobs):
        nb = min(2, sp.B)
        pitch = (sp.M + 15) // 16 * 16
        c = synth.region_codes(pool.codes, synth.JobSpec(sp.m0, sp.M, sp.s0, nb, sp.L,
                                                         sp.typed_every, sp.seed), pitch)
        s = device.RegionJob(idx[sp.L], sp.m0, sp.M, torch.from_numpy(c).cuda(), tr, sp.seed,
                             tables, outputs=("ds", "ap", "hap", "prob"))
        device.run_jobs([s.c_job()])
        subs.append((s, c[:, :sp.M]))
    torch.cuda.synchronize()
    for sp, j, (s, _) in zip(specs, jobs, subs):
        nb = min(2, sp.B)
        assert torch.equal(s.out["ds"], j.out["ds"][:nb]) and torch.equal(s.out["hap"], j.out["hap"][:nb])

    def one(k):
        sp, (s, c), tr = specs[k], subs[k], trs[k]
        a0 = np.ascontiguousarray(alleles[sp.L][0][sp.m0:sp.m0 + sp.M])
        a1 = np.ascontiguousarray(alleles[sp.L][1][sp.m0:sp.m0 + sp.M])
        Td, To = oracle_transitions(pan.pos[sp.m0:sp.m0 + sp.M],
                                    None if cm is None else cm[sp.m0:sp.m0 + sp.M], ne)
        assert np.array_equal(tr.Td, Td) and np.array_equal(tr.To, To)
        return check_against_f64(s.out["ds"].cpu().numpy(), s.out["ap"].cpu().numpy(),
                                 s.out["prob"].cpu().numpy(), s.out["hap"].cpu().numpy(), a0, a1,
                                 synth.HARDCALL_PALETTE[c], Td, To, False, sp.seed, 0, nthreads=1)

    with cf.ThreadPoolExecutor(16) as ex:
        errs = list(ex.map(one, range(len(specs))))
    longest = max(range(len(specs)), key=lambda k: specs[k].M)
    assert specs[longest].M > 70000
    print(f"512 jobs: max |DS - DS_fp64| = {max(errs):.3g} (longest region {specs[longest].M} "
          f"markers, L={specs[longest].L})")


# --------------------------------------------------------------------------
# 40,000 haplotypes: partition form of the sweep, serving
# --------------------------------------------------------------------------
def test_cfg5_partition_form_sweep():
    pan = synth.make_panel(40000, 8 * 512 - 100, seed=12345)
    prefix, div = device.pbwt_sweep(pan.cols, pan.H)
    torch.cuda.synchronize()
    p, d = oracle.pbwt_build(pan.alleles())
    assert np.array_equal(prefix.cpu().numpy().view(np.uint32), p)
    assert np.array_equal(div.cpu().numpy().view(np.uint32), d)


def _cfg5_requests(M, pool_B, n, seed, m_hi, b_hi=5000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        Mr = min(M, int(round(float(np.exp(rng.uniform(np.log(1000), np.log(m_hi)))))))
        B = min(pool_B, int(round(float(np.exp(rng.uniform(0, np.log(b_hi)))))))
        m0 = int(rng.integers(0, M - Mr + 1))
        s0 = int(rng.integers(0, pool_B - B + 1))
        out.append((m0, m0 + Mr, s0, B, 9000 + i))
    # the extremes of the workload's ranges are always in
    out[0] = (0, min(M, m_hi), 0, min(pool_B, b_hi), 9000)
    return out


def _serve_alone_and_mix(pan, reqs, codes):
    cm = pan.genetic_cM()
    with serving.Server(pan.cols, pan.H, pan.pos, cm, L=8, ne=10000, max_streams=1, max_batch=1,
                        hbm_budget_gb=24) as srv:
        alone = [srv.wait(srv.submit(r[0], r[1], c, seed=r[4])) for r, c in zip(reqs, codes)]
    with serving.Server(pan.cols, pan.H, pan.pos, cm, L=8, ne=10000, max_streams=3, max_batch=4,
                        hbm_budget_gb=24, max_queue_delay_us=100000) as srv:
        ids = [srv.submit(r[0], r[1], c, seed=r[4]) for r, c in zip(reqs, codes)]
        mix = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["max_batch_requests"] > 1, st
    for r, a, b in zip(reqs, alone, mix):
        assert np.array_equal(a.hap, b.hap) and np.array_equal(a.ds, b.ds), r[:4]
    return alone


def test_cfg5_serving_40000_haplotypes_with_f64_subset():
    H, M, pool_B = 40000, 30000, 5000
    pan = synth.make_panel(H, M, seed=12345)
    pool = synth.make_targets(pan, pool_B, 1, seed=12346, with_truth=False)
    reqs = _cfg5_requests(M, pool_B, 6, seed=5, m_hi=M)
    codes = []
    for r in reqs:
        c = np.full((r[3], r[1] - r[0]), synth.CODE_MISSING, np.uint8)
        c[:, ::20] = pool.codes[r[2]:r[2] + r[3], r[0]:r[1]:20]
        codes.append(c)
    del pool
    alone = _serve_alone_and_mix(pan, reqs, codes)
    idx = device.DeviceIndex(pan.cols, H, 8, keep_prefix=True, with_div=False)
    torch.cuda.synchronize()
    check_index_streamed(pan, idx.prefix, {8: idx})
    idx.drop_prefix()
    torch.cuda.empty_cache()
    a0, a1 = idx.state_alleles()
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    cm = pan.genetic_cM()
    for r, c, a in zip(reqs, codes, alone):
        m0, m1, _, B, seed = r
        hap, ds = _plain(pan, idx, tables, m0, m1, c, seed)
        assert np.array_equal(a.hap, hap) and np.array_equal(a.ds, ds), r[:4]
        # two samples through the plain path with probabilities, against FP64
        nb = min(2, B)
        tr = device.Transitions.build(pan.pos[m0:m1], None if cm is None else cm[m0:m1], 10000)
        Td, To = oracle_transitions(pan.pos[m0:m1], None if cm is None else cm[m0:m1], 10000)
        assert np.array_equal(tr.Td, Td) and np.array_equal(tr.To, To)
        Mr = m1 - m0
        pitch = (Mr + 15) // 16 * 16
        cc = np.full((nb, pitch), synth.CODE_MISSING, np.uint8)
        cc[:, :Mr] = c[:nb]
        s = device.RegionJob(idx, m0, Mr, torch.from_numpy(cc).cuda(), tr, seed, tables,
                             outputs=("ds", "ap", "hap", "prob"))
        device.run_jobs([s.c_job()])
        torch.cuda.synchronize()
        assert np.array_equal(s.out["ds"].cpu().numpy(), a.ds[:nb])
        assert np.array_equal(s.out["hap"].cpu().numpy(), a.hap[:nb])
        check_against_f64(s.out["ds"].cpu().numpy(), s.out["ap"].cpu().numpy(),
                          s.out["prob"].cpu().numpy(), s.out["hap"].cpu().numpy(),
                          np.ascontiguousarray(a0[m0:m1]), np.ascontiguousarray(a1[m0:m1]),
                          synth.HARDCALL_PALETTE[c[:nb]], Td, To, False, seed, 0)


def test_cfg5_serving_full_panel():
    # the serving workload's own panel, 40,000 x 500,000, resident; requests up
    # to 5,000 samples and 100,000 markers: alone = mix = plain
    H, M, pool_B = 40000, 500000, 5000
    pan = synth.make_panel(H, M, seed=12345)
    reqs = _cfg5_requests(M, pool_B, 5, seed=7, m_hi=100000)
    pool = synth.make_targets(pan, pool_B, 1, seed=12346, with_truth=False)
    codes = []
    for r in reqs:
        c = np.full((r[3], r[1] - r[0]), synth.CODE_MISSING, np.uint8)
        c[:, ::20] = pool.codes[r[2]:r[2] + r[3], r[0]:r[1]:20]
        codes.append(c)
    del pool
    alone = _serve_alone_and_mix(pan, reqs, codes)
    idx = device.DeviceIndex(pan.cols, H, 8, keep_prefix=False, with_div=False)
    torch.cuda.empty_cache()
    tables = device.SharedTables(synth.HARDCALL_PALETTE)
    for r, c, a in zip(reqs, codes, alone):
        hap, ds = _plain(pan, idx, tables, r[0], r[1], c, r[4])
        assert np.array_equal(a.hap, hap) and np.array_equal(a.ds, ds), r[:4]
