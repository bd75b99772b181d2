"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: PBWT prefix/divergence, selected states and state alleles bit-exact;
posterior dosage / allele probabilities within 1e-5 absolute of the FP64
oracle; sampled haplotypes bit-exact against the host replay of the same
Philox4x32-10 counters on the device's FP32 probabilities.
"""
import math

import numpy as np
import pytest
import torch

import oracle
from swiftimpute_amd import _native, device, synth
from test_oracle import fixture_panel

pytestmark = pytest.mark.gpu

DS_TOL = 1e-5


def pack_cols(ref):
    """[M][H] uint8 -> (cols [M][W] uint64 of nonzero alleles, miss or None)."""
    M, H = ref.shape
    W = (H + 63) // 64
    def pack(bits):
        b = np.zeros((M, W * 64), np.uint8)
        b[:, :H] = bits
        return np.packbits(b, axis=1, bitorder="little").view(np.uint64).reshape(M, W)
    miss = ref == 255
    return pack(ref != 0), (pack(miss) if miss.any() else None)


def gpu_index(ref, L, mode=0, gpu_compat=False):
    cols, miss = pack_cols(ref)
    idx = device.DeviceIndex(cols, ref.shape[1], L, mode=mode, miss=miss, gpu_compat=gpu_compat)
    torch.cuda.synchronize()
    return idx


def run_fb(idx, codes, palette, trans, seed=12345, argmax=False, sample_base=0, prob=True,
           logspace=False):
    B, M = codes.shape[0], idx.M
    pitch = (M + 15) // 16 * 16
    c = np.full((B, pitch), palette.shape[0] - 1, np.uint8)
    c[:, :M] = codes[:, :M]
    job = device.Job(idx, torch.from_numpy(c).cuda(), palette, trans, seed=seed, argmax=argmax,
                     sample_base=sample_base, logspace=logspace)
    job.alloc_outputs(ds=True, ap=True, hap=True, prob=prob, loglik=True)
    device.run_jobs([job.c_job()])
    torch.cuda.synchronize()
    o = job.out
    return {k: (getattr(o, k).cpu().numpy() if getattr(o, k) is not None else None)
            for k in ("ds", "ap", "hap", "prob", "loglik")}


# ----------------------------------------------------------------------- PBWT
def test_pbwt_fixture():
    _, ref = fixture_panel()
    idx = gpu_index(ref, 4)
    assert idx.prefix.cpu().numpy()[:2].tolist() == [[0, 1, 2, 5, 3, 4, 6, 7],
                                                      [0, 1, 2, 3, 5, 4, 6, 7]]
    p, d = oracle.pbwt_build(ref)
    assert np.array_equal(idx.prefix.cpu().numpy().view(np.uint32), p)
    assert np.array_equal(idx.div.cpu().numpy().view(np.uint32), d)


@pytest.mark.parametrize("M,H", [(1, 1), (2, 5), (40, 63), (33, 64), (50, 65), (20, 1000),
                                 (300, 4432), (7, 20000)])
def test_pbwt_random(M, H):
    rng = np.random.default_rng(M * 1000 + H)
    ref = (rng.random((M, H)) < rng.uniform(0.05, 0.9)).astype(np.uint8)
    idx = gpu_index(ref, 4)
    p, d = oracle.pbwt_build(ref)
    assert np.array_equal(idx.prefix.cpu().numpy().view(np.uint32), p)
    assert np.array_equal(idx.div.cpu().numpy().view(np.uint32), d)


def _dup_panel(rng, M, H, n_distinct=6, flip=0.01):
    """Panels of near-identical haplotypes: large groups in every chunk."""
    base = (rng.random((M, n_distinct)) < 0.5).astype(np.uint8)
    ref = base[:, rng.integers(0, n_distinct, H)]
    ref[rng.random((M, H)) < flip] ^= 1
    return ref


@pytest.mark.parametrize("M,H,chunk", [(1, 1, 1), (5, 2, 1), (40, 63, 7), (33, 64, 4),
                                       (50, 65, 5), (64, 300, 3), (100, 1100, 16), (37, 1025, 2),
                                       (300, 4432, 64), (30, 40000, 7), (9, 65536, 4),
                                       (700, 2000, 512), (1200, 500, 100), (1100, 20000, 512),
                                       (900, 24000, 200), (300, 32000, 96), (30, 65537, 7),
                                       (23, 100000, 5), (12, 300000, 3), (6, 524288, 2)])
@pytest.mark.parametrize("dup", [False, True])
def test_pbwt_chunked(M, H, chunk, dup):
    rng = np.random.default_rng(M * 7 + H + chunk + dup)
    if dup:
        ref = _dup_panel(rng, M, H)
    else:
        ref = (rng.random((M, H)) < rng.uniform(0.05, 0.9)).astype(np.uint8)
    ref[rng.random((M, H)) < 0.01] = 255      # missing alleles sort with the ones
    cols, _ = pack_cols(ref)
    prefix, div = device.pbwt_sweep(cols, H, chunk=chunk)
    torch.cuda.synchronize()
    p, d = oracle.pbwt_build(ref)
    assert np.array_equal(prefix.cpu().numpy().view(np.uint32), p)
    assert np.array_equal(div.cpu().numpy().view(np.uint32), d)


def test_pbwt_default_chunk_matches_all_chunk_lengths():
    # size-independent property: every chunk length gives the same rows
    pan = synth.make_panel(3000, 2100, seed=99)
    ref_p, _ = device.pbwt_sweep(pan.cols, pan.H, with_div=False)
    for chunk in (1, 3, 33, 256, 1024, 4096):
        p, _ = device.pbwt_sweep(pan.cols, pan.H, chunk=chunk, with_div=False)
        assert torch.equal(p, ref_p), chunk


def test_pbwt_config1_panel():
    pan = synth.make_panel(4432, 1644, seed=12345)
    ref = pan.alleles()
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    torch.cuda.synchronize()
    p, d = oracle.pbwt_build(ref)
    assert np.array_equal(idx.prefix.cpu().numpy().view(np.uint32), p)
    assert np.array_equal(idx.div.cpu().numpy().view(np.uint32), d)


# ---------------------------------------------------------------- state table
@pytest.mark.parametrize("H,L", [(4432, 8), (300, 16), (100, 64), (5, 8), (3, 3), (1, 4),
                                 (77, 5), (200, 12), (300, 100), (200, 128), (90, 128), (600, 256)])
def test_state_table(H, L):
    rng = np.random.default_rng(H + L)
    M = 40
    ref = (rng.random((M, H)) < 0.5).astype(np.uint8)
    ref[rng.random((M, H)) < 0.02] = 255
    idx = gpu_index(ref, L)
    p, _ = oracle.pbwt_build(ref)
    h0, _, a0, a1 = oracle.state_table(ref, p, L)
    assert np.array_equal(idx.states.cpu().numpy().view(np.uint32), h0)
    g0, g1 = idx.state_alleles()
    assert np.array_equal(g0, a0) and np.array_equal(g1, a1)
    # class-present mask and padding bytes
    sa = idx.sa.cpu().numpy()
    cls = np.minimum(a0.astype(int) + a1.astype(int), 2)
    assert np.array_equal((sa[:, :L] >> 2) & 3, cls)
    assert ((sa[:, L:] & 0x7f) == 0x4c).all()
    mm = idx.mmask.cpu().numpy()
    for m in range(M):
        assert mm[m] == sum(1 << c for c in set(cls[m].tolist()))
    # the class mask rides in bit 7 of bytes 0-2 of every 4-byte word
    k = np.arange(idx.Lp)
    want = np.where((k % 4)[None, :] < 3, (mm[:, None] >> (k % 4)[None, :]) & 1, 0)
    assert np.array_equal(sa >> 7, want)


@pytest.mark.parametrize("H,L", [(4432, 8), (200, 32), (128, 3), (10001, 8), (300, 40), (300, 64),
                                 (70, 12), (5, 8), (300, 65), (20000, 100), (90, 128), (600, 200)])
def test_state_table_gpu_compat(H, L):
    # the reference's GPU selection kernels, run lane by lane in the oracle on
    # the PBWT rows and divergences
    rng = np.random.default_rng(H * 3 + L)
    M = 40
    ref = (rng.random((M, H)) < 0.5).astype(np.uint8)
    ref[rng.random((M, H)) < 0.02] = 255
    idx = gpu_index(ref, L, gpu_compat=True)
    p, d = oracle.pbwt_build(ref)
    h0, _, a0, a1 = oracle.state_table_gpu(ref, p, d, L)
    assert np.array_equal(idx.states.cpu().numpy().view(np.uint32), h0)
    g0, g1 = idx.state_alleles()
    assert np.array_equal(g0, a0) and np.array_equal(g1, a1)


def test_state_table_simple_mode():
    rng = np.random.default_rng(4)
    ref = (rng.random((10, 40)) < 0.5).astype(np.uint8)
    idx = gpu_index(ref, 8, mode=1)
    h0, _, a0, a1 = oracle.state_table(ref, np.zeros((10, 40), np.uint32), 8, simple_mode=True)
    assert np.array_equal(idx.states.cpu().numpy().view(np.uint32), h0)
    g0, g1 = idx.state_alleles()
    assert np.array_equal(g0, a0) and np.array_equal(g1, a1)


# ------------------------------------------------------------ forward-backward
def check_fb(idx, codes, palette, trans, res, per_marker, fold=True, sample_base=0, seed=12345,
             argmax=False, tol=DS_TOL):
    a0, a1 = idx.state_alleles()
    gl = palette[codes[:, : idx.M]]
    Td, To = trans.Td, trans.To
    if trans.source is not None:
        # the oracle's own restatement of the map and constants
        # (imputer.cpp:596-605, transition.cu:83-121), which the library's
        # must equal bit for bit
        pos, cM, ne, rho = trans.source
        Td, To = oracle.transitions(oracle.genetic_map(pos, cM), ne, rho)
        assert np.array_equal(Td, trans.Td) and np.array_equal(To, trans.To)
    if fold:
        ref = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=per_marker)
    else:
        ref = oracle.fb_f64_fast(a0, a1, gl, Td, To, per_marker=per_marker)
    err = np.abs(res["ds"] - ref["ds"]).max()
    assert err <= tol, f"max |DS - DS_fp64| = {err:.3g}"
    assert np.abs(res["ap"] - ref["ap"]).max() <= tol
    if ref["prob"] is not None and res["prob"] is not None:
        assert np.abs(res["prob"] - ref["prob"]).max() <= tol
    ll_err = np.abs(res["loglik"] - ref["loglik"]) / np.maximum(1.0, np.abs(ref["loglik"]))
    assert ll_err.max() < 1e-5
    hap = oracle.sample_replay(res["prob"], a0, a1, seed, mode="argmax" if argmax else "philox",
                               sample_base=sample_base)
    assert np.array_equal(res["hap"], hap), "sampled haplotypes differ from the host replay"
    return err


def hardcall_codes(rng, B, M, typed_every=4):
    codes = np.full((B, M), 3, np.uint8)
    codes[:, ::typed_every] = rng.integers(0, 3, (B, (M + typed_every - 1) // typed_every))
    return codes


@pytest.mark.parametrize("logspace", [False, True])
def test_untyped_code_with_nonzero_likelihoods_is_ignored(logspace):
    # SI_FB_UNTYPED naming a palette entry that is not (0, 0, 0) (here the
    # hom-ref call): the code is ignored and every marker takes the exact
    # step, so the outputs equal the same job without the flag
    rng = np.random.default_rng(808)
    pan = synth.make_panel(300, 250, seed=41)
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = hardcall_codes(rng, 13, pan.M, typed_every=5)
    pitch = (pan.M + 15) // 16 * 16
    c = np.full((13, pitch), 3, np.uint8)
    c[:, :pan.M] = codes
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000)
    outs = []
    for code in (None, 0):
        job = device.Job(idx, torch.from_numpy(c).cuda(), synth.HARDCALL_PALETTE, tr, seed=5,
                         logspace=logspace, untyped=False)
        job.alloc_outputs(ds=True, ap=True, hap=True, loglik=True)
        cj = job.c_job()
        if code is not None:
            cj.flags |= _native.SI_FB_UNTYPED
            cj.untyped_code = code
        device.run_jobs([cj])
        torch.cuda.synchronize()
        outs.append({k: getattr(job.out, k).cpu().numpy() for k in ("ds", "ap", "hap", "loglik")})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 8, 12, 16, 32, 64, 65, 100, 128, 200, 256])
@pytest.mark.parametrize("per_marker", [False, True])
@pytest.mark.parametrize("logspace", [False, True])
def test_fb_states_and_modes(L, per_marker, logspace):
    rng = np.random.default_rng(L * 7 + per_marker)
    pan = synth.make_panel(max(4 * L, 64), 97, seed=L, spacing_mode=1)
    idx = device.DeviceIndex(pan.cols, pan.H, L)
    codes = hardcall_codes(rng, 9, pan.M)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000, per_marker=per_marker)
    res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, logspace=logspace)
    check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, per_marker)


@pytest.mark.parametrize("logspace", [False, True])
def test_fb_general_genotype_likelihoods(logspace):
    # a palette of soft likelihoods (not hard calls), including mixed -999s
    rng = np.random.default_rng(21)
    pal = np.concatenate([synth.HARDCALL_PALETTE[:3],
                          rng.uniform(-4, 0, (20, 3)).astype(np.float32),
                          np.array([[-999, -2.0, -999], [0, 0, 0]], np.float32)])
    pan = synth.make_panel(256, 300, seed=5, spacing_mode=1)
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = rng.integers(0, pal.shape[0], (6, pan.M)).astype(np.uint8)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 5000, per_marker=True)
    res = run_fb(idx, codes, pal, tr, logspace=logspace)
    check_fb(idx, codes, pal, tr, res, True)


@pytest.mark.parametrize("ne", [1000, 10000, 100000])
@pytest.mark.parametrize("logspace", [False, True])
def test_fb_long_chain_reaches_floor(ne, logspace):
    # 4000 markers: the backward vector falls to the -999 floor (about 2-4 per
    # marker) and the floor then shapes the posteriors; per-marker transitions
    # with heavy-tailed spacing keep the model non-degenerate.
    rng = np.random.default_rng(ne)
    pan = synth.make_panel(400, 4000, seed=7, spacing_mode=1)
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = hardcall_codes(rng, 4, pan.M, typed_every=20)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), ne, per_marker=True)
    res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, logspace=logspace)
    check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, True, fold=False)


@pytest.mark.parametrize("logspace", [False, True])
def test_fb_all_states_mismatch(logspace):
    # hom-alt observed where every selected pair is hom-ref: the emission
    # offset is -999 and the floor dominates the backward step.
    ref = np.zeros((30, 16), np.uint8)
    ref[:, 8:] = 1
    ref[:, 3] = 1
    idx = gpu_index(ref, 4)
    codes = np.full((2, 30), 3, np.uint8)
    codes[:, 10] = 2
    codes[1, 20] = 0
    g = np.arange(1, 31, dtype=np.uint64) * 3000
    tr = device.Transitions.build(g, None, 10000, per_marker=True)
    res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, logspace=logspace)
    check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, True)


def test_fb_coincident_markers_logspace():
    # Duplicate positions give rho = 0: stay = 1, switch = -999 (no switching
    # at all, the floor carries every path).  Only the log-space kernel
    # represents this; the linear kernel refuses it.
    rng = np.random.default_rng(17)
    pan = synth.make_panel(128, 120, seed=17)
    pos = pan.pos.copy()
    pos[40:60] = pos[40]           # a run of coincident markers
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = hardcall_codes(rng, 6, pan.M, typed_every=3)
    for per_marker in (True, False):
        p2 = pos.copy()
        if not per_marker:
            p2[1] = p2[0]          # the first interval drives every step
        tr = device.Transitions.build(p2, None, 10000, per_marker=per_marker)
        assert (tr.To == -999.0).any()
        assert tr.floor_scale()  # the job runs the banded kernels
        res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, logspace=True)
        # Reference-compatible mode with a zero-length first interval has
        # To = -999 at every step: no switching at all, and the posteriors are
        # set by terms e^-999 apart, which FP32 resolves to 6.1e-5 at |x| ~ 999
        # unless the floor is carried apart from the residual (SI_FB_BANDED).
        check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, per_marker, fold=True, tol=DS_TOL)
    with pytest.raises(Exception):
        run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, logspace=False)


@pytest.mark.parametrize("L", [3, 8, 16, 33, 100, 256])
def test_fb_banded_variants_and_batch_agree(L):
    # Banded jobs: every lane shape gives the same bits (the same trees over
    # the states), and the batched form equals the plain one.
    rng = np.random.default_rng(L + 70)
    pan = synth.make_panel(160, 150, seed=L + 3)
    pos = pan.pos.copy()
    pos[30:70] = pos[30]
    pos[1] = pos[0]
    idx = device.DeviceIndex(pan.cols, pan.H, L)
    codes = hardcall_codes(rng, 11, pan.M, typed_every=3)
    out = {}
    for per_marker in (False, True):
        tr = device.Transitions.build(pos, None, 10000, per_marker=per_marker)
        assert tr.floor_scale()
        ref = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, logspace=True)
        check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, ref, per_marker, fold=True, tol=DS_TOL)
        out[per_marker] = ref
        B, M = codes.shape[0], idx.M
        pitch = (M + 15) // 16 * 16
        c = np.full((B, pitch), 3, np.uint8)
        c[:, :M] = codes
        keys = ("ds", "ap", "hap", "prob", "loglik")
        for v in device.variants(idx.Lp):
            job = device.Job(idx, torch.from_numpy(c).cuda(), synth.HARDCALL_PALETTE, tr, seed=12345,
                             logspace=True)
            job.alloc_outputs(ds=True, ap=True, hap=True, prob=True, loglik=True)
            device.run_jobs([job.c_job()], variant=v)
            torch.cuda.synchronize()
            for k in keys:
                assert np.array_equal(getattr(job.out, k).cpu().numpy(), ref[k]), (v, k)
        # two cohorts of one region in one batch (a shared emission table)
        tables = device.SharedTables(synth.HARDCALL_PALETTE)
        rjs = [device.RegionJob(idx, 0, M, torch.from_numpy(c[lo:hi]).cuda(), tr, seed=5, tables=tables,
                                logspace=True, outputs=keys) for lo, hi in ((0, 4), (4, B))]
        plain = []
        for rj in rjs:
            device.run_jobs([rj.c_job()])
            torch.cuda.synchronize()
            plain.append({k: rj.out[k].cpu().numpy() for k in keys})
            rj.out["ds"].zero_()
            rj.out["hap"].zero_()
        b = device.Batch([rj.c_job() for rj in rjs])
        b.run(use_graph=False)
        torch.cuda.synchronize()
        for rj, pr in zip(rjs, plain):
            for k in keys:
                assert np.array_equal(rj.out[k].cpu().numpy(), pr[k]), ("batch", k)
        b.close()


@pytest.mark.parametrize("L", [4, 8, 16])
def test_fb_panel_with_missing_alleles(L):
    # Missing panel alleles reach the sampled haplotypes as 255 through the
    # block pass's packed state bits (or, for 16 states on 4 lanes, the state
    # table read by the sampler): the host replay over the same state
    # alleles, and every variant, must agree bit for bit.
    rng = np.random.default_rng(L + 71)
    M, H = 200, 120
    ref = (rng.random((M, H)) < 0.4).astype(np.uint8)
    ref[rng.random((M, H)) < 0.05] = 255
    idx = gpu_index(ref, L)
    a0, a1 = idx.state_alleles()
    assert (a0 == 255).any() and (a1 == 255).any()
    codes = hardcall_codes(rng, 8, M, typed_every=3)
    g = np.arange(1, M + 1, dtype=np.uint64) * 2000
    tr = device.Transitions.build(g, None, 10000, per_marker=True)
    res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr)
    check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, True)
    assert (res["hap"] == 255).any()
    pitch = (M + 15) // 16 * 16
    c = np.full((8, pitch), 3, np.uint8)
    c[:, :M] = codes
    for v in device.variants(L):
        job = device.Job(idx, torch.from_numpy(c).cuda(), synth.HARDCALL_PALETTE, tr, seed=12345)
        job.alloc_outputs(ds=True, ap=False, hap=True, prob=False, loglik=False)
        device.run_jobs([job.c_job()], variant=v)
        torch.cuda.synchronize()
        assert np.array_equal(job.out.hap.cpu().numpy(), res["hap"]), v
        assert np.array_equal(job.out.ds.cpu().numpy(), res["ds"]), v


def test_fb_logspace_switch_terms_around_linear_bound():
    # The log-space steps add the shared switch (and floor) terms in linear
    # space when they are at least 2^-30 (e^-20.8) and use the exact
    # three-term form below that.  Switch terms spread across that bound, per
    # marker, make one chain alternate between the two forms; both must stay
    # within the FP64 oracle's tolerance.
    rng = np.random.default_rng(31)
    M = 400
    pan = synth.make_panel(256, M, seed=31)
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = hardcall_codes(rng, 6, M, typed_every=3)
    To = rng.uniform(-24.0, -17.0, M - 1).astype(np.float32)
    Td = np.log1p(-8.0 * np.exp(To.astype(np.float64))).astype(np.float32)
    co = device.step_coefs(Td, To)
    full = np.zeros((M, 4), np.float32)
    full[1:] = co
    tr = device.Transitions(True, co[0], torch.from_numpy(full).cuda(), Td, To)
    assert (np.exp(To.astype(np.float64)) < 2.0 ** -30).any()
    assert (np.exp(To.astype(np.float64)) > 2.0 ** -30).any()
    res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, logspace=True)
    check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, True)


@pytest.mark.parametrize("M", [1, 2, 15, 16, 17, 33])
def test_fb_marker_count_edges(M):
    rng = np.random.default_rng(M)
    pan = synth.make_panel(64, M, seed=M)
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = hardcall_codes(rng, 3, M, typed_every=2)
    tr = device.Transitions.build(pan.pos, None, 10000, per_marker=False)
    if M >= 2:  # non-degenerate transition for the compat step
        tr.step = device.step_coefs(np.array([-0.2], np.float32), np.array([-3.0], np.float32))[0]
        tr.Td = np.full(M - 1, -0.2, np.float32)
        tr.To = np.full(M - 1, -3.0, np.float32)
        tr.source = None  # constants chosen here, not from the map
    res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr)
    check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, False)


def test_fb_argmax_mode():
    rng = np.random.default_rng(2)
    pan = synth.make_panel(128, 200, seed=2, spacing_mode=1)
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = hardcall_codes(rng, 5, pan.M)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000, per_marker=True)
    res = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr, argmax=True)
    check_fb(idx, codes, synth.HARDCALL_PALETTE, tr, res, True, argmax=True)


def test_fb_batch_split_invariance():
    # Draws depend on (seed, global sample, marker) only: splitting a cohort into
    # batches with the right sample_base reproduces the whole-cohort run.
    rng = np.random.default_rng(8)
    pan = synth.make_panel(256, 150, seed=8, spacing_mode=1)
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    codes = hardcall_codes(rng, 12, pan.M)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000, per_marker=True)
    whole = run_fb(idx, codes, synth.HARDCALL_PALETTE, tr)
    a = run_fb(idx, codes[:5], synth.HARDCALL_PALETTE, tr, sample_base=0)
    b = run_fb(idx, codes[5:], synth.HARDCALL_PALETTE, tr, sample_base=5)
    for k in ("ds", "ap", "hap", "prob"):
        assert np.array_equal(whole[k], np.concatenate([a[k], b[k]])), k


@pytest.mark.parametrize("logspace", [True, False])
def test_fb_config1_full(logspace):
    """Config 1 at full size: 4432 haplotypes x 1644 markers, 444 targets,
    L = 8, Ne = 1e4 (reference semantics: first interval's transitions), in
    the log-space FP32 form the workload names and in linear FP32."""
    w = synth.WORKLOADS["cfg1"]
    pan = synth.make_panel(w["H"], w["M"], seed=12345)
    tg = synth.make_targets(pan, w["B"], w["typed_every"], seed=12346)
    idx = device.DeviceIndex(pan.cols, pan.H, w["L"])
    p, d = oracle.pbwt_build(pan.alleles())
    assert np.array_equal(idx.prefix.cpu().numpy().view(np.uint32), p)
    h0, _, _, _ = oracle.state_table(pan.alleles(), p, w["L"])
    assert np.array_equal(idx.states.cpu().numpy().view(np.uint32), h0)
    tr = device.Transitions.build(pan.pos, None, w["ne"], per_marker=False)
    res = run_fb(idx, tg.codes, synth.HARDCALL_PALETTE, tr, logspace=logspace)
    check_fb(idx, tg.codes, synth.HARDCALL_PALETTE, tr, res, False, fold=False)


def test_fb_config3_shape_subset():
    """Config 3 shape (8000 haplotypes, heavy-tailed spacing with hotspots,
    per-marker transitions) on 20000 markers x 16 samples, Ne 1e3/1e4/1e5."""
    pan = synth.make_panel(8000, 20000, seed=33, spacing_mode=1)
    tg = synth.make_targets(pan, 16, 20, seed=34, geno_error=0.02)
    idx = device.DeviceIndex(pan.cols, pan.H, 8, keep_prefix=False)
    for ne in (1000, 10000, 100000):
        tr = device.Transitions.build(pan.pos, pan.genetic_cM(), ne, per_marker=True)
        for logspace in (False, True):
            res = run_fb(idx, tg.codes, synth.HARDCALL_PALETTE, tr, logspace=logspace)
            err = check_fb(idx, tg.codes, synth.HARDCALL_PALETTE, tr, res, True, fold=False)
            print(f"ne={ne} {'log' if logspace else 'linear'} max |DS err| = {err:.3g}")
