"""The C++ drop-in API (include/swiftimpute/imputer.hpp) through its test
driver tests/cpp/test_api.cpp: host-only checks on CPU; on the GPU, the
Imputer's PBWT arrays, selected states and posteriors against the oracle, and
its sampled haplotypes against the C-ABI path (itself replay-checked)."""
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle
from conftest import GOLDEN, ROOT, read_vcf_gt
from test_oracle import fixture_panel
from swiftimpute_amd import device, synth

EXE = os.path.join(ROOT, "swiftimpute_amd", "_build", "test_api")


def run_exe(*args, timeout=600):
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-2000:])
    return r


def test_cpp_api_host_checks():
    r = run_exe("cpu", GOLDEN)
    assert r.returncode == 0 and "PASS cpu checks" in r.stdout


@pytest.mark.gpu
def test_cpp_api_fixture(tmp_path):
    r = run_exe("fixture", GOLDEN, tmp_path)
    assert r.returncode == 0 and "PASS fixture" in r.stdout
    lines = [l for l in open(tmp_path / "fixture_out.vcf") if not l.startswith("#")]
    assert len(lines) == 10
    # the region run re-sweeps the PBWT from the identity at rs4 (marker 3)
    _, ref = fixture_panel()
    p, _ = oracle.pbwt_build(ref[3:8])
    rp = np.fromfile(tmp_path / "region_prefix.bin", np.uint32).reshape(5, 8)
    assert np.array_equal(rp, p)
    assert not np.array_equal(rp, oracle.pbwt_build(ref)[0][3:8])
    h0, _, _, _ = oracle.state_table(ref[3:8], p, 4)
    assert np.array_equal(np.fromfile(tmp_path / "region_states.bin", np.uint32).reshape(5, 4), h0)
    check_written_vcf(tmp_path / "fixture_out.vcf", tmp_path / "fixture_hap.bin")


def check_written_vcf(vcf, hap_bin):
    """The fixture run's VCF (default config: DS, GP, AP, INFO) against the FP64
    oracle on the fixtures: values written with 4 decimals, GT = the sampled
    haplotypes, INFO = 1 - Var(DS) / (2p(1-p)) over the posterior dosages."""
    pos, ref = fixture_panel()
    _, rows = read_vcf_gt(os.path.join(GOLDEN, "test_target.vcf"))
    al = np.array(rows, np.int32)                              # [M][B][2], 255 missing
    codes = np.where((al == 255).any(axis=2), 3, np.minimum(al, 1).sum(axis=2)).T.astype(np.uint8)
    gl = synth.HARDCALL_PALETTE[codes]                         # [B][M][3]
    p, _ = oracle.pbwt_build(ref)
    _, _, a0, a1 = oracle.state_table(ref, p, 4)
    Td, To = oracle.transitions(oracle.genetic_map(pos), 10000)
    fb = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=False)
    hap = np.fromfile(hap_bin, np.uint8).reshape(2, 2, 10)
    header = [l for l in open(vcf) if l.startswith("##")]
    for f in ("ID=DS", "ID=GP", "ID=AP", "ID=INFO"):
        assert any(f in l for l in header), f
    lines = [l.rstrip("\n").split("\t") for l in open(vcf) if not l.startswith("#")]
    assert len(lines) == 10
    q = 5e-5 + 1e-5  # half a unit of the 4th decimal, plus the dosage tolerance
    for m, cols in enumerate(lines):
        assert cols[8] == "GT:DS:GP:AP"
        info = float(cols[7].split("INFO=")[1])
        d = fb["ds"][:, m]
        mean = d.mean()
        pf = mean / 2
        want = 1.0 if pf in (0.0, 1.0) else min(1.0, max(0.0, 1 - (np.mean(d * d) - mean ** 2) / (2 * pf * (1 - pf))))
        assert abs(info - want) <= 1e-3, (m, info, want)
        for s in range(2):
            gt, ds, gp, ap = cols[9 + s].split(":")
            assert gt == f"{hap[s, 0, m]}|{hap[s, 1, m]}"
            assert abs(float(ds) - fb["ds"][s, m]) <= q
            apv = [float(x) for x in ap.split(",")]
            assert np.abs(np.array(apv) - fb["ap"][s, m]).max() <= q
            p0, p1 = fb["ap"][s, m]
            gpw = np.array([(1 - p0) * (1 - p1), (1 - p0) * p1 + p0 * (1 - p1), p0 * p1])
            assert np.abs(np.array([float(x) for x in gp.split(",")]) - gpw).max() <= 2 * q


def write_inputs(d, pan, tg, palette):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{pan.H} {pan.M} {tg.B}\n")
    pan.cols.tofile(os.path.join(d, "cols.bin"))
    pan.pos.astype(np.uint64).tofile(os.path.join(d, "pos.bin"))
    cm = pan.genetic_cM()
    (cm if cm is not None else np.zeros(pan.M)).astype(np.float64).tofile(os.path.join(d, "cm.bin"))
    np.ascontiguousarray(tg.codes[:, : pan.M]).tofile(os.path.join(d, "codes.bin"))
    palette.astype(np.float32).tofile(os.path.join(d, "palette.bin"))


def read_outputs(d, pan, B, L):
    M, H = pan.M, pan.H
    f = lambda n, t: np.fromfile(os.path.join(d, n), t)
    return dict(prefix=f("prefix.bin", np.uint32).reshape(M, H), div=f("div.bin", np.uint32).reshape(M, H),
                states=f("states.bin", np.uint32).reshape(M, L), hap=f("hap.bin", np.uint8).reshape(B, 2, M),
                ap=f("ap.bin", np.float32).reshape(B, M, 2), loglik=f("loglik.bin", np.float64))


@pytest.mark.gpu
@pytest.mark.parametrize("per_marker,spacing,L,batch", [(False, 0, 8, 100), (True, 1, 16, 7),
                                                         (True, 1, 5, 1000), (False, 1, 100, 9)])
def test_cpp_api_against_oracle(tmp_path, per_marker, spacing, L, batch):
    pan = synth.make_panel(1000, 600, seed=31, spacing_mode=spacing)
    tg = synth.make_targets(pan, 23, 4, seed=32)
    ne = 10000
    write_inputs(tmp_path / "in", pan, tg, synth.HARDCALL_PALETTE)
    os.makedirs(tmp_path / "out", exist_ok=True)
    r = run_exe("run", tmp_path / "in", tmp_path / "out", L, ne, int(per_marker), 0, batch, 99)
    assert r.returncode == 0, r.stdout
    o = read_outputs(tmp_path / "out", pan, tg.B, L)
    ref = pan.alleles()
    p, d = oracle.pbwt_build(ref)
    assert np.array_equal(o["prefix"], p) and np.array_equal(o["div"], d)
    h0, _, a0, a1 = oracle.state_table(ref, p, L)
    assert np.array_equal(o["states"], h0)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), ne, per_marker=per_marker)
    gl = synth.HARDCALL_PALETTE[tg.codes[:, : pan.M]]
    fb = oracle.fb_f64(a0, a1, gl, tr.Td, tr.To, per_marker=per_marker)
    assert np.abs(o["ap"] - fb["ap"]).max() <= 1e-5
    assert np.abs(o["loglik"] - fb["loglik"]).max() <= 1e-5 * np.abs(fb["loglik"]).max()
    # sampled haplotypes: identical to the C-ABI path on the same seed (whole
    # cohort in one job; the API's batches use the matching sample_base)
    idx = device.DeviceIndex(pan.cols, pan.H, L)
    job = device.Job(idx, torch.from_numpy(tg.codes).cuda(), synth.HARDCALL_PALETTE, tr, seed=99)
    job.alloc_outputs(ds=True, hap=True)
    device.run_jobs([job.c_job()])
    torch.cuda.synchronize()
    assert np.array_equal(job.out.hap.cpu().numpy(), o["hap"])


@pytest.mark.gpu
@pytest.mark.parametrize("H,L", [(600, 8), (12000, 16)])
def test_cpp_api_gpu_compat_selection(tmp_path, H, L):
    # SelectionMode::GpuCompat: the states the reference's GPU kernels select,
    # and the posteriors of the model over those states
    pan = synth.make_panel(H, 300, seed=51)
    tg = synth.make_targets(pan, 6, 4, seed=52)
    write_inputs(tmp_path / "in", pan, tg, synth.HARDCALL_PALETTE)
    os.makedirs(tmp_path / "out", exist_ok=True)
    r = run_exe("run", tmp_path / "in", tmp_path / "out", L, 10000, 1, 0, 100, 5, 1, 1)
    assert r.returncode == 0, r.stdout
    o = read_outputs(tmp_path / "out", pan, tg.B, L)
    ref = pan.alleles()
    p, d = oracle.pbwt_build(ref)
    h0, _, a0, a1 = oracle.state_table_gpu(ref, p, d, L)
    assert np.array_equal(o["states"], h0)
    assert not np.array_equal(h0, oracle.state_table(ref, p, L)[0])
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000, per_marker=True)
    gl = synth.HARDCALL_PALETTE[tg.codes[:, : pan.M]]
    fb = oracle.fb_f64(a0, a1, gl, tr.Td, tr.To, per_marker=True)
    assert np.abs(o["ap"] - fb["ap"]).max() <= 1e-5


@pytest.mark.gpu
def test_cpp_api_windowed(tmp_path):
    # window_size below M: the reference's windowed model (imputer.cpp:965-971)
    pan = synth.make_panel(500, 700, seed=61, spacing_mode=1)
    tg = synth.make_targets(pan, 9, 4, seed=62)
    write_inputs(tmp_path / "in", pan, tg, synth.HARDCALL_PALETTE)
    os.makedirs(tmp_path / "out", exist_ok=True)
    r = run_exe("run", tmp_path / "in", tmp_path / "out", 8, 10000, 1, 0, 4, 5, 1, 0, 150)
    assert r.returncode == 0, r.stdout
    o = read_outputs(tmp_path / "out", pan, tg.B, 8)
    ref = pan.alleles()
    p, _ = oracle.pbwt_build(ref)
    _, _, a0, a1 = oracle.state_table(ref, p, 8)
    tr = device.Transitions.build(pan.pos, pan.genetic_cM(), 10000, per_marker=True)
    gl = synth.HARDCALL_PALETTE[tg.codes[:, : pan.M]]
    fb = oracle.fb_f64_windowed(a0, a1, gl, tr.Td, tr.To, 150, per_marker=True)
    assert np.abs(o["ap"] - fb["ap"]).max() <= 1e-5
    assert np.abs(o["loglik"] - fb["loglik"]).max() <= 1e-5 * np.abs(fb["loglik"]).max()


@pytest.mark.gpu
def test_cpp_api_simple_selection(tmp_path):
    pan = synth.make_panel(200, 300, seed=41)
    tg = synth.make_targets(pan, 5, 4, seed=42)
    write_inputs(tmp_path / "in", pan, tg, synth.HARDCALL_PALETTE)
    os.makedirs(tmp_path / "out", exist_ok=True)
    r = run_exe("run", tmp_path / "in", tmp_path / "out", 8, 10000, 0, 1, 100, 1, 0)
    assert r.returncode == 0, r.stdout
    states = np.fromfile(tmp_path / "out" / "states.bin", np.uint32).reshape(300, 8)
    assert (states == np.arange(0, 16, 2)[None, :]).all()


@pytest.mark.gpu
def test_impute_vcf_end_to_end(tmp_path):
    # si_impute_vcf (the reference's free impute(ref_vcf, tgt_vcf, out_vcf,
    # config)) on synthetic VCF text: GT equals the C-ABI path's sampled
    # haplotypes, DS/AP the FP64 oracle to the 4 decimals written
    import ctypes

    from swiftimpute_amd import _native
    pan = synth.make_panel(600, 400, seed=71)
    tg = synth.make_targets(pan, 7, 4, seed=72)
    rv, tv, ov = (str(tmp_path / n) for n in ("p.vcf", "t.vcf", "o.vcf"))
    synth.write_panel_vcf(rv, pan)
    synth.write_targets_vcf(tv, pan, tg)
    lib = _native.api()
    o = _native.ImputeOpts()
    lib.si_impute_defaults(ctypes.byref(o))
    o.num_states, o.precision, o.seed = 8, 2, 5
    t = _native.ImputeTiming()
    rc = lib.si_impute_vcf(rv.encode(), tv.encode(), ov.encode(), ctypes.byref(o), ctypes.byref(t))
    assert rc == 0, lib.si_impute_last_error()
    assert (t.samples, t.markers) == (7, 400)
    assert t.untyped_genotypes == int((tg.codes[:, :400] == 3).sum())
    assert 0 < t.impute_s <= t.total_s
    lines = [l.rstrip("\n").split("\t") for l in open(ov) if not l.startswith("#")]
    assert len(lines) == 400 and lines[0][8] == "GT:DS:GP:AP"
    idx = device.DeviceIndex(pan.cols, pan.H, 8)
    tr = device.Transitions.build(pan.pos, None, 10000)
    job = device.Job(idx, torch.from_numpy(tg.codes).cuda(), synth.HARDCALL_PALETTE, tr, seed=5,
                     logspace=True)
    job.alloc_outputs(ds=True, ap=True, hap=True)
    device.run_jobs([job.c_job()])
    torch.cuda.synchronize()
    hap = job.out.hap.cpu().numpy()
    a0, a1 = idx.state_alleles()
    Td, To = oracle.transitions(oracle.genetic_map(pan.pos), 10000)
    fb = oracle.fb_f64_fast(a0, a1, synth.HARDCALL_PALETTE[tg.codes[:, :400]], Td, To)
    for m, cols in enumerate(lines):
        for s in range(7):
            gt, ds, _, ap = cols[9 + s].split(":")
            assert gt == f"{hap[s, 0, m]}|{hap[s, 1, m]}", (m, s)
            assert abs(float(ds) - fb["ds"][s, m]) <= 6e-5
            assert np.abs(np.array([float(x) for x in ap.split(",")]) - fb["ap"][s, m]).max() <= 6e-5
