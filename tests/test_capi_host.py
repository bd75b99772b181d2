"""CPU checks of the C ABI: the library loads, exports every declared symbol,
its host-side functions agree with the oracle bit for bit, and argument
errors come back as status codes (no GPU call is made by any of these)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from swiftimpute_amd import _native, device, synth
from conftest import ROOT


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(si_[a-z0-9_]+)\s*\(", txt)))


def test_hip_library_exports_every_declared_symbol():
    lib = _native.hip()
    names = declared("swiftimpute_hip.h")
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert {n for n, _, _ in _native.HIP_SYMBOLS} == set(names)


def test_api_library_exports_synth_symbols():
    lib = _native.api()
    for n in declared(os.path.join("swiftimpute", "synth.h")):
        assert hasattr(lib, n), n


def test_api_library_exports_impute_symbols():
    lib = _native.api()
    names = declared(os.path.join("swiftimpute", "impute_c.h"))
    assert set(names) == {"si_impute_defaults", "si_impute_vcf", "si_impute_last_error"}
    for n in names:
        assert hasattr(lib, n), n
    o = _native.ImputeOpts()
    lib.si_impute_defaults(ctypes.byref(o))
    # the reference's defaults (imputer.hpp:44-55, types.hpp:77-90, sampling.hpp:74)
    assert (o.num_states, o.ne, o.batch_size, o.window_size, o.precision, o.seed) == \
        (8, 10000, 100, 10000, 0, 12345)
    # a missing argument is refused before anything is loaded
    assert lib.si_impute_vcf(None, b"t.vcf", b"o.vcf", ctypes.byref(o), None) == _native.SI_ERR_ARG
    assert b"bad arguments" in lib.si_impute_last_error()


def test_api_library_exports_server_symbols():
    from swiftimpute_amd import serving
    lib = serving._lib()
    names = declared(os.path.join("swiftimpute", "server.h"))
    assert {n for n, _, _ in serving._SYMBOLS} == set(names)
    cfg = serving.ServerCfg()
    lib.si_server_defaults(ctypes.byref(cfg))
    assert (cfg.n_states, cfg.ne, cfg.max_streams, cfg.max_queue_delay_us) == (8, 10000, 4, 0)
    assert abs(cfg.rho_rate - 4.0) < 1e-7
    h = ctypes.c_void_p()
    one = np.zeros(1, np.uint64)
    # H = 0 is refused before any device call
    assert lib.si_server_create(ctypes.byref(cfg), one.ctypes.data_as(_native._u64p), 0, 1,
                                one.ctypes.data_as(_native._u64p), None, ctypes.byref(h)) != 0
    assert b"bad arguments" in lib.si_server_last_error()
    assert lib.si_server_submit(None, 0, 1, 1, None, 0, None) != 0


def test_pbwt_argument_errors():
    lib = _native.hip()
    one = ctypes.c_void_p(8)  # never dereferenced: the checks come first
    assert lib.si_pbwt_sweep(one, 0, 5, one, None, None, 0, None) == _native.SI_ERR_ARG
    assert lib.si_pbwt_sweep(one, 524289, 5, one, None, None, 0, None) == _native.SI_ERR_UNSUPPORTED
    assert b"524288" in lib.si_last_error()
    # above 65536 haplotypes the global-memory form with 512-marker chunks
    assert lib.si_pbwt_workspace_bytes(100000, 1000) == lib.si_pbwt_chunked_workspace_bytes(100000, 1000, 512)
    assert lib.si_pbwt_sweep_chunked(one, 10, 5, one, None, 0, None, 0, None) == _native.SI_ERR_ARG
    assert lib.si_pbwt_sweep(None, 10, 5, one, None, None, 0, None) == _native.SI_ERR_ARG
    assert lib.si_pbwt_sweep(one, 10, 0, one, None, None, 0, None) == 0  # nothing to do
    # workspace: three phases' buffers, growing with the number of chunks
    w512 = lib.si_pbwt_chunked_workspace_bytes(20000, 100000, 512)
    w64 = lib.si_pbwt_chunked_workspace_bytes(20000, 100000, 64)
    assert w64 > w512 >= 196 * 3 * 20000 * 4
    # default chunk: one round of chunks on 7/8 of the compute units (256 when
    # no device answers, as here: 224 slots), rounded up to a multiple of 128
    # above 64 markers; 512 in the partition form
    def ws(H, M, C):
        return lib.si_pbwt_chunked_workspace_bytes(H, M, C)
    assert lib.si_pbwt_workspace_bytes(20000, 100000) == ws(20000, 100000, 512)   # 447 -> 512
    assert lib.si_pbwt_workspace_bytes(20000, 200000) == ws(20000, 200000, 896)   # 893 -> 896
    assert lib.si_pbwt_workspace_bytes(20000, 10000) == ws(20000, 10000, 45)
    assert lib.si_pbwt_workspace_bytes(20000, 1000) == ws(20000, 1000, 16)
    assert lib.si_pbwt_workspace_bytes(40000, 100000) == lib.si_pbwt_chunked_workspace_bytes(40000, 100000, 512)


def test_abi_version_and_padding():
    lib = _native.hip()
    assert lib.si_abi_version() == 1
    assert [lib.si_padded_states(L) for L in (1, 4, 5, 8, 9, 16, 17, 33, 64, 65, 128, 129, 256)] == \
        [4, 4, 8, 8, 16, 16, 32, 64, 64, 128, 128, 256, 256]


@pytest.mark.parametrize("H,L", [(1, 8), (16, 16), (4432, 8), (20000, 16), (8000, 8), (40000, 8)])
def test_select_perm_matches_oracle(H, L):
    assert np.array_equal(device.select_perm(H, L), oracle.select_perm(H, L))


def _gather(prefix_row, perm, L):
    H = prefix_row.size
    out = []
    for k in range(L):
        e = int(perm[k]) if k < H else (_native.SI_PERM_FIXED | (k % H))
        out.append(e & 0x7FFFFFFF if e & _native.SI_PERM_FIXED else int(prefix_row[e]))
    return out


@pytest.mark.parametrize("H,L", [(128, 1), (128, 8), (200, 32), (4432, 8), (10000, 16), (10001, 8),
                                 (20000, 16), (300, 33), (300, 64), (40, 8), (70, 12), (5, 8),
                                 (1, 3), (300, 65), (20000, 100), (90, 128), (70000, 256)])
def test_gpu_compat_perm_matches_the_reference_kernels(H, L):
    # the permutation gathered from each PBWT row equals the oracle's lane-by-lane
    # run of the reference's GPU selection kernels on that row and its divergences
    rng = np.random.default_rng(H * 7 + L)
    ref = (rng.random((4, H)) < 0.45).astype(np.uint8)
    prefix, div = oracle.pbwt_build(ref)
    perm = device.select_perm(H, L, gpu_compat=True)
    assert perm.size == min(L, H)
    for m in range(ref.shape[0]):
        assert _gather(prefix[m], perm, L) == oracle.select_states_gpu(prefix[m], div[m], L).tolist()


@pytest.mark.parametrize("spacing_mode", [0, 1])
@pytest.mark.parametrize("ne", [1000, 10000, 20000, 100000])
def test_transition_constants_match_oracle(spacing_mode, ne):
    p = synth.make_panel(64, 300, seed=3, spacing_mode=spacing_mode)
    g1 = device.genetic_map(p.pos, p.genetic_cM())
    g2 = oracle.genetic_map(p.pos, p.genetic_cM())
    assert np.array_equal(g1, g2)
    Td1, To1 = device.transition_constants(g1, ne)
    Td2, To2 = oracle.transitions(g2, ne)
    assert np.array_equal(Td1, Td2) and np.array_equal(To1, To2)


def test_step_coefs():
    Td = np.array([-0.1, -4.3, 0.0, -2.0], np.float32)
    To = np.array([-5.0, -4.3, -999.0, -3.0], np.float32)
    c = device.step_coefs(Td, To)
    assert np.allclose(c[:, 0], np.exp(To.astype(np.float64)), rtol=1e-7)
    assert np.allclose(c[:, 1], np.exp(Td.astype(np.float64)) - np.exp(To.astype(np.float64)),
                       rtol=1e-6)
    assert c[1, 1] == 0.0 and c[1, 3] == -np.inf  # Td == To: no diagonal excess
    assert c[2, 0] == 0.0 and c[2, 1] == 1.0


def test_emission_table_hardcalls():
    eps, emax = device.emission_table(synth.HARDCALL_PALETTE)
    # hom-ref call with every class present: (1, 0, 0), offset 0
    assert eps[0, 7, :3].tolist() == [1.0, 0.0, 0.0] and emax[0, 7] == 0.0
    # hom-ref call where only het/hom-alt states exist: all mismatch, offset -999
    assert eps[0, 6, :3].tolist() == [0.0, 1.0, 1.0] and emax[0, 6] == -999.0
    # missing: uniform
    assert eps[3, 7, :3].tolist() == [1.0, 1.0, 1.0]


def test_sample_uniform_matches_oracle():
    lib = _native.hip()
    for seed in (0, 12345, 2**40 + 7):
        for s in (0, 3, 9999):
            for m in (0, 1, 2, 3, 4, 99999):
                assert lib.si_sample_uniform(seed, s, m, 0) == oracle.sample_uniform(seed, s, m, 0)


def test_synth_is_thread_count_independent():
    a = synth.make_panel(300, 500, seed=9, spacing_mode=1, threads=1)
    b = synth.make_panel(300, 500, seed=9, spacing_mode=1, threads=8)
    assert np.array_equal(a.cols, b.cols) and np.array_equal(a.pos, b.pos)
    ta = synth.make_targets(a, 7, 4, threads=1)
    tb = synth.make_targets(a, 7, 4, threads=8)
    assert np.array_equal(ta.codes, tb.codes) and np.array_equal(ta.truth, tb.truth)


def test_synth_targets_typed_pattern_and_truth():
    p = synth.make_panel(200, 400, seed=1)
    t = synth.make_targets(p, 20, 4, geno_error=0.0)
    typed = np.arange(400) % 4 == 0
    assert (t.codes[:, :400][:, ~typed] == 3).all()
    g = t.truth[:, 0, :] + t.truth[:, 1, :]
    assert np.array_equal(t.codes[:, :400][:, typed], g[:, typed])


def _job(**kw):
    j = _native.FbJob()
    j.n_samples, j.n_markers, j.n_states, j.code_pitch, j.n_palette = 1, 16, 8, 16, 4
    for k, v in kw.items():
        setattr(j, k, v)
    return j


@pytest.mark.parametrize("kw,code", [
    (dict(n_states=0), -1), (dict(n_states=257), -1), (dict(n_palette=0), -3),
    (dict(n_palette=65), -3), (dict(code_pitch=8), -1), (dict(code_pitch=17), -1),
    (dict(flags=4), -1),  # log-space needs the palette
])
def test_fb_argument_errors(kw, code):
    lib = _native.hip()
    j = _job(**kw)
    rc = lib.si_fb_run(ctypes.byref(j), 1, None, 0, None)
    assert rc == code
    assert lib.si_last_error()


def test_fb_missing_inputs_error():
    lib = _native.hip()
    j = _job()
    assert lib.si_fb_run(ctypes.byref(j), 1, None, 0, None) == -1  # sa/codes missing
    assert b"required" in lib.si_last_error()


def test_state_table_argument_errors():
    lib = _native.hip()
    perm = np.zeros(8, np.uint32)
    p32 = perm.ctypes.data_as(_native._u32p)
    assert lib.si_state_table(None, None, None, 0, 10, 8, 0, p32, None, None, None, None) == -1
    assert lib.si_state_table(None, None, None, 10, 10, 8, 1, p32, None, None, None, None) == -1
    assert lib.si_state_table(None, None, None, 10, 10, 8, 0, None, None, None, None, None) == -1
    fake = ctypes.c_void_p(256)  # never dereferenced: the perm check fails first
    for bad in (10, _native.SI_PERM_FIXED | 10, 0xFFFFFFFF):
        perm[3] = bad
        assert lib.si_state_table(fake, fake, None, 10, 10, 8, 0, p32, None, fake, fake, None) == -1
        assert b"outside the panel" in lib.si_last_error()
    q = np.zeros(4, np.uint32)
    assert lib.si_select_perm_gpu_compat(0, 8, q.ctypes.data_as(_native._u32p)) == -1
    assert lib.si_select_perm_gpu_compat(10, 257, q.ctypes.data_as(_native._u32p)) == -1
    assert lib.si_state_table(fake, fake, None, 10, 10, 257, 0, p32, None, fake, fake, None) == -1


def test_short_window_is_refused():
    lib = _native.hip()
    fake = ctypes.c_void_p(256)  # never dereferenced: the job check fails first
    j = _native.FbJob()
    j.n_samples, j.n_markers, j.n_states, j.code_pitch, j.n_palette = 2, 100, 8, 112, 4
    j.sa = j.mmask = j.codes = j.eps = j.emax = j.ds = fake
    j.step = _native.StepCoef(1e-3, 0.9, -6.9, -0.1)
    j.window = 20
    arr = (_native.FbJob * 1)(j)
    assert lib.si_fb_run(arr, 1, None, 0, None) == _native.SI_ERR_UNSUPPORTED
    assert b"window 20 shorter than 32" in lib.si_last_error()
    # the workspace grows by one vector per window and sample
    arr[0].window = 0
    base = lib.si_fb_workspace_bytes(arr, 1)
    arr[0].window = 32
    assert lib.si_fb_workspace_bytes(arr, 1) > base
    arr[0].window = 100  # M <= window: no windows
    assert lib.si_fb_workspace_bytes(arr, 1) == base


def test_variant_queries():
    lib = _native.hip()
    v = _native.FbVariant()
    assert lib.si_fb_default_variant(16, ctypes.byref(v)) == 0
    assert v.as_tuple() == (4, 4, 16)
    assert lib.si_fb_default_variant(0, ctypes.byref(v)) == _native.SI_ERR_ARG
    assert lib.si_fb_default_variant(257, ctypes.byref(v)) == _native.SI_ERR_ARG
    assert lib.si_fb_default_variant(200, ctypes.byref(v)) == 0
    assert v.as_tuple() == (64, 64, 16)
    assert lib.si_fb_default_variant(100, ctypes.byref(v)) == 0
    assert v.as_tuple() == (32, 32, 16)
    assert device.lane_candidates(16, _native.SI_FB_PASS_CHAIN) == [2, 4, 8]
    assert device.lane_candidates(16, _native.SI_FB_PASS_BLOCK) == [4, 8]
    assert device.lane_candidates(64, _native.SI_FB_PASS_CHAIN) == [8, 16, 32]
    assert device.lane_candidates(128, _native.SI_FB_PASS_CHAIN) == [16, 32]
    assert device.lane_candidates(128, _native.SI_FB_PASS_BLOCK) == [32]
    assert device.lane_candidates(256, _native.SI_FB_PASS_CHAIN) == [32, 64]
    assert device.lane_candidates(256, _native.SI_FB_PASS_BLOCK) == [64]
    assert device.lane_candidates(3, _native.SI_FB_PASS_BLOCK) == [1, 2]
    assert lib.si_fb_lane_candidates(16, 3, None, 0) == 0  # not a single pass
    for L in (1, 4, 5, 8, 9, 16, 17, 32, 33, 64, 65, 128, 129, 256):
        vs = device.variants(L)
        assert device.default_variant(L) in vs
        assert {cb for *_, cb in vs} == {16}


def test_fb_variant_argument_errors():
    lib = _native.hip()
    j = _job()
    bad = _native.FbVariant(3, 2, 16)
    assert lib.si_fb_run_variant(ctypes.byref(j), 1, ctypes.byref(bad), None, 0, 3, None) == -1
    out = _native.FbVariant()
    assert lib.si_fb_tune(None, 0, None, 0, 0, None, ctypes.byref(out)) == _native.SI_ERR_ARG
    assert lib.si_fb_batch_n_classes(None) == 0
    assert lib.si_fb_batch_set_variant(None, 0, ctypes.byref(out)) == _native.SI_ERR_ARG
    assert lib.si_fb_batch_tune(None, None, 0, 0, None) == _native.SI_ERR_ARG
