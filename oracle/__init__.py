"""CPU oracle for the SwiftImpute imputation hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this package, as the checker and as the timed CPU baseline.  The
product (``swiftimpute_amd``) never imports it.

The numerics live in ``si_oracle.cpp`` (C++, built by ``oracle/Makefile`` into
``libsi_oracle.so``); this module is a numpy-facing ctypes wrapper.  Each
wrapped function cites the reference file:line it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsi_oracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, no fast-math)."""
    cmd = ["make", "-C", _HERE]
    if force:
        cmd = ["make", "-B", "-C", _HERE]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_pbwt_build.argtypes = [_u8p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]
        L.or_pbwt_sweep_window.argtypes = [_u8p, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_uint32, _u32p, _u32p, _u32p]
        L.or_select_perm.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p]
        L.or_select_states_sort.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        L.or_state_table.argtypes = [_u8p, _u32p, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_int, _u32p, _u32p, _u8p, _u8p]
        for name in ("or_select_states_gpu_warp", "or_select_states_gpu_insert",
                     "or_select_states_gpu_block", "or_select_states_gpu"):
            getattr(L, name).argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        L.or_state_table_gpu.argtypes = [_u8p, _u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, _u32p, _u32p, _u8p, _u8p]
        L.or_genetic_map.argtypes = [_u64p, _f64p, ctypes.c_uint32, _f64p]
        L.or_genetic_map.restype = None
        L.or_transitions.argtypes = [_f64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                     _f32p, _f32p]
        L.or_transitions.restype = None
        L.or_fb_f64.argtypes = [_u8p, _u8p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                _f64p, _f64p, _f64p, _f64p, _f64p]
        L.or_fb_f64_windowed.argtypes = [_u8p, _u8p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                         ctypes.c_uint32, _f64p, _f64p, _f64p, _f64p]
        L.or_fb_f64_fast.argtypes = [_u8p, _u8p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                     _f64p, _f64p, _f64p, _f64p, ctypes.c_int]
        L.or_fb_f32_faithful.argtypes = [_u8p, _u8p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                         _f32p, _f32p]
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_philox4x32_10.restype = None
        L.or_sample_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32]
        L.or_sample_uniform.restype = ctypes.c_float
        L.or_sample_from_probs.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_float]
        L.or_sample_from_probs.restype = ctypes.c_uint32
        L.or_argmax.argtypes = [_f32p, ctypes.c_uint32]
        L.or_argmax.restype = ctypes.c_uint32
        L.or_sample_replay.argtypes = [_f32p, _u8p, _u8p, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                       ctypes.c_int, ctypes.c_uint32, _u8p]
        L.or_sample_replay.restype = None
        L.or_cpu_reference_impute.argtypes = [_u8p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_uint64, ctypes.c_int, _u8p]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(t)


def _chk(rc, name):
    if rc != 0:
        raise ValueError(f"{name} failed with status {rc}")


# --------------------------------------------------------------------------- A1
def pbwt_build(ref: np.ndarray):
    """PBWTBuilder::build (src/pbwt/pbwt_index.cpp:95-247). ref: [M][H] uint8."""
    ref = np.ascontiguousarray(ref, dtype=np.uint8)
    M, H = ref.shape
    prefix = np.empty((M, H), np.uint32)
    div = np.empty((M, H), np.uint32)
    _chk(lib().or_pbwt_build(_p(ref, _u8p), M, H, _p(prefix, _u32p), _p(div, _u32p)),
         "or_pbwt_build")
    return prefix, div


class PbwtStream:
    """or_pbwt_build over a panel fed window by window (pbwt_index.cpp:153-208):
    ``next(ref_window)`` returns the window's (prefix, div) rows, carrying the
    order into the next window, so full-size panels are checked without their
    [M][H] arrays on the host."""

    def __init__(self, H: int):
        self.H = H
        self.m = 0
        self.order = np.arange(H, dtype=np.uint32)

    def next(self, ref: np.ndarray, with_div: bool = True):
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        M, H = ref.shape
        assert H == self.H
        prefix = np.empty((M, H), np.uint32)
        div = np.empty((M, H), np.uint32) if with_div else None
        _chk(lib().or_pbwt_sweep_window(_p(ref, _u8p), M, H, self.m, _p(self.order, _u32p),
                                        _p(prefix, _u32p), _p(div, _u32p)),
             "or_pbwt_sweep_window")
        self.m += M
        return prefix, div


# --------------------------------------------------------------------------- A2/A4
def select_perm(H: int, L: int) -> np.ndarray:
    """libstdc++ std::sort permutation for H equal keys (pbwt_index.cpp:30-33)."""
    out = np.empty(min(L, H), np.uint32)
    _chk(lib().or_select_perm(H, L, _p(out, _u32p)), "or_select_perm")
    return out


def select_states_sort(prefix_row: np.ndarray, div_row: np.ndarray, L: int) -> np.ndarray:
    """PBWTIndex::select_states at one marker (pbwt_index.cpp:20-79)."""
    prefix_row = np.ascontiguousarray(prefix_row, np.uint32)
    div_row = np.ascontiguousarray(div_row, np.uint32)
    out = np.empty(L, np.uint32)
    _chk(lib().or_select_states_sort(_p(prefix_row, _u32p), _p(div_row, _u32p),
                                     prefix_row.size, L, _p(out, _u32p)), "or_select_states_sort")
    return out


def state_table(ref: np.ndarray, prefix: np.ndarray, L: int, simple_mode: bool = False):
    """States + pairing + alleles (pbwt_index.cpp:20-44; imputer.cpp:913-942).

    Returns (h0, h1, a0, a1), each [M][L]."""
    ref = np.ascontiguousarray(ref, np.uint8)
    prefix = np.ascontiguousarray(prefix, np.uint32)
    M, H = ref.shape
    h0 = np.empty((M, L), np.uint32)
    h1 = np.empty((M, L), np.uint32)
    a0 = np.empty((M, L), np.uint8)
    a1 = np.empty((M, L), np.uint8)
    _chk(lib().or_state_table(_p(ref, _u8p), _p(prefix, _u32p), M, H, L, int(simple_mode),
                              _p(h0, _u32p), _p(h1, _u32p), _p(a0, _u8p), _p(a1, _u8p)),
         "or_state_table")
    return h0, h1, a0, a1


def select_states_gpu(prefix_row: np.ndarray, div_row: np.ndarray, L: int,
                      kernel: str = "dispatch") -> np.ndarray:
    """The reference's GPU selection at one marker (pbwt_selector.cu:47-463):
    ``kernel`` is "warp", "insert", "block" or "dispatch" (launch_select_states'
    pick)."""
    prefix_row = np.ascontiguousarray(prefix_row, np.uint32)
    div_row = np.ascontiguousarray(div_row, np.uint32)
    out = np.empty(L, np.uint32)
    fn = {"warp": "or_select_states_gpu_warp", "insert": "or_select_states_gpu_insert",
          "block": "or_select_states_gpu_block", "dispatch": "or_select_states_gpu"}[kernel]
    _chk(getattr(lib(), fn)(_p(prefix_row, _u32p), _p(div_row, _u32p), prefix_row.size, L,
                            _p(out, _u32p)), fn)
    return out


def state_table_gpu(ref: np.ndarray, prefix: np.ndarray, div: np.ndarray, L: int):
    """State table from the reference's GPU selection over the PBWT arrays,
    pairs (h, (h+1) % H) as imputer.cpp:862-873.  Returns (h0, h1, a0, a1)."""
    ref = np.ascontiguousarray(ref, np.uint8)
    prefix = np.ascontiguousarray(prefix, np.uint32)
    div = np.ascontiguousarray(div, np.uint32)
    M, H = ref.shape
    h0 = np.empty((M, L), np.uint32)
    h1 = np.empty((M, L), np.uint32)
    a0 = np.empty((M, L), np.uint8)
    a1 = np.empty((M, L), np.uint8)
    _chk(lib().or_state_table_gpu(_p(ref, _u8p), _p(prefix, _u32p), _p(div, _u32p), M, H, L,
                                  _p(h0, _u32p), _p(h1, _u32p), _p(a0, _u8p), _p(a1, _u8p)),
         "or_state_table_gpu")
    return h0, h1, a0, a1


# --------------------------------------------------------------------------- A5
def genetic_map(pos: np.ndarray, cM: np.ndarray | None = None) -> np.ndarray:
    """Morgans per marker (imputer.cpp:596-605)."""
    pos = np.ascontiguousarray(pos, np.uint64)
    g = np.empty(pos.size, np.float64)
    cmv = None if cM is None else np.ascontiguousarray(cM, np.float64)
    lib().or_genetic_map(_p(pos, _u64p), _p(cmv, _f64p), pos.size, _p(g, _f64p))
    return g


def transitions(g: np.ndarray, ne: int, rho_rate: float = 4.0):
    """(Td, To) FP32 per interval (transition.cu:83-121)."""
    g = np.ascontiguousarray(g, np.float64)
    M = g.size
    n = max(M - 1, 0)
    Td = np.empty(max(n, 1), np.float32)
    To = np.empty(max(n, 1), np.float32)
    lib().or_transitions(_p(g, _f64p), M, int(ne), float(rho_rate), _p(Td, _f32p), _p(To, _f32p))
    return Td[:n], To[:n]


# --------------------------------------------------------------------------- A6-A9, A14
def fb_f64(a0, a1, gl, Td, To, per_marker=False, want_logpost=False):
    """FP64 fold-form forward-backward + sampler normalisation + dosage
    (forward_backward.cu:41-317, sampling.cu:77-100).  O(B*M*L^2).

    Returns dict(prob [B][M][L], ds [B][M], ap [B][M][2], loglik [B] (, logpost))."""
    a0 = np.ascontiguousarray(a0, np.uint8)
    a1 = np.ascontiguousarray(a1, np.uint8)
    gl = np.ascontiguousarray(gl, np.float32)
    B, M = gl.shape[0], gl.shape[1]
    L = a0.shape[1]
    Td = np.ascontiguousarray(Td if len(Td) else np.zeros(1), np.float32)
    To = np.ascontiguousarray(To if len(To) else np.zeros(1), np.float32)
    prob = np.empty((B, M, L), np.float64)
    lp = np.empty((B, M, L), np.float64) if want_logpost else None
    ds = np.empty((B, M), np.float64)
    ap = np.empty((B, M, 2), np.float64)
    ll = np.empty(B, np.float64)
    _chk(lib().or_fb_f64(_p(a0, _u8p), _p(a1, _u8p), _p(gl, _f32p), _p(Td, _f32p),
                         _p(To, _f32p), B, M, L, int(per_marker), _p(prob, _f64p), _p(lp, _f64p),
                         _p(ds, _f64p), _p(ap, _f64p), _p(ll, _f64p)), "or_fb_f64")
    out = dict(prob=prob, ds=ds, ap=ap, loglik=ll)
    if want_logpost:
        out["logpost"] = lp
    return out


def fb_f64_windowed(a0, a1, gl, Td, To, window, per_marker=False):
    """The reference's windowed forward-backward (imputer.cpp:1044-1174,
    forward_backward.cu:406-626) in FP64, with the sampler normalisation and
    dosage.  window = 0 or M <= window is fb_f64.  Returns dict(prob, ds, ap,
    loglik)."""
    a0 = np.ascontiguousarray(a0, np.uint8)
    a1 = np.ascontiguousarray(a1, np.uint8)
    gl = np.ascontiguousarray(gl, np.float32)
    B, M = gl.shape[0], gl.shape[1]
    L = a0.shape[1]
    Td = np.ascontiguousarray(Td if len(Td) else np.zeros(1), np.float32)
    To = np.ascontiguousarray(To if len(To) else np.zeros(1), np.float32)
    prob = np.empty((B, M, L), np.float64)
    ds = np.empty((B, M), np.float64)
    ap = np.empty((B, M, 2), np.float64)
    ll = np.empty(B, np.float64)
    _chk(lib().or_fb_f64_windowed(_p(a0, _u8p), _p(a1, _u8p), _p(gl, _f32p), _p(Td, _f32p),
                                  _p(To, _f32p), B, M, L, int(per_marker), int(window),
                                  _p(prob, _f64p), _p(ds, _f64p), _p(ap, _f64p), _p(ll, _f64p)),
         "or_fb_f64_windowed")
    return dict(prob=prob, ds=ds, ap=ap, loglik=ll)


def fb_f64_fast(a0, a1, gl, Td, To, per_marker=False, want_prob=True, nthreads=None):
    """FP64 rank-1-plus-diagonal form of fb_f64 (O(B*M*L)), threaded over samples."""
    a0 = np.ascontiguousarray(a0, np.uint8)
    a1 = np.ascontiguousarray(a1, np.uint8)
    gl = np.ascontiguousarray(gl, np.float32)
    B, M = gl.shape[0], gl.shape[1]
    L = a0.shape[1]
    Td = np.ascontiguousarray(Td if len(Td) else np.zeros(1), np.float32)
    To = np.ascontiguousarray(To if len(To) else np.zeros(1), np.float32)
    prob = np.empty((B, M, L), np.float64) if want_prob else None
    ds = np.empty((B, M), np.float64)
    ap = np.empty((B, M, 2), np.float64)
    ll = np.empty(B, np.float64)
    nt = nthreads or min(16, os.cpu_count() or 1)
    _chk(lib().or_fb_f64_fast(_p(a0, _u8p), _p(a1, _u8p), _p(gl, _f32p), _p(Td, _f32p),
                              _p(To, _f32p), B, M, L, int(per_marker), _p(prob, _f64p),
                              _p(ds, _f64p), _p(ap, _f64p), _p(ll, _f64p), nt), "or_fb_f64_fast")
    return dict(prob=prob, ds=ds, ap=ap, loglik=ll)


def fb_f32_faithful(a0, a1, gl, Td, To, per_marker=False):
    """The reference's FP32 operation order (drift demonstration only)."""
    a0 = np.ascontiguousarray(a0, np.uint8)
    a1 = np.ascontiguousarray(a1, np.uint8)
    gl = np.ascontiguousarray(gl, np.float32)
    B, M = gl.shape[0], gl.shape[1]
    L = a0.shape[1]
    Td = np.ascontiguousarray(Td if len(Td) else np.zeros(1), np.float32)
    To = np.ascontiguousarray(To if len(To) else np.zeros(1), np.float32)
    prob = np.empty((B, M, L), np.float32)
    ds = np.empty((B, M), np.float32)
    _chk(lib().or_fb_f32_faithful(_p(a0, _u8p), _p(a1, _u8p), _p(gl, _f32p), _p(Td, _f32p),
                                  _p(To, _f32p), B, M, L, int(per_marker), _p(prob, _f32p),
                                  _p(ds, _f32p)), "or_fb_f32_faithful")
    return dict(prob=prob, ds=ds)


# --------------------------------------------------------------------------- A11
def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().or_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(out, _u32p))
    return out


def sample_uniform(seed: int, s: int, m: int, stream: int = 0) -> float:
    return float(lib().or_sample_uniform(seed, s, m, stream))


def sample_from_probs(p: np.ndarray, u: float) -> int:
    p = np.ascontiguousarray(p, np.float32)
    return int(lib().or_sample_from_probs(_p(p, _f32p), p.size, u))


def argmax(p: np.ndarray) -> int:
    p = np.ascontiguousarray(p, np.float32)
    return int(lib().or_argmax(_p(p, _f32p), p.size))


def sample_replay(prob_f32, a0, a1, seed, stream=0, mode="philox", sample_base=0) -> np.ndarray:
    """Host replay of the device sampler from the device's FP32 probabilities."""
    prob_f32 = np.ascontiguousarray(prob_f32, np.float32)
    a0 = np.ascontiguousarray(a0, np.uint8)
    a1 = np.ascontiguousarray(a1, np.uint8)
    B, M, L = prob_f32.shape
    hap = np.empty((B, 2, M), np.uint8)
    lib().or_sample_replay(_p(prob_f32, _f32p), _p(a0, _u8p), _p(a1, _u8p), B, M, L, seed,
                           stream, 1 if mode == "argmax" else 0, sample_base, _p(hap, _u8p))
    return hap


def cpu_reference_impute(ref, gl, Td, To, L, seed, nthreads=1) -> np.ndarray:
    """The reference CPU path (PBWT + per-(s,m) std::sort + L^2 FP32 FB with
    sqrt(M) replay + sampling), used as the timed CPU baseline."""
    ref = np.ascontiguousarray(ref, np.uint8)
    gl = np.ascontiguousarray(gl, np.float32)
    M, H = ref.shape
    B = gl.shape[0]
    Td = np.ascontiguousarray(Td if len(Td) else np.zeros(1), np.float32)
    To = np.ascontiguousarray(To if len(To) else np.zeros(1), np.float32)
    hap = np.empty((B, 2, M), np.uint8)
    _chk(lib().or_cpu_reference_impute(_p(ref, _u8p), _p(gl, _f32p), _p(Td, _f32p),
                                       _p(To, _f32p), M, H, B, L, seed, nthreads, _p(hap, _u8p)),
         "or_cpu_reference_impute")
    return hap
