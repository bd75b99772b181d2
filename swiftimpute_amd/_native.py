"""ctypes bindings for the native libraries.

``libsi_hip.so``   -- HIP kernels behind the C ABI of ``include/swiftimpute_hip.h``.
``libswiftimpute.so`` -- the C++ API and the synthetic data generator.

Both are built in-tree by ``swiftimpute_amd.build``.  There is no fallback:
if a library is missing, loading raises, and every device operation of this
package goes through these libraries.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

SI_OK = 0
SI_ERR_ARG = -1
SI_ERR_DEVICE = -2
SI_ERR_UNSUPPORTED = -3
SI_ERR_WORKSPACE = -4
SI_FB_PER_MARKER = 1
SI_FB_ARGMAX = 2
SI_FB_LOGSPACE = 4
SI_FB_UNTYPED = 8
SI_FB_BANDED = 16
SI_FB_PASS_CHAIN = 1
SI_FB_PASS_BLOCK = 2
SI_PERM_FIXED = 0x80000000
SI_FB_NOALLOC_SCRATCH_BYTES = 64 << 20


class StepCoef(ctypes.Structure):
    _fields_ = [("t_o", ctypes.c_float), ("u", ctypes.c_float), ("lto", ctypes.c_float),
                ("lu", ctypes.c_float)]


class FbVariant(ctypes.Structure):
    """Mirror of ``si_fb_variant``: lanes per sample in the chain and block
    passes and the checkpoint block length (markers)."""
    _fields_ = [("chain_lanes", ctypes.c_uint32), ("block_lanes", ctypes.c_uint32),
                ("ckpt_block", ctypes.c_uint32)]

    def as_tuple(self):
        return (self.chain_lanes, self.block_lanes, self.ckpt_block)


class FbJob(ctypes.Structure):
    """Mirror of ``si_fb_job`` (include/swiftimpute_hip.h)."""
    _fields_ = [
        ("n_samples", ctypes.c_uint32), ("n_markers", ctypes.c_uint32),
        ("n_states", ctypes.c_uint32), ("flags", ctypes.c_uint32),
        ("code_pitch", ctypes.c_uint32), ("n_palette", ctypes.c_uint32),
        ("rng_stream", ctypes.c_uint32), ("sample_base", ctypes.c_uint32),
        ("seed", ctypes.c_uint64), ("step", StepCoef),
        ("steps", _vp), ("sa", _vp), ("mmask", _vp), ("codes", _vp), ("eps", _vp),
        ("emax", _vp), ("palette", _vp), ("ds", _vp), ("ap", _vp), ("hap", _vp), ("prob", _vp),
        ("loglik", _vp), ("window", ctypes.c_uint32), ("untyped_code", ctypes.c_uint32),
    ]


class SynthPanelCfg(ctypes.Structure):
    _fields_ = [("n_haplotypes", ctypes.c_uint32), ("n_markers", ctypes.c_uint32),
                ("n_founders", ctypes.c_uint32), ("spacing_mode", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("mutation", ctypes.c_double),
                ("spacing_bp", ctypes.c_double), ("lognorm_median_bp", ctypes.c_double),
                ("lognorm_sigma", ctypes.c_double), ("hotspot_frac", ctypes.c_double),
                ("hotspot_mult", ctypes.c_double)]


class SynthTargetCfg(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_uint32), ("typed_every", ctypes.c_uint32),
                ("typed_offset", ctypes.c_uint32), ("code_pitch", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("het_lo", ctypes.c_double),
                ("het_hi", ctypes.c_double), ("geno_error", ctypes.c_double)]


# Every symbol include/swiftimpute_hip.h declares: (name, restype, argtypes).
HIP_SYMBOLS = [
    ("si_last_error", ctypes.c_char_p, []),
    ("si_abi_version", ctypes.c_int, []),
    ("si_padded_states", ctypes.c_uint32, [ctypes.c_uint32]),
    ("si_pbwt_workspace_bytes", ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32]),
    ("si_pbwt_sweep", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp,
                                     ctypes.c_size_t, _vp]),
    ("si_pbwt_chunked_workspace_bytes", ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32,
                                                          ctypes.c_uint32]),
    ("si_pbwt_sweep_chunked", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp,
                                             ctypes.c_uint32, _vp, ctypes.c_size_t, _vp]),
    ("si_select_perm", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, _u32p]),
    ("si_select_perm_gpu_compat", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, _u32p]),
    ("si_state_table", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, _u32p, _vp, _vp, _vp,
                                      _vp]),
    ("si_genetic_map", ctypes.c_int, [_u64p, _f64p, ctypes.c_uint32, _f64p]),
    ("si_transition_constants", ctypes.c_int, [_f64p, ctypes.c_uint32, ctypes.c_uint32,
                                               ctypes.c_float, _f32p, _f32p]),
    ("si_step_coefs", ctypes.c_int, [_f32p, _f32p, ctypes.c_uint32, ctypes.POINTER(StepCoef)]),
    ("si_emission_table", ctypes.c_int, [_f32p, ctypes.c_uint32, _f32p, _f32p]),
    ("si_fb_workspace_bytes", ctypes.c_size_t, [ctypes.POINTER(FbJob), ctypes.c_uint32]),
    ("si_fb_run", ctypes.c_int, [ctypes.POINTER(FbJob), ctypes.c_uint32, _vp, ctypes.c_size_t,
                                 _vp]),
    ("si_fb_run_passes", ctypes.c_int, [ctypes.POINTER(FbJob), ctypes.c_uint32, _vp,
                                        ctypes.c_size_t, ctypes.c_uint32, _vp]),
    ("si_fb_batch_create", ctypes.c_int, [ctypes.POINTER(FbJob), ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    ("si_fb_batch_reset", ctypes.c_int, [_vp, ctypes.POINTER(FbJob), ctypes.c_uint32]),
    ("si_fb_batch_workspace_bytes", ctypes.c_size_t, [_vp]),
    ("si_fb_batch_run", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]),
    ("si_fb_batch_destroy", ctypes.c_int, [_vp]),
    ("si_fb_default_variant", ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(FbVariant)]),
    ("si_fb_lane_candidates", ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                                ctypes.c_uint32]),
    ("si_fb_run_variant", ctypes.c_int, [ctypes.POINTER(FbJob), ctypes.c_uint32,
                                         ctypes.POINTER(FbVariant), _vp, ctypes.c_size_t,
                                         ctypes.c_uint32, _vp]),
    ("si_fb_run_noalloc", ctypes.c_int, [ctypes.POINTER(FbJob), ctypes.c_uint32,
                                         ctypes.POINTER(FbVariant), _vp]),
    ("si_fb_tune", ctypes.c_int, [ctypes.POINTER(FbJob), ctypes.c_uint32, _vp, ctypes.c_size_t,
                                  ctypes.c_uint32, _vp, ctypes.POINTER(FbVariant)]),
    ("si_fb_batch_n_classes", ctypes.c_uint32, [_vp]),
    ("si_fb_batch_class_info", ctypes.c_int, [_vp, ctypes.c_uint32, _u32p, _u32p,
                                              ctypes.POINTER(FbVariant),
                                              ctypes.POINTER(ctypes.c_int)]),
    ("si_fb_batch_set_variant", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(FbVariant)]),
    ("si_fb_batch_tune", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_uint32, _vp]),
    ("si_fb_tune_cache_clear", None, []),
    ("si_sample_uniform", ctypes.c_float, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_uint32]),
]

API_SYMBOLS = [
    ("si_synth_panel_defaults", None, [ctypes.POINTER(SynthPanelCfg), ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.c_uint64]),
    ("si_synth_panel", ctypes.c_int, [ctypes.POINTER(SynthPanelCfg), _u64p, _u64p, _f64p, _f64p,
                                      ctypes.c_int]),
    ("si_synth_target_defaults", None, [ctypes.POINTER(SynthTargetCfg), ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]),
    ("si_synth_targets", ctypes.c_int, [ctypes.POINTER(SynthTargetCfg), _u64p, ctypes.c_uint32,
                                        ctypes.c_uint32, _f64p, _u8p, _u8p, ctypes.c_int]),
    ("si_hardcall_palette", None, [_f32p]),
    ("si_impute_defaults", None, [ctypes.c_void_p]),
    ("si_impute_vcf", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    ("si_impute_last_error", ctypes.c_char_p, []),
]


class ImputeOpts(ctypes.Structure):
    """Mirror of ``si_impute_opts`` (include/swiftimpute/impute_c.h)."""
    _fields_ = [("num_states", ctypes.c_uint32), ("ne", ctypes.c_uint32),
                ("batch_size", ctypes.c_uint32), ("window_size", ctypes.c_uint32),
                ("precision", ctypes.c_uint32), ("per_marker", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class ImputeTiming(ctypes.Structure):
    """Mirror of ``si_impute_timing``."""
    _fields_ = [(n, ctypes.c_double) for n in ("reference_load_s", "target_load_s", "index_s",
                                              "impute_s", "write_s", "total_s")] + \
               [(n, ctypes.c_uint64) for n in ("samples", "markers", "untyped_genotypes")]

_hip = None
_api = None


class NativeError(RuntimeError):
    pass


def _load(name, symbols):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise NativeError(
            f"{path} is missing: build it with `python -m swiftimpute_amd.build` "
            "(there is no non-native fallback)")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    for sym, res, args in symbols:
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
    return lib


def hip():
    global _hip
    if _hip is None:
        _hip = _load("libsi_hip.so", HIP_SYMBOLS)
        if _hip.si_abi_version() != 1:
            raise NativeError("libsi_hip.so ABI version mismatch")
    return _hip


def api():
    global _api
    if _api is None:
        hip()
        _api = _load("libswiftimpute.so", API_SYMBOLS)
    return _api


def check(rc: int, what: str):
    if rc != 0:
        msg = hip().si_last_error()
        raise NativeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
