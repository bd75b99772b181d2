"""Serving form: Python front end of ``include/swiftimpute/server.h``.

The scheduler, admission control and batching run natively (libswiftimpute.so,
``swiftimpute_amd/csrc/api/server.cpp``); this module only marshals arrays.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native

_vp = ctypes.c_void_p


class ServerCfg(ctypes.Structure):
    _fields_ = [("n_states", ctypes.c_uint32), ("ne", ctypes.c_uint32), ("rho_rate", ctypes.c_float),
                ("per_marker", ctypes.c_uint32), ("max_streams", ctypes.c_uint32),
                ("max_batch", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64),
                ("device_id", ctypes.c_int), ("max_queue_delay_us", ctypes.c_uint32),
                ("selection", ctypes.c_uint32), ("region_resweep", ctypes.c_uint32),
                ("plain", ctypes.c_uint32), ("staging_bytes", ctypes.c_uint64),
                ("window_size", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint64) for n in ("requests", "batches", "genotypes", "untyped_genotypes",
                                                "max_batch_requests", "peak_bytes",
                                                "deferred_admissions", "direct_copies",
                                                "h2d_bytes", "d2h_bytes")] +
                [(n, ctypes.c_double) for n in ("launch_host_s", "h2d_s", "kernel_s", "d2h_s",
                                                "copy_host_s")])


_SYMBOLS = [
    ("si_server_defaults", None, [ctypes.POINTER(ServerCfg)]),
    ("si_server_create", ctypes.c_int, [ctypes.POINTER(ServerCfg), _native._u64p, ctypes.c_uint32,
                                        ctypes.c_uint32, _native._u64p, _native._f64p,
                                        ctypes.POINTER(_vp)]),
    ("si_server_submit", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        _native._u8p, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_uint64)]),
    ("si_server_submit_into", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             _native._u8p, ctypes.c_uint64, _native._u8p,
                                             _native._f32p, ctypes.POINTER(ctypes.c_uint64)]),
    ("si_server_wait", ctypes.c_int, [_vp, ctypes.c_uint64, _native._u8p, _native._f32p,
                                      ctypes.POINTER(ctypes.c_double)]),
    ("si_server_get_stats", ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    ("si_server_destroy", ctypes.c_int, [_vp]),
    ("si_server_last_error", ctypes.c_char_p, []),
]

_bound = False


def _lib():
    global _bound
    lib = _native.api()
    if not _bound:
        for name, res, args in _SYMBOLS:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return lib


def _check(rc, what):
    if rc != 0:
        raise _native.NativeError(f"{what} failed ({rc}): {_lib().si_server_last_error().decode()}")


@dataclass
class Result:
    hap: np.ndarray      # [B][2][M] uint8
    ds: np.ndarray       # [B][M] float32
    latency_s: float


class Server:
    """Resident panel + continuous batching of (cohort, region) requests."""

    def __init__(self, cols: np.ndarray, H: int, pos: np.ndarray, cM: np.ndarray | None = None,
                 L: int = 8, ne: int = 10000, per_marker: bool = False, max_streams: int = 4,
                 max_batch: int = 256, hbm_budget_gb: float = 32.0, rho_rate: float = 4.0,
                 max_queue_delay_us: int = 0, gpu_compat_selection: bool = False,
                 region_resweep: bool = False, plain: bool = False, staging_gb: float = 4.0,
                 window_size: int = 0):
        lib = _lib()
        cfg = ServerCfg()
        lib.si_server_defaults(ctypes.byref(cfg))
        cfg.n_states, cfg.ne, cfg.per_marker = L, ne, int(per_marker)
        cfg.max_streams, cfg.max_batch = max_streams, max_batch
        cfg.hbm_budget_bytes = int(hbm_budget_gb * (1 << 30))
        cfg.rho_rate = rho_rate
        cfg.max_queue_delay_us = max_queue_delay_us
        cfg.selection = int(gpu_compat_selection)
        cfg.region_resweep = int(region_resweep)
        cfg.plain = int(plain)
        cfg.staging_bytes = int(staging_gb * (1 << 30))
        cfg.window_size = window_size
        cols = np.ascontiguousarray(cols, np.uint64)
        pos = np.ascontiguousarray(pos, np.uint64)
        cmv = None if cM is None else np.ascontiguousarray(cM, np.float64)
        h = _vp()
        _check(lib.si_server_create(ctypes.byref(cfg), cols.ctypes.data_as(_native._u64p), H,
                                    cols.shape[0], pos.ctypes.data_as(_native._u64p),
                                    cmv.ctypes.data_as(_native._f64p) if cmv is not None else None,
                                    ctypes.byref(h)), "si_server_create")
        self.h = h
        self.M = cols.shape[0]
        self._shapes = {}

    def submit(self, m0: int, m1: int, codes: np.ndarray, seed: int = 12345) -> int:
        codes = np.ascontiguousarray(codes, np.uint8)
        B, Mr = codes.shape
        assert Mr == m1 - m0
        rid = ctypes.c_uint64()
        _check(_lib().si_server_submit(self.h, m0, m1, B, codes.ctypes.data_as(_native._u8p), seed,
                                       ctypes.byref(rid)), "si_server_submit")
        self._shapes[rid.value] = (B, Mr)
        return rid.value

    def submit_into(self, m0: int, m1: int, codes: np.ndarray, hap: np.ndarray, ds: np.ndarray,
                    seed: int = 12345) -> int:
        """Submit with results written into the given arrays (hap [B][2][M]
        uint8, ds [B][M] float32, C-contiguous), which must stay alive until
        wait() returns for this request."""
        codes = np.ascontiguousarray(codes, np.uint8)
        B, Mr = codes.shape
        assert Mr == m1 - m0
        assert hap.shape == (B, 2, Mr) and hap.dtype == np.uint8 and hap.flags.c_contiguous
        assert ds.shape == (B, Mr) and ds.dtype == np.float32 and ds.flags.c_contiguous
        rid = ctypes.c_uint64()
        _check(_lib().si_server_submit_into(self.h, m0, m1, B, codes.ctypes.data_as(_native._u8p),
                                            seed, hap.ctypes.data_as(_native._u8p),
                                            ds.ctypes.data_as(_native._f32p), ctypes.byref(rid)),
               "si_server_submit_into")
        self._shapes[rid.value] = (hap, ds)
        return rid.value

    def wait(self, rid: int) -> Result:
        got = self._shapes.get(rid)
        if got is not None and isinstance(got[0], np.ndarray):  # submitted into caller arrays
            self._shapes.pop(rid)
            lat = ctypes.c_double()
            _check(_lib().si_server_wait(self.h, rid, None, None, ctypes.byref(lat)),
                   "si_server_wait")
            return Result(got[0], got[1], lat.value)
        B, Mr = self._shapes.pop(rid)
        hap = np.empty((B, 2, Mr), np.uint8)
        ds = np.empty((B, Mr), np.float32)
        lat = ctypes.c_double()
        _check(_lib().si_server_wait(self.h, rid, hap.ctypes.data_as(_native._u8p),
                                     ds.ctypes.data_as(_native._f32p), ctypes.byref(lat)),
               "si_server_wait")
        return Result(hap, ds, lat.value)

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib().si_server_get_stats(self.h, ctypes.byref(st)), "si_server_get_stats")
        return {n: (float if t is ctypes.c_double else int)(getattr(st, n)) for n, t in ServerStats._fields_}

    def close(self):
        if self.h:
            _lib().si_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
