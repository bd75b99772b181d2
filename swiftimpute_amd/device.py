"""Python driver of the C ABI: device buffers (torch allocations), the resident
panel index, and forward-backward jobs.

This mirrors what the C++ ``Imputer`` does (include/swiftimpute/imputer.hpp):
``DeviceIndex`` is ``Imputer::build_index()`` (PBWT sweep + state table on the
GPU), ``run_job`` is the per-batch ``impute`` step.  torch only provides
memory and the stream; all compute is in libsi_hip.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native
from ._native import SI_FB_LOGSPACE, FbJob, FbVariant, StepCoef, check, hip

DEV = "cuda"


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def padded_states(L: int) -> int:
    return int(hip().si_padded_states(L))


def select_perm(H: int, L: int, gpu_compat: bool = False) -> np.ndarray:
    """Selection permutation of the CPU path (std::sort over equal keys), or
    with ``gpu_compat`` the one the reference's GPU selection kernels make
    (entries flagged SI_PERM_FIXED name a haplotype directly)."""
    out = np.empty(min(L, H), np.uint32)
    name = "si_select_perm_gpu_compat" if gpu_compat else "si_select_perm"
    check(getattr(hip(), name)(H, L, out.ctypes.data_as(_native._u32p)), name)
    return out


# ------------------------------------------------------------------ transitions
def genetic_map(pos: np.ndarray, cM: np.ndarray | None = None) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.uint64)
    g = np.empty(pos.size, np.float64)
    cmv = None if cM is None else np.ascontiguousarray(cM, np.float64)
    check(hip().si_genetic_map(pos.ctypes.data_as(_native._u64p),
                               cmv.ctypes.data_as(_native._f64p) if cmv is not None else None,
                               pos.size, g.ctypes.data_as(_native._f64p)), "si_genetic_map")
    return g


def transition_constants(g: np.ndarray, ne: int, rho_rate: float = 4.0):
    g = np.ascontiguousarray(g, np.float64)
    n = max(g.size - 1, 0)
    Td = np.zeros(max(n, 1), np.float32)
    To = np.zeros(max(n, 1), np.float32)
    check(hip().si_transition_constants(g.ctypes.data_as(_native._f64p), g.size, ne, rho_rate,
                                        Td.ctypes.data_as(_native._f32p),
                                        To.ctypes.data_as(_native._f32p)),
          "si_transition_constants")
    return Td[:n], To[:n]


def step_coefs(Td: np.ndarray, To: np.ndarray) -> np.ndarray:
    """[n][4] float32 (t_o, u, lto, lu) per interval."""
    n = len(Td)
    arr = (StepCoef * max(n, 1))()
    Td = np.ascontiguousarray(Td, np.float32)
    To = np.ascontiguousarray(To, np.float32)
    check(hip().si_step_coefs(Td.ctypes.data_as(_native._f32p), To.ctypes.data_as(_native._f32p),
                              n, arr), "si_step_coefs")
    return np.frombuffer(arr, np.float32).reshape(-1, 4)[:n].copy()


def untyped_code(palette: np.ndarray):
    """The first palette code whose likelihoods are all 0 (an untyped or
    missing genotype: every state's emission is 1), or None."""
    z = np.flatnonzero((np.asarray(palette, np.float32) == 0.0).all(axis=1))
    return int(z[0]) if len(z) else None


def emission_table(palette: np.ndarray):
    palette = np.ascontiguousarray(palette, np.float32)
    P = palette.shape[0]
    eps = np.empty((P, 8, 4), np.float32)
    emax = np.empty((P, 8), np.float32)
    check(hip().si_emission_table(palette.ctypes.data_as(_native._f32p), P,
                                  eps.ctypes.data_as(_native._f32p),
                                  emax.ctypes.data_as(_native._f32p)), "si_emission_table")
    return eps, emax


@dataclass
class Transitions:
    """Per-step coefficients for one job.  per_marker=False reproduces the
    reference, whose kernels read only the first interval's matrix."""
    per_marker: bool
    step: np.ndarray                 # [4] coefficient used for every step
    steps_dev: torch.Tensor | None   # [M][4] when per_marker
    Td: np.ndarray
    To: np.ndarray
    # what the constants were built from: (pos, cM, ne, rho_rate), or None
    # when a caller supplied Td/To directly
    source: tuple | None = None

    @staticmethod
    def build(pos, cM, ne: int, rho_rate: float = 4.0, per_marker: bool = False,
              device=DEV) -> "Transitions":
        g = genetic_map(pos, cM)
        Td, To = transition_constants(g, ne, rho_rate)
        co = step_coefs(Td, To)
        step = co[0] if len(co) else np.array([1.0, 0.0, 0.0, -np.inf], np.float32)
        steps_dev = None
        if per_marker:
            full = np.zeros((len(g), 4), np.float32)
            full[1:] = co
            steps_dev = torch.from_numpy(full).to(device)
        return Transitions(per_marker, step, steps_dev, Td, To, (pos, cM, ne, rho_rate))

    def min_switch(self) -> float:
        """Smallest e^To over the steps the kernel will use."""
        if len(self.To) == 0:
            return 1.0
        to = self.To if self.per_marker else self.To[:1]
        return float(np.exp(to.astype(np.float64)).min())

    def floor_scale(self) -> bool:
        """Whether a switch probability the kernel uses is at the -999 floor's
        scale (below 2^-30, coincident markers): a log-space job then carries
        the floor exactly (SI_FB_BANDED)."""
        return self.min_switch() < 2.0 ** -30


# --------------------------------------------------------------------- index
def pbwt_sweep(cols: np.ndarray, H: int, chunk: int | None = None, with_div: bool = True,
               stream=None, device=DEV):
    """PBWT prefix (and divergence) rows [M][H] int32 on the device from a
    bit-packed panel; chunk=None uses the library's default chunk length."""
    M = cols.shape[0]
    c = torch.from_numpy(np.ascontiguousarray(cols).view(np.int64)).to(device)
    prefix = torch.empty((M, H), dtype=torch.int32, device=device)
    div = torch.empty((M, H), dtype=torch.int32, device=device) if with_div else None
    sh = _stream_handle(stream)
    if chunk is None:
        check(hip().si_pbwt_sweep(_ptr(c), H, M, _ptr(prefix), _ptr(div), None, 0, sh),
              "si_pbwt_sweep")
    else:
        check(hip().si_pbwt_sweep_chunked(_ptr(c), H, M, _ptr(prefix), _ptr(div), chunk, None, 0,
                                          sh), "si_pbwt_sweep_chunked")
    return prefix, div


class DeviceIndex:
    """Resident panel, PBWT index and state table (Imputer::build_index)."""

    def __init__(self, cols: np.ndarray, H: int, L: int, mode: int = 0, miss: np.ndarray | None = None,
                 keep_prefix: bool = True, with_div: bool = True, stream=None, device=DEV,
                 gpu_compat: bool = False):
        M = cols.shape[0]
        self.H, self.M, self.L, self.Lp = H, M, L, padded_states(L)
        self.cols = torch.from_numpy(np.ascontiguousarray(cols).view(np.int64)).to(device)
        self.miss = (torch.from_numpy(np.ascontiguousarray(miss).view(np.int64)).to(device)
                     if miss is not None else None)
        sh = _stream_handle(stream)
        lib = hip()
        self.prefix = torch.empty((M, H), dtype=torch.int32, device=device)
        self.div = torch.empty((M, H), dtype=torch.int32, device=device) if with_div else None
        if mode == 0:
            check(lib.si_pbwt_sweep(_ptr(self.cols), H, M, _ptr(self.prefix), _ptr(self.div), None, 0,
                                    sh), "si_pbwt_sweep")
        self.perm = select_perm(H, L, gpu_compat)
        self.states = torch.empty((M, L), dtype=torch.int32, device=device)
        self.sa = torch.empty((M, self.Lp), dtype=torch.uint8, device=device)
        self.mmask = torch.empty((M,), dtype=torch.uint8, device=device)
        check(lib.si_state_table(_ptr(self.prefix), _ptr(self.cols), _ptr(self.miss), H, M, L, mode,
                                 self.perm.ctypes.data_as(_native._u32p), _ptr(self.states),
                                 _ptr(self.sa), _ptr(self.mmask), sh), "si_state_table")
        if not keep_prefix:
            self.prefix = None
            self.div = None

    def derive(self, L: int, mode: int = 0, stream=None, gpu_compat: bool = False) -> "DeviceIndex":
        """Another state table (L states) over this index's PBWT arrays (which
        must have been kept), sharing the resident panel and prefix rows."""
        if self.prefix is None and mode == 0:
            raise ValueError("derive() needs the prefix rows (keep_prefix=True)")
        d = object.__new__(DeviceIndex)
        d.H, d.M, d.L, d.Lp = self.H, self.M, L, padded_states(L)
        d.cols, d.miss, d.prefix, d.div = self.cols, self.miss, self.prefix, self.div
        dev = self.cols.device
        d.perm = select_perm(self.H, L, gpu_compat)
        d.states = torch.empty((self.M, L), dtype=torch.int32, device=dev)
        d.sa = torch.empty((self.M, d.Lp), dtype=torch.uint8, device=dev)
        d.mmask = torch.empty((self.M,), dtype=torch.uint8, device=dev)
        check(hip().si_state_table(_ptr(self.prefix), _ptr(self.cols), _ptr(self.miss), self.H,
                                   self.M, L, mode, d.perm.ctypes.data_as(_native._u32p),
                                   _ptr(d.states), _ptr(d.sa), _ptr(d.mmask),
                                   _stream_handle(stream)), "si_state_table")
        return d

    def drop_prefix(self):
        self.prefix = None
        self.div = None

    def state_alleles(self):
        """(a0, a1) [M][L] uint8 raw alleles of the selected state pairs, from
        the device state table (for the oracle)."""
        sa = self.sa.cpu().numpy()[:, : self.L]
        a0 = np.where(sa & 16, 255, sa & 1).astype(np.uint8)
        a1 = np.where(sa & 32, 255, (sa >> 1) & 1).astype(np.uint8)
        return a0, a1


# ----------------------------------------------------------------------- jobs
@dataclass
class JobOutputs:
    ds: torch.Tensor | None = None
    ap: torch.Tensor | None = None
    hap: torch.Tensor | None = None
    prob: torch.Tensor | None = None
    loglik: torch.Tensor | None = None


@dataclass
class Job:
    """One imputation job (B samples x M markers against a DeviceIndex)."""
    index: DeviceIndex
    codes: torch.Tensor          # [B][pitch] uint8 on device
    palette: np.ndarray          # [P][3]
    trans: Transitions
    seed: int = 12345
    rng_stream: int = 0
    sample_base: int = 0
    argmax: bool = False
    logspace: bool = False
    window: int = 0              # the reference's windowed model when M > window > 0
    untyped: bool = True         # untyped markers take the composed step (SI_FB_UNTYPED)
    banded: bool | None = None   # SI_FB_BANDED (None: for log-space jobs at the floor's scale)
    out: JobOutputs = field(default_factory=JobOutputs)
    _keep: list = field(default_factory=list)

    @property
    def B(self):
        return self.codes.shape[0]

    def alloc_outputs(self, ds=True, ap=False, hap=True, prob=False, loglik=False):
        B, M, L = self.B, self.index.M, self.index.L
        dev = self.codes.device
        self.out = JobOutputs(
            ds=torch.empty((B, M), dtype=torch.float32, device=dev) if ds else None,
            ap=torch.empty((B, M, 2), dtype=torch.float32, device=dev) if ap else None,
            hap=torch.empty((B, 2, M), dtype=torch.uint8, device=dev) if hap else None,
            prob=torch.empty((B, M, L), dtype=torch.float32, device=dev) if prob else None,
            loglik=torch.empty((B,), dtype=torch.float64, device=dev) if loglik else None,
        )
        return self.out

    def c_job(self) -> FbJob:
        eps, emax = emission_table(self.palette)
        dev = self.codes.device
        eps_t = torch.from_numpy(eps).to(dev)
        emax_t = torch.from_numpy(emax).to(dev)
        pal_t = torch.from_numpy(np.ascontiguousarray(self.palette, np.float32)).to(dev)
        self._keep = [eps_t, emax_t, pal_t]
        if not self.logspace and self.trans.min_switch() < 1e-30:
            raise ValueError("a switch probability e^To is below the linear-space kernel's range "
                             "(coincident markers); use logspace=True")
        j = FbJob()
        j.n_samples = self.B
        j.n_markers = self.index.M
        j.n_states = self.index.L
        banded = self.trans.floor_scale() if self.banded is None else self.banded
        j.flags = ((_native.SI_FB_PER_MARKER if self.trans.per_marker else 0)
                   | (_native.SI_FB_ARGMAX if self.argmax else 0)
                   | (_native.SI_FB_LOGSPACE if self.logspace else 0)
                   | (_native.SI_FB_BANDED if self.logspace and banded else 0))
        j.code_pitch = self.codes.shape[1]
        j.n_palette = self.palette.shape[0]
        j.rng_stream = self.rng_stream
        j.sample_base = self.sample_base
        j.window = self.window
        uc = untyped_code(self.palette) if self.untyped else None
        if uc is not None:
            j.flags |= _native.SI_FB_UNTYPED
            j.untyped_code = uc
        j.seed = self.seed
        s = self.trans.step
        j.step = StepCoef(float(s[0]), float(s[1]), float(s[2]), float(s[3]))
        j.steps = _ptr(self.trans.steps_dev)
        j.sa = _ptr(self.index.sa)
        j.mmask = _ptr(self.index.mmask)
        j.codes = _ptr(self.codes)
        j.eps = _ptr(eps_t)
        j.emax = _ptr(emax_t)
        j.palette = _ptr(pal_t)
        o = self.out
        j.ds, j.ap, j.hap, j.prob, j.loglik = (_ptr(o.ds), _ptr(o.ap), _ptr(o.hap), _ptr(o.prob),
                                               _ptr(o.loglik))
        return j


def workspace_bytes(jobs) -> int:
    arr = (FbJob * len(jobs))(*[j if isinstance(j, FbJob) else j.c_job() for j in jobs])
    return int(hip().si_fb_workspace_bytes(arr, len(jobs)))


def run_jobs(cjobs, workspace: torch.Tensor | None = None, stream=None, variant=None):
    """si_fb_run over prepared FbJob structs (plain form: one job after another).
    variant: (chain lanes, block lanes, checkpoint block) or None for the default."""
    arr = (FbJob * len(cjobs))(*cjobs)
    need = int(hip().si_fb_workspace_bytes(arr, len(cjobs)))
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=DEV)
    var = None if variant is None else ctypes.byref(FbVariant(*variant))
    check(hip().si_fb_run_variant(arr, len(cjobs), var, _ptr(workspace), workspace.numel(),
                                  3, _stream_handle(stream)), "si_fb_run")
    return workspace


def default_variant(L: int):
    v = FbVariant()
    check(hip().si_fb_default_variant(L, ctypes.byref(v)), "si_fb_default_variant")
    return v.as_tuple()


def lane_candidates(L: int, pass_: int):
    """Lane counts built for a pass (SI_FB_PASS_CHAIN or SI_FB_PASS_BLOCK) at L states."""
    buf = (ctypes.c_uint32 * 8)()
    n = int(hip().si_fb_lane_candidates(L, pass_, buf, 8))
    return [int(buf[i]) for i in range(min(n, 8))]


def variants(L: int):
    """Every (chain lanes, block lanes, checkpoint block) built for L states."""
    return [(gc, gb, cb) for cb in (16,) for gc in lane_candidates(L, 1)
            for gb in lane_candidates(L, 2)]


def tune_jobs(cjobs, max_markers=0, stream=None, workspace=None):
    """si_fb_tune: the fastest variant for these jobs (outputs untouched)."""
    arr = (FbJob * len(cjobs))(*cjobs)
    need = int(hip().si_fb_workspace_bytes(arr, len(cjobs)))
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=DEV)
    v = FbVariant()
    check(hip().si_fb_tune(arr, len(cjobs), _ptr(workspace), workspace.numel(), max_markers,
                           _stream_handle(stream), ctypes.byref(v)), "si_fb_tune")
    return v.as_tuple()


class Batch:
    """Batched form: many jobs in persistent work-queue launches (one per pass
    and padded-state class), optionally replayed from a captured hipGraph.
    All jobs must share one emission table (build them with the same
    ``tables``)."""

    def __init__(self, cjobs):
        self._arr = (FbJob * len(cjobs))(*cjobs)
        h = ctypes.c_void_p()
        check(hip().si_fb_batch_create(self._arr, len(cjobs), ctypes.byref(h)), "si_fb_batch_create")
        self.handle = h
        self.ws_bytes = int(hip().si_fb_batch_workspace_bytes(h))
        self.workspace = None

    def reset(self, cjobs):
        """Plan the batch again for other jobs, reusing its device tables
        (si_fb_batch_reset); variants go back to untuned defaults."""
        self._arr = (FbJob * len(cjobs))(*cjobs)
        check(hip().si_fb_batch_reset(self.handle, self._arr, len(cjobs)), "si_fb_batch_reset")
        self.ws_bytes = int(hip().si_fb_batch_workspace_bytes(self.handle))

    def run(self, use_graph=False, stream=None, workspace=None):
        if workspace is not None:
            self.workspace = workspace
        if self.workspace is None or self.workspace.numel() < self.ws_bytes:
            self.workspace = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=DEV)
        check(hip().si_fb_batch_run(self.handle, _ptr(self.workspace), self.workspace.numel(),
                                    int(bool(use_graph)), _stream_handle(stream)), "si_fb_batch_run")

    def _ensure_ws(self):
        if self.workspace is None or self.workspace.numel() < self.ws_bytes:
            self.workspace = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=DEV)

    def tune(self, max_markers=0, stream=None):
        """Tune every class without a fixed variant (si_fb_batch_tune)."""
        self._ensure_ws()
        check(hip().si_fb_batch_tune(self.handle, _ptr(self.workspace), self.workspace.numel(),
                                     max_markers, _stream_handle(stream)), "si_fb_batch_tune")

    def classes(self):
        """[(padded states, logspace, (chain lanes, block lanes, ckpt block), tuned)]"""
        out = []
        for k in range(int(hip().si_fb_batch_n_classes(self.handle))):
            lp, fl, t = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int()
            v = FbVariant()
            check(hip().si_fb_batch_class_info(self.handle, k, ctypes.byref(lp), ctypes.byref(fl),
                                               ctypes.byref(v), ctypes.byref(t)),
                  "si_fb_batch_class_info")
            out.append((lp.value, bool(fl.value & SI_FB_LOGSPACE), v.as_tuple(), bool(t.value)))
        return out

    def set_variant(self, k, variant):
        check(hip().si_fb_batch_set_variant(self.handle, k, ctypes.byref(FbVariant(*variant))),
              "si_fb_batch_set_variant")

    def close(self):
        if self.handle:
            hip().si_fb_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SharedTables:
    """One emission table on the device, shared by the jobs of a batch."""

    def __init__(self, palette, device=DEV, untyped: bool = True):
        self.palette = np.ascontiguousarray(palette, np.float32)
        self.ucode = untyped_code(self.palette) if untyped else None
        eps, emax = emission_table(self.palette)
        self.eps = torch.from_numpy(eps).to(device)
        self.emax = torch.from_numpy(emax).to(device)
        self.pal = torch.from_numpy(self.palette).to(device)

    def apply(self, j: FbJob) -> FbJob:
        j.eps, j.emax, j.palette = _ptr(self.eps), _ptr(self.emax), _ptr(self.pal)
        j.n_palette = self.palette.shape[0]
        if self.ucode is not None:
            j.flags |= _native.SI_FB_UNTYPED
            j.untyped_code = self.ucode
        return j


class RegionJob:
    """A job over the marker region [m0, m0 + M) of a resident index: its
    state table and transition steps are slices of the panel-wide ones, its
    codes a [B][pitch] device tensor.  Outputs (device tensors) are allocated
    for the names in `outputs`.  logspace=None picks the log-space kernel when
    a switch probability the job uses is below the linear kernel's range."""

    def __init__(self, index: DeviceIndex, m0: int, M: int, codes: torch.Tensor,
                 trans: Transitions, seed: int, tables: "SharedTables", logspace=None,
                 outputs=("ds", "hap"), rng_stream: int = 0, sample_base: int = 0):
        if m0 < 0 or M < 1 or m0 + M > index.M:
            raise ValueError("region outside the panel")
        if codes.dim() != 2 or codes.shape[1] < M or codes.shape[1] % 16:
            raise ValueError("codes must be [B][pitch] with pitch >= M, a multiple of 16")
        self.index, self.m0, self.M, self.codes = index, m0, M, codes
        self.B = codes.shape[0]
        self.trans, self.seed, self.tables = trans, seed, tables
        self.logspace = (trans.min_switch() < 1e-30) if logspace is None else bool(logspace)
        self.rng_stream, self.sample_base = rng_stream, sample_base
        dev = codes.device
        B, L = self.B, index.L
        shapes = {"ds": ((B, M), torch.float32), "ap": ((B, M, 2), torch.float32),
                  "hap": ((B, 2, M), torch.uint8), "prob": ((B, M, L), torch.float32),
                  "loglik": ((B,), torch.float64)}
        self.out = {k: torch.empty(shapes[k][0], dtype=shapes[k][1], device=dev) for k in outputs}

    def c_job(self) -> FbJob:
        if not self.logspace and self.trans.min_switch() < 1e-30:
            raise ValueError("a switch probability e^To is below the linear-space kernel's range; "
                             "use logspace=True")
        ix = self.index
        j = FbJob()
        j.n_samples, j.n_markers, j.n_states = self.B, self.M, ix.L
        j.flags = ((_native.SI_FB_PER_MARKER if self.trans.per_marker else 0)
                   | (_native.SI_FB_LOGSPACE if self.logspace else 0)
                   | (_native.SI_FB_BANDED if self.logspace and self.trans.floor_scale() else 0))
        j.code_pitch = self.codes.shape[1]
        j.rng_stream, j.sample_base, j.seed = self.rng_stream, self.sample_base, self.seed
        s = self.trans.step
        j.step = StepCoef(float(s[0]), float(s[1]), float(s[2]), float(s[3]))
        j.steps = _ptr(self.trans.steps_dev)
        j.sa = ctypes.c_void_p(ix.sa.data_ptr() + self.m0 * ix.Lp)
        j.mmask = ctypes.c_void_p(ix.mmask.data_ptr() + self.m0)
        j.codes = _ptr(self.codes)
        self.tables.apply(j)
        for k in ("ds", "ap", "hap", "prob", "loglik"):
            setattr(j, k, _ptr(self.out.get(k)))
        return j
