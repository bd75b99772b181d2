"""Synthetic workloads (numpy front end of ``include/swiftimpute/synth.h``).

The workload shapes are the imputation benchmarks this package is measured
on; every generator is seeded and thread-count independent.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _native

HARDCALL_PALETTE = np.array([[0.0, -999.0, -999.0], [-999.0, 0.0, -999.0],
                             [-999.0, -999.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
CODE_MISSING = 3


def _threads(n):
    return n or min(16, os.cpu_count() or 1)


@dataclass
class Panel:
    H: int
    M: int
    cols: np.ndarray      # [M][W] uint64 bit-packed
    pos: np.ndarray       # [M] uint64 bp
    cM: np.ndarray        # [M] centiMorgans (copying map)
    morgans: np.ndarray   # [M] Morgans (copying map)
    spacing_mode: int

    @property
    def W(self):
        return (self.H + 63) // 64

    def alleles(self) -> np.ndarray:
        """[M][H] uint8 panel (the reference's ReferencePanel layout)."""
        b = self.cols.view(np.uint8).reshape(self.M, self.W * 8)
        return np.unpackbits(b, axis=1, bitorder="little")[:, : self.H].copy()

    def genetic_cM(self):
        """cM values to hand the imputer: None for the constant-spacing map
        (positions alone give 1 cM/Mb, as the reference computes), the
        hotspot-aware map otherwise."""
        return None if self.spacing_mode == 0 else self.cM


def make_panel(H: int, M: int, seed: int = 12345, spacing_mode: int = 0, n_founders: int = 40,
               mutation: float = 1e-3, threads: int | None = None) -> Panel:
    lib = _native.api()
    cfg = _native.SynthPanelCfg()
    lib.si_synth_panel_defaults(ctypes.byref(cfg), H, M, seed)
    cfg.spacing_mode = spacing_mode
    cfg.n_founders = n_founders
    cfg.mutation = mutation
    W = (H + 63) // 64
    cols = np.empty((M, W), np.uint64)
    pos = np.empty(M, np.uint64)
    cM = np.empty(M, np.float64)
    mg = np.empty(M, np.float64)
    rc = lib.si_synth_panel(ctypes.byref(cfg), cols.ctypes.data_as(_native._u64p),
                            pos.ctypes.data_as(_native._u64p), cM.ctypes.data_as(_native._f64p),
                            mg.ctypes.data_as(_native._f64p), _threads(threads))
    if rc != 0:
        raise ValueError(f"si_synth_panel failed ({rc})")
    return Panel(H, M, cols, pos, cM, mg, spacing_mode)


@dataclass
class Targets:
    B: int
    M: int
    pitch: int
    codes: np.ndarray         # [B][pitch] uint8, palette indices
    truth: np.ndarray | None  # [B][2][M] uint8
    typed_every: int

    def untyped_mask(self) -> np.ndarray:
        return self.codes[:, : self.M] == CODE_MISSING

    def gl(self) -> np.ndarray:
        """[B][M][3] float32 log10 genotype likelihoods (TargetData layout)."""
        return HARDCALL_PALETTE[self.codes[:, : self.M]]


def make_targets(panel: Panel, B: int, typed_every: int, seed: int = 12346,
                 geno_error: float = 0.005, with_truth: bool = True,
                 threads: int | None = None) -> Targets:
    lib = _native.api()
    cfg = _native.SynthTargetCfg()
    lib.si_synth_target_defaults(ctypes.byref(cfg), B, panel.M, typed_every, seed)
    cfg.geno_error = geno_error
    codes = np.empty((B, cfg.code_pitch), np.uint8)
    truth = np.empty((B, 2, panel.M), np.uint8) if with_truth else None
    rc = lib.si_synth_targets(ctypes.byref(cfg), panel.cols.ctypes.data_as(_native._u64p),
                              panel.H, panel.M, panel.morgans.ctypes.data_as(_native._f64p),
                              codes.ctypes.data_as(_native._u8p),
                              truth.ctypes.data_as(_native._u8p) if truth is not None else None,
                              _threads(threads))
    if rc != 0:
        raise ValueError(f"si_synth_targets failed ({rc})")
    return Targets(B, panel.M, cfg.code_pitch, codes, truth, typed_every)


# Benchmark workloads: (H, M, B, L, Ne, typed_every, spacing_mode)
WORKLOADS = {
    "cfg1": dict(H=4432, M=1644, B=444, L=8, ne=10000, typed_every=4, spacing_mode=0),
    "cfg2": dict(H=20000, M=100000, B=10000, L=16, ne=20000, typed_every=20, spacing_mode=0),
    "cfg3": dict(H=8000, M=100000, B=2000, L=8, ne=10000, typed_every=20, spacing_mode=1),
}


# --------------------------------------------------------------------------
# Uneven job sets and request streams over one shared panel.  A job is a
# contiguous marker region of the panel and a contiguous slice of a target
# pool typed at every marker; the job's own typed density keeps every
# typed_every-th marker of its region (counted from the region start) and
# masks the rest.
# --------------------------------------------------------------------------
@dataclass
class JobSpec:
    m0: int           # first marker of the region
    M: int            # region length
    s0: int           # first sample of the cohort slice of the pool
    B: int            # cohort size
    L: int            # states
    typed_every: int  # typed-marker density 1/typed_every
    seed: int         # Philox seed of the job's sampling


def _log_uniform(rng, lo, hi):
    return int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))


def uneven_jobs(n_jobs: int, panel_M: int, pool_B: int, m_lo: int = 2000, m_hi: int = 80000,
                b_lo: int = 1, b_hi: int = 2000, Ls=(4, 8, 16), densities=(10, 20, 50),
                seed: int = 12345) -> list:
    """Config-4 style job set: region length and cohort size log-uniform,
    L and typed density uniform over the given choices, region start uniform
    in the panel; job j draws from seed + j."""
    out = []
    for j in range(n_jobs):
        rng = np.random.default_rng(seed + j)
        M = min(_log_uniform(rng, m_lo, m_hi), panel_M)
        B = min(_log_uniform(rng, b_lo, b_hi), pool_B)
        m0 = int(rng.integers(0, panel_M - M + 1))
        s0 = int(rng.integers(0, pool_B - B + 1))
        L = int(rng.choice(Ls))
        te = int(rng.choice(densities))
        out.append(JobSpec(m0, M, s0, B, L, te, seed + j))
    return out


def poisson_requests(rate: float, duration_s: float, panel_M: int, pool_B: int,
                     m_lo: int = 1000, m_hi: int = 100000, b_lo: int = 1, b_hi: int = 5000,
                     L: int = 8, typed_every: int = 20, seed: int = 12345) -> list:
    """Config-5 style request stream: Poisson arrivals at `rate` per second
    over `duration_s`; cohort and region log-uniform.  Returns
    [(arrival_s, JobSpec)]."""
    rng = np.random.default_rng(seed)
    t, out, j = 0.0, [], 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration_s:
            return out
        M = min(_log_uniform(rng, m_lo, m_hi), panel_M)
        B = min(_log_uniform(rng, b_lo, b_hi), pool_B)
        m0 = int(rng.integers(0, panel_M - M + 1))
        s0 = int(rng.integers(0, pool_B - B + 1))
        out.append((t, JobSpec(m0, M, s0, B, L, typed_every, seed + 1 + j)))
        j += 1


def region_codes(pool: np.ndarray, spec: JobSpec, pitch: int | None = None) -> np.ndarray:
    """[B][pitch] codes of a job from a pool typed at every marker: markers of
    the region with (m - m0) % typed_every != 0 become missing (code 3)."""
    M = spec.M
    if pitch is None:
        pitch = M
    c = np.full((spec.B, pitch), CODE_MISSING, np.uint8)
    c[:, :M:spec.typed_every] = pool[spec.s0:spec.s0 + spec.B, spec.m0:spec.m0 + M:spec.typed_every]
    return c


# --------------------------------------------------------------------------
# VCF text of a synthetic panel and cohort (the inputs an end-to-end run of
# the C++ API reads): biallelic markers on chr1 at the panel's positions,
# REF A / ALT G, panel genotypes phased "a|b" (haplotypes 2s, 2s+1), target
# genotypes unphased from the codes ("./." for untyped markers).
# --------------------------------------------------------------------------
_GT_PHASED = np.array([b"0|0", b"0|1", b"1|0", b"1|1"])
_GT_CODES = np.array([b"0/0", b"0/1", b"1/1", b"./."])


def _vcf_header(sample_names) -> str:
    return ("##fileformat=VCFv4.2\n##source=swiftimpute_amd.synth\n##contig=<ID=chr1>\n"
            "##FORMAT=<ID=GT,Number=1,Type=String,Description=\"Genotype\">\n"
            "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(sample_names) + "\n")


def _write_rows(f, pos, fields):
    """fields: [M][n] array of byte strings (one genotype field each)."""
    for m in range(fields.shape[0]):
        f.write(f"chr1\t{int(pos[m])}\trs{m + 1}\tA\tG\t.\tPASS\t.\tGT\t".encode())
        f.write(b"\t".join(fields[m].tolist()))
        f.write(b"\n")


def write_panel_vcf(path: str, panel: Panel) -> None:
    """The panel as a VCF of H/2 diploid samples (H must be even)."""
    if panel.H % 2:
        raise ValueError("a VCF panel needs an even haplotype count")
    a = panel.alleles().astype(np.int64)               # [M][H]
    fields = _GT_PHASED[2 * a[:, 0::2] + a[:, 1::2]]   # [M][H/2]
    with open(path, "wb") as f:
        f.write(_vcf_header([f"REF{s}" for s in range(panel.H // 2)]).encode())
        _write_rows(f, panel.pos, fields)


def write_targets_vcf(path: str, panel: Panel, targets: Targets) -> None:
    """The cohort's hard calls (codes 0/1/2, 3 = untyped) as a VCF."""
    fields = _GT_CODES[targets.codes[:, : panel.M].T]  # [M][B]
    with open(path, "wb") as f:
        f.write(_vcf_header([f"TGT{s}" for s in range(targets.B)]).encode())
        _write_rows(f, panel.pos, fields)
