"""Seeded synthetic DIA runs and spectral libraries for the selection path.

Produces exactly the arrays the reference's ``DIAData.from_arrays`` and
``SpecLibFlat.from_arrays`` take (scripts/test_search.py:67-132 shows the
real-data call), for five workloads:

  0  1,000 cycles x 8 windows (75 Th), 2,000 peaks/spectrum, 2k precursors,
     12 fragments, +-60 cycles, 30% planted, 20 ppm, top-3
  1  6,000 x 60 (10 Th), 1,500 peaks/spectrum (540M peaks), 400k precursors
     (200k targets + 200k shifted decoys), 12 fragments, +-150, 25% of
     targets planted, 10 ppm, top-3
  2  as 1 with 6-24 fragments, per-precursor half-window 30-200 cycles and
     precursor m/z from a 20-component Gaussian mixture (sigma 15 Th)
  3  the data of 2 (used with a backward pass)
  4  20,000 x 100 (6 Th), 1,000 peaks/spectrum (2B peaks), 2M precursors,
     6-24 fragments, half-window 50-400, 20% planted, 5 ppm, top-5

Spectra are cycle-major: spectrum (c, w) has index c*W + w, delta_scan_idx w,
cycle_idx c and rt = c, so the RT index is 0..C-1 with delta_t = 1 and
tolerances in seconds equal tolerances in cycles.  Each spectrum's peaks are
sorted by m/z after planted elution peaks are inserted.  Every random stream
is numpy.random.default_rng(SeedSequence([config_id, stream_id])).
"""
from __future__ import annotations

import dataclasses
import numpy as np

FRAG_B = 98
FRAG_Y = 121


@dataclasses.dataclass
class RunSpec:
    config_id: int
    n_cycles: int
    n_windows: int
    window_lo: float
    window_hi: float
    peaks_per_spectrum: int
    n_targets: int
    with_decoys: bool
    frag_min: int
    frag_max: int
    half_window_min: float  # RT tolerance (cycles); min == max -> global
    half_window_max: float
    planted_fraction: float
    mass_tolerance: float
    candidate_count: int
    top_k_fragments: int
    clustered_mz: bool = False
    fwhm_rt: float = 4.7096  # sigma = 2.0 cycles exactly
    kernel_size: int = 20
    peak_length: int = 5
    plant_sigma: float = 3.0

    @property
    def n_precursors(self) -> int:
        return self.n_targets * (2 if self.with_decoys else 1)

    @property
    def per_precursor_rt(self) -> bool:
        return self.half_window_min != self.half_window_max

    def selection_config(self) -> dict:
        return {
            "fwhm_rt": self.fwhm_rt,
            "kernel_size": self.kernel_size,
            "peak_length": self.peak_length,
            "mass_tolerance": self.mass_tolerance,
            "rt_tolerance": float(self.half_window_max),
            "candidate_count": self.candidate_count,
            "top_k_fragments": self.top_k_fragments,
        }


def config(config_id: int, **overrides) -> RunSpec:
    """The workload of a numbered configuration, optionally resized."""
    if config_id == 0:
        s = RunSpec(0, 1000, 8, 400.0, 1000.0, 2000, 2000, False, 12, 12,
                    60, 60, 0.30, 20.0, 3, 12)
    elif config_id == 1:
        s = RunSpec(1, 6000, 60, 400.0, 1000.0, 1500, 200_000, True, 12, 12,
                    150, 150, 0.25, 10.0, 3, 12)
    elif config_id in (2, 3):
        s = RunSpec(config_id, 6000, 60, 400.0, 1000.0, 1500, 200_000, True,
                    6, 24, 30, 200, 0.25, 10.0, 3, 24, clustered_mz=True)
    elif config_id == 4:
        s = RunSpec(4, 20000, 100, 400.0, 1000.0, 1000, 1_000_000, True, 6,
                    24, 50, 400, 0.20, 5.0, 5, 24, clustered_mz=True)
    else:
        raise ValueError(f"unknown config {config_id}")
    return dataclasses.replace(s, **overrides)


def _rng(cid: int, stream: int) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([cid, stream]))


def generate(spec: RunSpec, progress=None) -> dict:
    """Returns {'dia': (10 arrays), 'lib': (17 arrays), ...metadata}."""
    cid = spec.config_id
    C, W = spec.n_cycles, spec.n_windows
    P_t = spec.n_targets
    P = spec.n_precursors
    width = (spec.window_hi - spec.window_lo) / W
    win_lo = (spec.window_lo + width * np.arange(W)).astype(np.float32)
    win_hi = (spec.window_lo + width * (np.arange(W) + 1)).astype(np.float32)

    # ---------------- library ----------------
    r = _rng(cid, 1)
    if spec.clustered_mz:
        centers = r.uniform(spec.window_lo, spec.window_hi, 20)
        comp = r.integers(0, 20, P_t)
        pmz = r.normal(centers[comp], 15.0)
        pmz = np.clip(pmz, spec.window_lo, spec.window_hi)
    else:
        pmz = r.uniform(spec.window_lo, spec.window_hi, P_t)
    pmz = pmz.astype(np.float32)
    prt = _rng(cid, 2).uniform(0, C, P_t).astype(np.float32)
    nf = _rng(cid, 3).integers(spec.frag_min, spec.frag_max + 1, P_t)
    fstart_t = np.concatenate([[0], np.cumsum(nf)[:-1]]).astype(np.uint64)
    NF_t = int(nf.sum())
    fmz_t = _rng(cid, 4).uniform(200.0, 1800.0, NF_t).astype(np.float32)
    # Dirichlet(1) per precursor = normalised exponentials
    e = _rng(cid, 5).exponential(1.0, NF_t)
    seg_sum = np.add.reduceat(e, fstart_t.astype(np.int64))
    fint_t = (e / np.repeat(seg_sum, nf)).astype(np.float32)
    if spec.per_precursor_rt:
        hw_t = _rng(cid, 6).integers(int(spec.half_window_min),
                                     int(spec.half_window_max) + 1,
                                     P_t).astype(np.float32)
    else:
        hw_t = np.full(P_t, spec.half_window_max, dtype=np.float32)

    if spec.with_decoys:
        shift = _rng(cid, 7).uniform(5.0, 50.0, NF_t)
        sign = np.where(_rng(cid, 8).random(NF_t) < 0.5, -1.0, 1.0)
        fmz_d = (fmz_t.astype(np.float64) + sign * shift).astype(np.float32)
        precursor_mz = np.concatenate([pmz, pmz])
        precursor_rt = np.concatenate([prt, prt])
        nf_all = np.concatenate([nf, nf])
        fragment_mz = np.concatenate([fmz_t, fmz_d])
        fragment_int = np.concatenate([fint_t, fint_t])
        hw = np.concatenate([hw_t, hw_t])
    else:
        precursor_mz, precursor_rt, nf_all = pmz, prt, nf
        fragment_mz, fragment_int, hw = fmz_t, fint_t, hw_t
    NF = int(nf_all.sum())
    fstart = np.concatenate([[0], np.cumsum(nf_all)[:-1]]).astype(np.uint64)
    fstop = (fstart + nf_all.astype(np.uint64)).astype(np.uint64)
    local = np.arange(NF) - np.repeat(fstart.astype(np.int64), nf_all)
    ftype = np.where(local % 2 == 0, FRAG_B, FRAG_Y).astype(np.uint8)
    fnum = (2 + local // 2).astype(np.uint8)
    ones_f = np.ones(NF, dtype=np.uint8)
    lib = (
        np.arange(P, dtype=np.uint64),      # precursor_idx
        precursor_mz.astype(np.float32),     # precursor_mz_library
        precursor_mz.astype(np.float32),     # precursor_mz
        precursor_rt.astype(np.float32),     # precursor_rt_library
        precursor_rt.astype(np.float32),     # precursor_rt
        np.full(P, 10, dtype=np.uint8),      # precursor_naa
        fstart, fstop,
        fragment_mz.astype(np.float32),      # fragment_mz_library
        fragment_mz.astype(np.float32),      # fragment_mz
        fragment_int.astype(np.float32),     # fragment_intensity
        ones_f, ones_f.copy(), np.zeros(NF, dtype=np.uint8),
        fnum, ones_f.copy(), ftype,
    )

    # ---------------- planted elution profiles ----------------
    rp = _rng(cid, 9)
    n_plant = int(round(spec.planted_fraction * P_t))
    planted = np.sort(rp.choice(P_t, n_plant, replace=False)) if n_plant else \
        np.zeros(0, dtype=np.int64)
    amp = rp.lognormal(8.0, 1.0, n_plant)
    span = int(np.floor(4 * spec.plant_sigma))
    offs = np.arange(-span, span + 1)
    # window of each planted precursor (first window containing it)
    pw = np.searchsorted(win_hi, pmz[planted], side="left")
    pw = np.minimum(pw, W - 1)
    pl_spec, pl_mz, pl_int = [], [], []
    jit = _rng(cid, 10)
    for k, t in enumerate(planted):
        c_center = float(prt[t])
        cyc = np.floor(c_center).astype(np.int64) + offs
        cyc = cyc[(cyc >= 0) & (cyc < C) & (np.abs(cyc - c_center) <= 4 * spec.plant_sigma)]
        if cyc.size == 0:
            continue
        a, b = int(fstart_t[t]), int(fstart_t[t]) + int(nf[t])
        fm = fmz_t[a:b].astype(np.float64)
        fi = fint_t[a:b].astype(np.float64)
        prof = np.exp(-((cyc - c_center) ** 2) / (2 * spec.plant_sigma ** 2))
        mzs = fm[None, :] * (1.0 + jit.uniform(-3e-6, 3e-6, (cyc.size, fm.size)))
        ints = amp[k] * fi[None, :] * prof[:, None]
        pl_spec.append(np.repeat(cyc * W + int(pw[k]), fm.size))
        pl_mz.append(mzs.reshape(-1))
        pl_int.append(ints.reshape(-1))
    if pl_spec:
        pl_spec = np.concatenate(pl_spec)
        pl_mz = np.concatenate(pl_mz).astype(np.float32)
        pl_int = np.concatenate(pl_int).astype(np.float32)
        order = np.argsort(pl_spec, kind="stable")
        pl_spec, pl_mz, pl_int = pl_spec[order], pl_mz[order], pl_int[order]
    else:
        pl_spec = np.zeros(0, dtype=np.int64)
        pl_mz = np.zeros(0, dtype=np.float32)
        pl_int = np.zeros(0, dtype=np.float32)

    # ---------------- peaks, chunked over cycles ----------------
    n_spec = C * W
    K = spec.peaks_per_spectrum
    pl_count = np.bincount(pl_spec, minlength=n_spec).astype(np.int64)
    per_spec = K + pl_count
    stop = np.cumsum(per_spec)
    start = stop - per_spec
    total = int(stop[-1]) if n_spec else 0
    peak_mz = np.empty(total, dtype=np.float32)
    peak_int = np.empty(total, dtype=np.float32)
    rmz = _rng(cid, 11)
    rin = _rng(cid, 12)
    pl_start = np.concatenate([[0], np.cumsum(pl_count)])
    chunk = max(1, 4_000_000 // max(K * W, 1))
    for c0 in range(0, C, chunk):
        c1 = min(C, c0 + chunk)
        s0, s1 = c0 * W, c1 * W
        ns = s1 - s0
        extra = int(pl_count[s0:s1].max()) if ns else 0
        base_mz = rmz.uniform(150.0, 2000.0, (ns, K)).astype(np.float32)
        base_in = rin.lognormal(6.0, 1.5, (ns, K)).astype(np.float32)
        if extra:
            mz2 = np.full((ns, K + extra), np.inf, dtype=np.float32)
            in2 = np.zeros((ns, K + extra), dtype=np.float32)
            mz2[:, :K] = base_mz
            in2[:, :K] = base_in
            a, b = pl_start[s0], pl_start[s1]
            rows = pl_spec[a:b] - s0
            # column of each planted peak within its spectrum
            first = np.searchsorted(rows, np.arange(ns), side="left")
            col = K + (np.arange(b - a) - first[rows])
            mz2[rows, col] = pl_mz[a:b]
            in2[rows, col] = pl_int[a:b]
        else:
            mz2, in2 = base_mz, base_in
        order = np.argsort(mz2, axis=1, kind="stable")
        mz_s = np.take_along_axis(mz2, order, axis=1)
        in_s = np.take_along_axis(in2, order, axis=1)
        keep = np.isfinite(mz_s)
        o0, o1 = int(start[s0]), int(stop[s1 - 1])
        peak_mz[o0:o1] = mz_s[keep]
        peak_int[o0:o1] = in_s[keep]
        if progress:
            progress(c1, C)
    spec_idx = np.arange(n_spec)
    dia = (
        (spec_idx % W).astype(np.int64),                  # delta_scan_idx
        win_lo[spec_idx % W].astype(np.float32),          # isolation_lower_mz
        win_hi[spec_idx % W].astype(np.float32),          # isolation_upper_mz
        start.astype(np.int64), stop.astype(np.int64),
        (spec_idx // W).astype(np.int64),                 # cycle_idx
        (spec_idx // W).astype(np.float32),               # rt
        peak_mz, peak_int,
        np.zeros((W, 1, 1, 1), dtype=np.float32),
    )
    pairs = (np.stack([np.arange(P_t), np.arange(P_t) + P_t], axis=1)
             if spec.with_decoys else np.zeros((0, 2), dtype=np.int64))
    return {
        "spec": spec,
        "dia": dia,
        "lib": lib,
        "rt_tolerance_per_precursor": hw.astype(np.float32) if spec.per_precursor_rt else None,
        "planted": planted,
        "pairs": pairs,
        "n_peaks": total,
    }
