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
is numpy.random.default_rng(SeedSequence([config_id, stream_id])); the peaks
of chunk q of cycles draw from SeedSequence([config_id, stream_id, q]), so
chunks are generated on threads and the arrays do not depend on the thread
count.
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


def _cpus() -> int:
    try:
        import os
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        import os
        return os.cpu_count() or 1


def generate(spec: RunSpec, progress=None, threads: int | None = None) -> dict:
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
    # Each planted target gets, at every cycle within 4 sigma of its RT
    # centre, one peak per fragment in its first isolation window:
    # m/z = fragment m/z * (1 + U(-3, 3) ppm), intensity = amp * library
    # intensity * exp(-(cycle - centre)^2 / (2 sigma^2)).
    rp = _rng(cid, 9)
    n_plant = int(round(spec.planted_fraction * P_t))
    planted = np.sort(rp.choice(P_t, n_plant, replace=False)) if n_plant else \
        np.zeros(0, dtype=np.int64)
    amp = rp.lognormal(8.0, 1.0, n_plant)
    span = int(np.floor(4 * spec.plant_sigma))
    offs = np.arange(-span, span + 1)
    pw = np.searchsorted(win_hi, pmz[planted], side="left")
    pw = np.minimum(pw, W - 1)
    centre = prt[planted].astype(np.float64)
    cyc = np.floor(centre).astype(np.int64)[:, None] + offs[None, :]
    ok = (cyc >= 0) & (cyc < C) & (np.abs(cyc - centre[:, None]) <= 4 * spec.plant_sigma)
    kk, oo = np.nonzero(ok)                      # (planted k, cycle) pairs
    pc = cyc[kk, oo]
    prof = np.exp(-((pc - centre[kk]) ** 2) / (2 * spec.plant_sigma ** 2))
    nfk = nf[planted][kk]                        # fragments of each pair
    pair = np.repeat(np.arange(kk.size), nfk)
    fofs = np.arange(int(nfk.sum())) - np.repeat(np.cumsum(nfk) - nfk, nfk)
    fidx = fstart_t[planted][kk][pair].astype(np.int64) + fofs
    jit = _rng(cid, 10).uniform(-3e-6, 3e-6, fidx.size)
    pl_cyc = pc[pair]
    pl_spec = pl_cyc * W + pw[kk][pair]
    pl_mz = (fmz_t[fidx].astype(np.float64) * (1.0 + jit)).astype(np.float32)
    pl_int = (amp[kk][pair] * fint_t[fidx].astype(np.float64) * prof[pair]).astype(np.float32)
    # grouped by cycle here; sorted by (spectrum, m/z) per chunk below
    order = np.argsort(pl_cyc, kind="stable")
    pl_cyc, pl_spec = pl_cyc[order], pl_spec[order]
    pl_mz, pl_int = pl_mz[order], pl_int[order]

    # ---------------- peaks, in chunks of cycles on threads ----------------
    # Chunk q of cycles draws from its own streams (SeedSequence([cid, 11,
    # q]) for m/z, [cid, 12, q] for intensity), so chunks are independent
    # and the result does not depend on the thread count.  Background m/z is
    # sorted per spectrum; intensities are i.i.d. and independent of m/z, so
    # they stay in draw order.  Planted peaks are merged into the sorted
    # rows (after background peaks of equal m/z).
    n_spec = C * W
    K = spec.peaks_per_spectrum
    pl_count = np.bincount(pl_spec, minlength=n_spec).astype(np.int64)
    per_spec = K + pl_count
    stop = np.cumsum(per_spec)
    start = stop - per_spec
    total = int(stop[-1]) if n_spec else 0
    peak_mz = np.empty(total, dtype=np.float32)
    peak_int = np.empty(total, dtype=np.float32)
    chunk = max(1, 2_000_000 // max(K * W, 1))
    chunks = [(q, c0, min(C, c0 + chunk)) for q, c0 in enumerate(range(0, C, chunk))]
    assert not pl_mz.size or float(pl_mz.max()) < 4096.0

    def fill(job):
        q, c0, c1 = job
        s0, s1 = c0 * W, c1 * W
        ns = s1 - s0
        rmz = np.random.default_rng(np.random.SeedSequence([cid, 11, q]))
        rin = np.random.default_rng(np.random.SeedSequence([cid, 12, q]))
        base_mz = rmz.random((ns, K), dtype=np.float32)
        base_mz *= np.float32(1850.0)
        base_mz += np.float32(150.0)
        base_mz.sort(axis=1)
        base_in = rin.standard_normal((ns, K), dtype=np.float32)
        base_in *= np.float32(1.5)
        base_in += np.float32(6.0)
        np.exp(base_in, out=base_in)
        o0, o1 = int(start[s0]), int(stop[s1 - 1])
        a, b = np.searchsorted(pl_cyc, [c0, c1], side="left").tolist()
        if b == a:
            peak_mz[o0:o1] = base_mz.reshape(-1)
            peak_int[o0:o1] = base_in.reshape(-1)
        else:
            # merge by the exact f64 key (row * 4096 + m/z)
            pkey = (pl_spec[a:b] - s0).astype(np.float64) * 4096.0 + pl_mz[a:b]
            po = np.argsort(pkey, kind="stable")
            pkey = pkey[po]
            bkey = (np.arange(ns, dtype=np.float64)[:, None] * 4096.0
                    + base_mz.astype(np.float64)).reshape(-1)
            pos = np.searchsorted(bkey, pkey, side="right") + np.arange(b - a)
            is_pl = np.zeros(o1 - o0, dtype=bool)
            is_pl[pos] = True
            seg_mz = peak_mz[o0:o1]
            seg_in = peak_int[o0:o1]
            seg_mz[pos] = pl_mz[a:b][po]
            seg_in[pos] = pl_int[a:b][po]
            seg_mz[~is_pl] = base_mz.reshape(-1)
            seg_in[~is_pl] = base_in.reshape(-1)
        return c1

    if chunks:
        import concurrent.futures as cf
        nthr = max(1, min(threads or _cpus(), 16, len(chunks)))
        with cf.ThreadPoolExecutor(nthr) as ex:
            for c1 in ex.map(fill, chunks):
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
