"""ctypes binding of the CPU restatement in oracle/adsr_oracle.c.

TEST INFRASTRUCTURE: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product (libadsr.so and the
alphadia_search_rs package) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ADSR_ORACLE_LIBRARY selects another build of the oracle (the sanitizer
# build of the CPU tests); by default oracle/liboracle.so.
LIB_PATH = os.environ.get("ADSR_ORACLE_LIBRARY") or os.path.join(HERE, "liboracle.so")

vp = C.c_void_p
sz = C.c_size_t
f32 = C.c_float


class OrcParams(C.Structure):
    _fields_ = [
        ("fwhm_rt", f32), ("kernel_size", sz), ("peak_length", sz),
        ("mass_tolerance", f32), ("rt_tolerance", f32),
        ("candidate_count", sz), ("top_k_fragments", sz),
        ("rt_tolerance_per_precursor", vp), ("sigma_per_window", vp),
    ]


class OrcCands(C.Structure):
    _fields_ = [("n", sz), ("precursor_idx", vp), ("rank", vp),
                ("cycle_center", vp), ("cycle_start", vp), ("cycle_stop", vp),
                ("score", vp)]


class OrcObs(C.Structure):
    _fields_ = [("isolation_window", f32 * 2), ("num_cycles", sz),
                ("slice_starts", vp), ("cycle_indices", vp),
                ("intensities", vp), ("n_entries", sz)]


class OrcDia(C.Structure):
    _fields_ = [("n_obs", sz), ("obs", C.POINTER(OrcObs)), ("rt", vp),
                ("n_rt", sz), ("delta_t", f32), ("grid", vp), ("n_grid", sz)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)

        def sig(name, res, *args):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = list(args)

        sig("orc_ppm_index", sz, f32, f32, f32, vp, sz)
        sig("orc_global_grid", vp, C.POINTER(sz))
        sig("orc_find_closest_index", sz, vp, sz, f32)
        sig("orc_mz_range_indices", None, vp, sz, f32, f32, C.POINTER(sz), C.POINTER(sz))
        sig("orc_rust_bsearch_f32", C.c_int, vp, sz, f32, C.POINTER(sz))
        sig("orc_rust_bsearch_u16", C.c_int, vp, sz, C.c_uint16, C.POINTER(sz))
        sig("orc_mean_delta_t", f32, vp, sz)
        sig("orc_cycle_idx_limits", None, vp, sz, f32, f32, f32, sz, C.POINTER(sz), C.POINTER(sz))
        sig("orc_dia_build", C.POINTER(OrcDia), vp, vp, vp, vp, vp, vp, vp, sz, vp, vp, sz, C.c_int)
        sig("orc_dia_free", None, C.POINTER(OrcDia))
        sig("orc_valid_observations", sz, C.POINTER(OrcDia), f32, vp)
        sig("orc_fill_xic_slice_raw", None, vp, sz, vp, vp, vp, vp, sz, sz, f32, f32)
        sig("orc_dense_xic", None, C.POINTER(OrcDia), f32, sz, sz, f32, vp, sz, vp)
        sig("orc_gaussian_kernel", sz, f32, f32, sz, f32, vp, sz)
        sig("orc_gaussian_kernel_sigma", sz, f32, sz, vp, sz)
        sig("orc_ndarray_sum", f32, vp, sz)
        sig("orc_convolution", None, vp, sz, vp, sz, sz, vp)
        sig("orc_axis_log_dot_product", None, vp, sz, sz, vp, vp)
        sig("orc_find_local_maxima", sz, vp, sz, sz, vp, vp)
        sig("orc_filter_sort_fragments", sz, sz, vp, vp, vp, vp, C.c_int, C.c_int, sz, vp)
        sig("orc_lib_build", vp, sz, vp, vp, vp, vp, vp, sz, vp, vp, vp, vp)
        sig("orc_lib_free", None, vp)
        sig("orc_select", C.POINTER(OrcCands), C.POINTER(OrcDia), vp, C.POINTER(OrcParams), sz, sz, C.c_int)
        sig("orc_cands_free", None, C.POINTER(OrcCands))
        sig("orc_select_one", C.c_int, C.POINTER(OrcDia), vp, C.POINTER(OrcParams), sz,
            C.POINTER(sz), C.POINTER(sz), C.POINTER(sz), vp, vp, vp, vp, vp, sz)
        sig("orc_logf", f32, f32)
        sig("orc_logf_check_range", C.c_uint64, C.c_uint32, sz, vp, C.c_int, C.POINTER(sz))
        sig("orc_lib_build_full", vp, sz, vp, vp, vp, vp, vp, vp, sz, vp, vp, vp, vp, vp)
        u8 = C.c_uint8
        sig("orc_correlation_safe", f32, vp, vp, sz)
        sig("orc_gamma_ln", f32, f32)
        sig("orc_hyperscore_weighted", f32, vp, vp, vp, vp, sz)
        sig("orc_longest_ion_series", None, vp, vp, vp, sz, C.POINTER(u8), C.POINTER(u8))
        sig("orc_intensity_ion_series", f32, vp, vp, vp, sz, u8)
        sig("orc_dot", f32, vp, vp, sz)
        sig("orc_fwhm_rt", f32, vp, sz, C.c_uint64, vp)
        sig("orc_normalize_profiles", None, vp, sz, sz, sz, vp)
        sig("orc_filter_non_zero", sz, vp, sz, sz, vp)
        sig("orc_median_axis_0", None, vp, sz, sz, vp)
        sig("orc_median", f32, vp, sz)
        sig("orc_std", f32, vp, sz)
        sig("orc_ranked_mask", None, vp, sz, sz, sz, vp)
        sig("orc_count_above", sz, vp, sz, f32, vp)
        sig("orc_weighted_mae", f32, vp, vp, sz)
        sig("orc_fragment_mz_errors", None, vp, vp, sz, sz, vp, vp, vp)
        sig("orc_fill_xic_and_mz_slice", None, C.POINTER(OrcDia), sz, vp, vp,
            C.c_uint64, C.c_uint64, f32, f32)
        sig("orc_dense_xic_mz", None, C.POINTER(OrcDia), f32, C.c_uint64, C.c_uint64, f32,
            vp, sz, vp, vp)
        sig("orc_idf_counts", C.POINTER(C.c_uint32), vp, sz)
        sig("orc_idf", f32, vp, f32, f32)
        sig("orc_score", None, C.POINTER(OrcDia), vp, f32, sz, sz, vp, vp, vp, vp, vp,
            vp, vp, C.c_int)
        sig("orc_quantify", None, C.POINTER(OrcDia), vp, f32, sz, sz, vp, vp, vp, vp, sz,
            vp, vp, vp, vp, vp, vp, vp, vp, C.c_int)
        _lib = L
    return _lib


def use_library(path):
    """Load the oracle from another build of adsr_oracle.c from now on
    (the planted-mistake builds of tests/test_oracle_mutants.py); returns
    the previous path."""
    global _lib, LIB_PATH
    prev = LIB_PATH
    LIB_PATH = path
    _lib = None
    return prev


def p(a):
    return None if a is None else a.ctypes.data_as(vp)


def c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def grid() -> np.ndarray:
    n = sz()
    ptr_ = lib().orc_global_grid(C.byref(n))
    return np.ctypeslib.as_array(C.cast(ptr_, C.POINTER(f32)), (n.value,)).copy()


class Dia:
    """Oracle DIAData built from the 10 from_arrays inputs."""

    def __init__(self, arrays, nthreads: int = 0):
        d, lo, hi, ps, pe, cy, rt, mz, it, _cyc = arrays
        self._keep = [np.ascontiguousarray(a) for a in (d, lo, hi, ps, pe, cy, rt, mz, it)]
        k = self._keep
        nt = nthreads or (os.cpu_count() or 1)
        self.h = lib().orc_dia_build(p(k[0]), p(k[1]), p(k[2]), p(k[3]), p(k[4]), p(k[5]),
                                     p(k[6]), len(k[0]), p(k[7]), p(k[8]), len(k[7]), nt)
        self._keep = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dia_free(self.h)
            self.h = None

    @property
    def n_obs(self):
        return self.h.contents.n_obs

    def rt(self):
        d = self.h.contents
        return np.ctypeslib.as_array(C.cast(d.rt, C.POINTER(f32)), (d.n_rt,)).copy() if d.n_rt else np.zeros(0, np.float32)

    def observation(self, i):
        d = self.h.contents
        o = d.obs[i]
        ss = np.ctypeslib.as_array(C.cast(o.slice_starts, C.POINTER(C.c_uint32)), (d.n_grid + 1,)).copy()
        n = o.n_entries
        ci = np.ctypeslib.as_array(C.cast(o.cycle_indices, C.POINTER(C.c_uint16)), (n,)).copy() if n else np.zeros(0, np.uint16)
        it = np.ctypeslib.as_array(C.cast(o.intensities, C.POINTER(f32)), (n,)).copy() if n else np.zeros(0, np.float32)
        return {"isolation_window": np.array(o.isolation_window[:], dtype=np.float32),
                "num_cycles": o.num_cycles, "slice_starts": ss,
                "cycle_indices": ci, "intensities": it}

    def valid_observations(self, mz):
        out = np.zeros(max(self.n_obs, 1), dtype=np.uint64)
        n = lib().orc_valid_observations(self.h, float(mz), p(out))
        return [int(x) for x in out[:n]]

    def dense_xic_mz(self, precursor_mz, c0, c1, tol, frag_mz):
        """DenseXICMZObservation::new: (xic, mz), each [n_frag][c1 - c0]."""
        fm = c32(frag_mz)
        L = (int(c1) - int(c0)) % (1 << 64)
        xic = np.zeros((len(fm), L), dtype=np.float32)
        mz = np.zeros((len(fm), L), dtype=np.float32)
        lib().orc_dense_xic_mz(self.h, float(precursor_mz), int(c0) % (1 << 64), int(c1),
                               float(tol), p(fm), len(fm), p(xic), p(mz))
        return xic, mz

    def dense_xic(self, precursor_mz, c0, c1, tol, frag_mz):
        fm = c32(frag_mz)
        out = np.zeros((len(fm), c1 - c0), dtype=np.float32)
        lib().orc_dense_xic(self.h, float(precursor_mz), c0, c1, float(tol), p(fm), len(fm), p(out))
        return out


class Lib:
    def __init__(self, arrays):
        (pidx, _pmzl, pmz, _prtl, prt, naa, fs, fe, fmzl, fmz, fint,
         _card, _chg, _loss, fnum, _pos, ftype) = arrays
        # naa and fragment_mz_library are optional: None is passed as NULL
        k = [None if a is None else np.ascontiguousarray(a)
             for a in (pidx, pmz, prt, naa, fs, fe, fmz, fmzl, fint, fnum, ftype)]
        self.h = lib().orc_lib_build_full(len(k[0]), p(k[0]), p(k[1]), p(k[2]), p(k[3]),
                                          p(k[4]), p(k[5]), len(k[6]), p(k[6]), p(k[7]),
                                          p(k[8]), p(k[9]), p(k[10]))
        self.n_prec = len(k[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_lib_free(self.h)
            self.h = None


def params(fwhm_rt=3.0, kernel_size=20, peak_length=5, mass_tolerance=7.0,
           rt_tolerance=200.0, candidate_count=3, top_k_fragments=12,
           rt_tolerance_per_precursor=None, sigma_per_window=None):
    prm = OrcParams(fwhm_rt, kernel_size, peak_length, mass_tolerance, rt_tolerance,
                    candidate_count, top_k_fragments, None, None)
    keep = []
    if rt_tolerance_per_precursor is not None:
        a = c32(rt_tolerance_per_precursor)
        keep.append(a)
        prm.rt_tolerance_per_precursor = p(a)
    if sigma_per_window is not None:
        a = c32(sigma_per_window)
        keep.append(a)
        prm.sigma_per_window = p(a)
    prm._keep = keep
    return prm


def select(dia: Dia, library: Lib, prm: OrcParams, row_begin=0, row_end=None, nthreads=0):
    """Candidate table (9 columns like CandidateCollection.to_arrays)."""
    if row_end is None:
        row_end = library.n_prec
    nt = nthreads or (os.cpu_count() or 1)
    c = lib().orc_select(dia.h, library.h, C.byref(prm), row_begin, row_end, nt)
    n = c.contents.n

    def col(ptr_, ct, dt):
        if n == 0:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(C.cast(ptr_, C.POINTER(ct)), (n,)).astype(dt, copy=True)

    cc = c.contents
    out = (col(cc.precursor_idx, C.c_uint64, np.uint64), col(cc.rank, C.c_uint64, np.uint64),
           col(cc.score, f32, np.float32), np.zeros(n, np.uint64), np.zeros(n, np.uint64),
           np.zeros(n, np.uint64), col(cc.cycle_center, C.c_uint64, np.uint64),
           col(cc.cycle_start, C.c_uint64, np.uint64), col(cc.cycle_stop, C.c_uint64, np.uint64))
    lib().orc_cands_free(c)
    return out


def select_one(dia: Dia, library: Lib, prm: OrcParams, row: int, max_cells=1 << 20):
    """Intermediate arrays of one precursor: dict or None (empty RT range)."""
    c0, c1, nf = sz(), sz(), sz()
    k = max(int(prm.top_k_fragments) if prm.top_k_fragments < (1 << 20) else 1024, 1)
    fm = np.zeros(k + 1024, np.float32)
    fw = np.zeros(k + 1024, np.float32)
    xic = np.zeros(max_cells, np.float32)
    conv = np.zeros(max_cells, np.float32)
    score = np.zeros(max_cells, np.float32)
    ok = lib().orc_select_one(dia.h, library.h, C.byref(prm), row, C.byref(c0), C.byref(c1),
                              C.byref(nf), p(fm), p(fw), p(xic), p(conv), p(score), max_cells)
    if ok == 0:
        return None
    if ok < 0:
        raise ValueError("intermediate arrays exceed max_cells")
    L = c1.value - c0.value
    F = nf.value
    return {"cycle_start": c0.value, "cycle_stop": c1.value, "fragment_mz": fm[:F].copy(),
            "fragment_intensity": fw[:F].copy(), "xic": xic[:F * L].reshape(F, L).copy(),
            "conv": conv[:F * L].reshape(F, L).copy(), "score": score[:L].copy()}


def filter_sort_fragments(mz, intensity, number, ftype, non_zero=True, filter_y1=True, top_k=None):
    mz, it = c32(mz), c32(intensity)
    num = np.ascontiguousarray(number, dtype=np.uint8)
    ty = np.ascontiguousarray(ftype, dtype=np.uint8)
    out = np.zeros(max(len(mz), 1), dtype=np.uint64)
    k = (1 << 64) - 1 if top_k is None else int(top_k)
    n = lib().orc_filter_sort_fragments(len(mz), p(mz), p(it), p(num), p(ty),
                                        int(non_zero), int(filter_y1), k, p(out))
    return out[:n].astype(np.int64)


def gaussian_kernel(fwhm_rt, sigma_scale_rt, kernel_width, rt_resolution):
    out = np.zeros(kernel_width + 2, np.float32)
    n = lib().orc_gaussian_kernel(fwhm_rt, sigma_scale_rt, kernel_width, rt_resolution, p(out), len(out))
    return out[:n].copy()


def convolution(kernel, xic):
    k, x = c32(kernel), c32(xic)
    rows, cols = x.shape
    out = np.zeros_like(x)
    lib().orc_convolution(p(k), len(k), p(x), rows, cols, p(out))
    return out


def axis_log_dot_product(a, w):
    a, w = c32(a), c32(w)
    rows, cols = a.shape
    out = np.zeros(cols, np.float32)
    lib().orc_axis_log_dot_product(p(a), rows, cols, p(w), p(out))
    return out


def find_local_maxima(a, offset):
    a = c32(a)
    idx = np.zeros(max(len(a), 1), np.uint64)
    val = np.zeros(max(len(a), 1), np.float32)
    n = lib().orc_find_local_maxima(p(a), len(a), offset, p(idx), p(val))
    return idx[:n].astype(np.int64), val[:n].copy()


# CandidateFeature f32 columns in to_dict_arrays order
# (src/candidate/features.rs:220-410), after precursor_idx and rank
FEATURES = [
    "score", "mean_correlation", "median_correlation", "correlation_std",
    "intensity_correlation", "num_fragments", "num_scans", "num_over_95",
    "num_over_90", "num_over_80", "num_over_50", "num_over_0",
    "num_over_0_rank_0_5", "num_over_0_rank_6_11", "num_over_0_rank_12_17",
    "num_over_0_rank_18_23", "num_over_50_rank_0_5", "num_over_50_rank_6_11",
    "num_over_50_rank_12_17", "num_over_50_rank_18_23",
    "hyperscore_intensity_observation", "hyperscore_intensity_library",
    "hyperscore_inverse_mass_error", "rt_observed", "delta_rt",
    "longest_b_series", "longest_y_series", "naa", "weighted_mass_error",
    "log10_b_ion_intensity", "log10_y_ion_intensity", "fwhm_rt",
    "idf_hyperscore", "idf_xic_dot_product", "idf_intensity_dot_product",
    "median_profile_sum", "median_profile_sum_filtered", "num_profiles",
    "num_profiles_filtered", "num_over_0_top6_idf", "num_over_50_top6_idf",
]


def score(dia: Dia, library: Lib, candidates, mass_tolerance=7.0, top_k_fragments=100,
          nthreads=0):
    """PeakGroupScoring.score of a candidate table (the 9 to_arrays columns).
    Returns a dict of the 43 to_dict_arrays columns, rows whose precursor is
    not in the library dropped."""
    pidx, rank, sc, _s0, _s1, _s2, cc, cs, ce = candidates
    n = len(pidx)
    keep = [np.ascontiguousarray(pidx, np.uint64), c32(sc),
            np.ascontiguousarray(cs, np.uint64), np.ascontiguousarray(ce, np.uint64),
            np.ascontiguousarray(cc, np.uint64)]
    feat = np.zeros((len(FEATURES), max(n, 1)), np.float32)
    found = np.zeros(max(n, 1), np.uint8)
    nt = nthreads or (os.cpu_count() or 1)
    lib().orc_score(dia.h, library.h, float(mass_tolerance), int(top_k_fragments), n,
                    p(keep[0]), p(keep[1]), p(keep[2]), p(keep[3]), p(keep[4]),
                    p(feat), p(found), nt)
    m = found[:n].astype(bool)
    out = {"precursor_idx": np.asarray(pidx, np.uint64)[m],
           "rank": np.asarray(rank, np.uint64)[m]}
    for i, name in enumerate(FEATURES):
        out[name] = feat[i, :n][m].copy()
    return out


# A planted mistake in the quantification layout below, for the tests that
# show the layout's pins catch it (0: none; set only by those tests).
LAYOUT_MUTANT = 0

_PQ_FRAGMENT_KEYS = ("fragment_mz_library", "fragment_mz", "fragment_intensity",
                     "fragment_cardinality", "fragment_charge", "fragment_loss_type",
                     "fragment_number", "fragment_position", "fragment_type",
                     "fragment_mz_observed", "fragment_correlation_observed",
                     "fragment_mass_error_observed")
_PQ_PRECURSOR_KEYS = ("precursor_idx", "mz_library", "mz", "rt_library", "rt", "naa",
                      "rank", "rt_observed")


def filter_fragments_by_intensity(pq: dict, intensity_threshold: float):
    """PrecursorQuantified::filter_fragments_by_intensity
    (src/precursor_quantified/mod.rs:29-94): the fragments whose intensity is
    above the threshold (strictly), in order, every fragment column alike;
    None when no fragment is left.  pq maps the field names of
    PrecursorQuantified (src/precursor_quantified/mod.rs:6-27) to scalars
    and per-fragment arrays."""
    it = np.asarray(pq["fragment_intensity"], np.float32)
    keep = it >= np.float32(intensity_threshold) if LAYOUT_MUTANT == 1 else \
        it > np.float32(intensity_threshold)
    if not keep.any():
        return None
    out = {k: pq[k] for k in _PQ_PRECURSOR_KEYS}
    for k in _PQ_FRAGMENT_KEYS:
        out[k] = np.asarray(pq[k])[keep]
    return out


def speclib_flat_quantified(pqs):
    """SpecLibFlatQuantified::from_precursor_quantified_vec
    (src/speclib_flat_quantified/mod.rs:213-330) and to_dict_arrays
    (:126-211): fragments concatenated in input order with each precursor's
    [start, stop) and its index and rank repeated per fragment; then the
    precursor columns (start and stop included) reordered by a stable sort on
    precursor_idx, the fragment columns left as they are.  Returns the two
    dicts of to_dict_arrays."""
    ptypes = {"precursor_idx": np.uint64, "mz_library": np.float32, "mz": np.float32,
              "rt_library": np.float32, "rt": np.float32, "naa": np.uint8,
              "rank": np.uint64, "rt_observed": np.float32,
              "flat_frag_start_idx": np.uint64, "flat_frag_stop_idx": np.uint64}
    ftypes = {"mz_library": np.float32, "mz": np.float32, "intensity": np.float32,
              "cardinality": np.uint8, "charge": np.uint8, "loss_type": np.uint8,
              "number": np.uint8, "position": np.uint8, "type": np.uint8,
              "mz_observed": np.float32, "correlation_observed": np.float32,
              "mass_error_observed": np.float32, "precursor_idx": np.uint64,
              "rank": np.uint64}
    prec = {k: [] for k in ptypes}
    frag = {k: [] for k in ftypes}
    pos = 0
    for pq in pqs:
        for k in _PQ_PRECURSOR_KEYS:
            prec[k].append(pq[k])
        nfr = len(pq["fragment_mz"])
        prec["flat_frag_start_idx"].append(pos)
        pos += nfr
        prec["flat_frag_stop_idx"].append(pos)
        for k in _PQ_FRAGMENT_KEYS:
            frag[k[len("fragment_"):]].append(np.asarray(pq[k]))
        frag["precursor_idx"].append(np.full(nfr, pq["precursor_idx"], np.uint64))
        frag["rank"].append(np.full(nfr, pq["rank"], np.uint64))
    pout = {k: np.asarray(v, dtype=ptypes[k]) for k, v in prec.items()}
    fout = {k: (np.concatenate(v).astype(ftypes[k]) if v else np.zeros(0, ftypes[k]))
            for k, v in frag.items()}
    if LAYOUT_MUTANT == 2:  # planted: fragments reordered with their precursors
        srt = np.argsort(pout["precursor_idx"], kind="stable")
        parts = [np.arange(int(a), int(b)) for a, b in
                 zip(pout["flat_frag_start_idx"][srt], pout["flat_frag_stop_idx"][srt])]
        idx = np.concatenate(parts) if parts else np.zeros(0, np.int64)
        fout = {k: v[idx] for k, v in fout.items()}
    srt = np.argsort(pout["precursor_idx"], kind="stable")  # sort_by_key, stable
    pout = {k: v[srt] for k, v in pout.items()}
    return pout, fout


def quantify(dia: Dia, library: Lib, lib_arrays, candidates, tolerance_ppm=7.0,
             top_k_fragments=10000, nthreads=0):
    """PeakGroupQuantification.quantify (src/peak_group_quantification/mod.rs:
    39-168): (precursor dict, fragment dict) as SpecLibFlatQuantified.
    to_dict_arrays (src/speclib_flat_quantified/mod.rs:126-211).  The
    per-candidate work (:86-166) is the C restatement orc_quantify; the
    PrecursorQuantified of each candidate, the intensity > 0 filter and the
    flat layout are filter_fragments_by_intensity and
    speclib_flat_quantified above."""
    (pidx, pmzl, pmz, prtl, prt, naa, fs, fe, fmzl, fmz, fint, card, chg, loss,
     fnum, fpos, ftype) = lib_arrays
    order = np.argsort(pidx, kind="stable")  # SpecLibFlat's stable sort
    cp, _rank, _sc, _a, _b, _c, cc, cs, ce = candidates
    rank = np.asarray(candidates[1], np.uint64)
    n = len(cp)
    ks = int(max(1, (np.asarray(fe, np.int64) - np.asarray(fs, np.int64)).max() if len(fs) else 1))
    k = [np.ascontiguousarray(cp, np.uint64), np.ascontiguousarray(cs, np.uint64),
         np.ascontiguousarray(ce, np.uint64), np.ascontiguousarray(cc, np.uint64)]
    row = np.zeros(max(n, 1), np.uintp)
    nf = np.zeros(max(n, 1), np.uintp)
    raw = np.zeros(max(n * ks, 1), np.uintp)
    inten, mzo, corr, err = (np.zeros(max(n * ks, 1), np.float32) for _ in range(4))
    rto = np.zeros(max(n, 1), np.float32)
    lib().orc_quantify(dia.h, library.h, float(tolerance_ppm), int(top_k_fragments), n,
                       p(k[0]), p(k[1]), p(k[2]), p(k[3]), ks, p(row), p(nf), p(raw),
                       p(inten), p(mzo), p(corr), p(err), p(rto),
                       nthreads or (os.cpu_count() or 1))
    none = np.iinfo(np.uintp).max
    pqs = []
    for i in range(n):
        if row[i] == none:
            continue
        sl = slice(i * ks, i * ks + int(nf[i]))
        r = int(order[int(row[i])])  # the row in the caller's array order
        rw = raw[sl].astype(np.int64)
        pq = {"precursor_idx": pidx[r], "mz_library": pmzl[r], "mz": pmz[r],
              "rt_library": prtl[r], "rt": prt[r], "naa": naa[r], "rank": rank[i],
              "rt_observed": rto[i]}
        for key, src in (("fragment_mz_library", fmzl), ("fragment_mz", fmz),
                         ("fragment_cardinality", card), ("fragment_charge", chg),
                         ("fragment_loss_type", loss), ("fragment_number", fnum),
                         ("fragment_position", fpos), ("fragment_type", ftype)):
            pq[key] = np.asarray(src)[rw]
        pq["fragment_intensity"] = inten[sl]
        pq["fragment_mz_observed"] = mzo[sl]
        pq["fragment_correlation_observed"] = corr[sl]
        pq["fragment_mass_error_observed"] = err[sl]
        kept = filter_fragments_by_intensity(pq, 0.0)  # mod.rs:166
        if kept is not None:
            pqs.append(kept)
    return speclib_flat_quantified(pqs)
