"""alphadia_search_rs on MI355X (gfx950).

Drop-in for the candidate-selection path of the PyO3 module
``alphadia_search_rs`` (src/lib.rs:74-93): the same class names, method
names, argument order, dtypes and return values for ``DIAData``,
``SpecLibFlat``, ``SelectionParameters``, ``PeakGroupSelection`` and
``CandidateCollection``, plus the SIMD-backend registry functions.  The work
runs in libadsr.so: host builders in C++ and the selection pipeline as
gfx950 HIP kernels.  There is no CPU selection backend.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import time

import numpy as np

from . import _native as N
from ._native import check, ptr

__all__ = [
    "DIAData",
    "SpecLibFlat",
    "SelectionParameters",
    "PeakGroupSelection",
    "CandidateCollection",
    "DeviceCandidateCollection",
    "ScoringParameters",
    "PeakGroupScoring",
    "CandidateFeatureCollection",
    "QuantificationParameters",
    "PeakGroupQuantification",
    "SpecLibFlatQuantified",
    "get_optimal_simd_backend",
    "set_simd_backend",
    "clear_simd_backend",
    "get_current_simd_backend",
]


# ---------------------------------------------------------------------------
# argument conversion (PyReadonlyArray1<T> semantics: exact dtype, 1-D)
# ---------------------------------------------------------------------------
def _arr(name: str, a, dtype, ndim: int = 1) -> np.ndarray:
    if not isinstance(a, np.ndarray):
        raise TypeError(
            f"argument '{name}': '{type(a).__name__}' object cannot be "
            f"converted to 'PyArray<{np.dtype(dtype).name}, {ndim}>'"
        )
    if a.dtype != np.dtype(dtype) or a.ndim != ndim:
        raise TypeError(
            f"argument '{name}': type mismatch: expected ndarray of "
            f"{np.dtype(dtype).name} with {ndim} dimension(s), got "
            f"{a.dtype.name} with {a.ndim}"
        )
    return np.ascontiguousarray(a)


def _same_len(names, arrays):
    n = len(arrays[0])
    for nm, a in zip(names, arrays):
        if len(a) != n:
            raise ValueError(
                f"array '{nm}' has length {len(a)}, expected {n} "
                f"(same as '{names[0]}')"
            )
    return n


# ---------------------------------------------------------------------------
# SIMD backend registry (src/lib.rs:53-72, src/simd/mod.rs:80-134)
# ---------------------------------------------------------------------------
def _name(fn) -> str:
    buf = C.create_string_buffer(64)
    check(fn(buf, 64))
    return buf.value.decode()


def get_optimal_simd_backend() -> str:
    return _name(N.lib.adsr_backend_optimal)


def get_current_simd_backend() -> str:
    return _name(N.lib.adsr_backend_current)


def set_simd_backend(backend_name: str) -> None:
    check(N.lib.adsr_backend_set(str(backend_name).encode()))


def clear_simd_backend() -> None:
    N.lib.adsr_backend_clear()


# ---------------------------------------------------------------------------
# DIAData (src/dia_data/mod.rs:15-135)
# ---------------------------------------------------------------------------
class DIAData:
    """Per-quadrupole-window 1-ppm binned peak index, resident in HBM."""

    def __init__(self):
        self._h = None
        self._n_obs = 0

    def __del__(self, _free=N.lib.adsr_dia_free):
        # the free function is bound at definition: module globals may be
        # gone when the interpreter shuts down
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    @staticmethod
    def from_arrays(
        spectrum_delta_scan_idx,
        isolation_lower_mz,
        isolation_upper_mz,
        spectrum_peak_start_idx,
        spectrum_peak_stop_idx,
        spectrum_cycle_idx,
        spectrum_rt,
        peak_mz,
        peak_intensity,
        cycle,
        device_build=None,
    ) -> "DIAData":
        """src/dia_data/mod.rs:43-72.

        ``device_build`` picks where the binned CSR is built: True on the
        gfx950 device (the CSR then lives in HBM only), False on host
        threads, None (default) on the device when one is present.  Both
        give the same CSR, bit for bit.
        """
        d = _arr("spectrum_delta_scan_idx", spectrum_delta_scan_idx, np.int64)
        lo = _arr("isolation_lower_mz", isolation_lower_mz, np.float32)
        hi = _arr("isolation_upper_mz", isolation_upper_mz, np.float32)
        ps = _arr("spectrum_peak_start_idx", spectrum_peak_start_idx, np.int64)
        pe = _arr("spectrum_peak_stop_idx", spectrum_peak_stop_idx, np.int64)
        cy = _arr("spectrum_cycle_idx", spectrum_cycle_idx, np.int64)
        rt = _arr("spectrum_rt", spectrum_rt, np.float32)
        mz = _arr("peak_mz", peak_mz, np.float32)
        it = _arr("peak_intensity", peak_intensity, np.float32)
        cyc = _arr("cycle", cycle, np.float32, ndim=4)
        names = ["spectrum_delta_scan_idx", "isolation_lower_mz",
                 "isolation_upper_mz", "spectrum_peak_start_idx",
                 "spectrum_peak_stop_idx", "spectrum_cycle_idx", "spectrum_rt"]
        ns = _same_len(names, [d, lo, hi, ps, pe, cy, rt])
        npk = _same_len(["peak_mz", "peak_intensity"], [mz, it])
        shape = (C.c_size_t * 4)(*cyc.shape)
        out = C.c_void_p()
        on_device = N.device_available() if device_build is None else bool(device_build)
        create = N.lib.adsr_dia_create_device if on_device else N.lib.adsr_dia_create
        check(create(ptr(d), ptr(lo), ptr(hi), ptr(ps), ptr(pe), ptr(cy), ptr(rt),
                     ns, ptr(mz), ptr(it), npk, ptr(cyc), shape, C.byref(out)))
        obj = DIAData()
        obj._h = out.value
        obj._n_obs = N.lib.adsr_dia_num_observations(obj._h)
        if not on_device and N.device_available():
            check(N.lib.adsr_dia_upload(obj._h))
        return obj

    def build_timings(self) -> dict:
        """Milliseconds of a device build: host preparation, peak copy to
        the device, device kernels (all 0 for a host-built CSR)."""
        v = [C.c_double(0.0) for _ in range(3)]
        check(N.lib.adsr_dia_build_timings(self._h, *(C.byref(x) for x in v)))
        return {"prepare_ms": v[0].value, "copy_ms": v[1].value,
                "kernels_ms": v[2].value}

    @property
    def num_observations(self) -> int:
        return int(self._n_obs)

    def get_valid_observations(self, precursor_mz: float) -> list:
        if not self._h:
            return []
        n = C.c_size_t()
        buf = np.zeros(max(self._n_obs, 1), dtype=np.uint64)
        check(N.lib.adsr_dia_valid_observations(self._h, float(precursor_mz),
                                                ptr(buf), len(buf), C.byref(n)))
        return [int(x) for x in buf[: n.value]]

    def memory_footprint_bytes(self) -> int:
        if not self._h:
            return 24  # an empty Vec<QuadrupoleObservation>
        return int(N.lib.adsr_dia_memory_footprint_bytes(self._h))

    def memory_footprint_mb(self) -> float:
        return self.memory_footprint_bytes() / (1024.0 * 1024.0)

    @property
    def has_mobility(self) -> bool:
        return False

    @property
    def has_ms1(self) -> bool:
        return False

    @property
    def mobility_values(self) -> list:
        return [1e-6, 0.0]

    @property
    def rt_values(self) -> np.ndarray:
        if not self._h:
            return np.zeros(0, dtype=np.float32)
        n = N.lib.adsr_dia_rt_values(self._h, None, 0)
        out = np.empty(n, dtype=np.float32)
        N.lib.adsr_dia_rt_values(self._h, ptr(out), n)
        return out

    @property
    def cycle(self) -> np.ndarray:
        if not self._h:
            return np.zeros((0, 0, 0, 0), dtype=np.float32)
        shape = (C.c_size_t * 4)()
        N.lib.adsr_dia_cycle_shape(self._h, shape)
        out = np.empty(tuple(shape), dtype=np.float32)
        N.lib.adsr_dia_cycle_copy(self._h, ptr(out))
        return out

    # -- internal views used by tests and tools ------------------------------
    def _rt_index(self):
        n = N.lib.adsr_dia_rt_index(self._h, None, 0, None)
        out = np.empty(n, dtype=np.float32)
        dt = C.c_float()
        N.lib.adsr_dia_rt_index(self._h, ptr(out), n, C.byref(dt))
        return out, dt.value

    def _observation(self, i: int) -> dict:
        win = np.zeros(2, dtype=np.float32)
        nc = C.c_uint64()
        nss = C.c_size_t()
        ne = C.c_size_t()
        check(N.lib.adsr_dia_observation(self._h, i, ptr(win), C.byref(nc),
                                         None, C.byref(nss), None, None,
                                         C.byref(ne)))
        ss = np.empty(nss.value, dtype=np.uint32)
        ci = np.empty(ne.value, dtype=np.uint16)
        it = np.empty(ne.value, dtype=np.float32)
        check(N.lib.adsr_dia_observation(self._h, i, None, None, ptr(ss),
                                         C.byref(nss), ptr(ci), ptr(it),
                                         C.byref(ne)))
        return {"isolation_window": win, "num_cycles": nc.value,
                "slice_starts": ss, "cycle_indices": ci, "intensities": it}


# ---------------------------------------------------------------------------
# SpecLibFlat (src/speclib_flat/mod.rs:6-199)
# ---------------------------------------------------------------------------
class SpecLibFlat:
    """Flat spectral library, precursor arrays sorted by precursor_idx."""

    def __init__(self):
        self._h = None
        self._np = 0
        self._nf = 0

    def __del__(self, _free=N.lib.adsr_lib_free):
        # the free function is bound at definition: module globals may be
        # gone when the interpreter shuts down
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    @staticmethod
    def from_arrays(
        precursor_idx,
        precursor_mz_library,
        precursor_mz,
        precursor_rt_library,
        precursor_rt,
        precursor_naa,
        flat_frag_start_idx,
        flat_frag_stop_idx,
        fragment_mz_library,
        fragment_mz,
        fragment_intensity,
        fragment_cardinality,
        fragment_charge,
        fragment_loss_type,
        fragment_number,
        fragment_position,
        fragment_type,
    ) -> "SpecLibFlat":
        """src/speclib_flat/mod.rs:98-188."""
        u64, f32, u8 = np.uint64, np.float32, np.uint8
        p = [
            _arr("precursor_idx", precursor_idx, u64),
            _arr("precursor_mz_library", precursor_mz_library, f32),
            _arr("precursor_mz", precursor_mz, f32),
            _arr("precursor_rt_library", precursor_rt_library, f32),
            _arr("precursor_rt", precursor_rt, f32),
            _arr("precursor_naa", precursor_naa, u8),
            _arr("flat_frag_start_idx", flat_frag_start_idx, u64),
            _arr("flat_frag_stop_idx", flat_frag_stop_idx, u64),
        ]
        f = [
            _arr("fragment_mz_library", fragment_mz_library, f32),
            _arr("fragment_mz", fragment_mz, f32),
            _arr("fragment_intensity", fragment_intensity, f32),
            _arr("fragment_cardinality", fragment_cardinality, u8),
            _arr("fragment_charge", fragment_charge, u8),
            _arr("fragment_loss_type", fragment_loss_type, u8),
            _arr("fragment_number", fragment_number, u8),
            _arr("fragment_position", fragment_position, u8),
            _arr("fragment_type", fragment_type, u8),
        ]
        n_prec = _same_len(["precursor_idx"] * len(p), p)
        n_frag = _same_len(["fragment_mz_library"] * len(f), f)
        out = C.c_void_p()
        check(N.lib.adsr_lib_create(n_prec, *[ptr(a) for a in p], n_frag,
                                    *[ptr(a) for a in f], C.byref(out)))
        obj = SpecLibFlat()
        obj._h = out.value
        obj._np = n_prec
        obj._nf = n_frag
        if N.device_available():
            check(N.lib.adsr_lib_upload(obj._h))
        return obj

    @property
    def num_precursors(self) -> int:
        return int(self._np)

    @property
    def num_fragments(self) -> int:
        return int(self._nf)


# ---------------------------------------------------------------------------
# SelectionParameters (src/peak_group_selection/parameters.rs:4-75)
# ---------------------------------------------------------------------------
def _extract_f32(key, v):
    if isinstance(v, (bool, np.bool_)) or not isinstance(
            v, (int, float, np.integer, np.floating)):
        raise TypeError(f"'{type(v).__name__}' object cannot be converted to "
                        f"'PyFloat' (key '{key}')")
    return float(np.float32(v))


def _extract_usize(key, v):
    if isinstance(v, (float, np.floating)) or not isinstance(
            v, (int, np.integer)):
        raise TypeError(f"'{type(v).__name__}' object cannot be interpreted "
                        f"as an integer (key '{key}')")
    v = int(v)
    if v < 0 or v >= 2 ** 64:
        raise OverflowError(f"can't convert {v} to usize (key '{key}')")
    return v


class SelectionParameters:
    """Parameters of PeakGroupSelection with the reference defaults.

    Attributes are read-only; ``update(dict)`` changes them.  Two extension
    keys are accepted by ``update``: ``rt_tolerance_per_precursor`` (float32
    array in library order, i.e. sorted by precursor_idx) and
    ``sigma_per_window`` (float32 array, one Gaussian sigma in cycles per
    quadrupole observation).  Without them results equal the reference.
    """

    _F32 = ("fwhm_rt", "mass_tolerance", "rt_tolerance")
    _USIZE = ("kernel_size", "peak_length", "candidate_count",
              "top_k_fragments")

    def __init__(self):
        object.__setattr__(self, "_v", {
            "fwhm_rt": 3.0,
            "kernel_size": 20,
            "peak_length": 5,
            "mass_tolerance": 7.0,
            "rt_tolerance": 200.0,
            "candidate_count": 3,
            "top_k_fragments": 12,
            "rt_tolerance_per_precursor": None,
            "sigma_per_window": None,
        })

    def __getattr__(self, name):
        v = object.__getattribute__(self, "_v")
        if name in v:
            return v[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        raise AttributeError(
            f"attribute '{name}' of 'SelectionParameters' objects is not "
            "writable")

    def update(self, config: dict) -> None:
        if not isinstance(config, dict):
            raise TypeError("update() expects a dict")
        v = object.__getattribute__(self, "_v")
        new = dict(v)
        for k in self._F32:
            if k in config:
                new[k] = _extract_f32(k, config[k])
        for k in self._USIZE:
            if k in config:
                new[k] = _extract_usize(k, config[k])
        for k in ("rt_tolerance_per_precursor", "sigma_per_window"):
            if k in config:
                a = config[k]
                new[k] = None if a is None else np.ascontiguousarray(
                    np.asarray(a, dtype=np.float32).reshape(-1))
        v.update(new)

    def _c(self) -> N.SelectionParamsC:
        v = object.__getattribute__(self, "_v")
        c = N.SelectionParamsC()
        c.fwhm_rt = v["fwhm_rt"]
        c.kernel_size = v["kernel_size"]
        c.peak_length = v["peak_length"]
        c.mass_tolerance = v["mass_tolerance"]
        c.rt_tolerance = v["rt_tolerance"]
        c.candidate_count = v["candidate_count"]
        c.top_k_fragments = v["top_k_fragments"]
        rp = v["rt_tolerance_per_precursor"]
        sw = v["sigma_per_window"]
        c.rt_tolerance_per_precursor = ptr(rp) if rp is not None else None
        c.sigma_per_window = ptr(sw) if sw is not None else None
        c.n_sigma_per_window = 0 if sw is None else len(sw)
        return c

    def __repr__(self):
        v = object.__getattribute__(self, "_v")
        keys = self._F32 + self._USIZE
        return "SelectionParameters(" + ", ".join(
            f"{k}={v[k]}" for k in keys) + ")"


# ---------------------------------------------------------------------------
# CandidateCollection (src/candidate/entry.rs:41-193)
# ---------------------------------------------------------------------------
class CandidateCollection:
    def __init__(self):
        self._h = None

    def __del__(self, _free=N.lib.adsr_candidates_free):
        # the free function is bound at definition: module globals may be
        # gone when the interpreter shuts down
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    @staticmethod
    def _wrap(handle) -> "CandidateCollection":
        c = CandidateCollection()
        c._h = handle
        return c

    def len(self) -> int:
        return int(N.lib.adsr_candidates_len(self._h)) if self._h else 0

    def __len__(self) -> int:
        return self.len()

    def is_empty(self) -> bool:
        return self.len() == 0

    @staticmethod
    def from_arrays(precursor_idxs, ranks, scores, scan_center, scan_start,
                    scan_stop, cycle_center, cycle_start, cycle_stop
                    ) -> "CandidateCollection":
        cols = [precursor_idxs, ranks, scores, scan_center, scan_start,
                scan_stop, cycle_center, cycle_start, cycle_stop]
        arrs = [np.ascontiguousarray(np.asarray(c, dtype=np.float32 if i == 2
                                                else np.uint64))
                for i, c in enumerate(cols)]
        n = len(arrs[0])
        if not all(len(a) == n for a in arrs):
            raise ValueError("All input arrays must have the same length")
        out = C.c_void_p()
        check(N.lib.adsr_candidates_from_arrays(n, *[ptr(a) for a in arrs],
                                                C.byref(out)))
        return CandidateCollection._wrap(out.value)

    def to_arrays(self):
        """(precursor_idx, rank, score, scan_center, scan_start, scan_stop,
        cycle_center, cycle_start, cycle_stop): u64 columns, score f32."""
        n = self.len()
        cols = [np.empty(n, dtype=np.float32 if i == 2 else np.uint64)
                for i in range(9)]
        if self._h:
            check(N.lib.adsr_candidates_export(self._h, *[ptr(a) for a in cols]))
        return tuple(cols)

    # Two file hand-offs.  to_arrow/to_parquet keep the nine to_arrays
    # columns, names and dtypes, so that a table read back is the same
    # CandidateCollection.  to_candidates_frame/to_candidates_parquet (below)
    # write the table of scripts/candidate_selection.py:121-209 instead:
    # frame_* names, rescaled by the run's cycle length, with the library's
    # elution group and decoy columns merged in.
    COLUMNS = ("precursor_idx", "rank", "score", "scan_center", "scan_start",
               "scan_stop", "cycle_center", "cycle_start", "cycle_stop")

    def to_arrow(self):
        """The nine columns as a pyarrow Table (u64 columns, score f32)."""
        import pyarrow as pa
        return pa.table(dict(zip(self.COLUMNS, self.to_arrays())))

    def to_parquet(self, path) -> None:
        """Writes ``to_arrow()`` to a parquet file."""
        import pyarrow.parquet as pq
        pq.write_table(self.to_arrow(), path)

    @staticmethod
    def from_arrow(table) -> "CandidateCollection":
        """The collection of a table with the nine columns (extra columns
        are ignored); ValueError when one is missing."""
        missing = [c for c in CandidateCollection.COLUMNS if c not in table.column_names]
        if missing:
            raise ValueError(f"candidate table lacks columns {missing}")
        return CandidateCollection.from_arrays(
            *[table.column(c).to_numpy() for c in CandidateCollection.COLUMNS])

    @staticmethod
    def read_parquet(path) -> "CandidateCollection":
        import pyarrow.parquet as pq
        return CandidateCollection.from_arrow(pq.read_table(path))

    # scripts/candidate_selection.py:121-168 (parse_candidates): the table
    # alphaDIA reads.  The cycle columns become frame_*, multiplied by the
    # run's cycle length (spectra per cycle); scan_start/center are 0 and
    # scan_stop 1; the library's elution_group_idx and decoy columns are
    # joined by precursor_idx (a left merge: every candidate row kept, in
    # order).
    SCRIPT_COLUMNS = ("precursor_idx", "rank", "score", "scan_center", "scan_start",
                      "scan_stop", "frame_center", "frame_start", "frame_stop")

    def to_candidates_frame(self, precursor_df, cycle_len):
        """The pandas DataFrame of parse_candidates: the nine columns named
        as in SCRIPT_COLUMNS, merged (how="left", on precursor_idx) with
        ``precursor_df[["precursor_idx", "elution_group_idx", "decoy"]]``,
        frame_* multiplied by ``cycle_len``, scan_start = scan_center = 0 and
        scan_stop = 1.  ``precursor_df`` is a DataFrame or a mapping of
        columns; ``cycle_len`` is ``spectrum_df["cycle_idx"].max() + 1``.
        The arithmetic is the script's own pandas expressions, so the
        resulting dtypes follow the type of ``cycle_len`` as they do there."""
        import pandas as pd
        df = pd.DataFrame(dict(zip(self.SCRIPT_COLUMNS, self.to_arrays())))
        lib = precursor_df if isinstance(precursor_df, pd.DataFrame) \
            else pd.DataFrame(dict(precursor_df))
        missing = [c for c in ("precursor_idx", "elution_group_idx", "decoy")
                   if c not in lib.columns]
        if missing:
            raise ValueError(f"precursor_df lacks columns {missing}")
        df = df.merge(lib[["precursor_idx", "elution_group_idx", "decoy"]],
                      on="precursor_idx", how="left")
        df["frame_start"] = df["frame_start"] * cycle_len
        df["frame_stop"] = df["frame_stop"] * cycle_len
        df["frame_center"] = df["frame_center"] * cycle_len
        df["scan_start"] = 0
        df["scan_stop"] = 1
        df["scan_center"] = 0
        return df

    def to_candidates_parquet(self, path, precursor_df, cycle_len) -> None:
        """Writes ``to_candidates_frame(precursor_df, cycle_len)`` to a
        parquet file, as scripts/candidate_selection.py:203-206 does."""
        self.to_candidates_frame(precursor_df, cycle_len).to_parquet(path)


class DeviceCandidateCollection:
    """The candidate table of ``PeakGroupSelection.search_device``, kept in
    device memory: the same rows, in the same order, as ``search``.  Hands
    the table to device-side consumers (scoring) without a host round trip.
    """

    def __init__(self):
        self._h = None

    def __del__(self, _free=N.lib.adsr_device_candidates_free):
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    def len(self) -> int:
        return int(N.lib.adsr_device_candidates_len(self._h)) if self._h else 0

    def __len__(self) -> int:
        return self.len()

    def is_empty(self) -> bool:
        return self.len() == 0

    def to_torch(self, stream=None):
        """The nine ``to_arrays`` columns (src/candidate/entry.rs:124-173) as
        torch tensors on the current device, written by one kernel on
        ``stream`` (default: torch's current stream).  The u64 columns are
        int64 tensors holding the same bits; score is float32."""
        import torch
        n = self.len()
        dev = torch.device("cuda", torch.cuda.current_device())
        cols = [torch.empty(n, dtype=torch.float32 if i == 2 else torch.int64,
                            device=dev) for i in range(9)]
        s = torch.cuda.current_stream(dev) if stream is None else stream
        if self._h and n:
            check(N.lib.adsr_device_candidates_export(
                self._h, *[C.c_void_p(t.data_ptr()) for t in cols],
                C.c_void_p(s.cuda_stream)))
        return tuple(cols)

    def to_dlpack(self, stream=None):
        """The nine columns of ``to_torch`` as DLPack capsules, for consumers
        outside torch; the tensors stay alive through the capsules."""
        from torch.utils.dlpack import to_dlpack
        return tuple(to_dlpack(t) for t in self.to_torch(stream))

    def to_collection(self) -> "CandidateCollection":
        """The host ``CandidateCollection`` of the same table."""
        if not self._h:
            return CandidateCollection()
        out = C.c_void_p()
        check(N.lib.adsr_device_candidates_to_host(self._h, C.byref(out)))
        return CandidateCollection._wrap(out.value)


# ---------------------------------------------------------------------------
# PeakGroupSelection (src/peak_group_selection/mod.rs:66-128)
# ---------------------------------------------------------------------------
class PeakGroupSelection:
    def __init__(self, params: SelectionParameters):
        if not isinstance(params, SelectionParameters):
            raise TypeError("PeakGroupSelection expects SelectionParameters")
        # the reference keeps its own copy of the parameters (mod.rs:75-85)
        p = SelectionParameters()
        object.__getattribute__(p, "_v").update(
            dict(object.__getattribute__(params, "_v")))
        self._params = p

    def search(self, dia_data: DIAData, lib: SpecLibFlat) -> CandidateCollection:
        """Runs on the gfx950 device; raises if no device is available."""
        if not isinstance(dia_data, DIAData) or not isinstance(lib, SpecLibFlat):
            raise TypeError("search(dia_data: DIAData, lib: SpecLibFlat)")
        if not dia_data._h or not lib._h:
            return CandidateCollection()
        c = self._params._c()
        out = C.c_void_p()
        t0 = time.perf_counter()
        check(N.lib.adsr_select(dia_data._h, lib._h, C.byref(c), C.byref(out)))
        dt = time.perf_counter() - t0
        cand = CandidateCollection._wrap(out.value)
        if os.environ.get("ADSR_QUIET", "") in ("", "0"):
            # mod.rs:120-125
            rate = lib.num_precursors / dt if dt > 0 else float("inf")
            print(f"Precursors per second: {rate}")
            print(f"Found {cand.len()} candidates")
            sys.stdout.flush()
        return cand

    def search_device(self, dia_data: DIAData,
                      lib: SpecLibFlat) -> DeviceCandidateCollection:
        """``search`` with the candidate table left in device memory."""
        if not isinstance(dia_data, DIAData) or not isinstance(lib, SpecLibFlat):
            raise TypeError("search_device(dia_data: DIAData, lib: SpecLibFlat)")
        out = DeviceCandidateCollection()
        if not dia_data._h or not lib._h:
            return out
        c = self._params._c()
        h = C.c_void_p()
        check(N.lib.adsr_select_device(dia_data._h, lib._h, C.byref(c), C.byref(h)))
        out._h = h.value
        return out


# ---------------------------------------------------------------------------
# PeakGroupScoring (src/peak_group_scoring/mod.rs:30-102)
# ---------------------------------------------------------------------------
class ScoringParameters:
    """src/peak_group_scoring/parameters.rs:4-40: read-only attributes with
    the reference defaults, changed through ``update(dict)``."""

    def __init__(self):
        object.__setattr__(self, "_v", {"mass_tolerance": 7.0, "top_k_fragments": 100})

    def __getattr__(self, name):
        v = object.__getattribute__(self, "_v")
        if name in v:
            return v[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        raise AttributeError(
            f"attribute '{name}' of 'ScoringParameters' objects is not writable")

    def update(self, config: dict) -> None:
        if not isinstance(config, dict):
            raise TypeError("update() expects a dict")
        v = object.__getattribute__(self, "_v")
        new = dict(v)
        if "mass_tolerance" in config:
            new["mass_tolerance"] = _extract_f32("mass_tolerance", config["mass_tolerance"])
        if "top_k_fragments" in config:
            new["top_k_fragments"] = _extract_usize("top_k_fragments",
                                                    config["top_k_fragments"])
        v.update(new)

    def _c(self) -> N.ScoringParamsC:
        v = object.__getattribute__(self, "_v")
        return N.ScoringParamsC(v["mass_tolerance"], v["top_k_fragments"])


FEATURE_NAMES = [N.lib.adsr_feature_name(i).decode() for i in range(41)]


class CandidateFeatureCollection:
    """src/candidate/features.rs:196-430: one row of features per scored
    candidate."""

    def __init__(self):
        self._h = None

    def __del__(self, _free=N.lib.adsr_features_free):
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    def len(self) -> int:
        return int(N.lib.adsr_features_len(self._h)) if self._h else 0

    def __len__(self) -> int:
        return self.len()

    def is_empty(self) -> bool:
        return self.len() == 0

    def to_dict_arrays(self) -> dict:
        """features.rs:220-410: precursor_idx and rank (uint64) and the 41
        float32 feature columns, keyed by name."""
        n = self.len()
        pidx = np.zeros(n, np.uint64)
        rank = np.zeros(n, np.uint64)
        cols = np.zeros((len(FEATURE_NAMES), n), np.float32)
        if self._h and n:
            check(N.lib.adsr_features_export(self._h, ptr(pidx), ptr(rank), ptr(cols)))
        out = {"precursor_idx": pidx, "rank": rank}
        for i, name in enumerate(FEATURE_NAMES):
            out[name] = cols[i].copy()
        return out

    def device_timings(self) -> dict:
        """Device time of the fragment plan and of the scoring kernel (ms)."""
        a, b = C.c_double(), C.c_double()
        check(N.lib.adsr_features_timings(self._h, C.byref(a), C.byref(b)))
        return {"plan_ms": a.value, "kernel_ms": b.value}


class PeakGroupScoring:
    def __init__(self, params: ScoringParameters):
        if not isinstance(params, ScoringParameters):
            raise TypeError("PeakGroupScoring expects ScoringParameters")
        p = ScoringParameters()
        object.__getattribute__(p, "_v").update(dict(object.__getattribute__(params, "_v")))
        self._params = p

    def score(self, dia_data: DIAData, lib: SpecLibFlat,
              candidates) -> CandidateFeatureCollection:
        """mod.rs:43-102 on the gfx950 device.  ``candidates`` is a
        CandidateCollection or a DeviceCandidateCollection (the table left
        in HBM by search_device)."""
        if not isinstance(dia_data, DIAData) or not isinstance(lib, SpecLibFlat):
            raise TypeError("score(dia_data: DIAData, lib: SpecLibFlat, candidates)")
        out = CandidateFeatureCollection()
        if not dia_data._h or not lib._h or not candidates._h:
            return out
        c = self._params._c()
        h = C.c_void_p()
        t0 = time.perf_counter()
        if isinstance(candidates, DeviceCandidateCollection):
            check(N.lib.adsr_score_device(dia_data._h, lib._h, candidates._h, C.byref(c),
                                          C.byref(h)))
        elif isinstance(candidates, CandidateCollection):
            check(N.lib.adsr_score(dia_data._h, lib._h, candidates._h, C.byref(c), C.byref(h)))
        else:
            raise TypeError("candidates must be a CandidateCollection")
        dt = time.perf_counter() - t0
        out._h = h.value
        if os.environ.get("ADSR_QUIET", "") in ("", "0"):
            # mod.rs:94-99
            rate = len(candidates) / dt if dt > 0 else float("inf")
            print(f"Scored {len(candidates)} candidates at {rate:.2f} candidates/second")
            sys.stdout.flush()
        return out


# ---------------------------------------------------------------------------
# PeakGroupQuantification (src/peak_group_quantification/mod.rs:27-168)
# ---------------------------------------------------------------------------
class QuantificationParameters:
    """src/peak_group_quantification/parameters.rs: ``tolerance_ppm`` and
    ``top_k_fragments``, readable and writable, plus ``update(dict)``."""

    def __init__(self):
        self.tolerance_ppm = 7.0
        self.top_k_fragments = 10000

    def __setattr__(self, name, value):
        if name == "tolerance_ppm":
            value = _extract_f32(name, value)
        elif name == "top_k_fragments":
            value = _extract_usize(name, value)
        else:
            raise AttributeError(f"'QuantificationParameters' object has no attribute '{name}'")
        object.__setattr__(self, name, value)

    def update(self, config: dict) -> None:
        if not isinstance(config, dict):
            raise TypeError("update() expects a dict")
        for k in ("tolerance_ppm", "top_k_fragments"):
            if k in config:
                setattr(self, k, config[k])


class SpecLibFlatQuantified:
    """src/speclib_flat_quantified/mod.rs: the quantified library."""

    def __init__(self):
        self._h = None

    def __del__(self, _free=N.lib.adsr_quantified_free):
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    @property
    def num_precursors(self) -> int:
        return int(N.lib.adsr_quantified_num_precursors(self._h)) if self._h else 0

    @property
    def num_fragments(self) -> int:
        return int(N.lib.adsr_quantified_num_fragments(self._h)) if self._h else 0

    def device_timings(self) -> dict:
        """Device time of the fragment plan and of the quantification
        kernel (ms)."""
        a, b = C.c_double(), C.c_double()
        check(N.lib.adsr_quantified_timings(self._h, C.byref(a), C.byref(b)))
        return {"plan_ms": a.value, "kernel_ms": b.value}

    def to_dict_arrays(self):
        """mod.rs:126-211: (precursor dict, fragment dict)."""
        n, m = self.num_precursors, self.num_fragments
        pk = [("precursor_idx", np.uint64), ("mz_library", np.float32), ("mz", np.float32),
              ("rt_library", np.float32), ("rt", np.float32), ("naa", np.uint8),
              ("rank", np.uint64), ("rt_observed", np.float32),
              ("flat_frag_start_idx", np.uint64), ("flat_frag_stop_idx", np.uint64)]
        fk = [("mz_library", np.float32), ("mz", np.float32), ("intensity", np.float32),
              ("cardinality", np.uint8), ("charge", np.uint8), ("loss_type", np.uint8),
              ("number", np.uint8), ("position", np.uint8), ("type", np.uint8),
              ("mz_observed", np.float32), ("correlation_observed", np.float32),
              ("mass_error_observed", np.float32), ("precursor_idx", np.uint64),
              ("rank", np.uint64)]
        prec = {k: np.zeros(n, t) for k, t in pk}
        frag = {k: np.zeros(m, t) for k, t in fk}
        if self._h:
            check(N.lib.adsr_quantified_export_precursors(self._h,
                                                          *[ptr(prec[k]) for k, _ in pk]))
            check(N.lib.adsr_quantified_export_fragments(self._h,
                                                         *[ptr(frag[k]) for k, _ in fk]))
        return prec, frag


class PeakGroupQuantification:
    def __init__(self, params: QuantificationParameters):
        if not isinstance(params, QuantificationParameters):
            raise TypeError("PeakGroupQuantification expects QuantificationParameters")
        p = QuantificationParameters()
        p.tolerance_ppm = params.tolerance_ppm
        p.top_k_fragments = params.top_k_fragments
        self._params = p

    def quantify(self, dia_data: DIAData, lib: SpecLibFlat,
                 candidates: CandidateCollection) -> SpecLibFlatQuantified:
        """mod.rs:39-82 on the gfx950 device."""
        if not isinstance(dia_data, DIAData) or not isinstance(lib, SpecLibFlat) or \
                not isinstance(candidates, CandidateCollection):
            raise TypeError("quantify(dia_data: DIAData, lib: SpecLibFlat, "
                            "candidates: CandidateCollection)")
        out = SpecLibFlatQuantified()
        if not dia_data._h or not lib._h or not candidates._h:
            return out
        c = N.QuantParamsC(self._params.tolerance_ppm, self._params.top_k_fragments)
        h = C.c_void_p()
        t0 = time.perf_counter()
        check(N.lib.adsr_quantify(dia_data._h, lib._h, candidates._h, C.byref(c), C.byref(h)))
        out._h = h.value
        if os.environ.get("ADSR_QUIET", "") in ("", "0"):
            # mod.rs:72-77
            print(f"Peak group quantification completed in {time.perf_counter() - t0:.2f}s "
                  f"for {out.num_precursors} precursors")
            sys.stdout.flush()
        return out
