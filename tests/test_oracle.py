"""Pins the CPU oracle (oracle/adsr_oracle.c) before it is trusted.

Three kinds of checks:
  * the reference's own known-answer tests (tests/golden/reference_kats.json),
  * brute force on small inputs (nearest bin, bin ranges, XIC sums, maxima,
    Rust binary search semantics, stable sorts),
  * closed forms and invariants (kernel symmetry and sum, convolution of a
    constant row, score of ln(1+x) weighted sums).
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_kats.json")))
L = oracle.lib


def ppm_index(res, start, end):
    n = L().orc_ppm_index(res, start, end, None, 0)
    out = np.zeros(n, np.float32)
    L().orc_ppm_index(res, start, end, oracle.p(out), n)
    return out


# ----------------------------------------------------------------- grid
@pytest.mark.parametrize("case", GOLD["ppm_index_len"])
def test_ppm_index_len(case):
    assert len(ppm_index(*case["args"])) == case["len"]


def test_ppm_index_first_clamped():
    case = GOLD["ppm_index_first"][0]
    assert ppm_index(*case["args"])[0] == case["first"]


@pytest.mark.parametrize("case", GOLD["ppm_index_props"])
def test_ppm_index_properties(case):
    g = ppm_index(*case["args"])
    res = case["args"][0]
    if "min_first" in case:
        assert g[0] >= case["min_first"]
    if "max_last" in case:
        assert g[-1] <= case["max_last"]
    if "min_len" in case:
        assert len(g) >= case["min_len"]
    if case.get("ratio"):
        # src/mz_index/tests.rs:66-78 verify_geometric_progression
        expected = np.float32(1.0) + np.float32(res) / np.float32(1e6)
        for i in range(1, min(10, len(g))):
            assert abs(np.float32(g[i] / g[i - 1]) - expected) < 1e-6


def test_global_grid_is_the_exact_f32_recurrence():
    g = oracle.grid()
    assert len(g) == 2590781 and g[0] == np.float32(150.0) and g[-1] >= 2000.0
    # recompute with numpy float32 scalars, one rounding per operation
    m = np.float32(150.0)
    r = np.float32(1.0) / np.float32(1e6)
    for i in range(1, 2000):
        m = np.float32(m + np.float32(m * r))
        assert g[i] == m
    assert np.all(np.diff(g.astype(np.float64)) > 0)


def test_find_closest_and_range_brute_force():
    g = oracle.grid()
    gp = oracle.p(g)
    n = len(g)
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(140, 2010, 3000).astype(np.float32),
                         g[rng.integers(0, n, 500)],
                         ((g[:-1][rng.integers(0, n - 1, 500)].astype(np.float64) +
                           g[1:][rng.integers(0, n - 1, 500)].astype(np.float64)) / 2
                          ).astype(np.float32),
                         np.array([0.0, -1.0, 3000.0, g[0], g[-1]], np.float32)])
    for x in xs:
        got = L().orc_find_closest_index(gp, n, float(x))
        lb = int(np.searchsorted(g, x, side="left"))
        cand = [i for i in (lb - 1, lb) if 0 <= i < n]
        d = [abs(np.float32(g[i] - x)) for i in cand]
        # nearest grid point, ties to the lower index (src/mz_index/mod.rs:82-89)
        want = cand[int(np.argmin(d))] if len(cand) == 2 and d[0] <= d[1] or len(cand) == 1 else cand[-1]
        if len(cand) == 2 and d[1] < d[0]:
            want = cand[1]
        if lb < n and g[lb] == x:
            want = lb
        assert got == want, (x, got, want)
    # mz_range_indices: first >= lo, last <= hi (inclusive), src/mz_index/tests.rs:215-280
    s, e = C.c_size_t(), C.c_size_t()
    for lo, hi in [(g[100], g[105]), (g[100], g[100]), (g[0] - 10, g[0] - 10),
                   (g[-1] + 10, g[-1] + 20), (g[0] - 10, g[-1] + 10),
                   (g[0] - 10, g[50]), (g[n - 50], g[-1] + 10)]:
        L().orc_mz_range_indices(gp, n, float(lo), float(hi), C.byref(s), C.byref(e))
        want = np.flatnonzero((g >= lo) & (g <= hi))
        if want.size:
            assert (s.value, e.value) == (want[0], want[-1] + 1)
        else:
            assert s.value == e.value
    assert (GOLD["ppm_index_len"][0]["len"] == 2995974)


# ----------------------------------------------------------------- binary search
def _rust_bsearch_model(v, key):
    """Rust 1.88 slice::binary_search_by, written out from its definition."""
    size = len(v)
    if size == 0:
        return False, 0
    base = 0
    while size > 1:
        half = size // 2
        mid = base + half
        if not v[mid] > key:
            base = mid
        size -= half
    if v[base] == key:
        return True, base
    return False, base + (1 if v[base] < key else 0)


def test_rust_bsearch_last_equal_and_insertion_point():
    rng = np.random.default_rng(5)
    for _ in range(400):
        n = int(rng.integers(0, 20))
        v = np.sort(rng.integers(0, 8, n)).astype(np.uint16)
        key = int(rng.integers(0, 9))
        idx = C.c_size_t()
        ok = L().orc_rust_bsearch_u16(oracle.p(v), n, key, C.byref(idx))
        m_ok, m_idx = _rust_bsearch_model(list(v), key)
        assert (bool(ok), idx.value) == (m_ok, m_idx)
        if ok:  # the LAST element equal to key
            assert v[idx.value] == key and (idx.value == n - 1 or v[idx.value + 1] != key)
        else:   # insertion point
            assert np.all(v[: idx.value] < key) and np.all(v[idx.value:] > key)


# ----------------------------------------------------------------- RT index
def _rt_axis():
    return (np.arange(101, dtype=np.float32) * np.float32(0.5)).astype(np.float32)


def _limits(args):
    rt = _rt_axis()
    dt = L().orc_mean_delta_t(oracle.p(rt), len(rt))
    lo, hi = C.c_size_t(), C.c_size_t()
    L().orc_cycle_idx_limits(oracle.p(rt), len(rt), dt, args[0], args[1], args[2],
                             C.byref(lo), C.byref(hi))
    return lo.value, hi.value


@pytest.mark.parametrize("case", GOLD["rt_limits_exact"])
def test_rt_limits_exact(case):
    assert list(_limits(case["args"])) == case["expect"]


@pytest.mark.parametrize("case", GOLD["rt_limits_range"])
def test_rt_limits_range(case):
    lo, hi = _limits(case["args"])
    assert case["lower"][0] <= lo <= case["lower"][1]
    assert case["upper"][0] <= hi <= case["upper"][1]


@pytest.mark.parametrize("case", GOLD["rt_limits_width"])
def test_rt_limits_width(case):
    lo, hi = _limits(case["args"])
    assert case["width"][0] <= hi - lo <= case["width"][1]


def test_rt_limits_empty_index():
    lo, hi = C.c_size_t(), C.c_size_t()
    L().orc_cycle_idx_limits(None, 0, 1.0, 25.0, 5.0, 5, C.byref(lo), C.byref(hi))
    assert (lo.value, hi.value) == (0, 0)


def _dia_from(arrays, cycle_shape=(1, 1, 1, 1)):
    a = arrays
    return (np.array(a["spectrum_delta_scan_idx"], np.int64),
            np.array(a["isolation_lower_mz"], np.float32),
            np.array(a["isolation_upper_mz"], np.float32),
            np.array(a["spectrum_peak_start_idx"], np.int64),
            np.array(a["spectrum_peak_stop_idx"], np.int64),
            np.array(a["spectrum_cycle_idx"], np.int64),
            np.array(a["spectrum_rt"], np.float32),
            np.array(a["peak_mz"], np.float32),
            np.array(a["peak_intensity"], np.float32),
            np.zeros(cycle_shape, np.float32))


def _rt_spectra(total, cycle_len):
    return {"spectrum_delta_scan_idx": [i % cycle_len for i in range(total)],
            "isolation_lower_mz": [400.0] * total, "isolation_upper_mz": [500.0] * total,
            "spectrum_peak_start_idx": [0] * total, "spectrum_peak_stop_idx": [0] * total,
            "spectrum_cycle_idx": [0] * total,
            "spectrum_rt": list((np.arange(total, dtype=np.float32) * np.float32(0.1))),
            "peak_mz": [], "peak_intensity": []}


def test_rt_from_spectra():
    c0, c1 = GOLD["rt_from_spectra"]
    d = oracle.Dia(_dia_from(_rt_spectra(c0["total_spectra"], c0["cycle_len"])))
    assert np.allclose(d.rt(), c0["rt"])
    d = oracle.Dia(_dia_from(_rt_spectra(c1["total_spectra"], c1["cycle_len"])))
    assert abs(d.h.contents.delta_t - c1["delta_t"]) < c1["tol"]


# ----------------------------------------------------------------- CSR / XIC
def _raw_csr(slices, n_slices, n_grid):
    ss = np.zeros(n_grid + 1, np.uint32)
    cyc, it = [], []
    for b in range(n_slices):
        if b < len(slices):
            cyc += slices[b][0]
            it += slices[b][1]
        ss[b + 1] = len(cyc)
    ss[n_slices + 1:] = len(cyc)
    return ss, np.array(cyc, np.uint16), np.array(it, np.float32)


@pytest.mark.parametrize("case", GOLD["fill_xic"])
def test_fill_xic_slice(case):
    g = oracle.grid()
    ss, cyc, it = _raw_csr(case["slices"], case["n_slices"], len(g))
    c0, c1 = case["c"]
    x = np.zeros(c1 - c0, np.float32)
    L().orc_fill_xic_slice_raw(oracle.p(g), len(g), oracle.p(ss), oracle.p(cyc),
                               oracle.p(it), oracle.p(x), c0, c1, case["tol"],
                               float(g[case["grid_idx"]]))
    if "expect" in case:
        assert x.tolist() == case["expect"]
    for k, v in case.get("check", {}).items():
        assert x[int(k)] == v


def test_slice_layout():
    case = GOLD["slice_layout"]
    ss, cyc, it = _raw_csr(case["slices"], 2, 2)
    assert ss[:3].tolist() == case["slice_starts"]
    assert cyc.tolist() == case["cycle_indices"] and it.tolist() == case["intensities"]


def test_builder_three_windows():
    case = GOLD["builder_three_windows"]
    d = oracle.Dia(_dia_from(case["arrays"]))
    assert d.n_obs == case["num_observations"]
    for o, w in enumerate(case["windows"]):
        assert d.observation(o)["isolation_window"].tolist() == w
    for mz, want in case["valid"]:
        assert d.valid_observations(mz) == want


def test_builder_cycle_order():
    case = GOLD["builder_cycle_order"]
    d = oracle.Dia(_dia_from(case["arrays"]))
    g = oracle.grid()
    b = L().orc_find_closest_index(oracle.p(g), len(g), case["mz"])
    o = d.observation(0)
    s, e = o["slice_starts"][b], o["slice_starts"][b + 1]
    assert o["cycle_indices"][s:e].tolist() == case["cycles"]
    assert o["intensities"][s:e].tolist() == case["intensities"]


@pytest.mark.parametrize("run", GOLD["dense_xic_runs"], ids=lambda r: r["source"])
def test_dense_xic(run):
    d = oracle.Dia(_dia_from(run["arrays"]))
    for case in run["cases"]:
        c0, c1 = case["c"]
        x = d.dense_xic(case["precursor_mz"], c0, c1, case["tol"], case["frags"])
        if "shape" in case:
            assert list(x.shape) == case["shape"], case["source"]
        if "values" in case:
            assert x.tolist() == case["values"], case["source"]


def test_dense_xic_brute_force():
    """XIC via the CSR equals snapping every peak and testing its bin's grid
    value against [mz - d, mz + d] (duplicate cycles summed in peak order)."""
    rng = np.random.default_rng(11)
    g = oracle.grid()
    n_spec, W = 60, 3
    peaks = 40
    mz = np.sort(rng.uniform(499.9, 500.1, (n_spec, peaks)).astype(np.float32), axis=1)
    it = rng.integers(1, 50, (n_spec, peaks)).astype(np.float32)
    arrays = {"spectrum_delta_scan_idx": [s % W for s in range(n_spec)],
              "isolation_lower_mz": [450.0] * n_spec, "isolation_upper_mz": [550.0] * n_spec,
              "spectrum_peak_start_idx": [s * peaks for s in range(n_spec)],
              "spectrum_peak_stop_idx": [(s + 1) * peaks for s in range(n_spec)],
              "spectrum_cycle_idx": [s // W for s in range(n_spec)],
              "spectrum_rt": [float(s // W) for s in range(n_spec)],
              "peak_mz": mz.reshape(-1).tolist(), "peak_intensity": it.reshape(-1).tolist()}
    d = oracle.Dia(_dia_from(arrays))
    gp = oracle.p(g)
    skipped = []
    for frag in [500.0, 500.03, 499.95]:
        tol = 50.0
        c0, c1 = 3, 17
        x = d.dense_xic(500.0, c0, c1, tol, [frag])[0]
        fm = np.float32(frag)
        delta = np.float32(np.float32(fm * np.float32(tol)) * np.float32(1e-6))
        lo, hi = np.float32(fm - delta), np.float32(fm + delta)
        want = np.zeros(c1 - c0, np.float32)
        for w in range(W):  # windows ascending, then bins, then cycles, then peak order
            ent = []
            for s in range(w, n_spec, W):
                cyc = s // W
                for k in range(peaks):
                    b = L().orc_find_closest_index(gp, len(g), float(mz[s, k]))
                    if lo <= g[b] <= hi and c0 <= cyc < c1:
                        ent.append((b, cyc, s * peaks + k, it[s, k]))
            ent.sort(key=lambda e: (e[0], e[1], e[2]))
            # fill_xic_slice starts each bin at binary_search(c0): when c0 is
            # present that is the LAST equal entry, so earlier c0 entries of
            # the same bin are not summed (src/quadrupole_observation/mod.rs:92)
            for i, (b, cyc, _, v) in enumerate(ent):
                if cyc == c0 and i + 1 < len(ent) and ent[i + 1][0] == b and \
                        ent[i + 1][1] == c0:
                    skipped.append(v)
                    continue
                want[cyc - c0] = np.float32(want[cyc - c0] + v)
        assert np.array_equal(x, want)
    assert skipped, "the data should exercise duplicate entries at cycle_start"


# ----------------------------------------------------------------- kernel
def test_kernel_kats():
    for case in GOLD["kernel"]:
        k = oracle.gaussian_kernel(*case["args"])
        assert len(k) == case["len"]
        if "sum" in case:
            assert abs(float(np.sum(k, dtype=np.float64)) - case["sum"]) <= case["sum_tol"]
            assert np.all(k <= k[case["center"]])


def test_kernel_closed_form_and_symmetry():
    k = oracle.gaussian_kernel(4.7096, 1.0, 20, 1.0)  # sigma exactly 2.0
    x = np.arange(-10, 10, dtype=np.float64)
    ref = np.exp(-0.5 * x * x / 4.0)
    ref /= ref.sum()
    assert np.allclose(k, ref, rtol=2e-6, atol=0)
    for d in range(1, 10):
        assert k[10 - d] == k[10 + d]
    # ndarray unrolled_fold: 8 partial sums, combined (0+4),(1+5),(2+6),(3+7)
    v = np.arange(1, 20, dtype=np.float32) * np.float32(0.1)
    p = [np.float32(0)] * 8
    for i in range(16):
        p[i % 8] = np.float32(p[i % 8] + v[i])
    acc = np.float32(0)
    for a, b in [(0, 4), (1, 5), (2, 6), (3, 7)]:
        acc = np.float32(acc + np.float32(p[a] + p[b]))
    for i in range(16, 19):
        acc = np.float32(acc + v[i])
    assert L().orc_ndarray_sum(oracle.p(v), len(v)) == acc


# ----------------------------------------------------------------- convolution
@pytest.mark.parametrize("case", GOLD["convolution_shapes"], ids=lambda c: c["source"])
def test_convolution_shapes_and_edges(case):
    k = oracle.gaussian_kernel(*case["kernel"])
    x = np.array(case["xic"], np.float32)
    y = oracle.convolution(k, x)
    assert y.shape == x.shape
    half = len(k) // 2
    if x.shape[1] < len(k):
        assert np.all(y == 0)
    else:
        assert np.all(y[:, :half] == 0) and np.all(y[:, x.shape[1] - half:] == 0)


def test_convolution_constant_row_closed_form():
    k = oracle.gaussian_kernel(4.7096, 1.0, 20, 1.0)
    c = np.float32(3.0)
    x = np.full((1, 40), c, np.float32)
    y = oracle.convolution(k, x)[0]
    s = np.float32(0.0)
    s = np.float32(s + np.float32(c * k[10]))
    for j in range(10):
        s = np.float32(s + np.float32(np.float32(c + c) * k[j]))
    assert np.all(y[10:30] == s) and np.all(y[:10] == 0) and np.all(y[30:] == 0)
    # effective taps sum to 1 + k[0] (the +-10 pair both use k[0])
    assert abs(float(s) / 3.0 - (1.0 + float(k[0]))) < 1e-6


def test_convolution_impulse_is_the_mirrored_kernel():
    k = oracle.gaussian_kernel(4.7096, 1.0, 20, 1.0)
    x = np.zeros((1, 60), np.float32)
    x[0, 30] = 1.0
    y = oracle.convolution(k, x)[0]
    for d in range(-10, 11):
        want = k[10] if d == 0 else k[10 - abs(d)]
        assert y[30 + d] == np.float32(np.float32(0.0) + want)


# ----------------------------------------------------------------- score
def test_score_kat_and_closed_form():
    case = GOLD["score"][0]
    a = np.array(case["array"], np.float32)
    w = np.array(case["weights"], np.float32)
    r = oracle.axis_log_dot_product(a, w)
    assert np.allclose(r, case["expect"], rtol=0, atol=case["tol"])
    ref = (np.log1p(a.astype(np.float64)) * w[:, None]).sum(axis=0)
    assert np.allclose(r, ref, rtol=1e-6)
    # glibc logf agrees with a float64 log to within one float ulp
    xs = np.random.default_rng(1).uniform(1, 1e6, 2000).astype(np.float32)
    for x in xs[:200]:
        assert abs(L().orc_logf(float(x)) - math.log(float(x))) <= 2 ** -22 * max(1.0, math.log(float(x)))


def test_logf_range_checker_counts_planted_mismatches():
    """orc_logf_check_range (the checker of the exhaustive device-ln test):
    on glibc's own values it finds nothing; planted bit flips are counted
    and the first one located; a NaN with another payload is no mismatch."""
    import ctypes as C
    L = oracle.lib()
    first, n = 0x3f7ff000, 8192          # around 1.0, both sides
    x = np.arange(first, first + n, dtype=np.uint32).view(np.float32)
    y = np.array([L.orc_logf(float(v)) for v in x], np.float32)
    fb = C.c_size_t(0)
    yp = y.ctypes.data_as(C.c_void_p)
    assert L.orc_logf_check_range(first, n, yp, 3, C.byref(fb)) == 0 and fb.value == n
    yb = y.view(np.uint32)
    for i in (17, 4000, 8191):
        yb[i] ^= 1
    assert L.orc_logf_check_range(first, n, yp, 4, C.byref(fb)) == 3 and fb.value == 17
    # NaN inputs: glibc's NaN and another payload compare equal
    nan_first = 0x7fc00000
    yn = np.full(64, np.float32(np.nan))
    yn.view(np.uint32)[5] = 0x7fc0dead
    assert L.orc_logf_check_range(nan_first, 64, yn.ctypes.data_as(C.c_void_p), 2,
                                  C.byref(fb)) == 0
    # ... but a number where glibc gives NaN is one
    yn[9] = 1.0
    assert L.orc_logf_check_range(nan_first, 64, yn.ctypes.data_as(C.c_void_p), 2,
                                  C.byref(fb)) == 1 and fb.value == 9


# ----------------------------------------------------------------- maxima
@pytest.mark.parametrize("case", GOLD["local_maxima"], ids=lambda c: c["source"])
def test_local_maxima_kats(case):
    idx, val = oracle.find_local_maxima(np.array(case["array"], np.float32), case["offset"])
    assert idx.tolist() == case["indices"] and val.tolist() == case["values"]


def test_local_maxima_brute_force_and_stable_order():
    rng = np.random.default_rng(2)
    for _ in range(200):
        n = int(rng.integers(0, 40))
        a = rng.integers(0, 6, n).astype(np.float32)
        idx, val = oracle.find_local_maxima(a, 7)
        found = [i for i in range(2, n - 2) if a[i - 2] < a[i - 1] < a[i] > a[i + 1] > a[i + 2]]
        want = sorted(found, key=lambda i: -a[i])  # Python sort is stable
        assert idx.tolist() == [i + 7 for i in want]
        assert val.tolist() == [float(a[i]) for i in want]


# ----------------------------------------------------------------- fragments
@pytest.mark.parametrize("case", GOLD["fragments"], ids=lambda c: c["source"])
def test_fragment_kats(case):
    mz = np.array(case["mz"], np.float32)
    it = np.array(case["intensity"], np.float32)
    order = oracle.filter_sort_fragments(mz, it, case["number"], case["type"],
                                         case["non_zero"], case["y1"], case["top_k"])
    assert mz[order].tolist() == case["mz_out"]
    assert it[order].tolist() == case["int_out"]


def test_fragment_tie_multiset():
    case = GOLD["fragments_tie_multiset"][0]
    mz = np.array(case["mz"], np.float32)
    it = np.array(case["intensity"], np.float32)
    order = oracle.filter_sort_fragments(mz, it, case["number"], case["type"], False, False,
                                         case["top_k"])
    assert (it[order] == 15.0).sum() == case["count_15"]
    assert (it[order] == 10.0).sum() == case["count_10"]
    assert np.all(np.diff(mz[order]) >= 0)


def test_fragment_invariants():
    case = GOLD["fragments_invariants"]
    mz = np.array(case["mz"], np.float32)
    it = np.array(case["intensity"], np.float32)
    for nz in case["non_zero"]:
        for k in case["top_k"]:
            order = oracle.filter_sort_fragments(mz, it, [1] * 5, [1] * 5, nz, False, k)
            assert len(order) <= k
            if nz:
                assert np.all(it[order] > 0)
            assert np.all(np.diff(order) > 0)  # m/z ascending == original order here


def test_fragment_selection_brute_force():
    """Top-k by intensity (ties: earlier first) then stable m/z sort, for the
    insertion-sort regime (<= 16 fragments) and k == n-1 (max swap)."""
    rng = np.random.default_rng(9)
    for _ in range(300):
        n = int(rng.integers(1, 16))
        mz = rng.integers(200, 206, n).astype(np.float32)    # many equal m/z
        it = rng.integers(0, 5, n).astype(np.float32)         # many equal intensities
        k = int(rng.integers(0, n + 2))
        order = oracle.filter_sort_fragments(mz, it, [2] * n, [98] * n, True, True, k)
        kept = [j for j in range(n) if it[j] > 0]
        kk = min(k, len(kept))
        if kk < len(kept):
            if kk == 0:
                sel = []
            elif kk == len(kept) - 1:
                m = min(range(len(kept)), key=lambda q: (it[kept[q]], q))
                v = list(kept)
                v[m], v[-1] = v[-1], v[m]
                sel = v[:kk]
            else:
                sel = sorted(kept, key=lambda j: -it[j])[:kk]
        else:
            sel = kept
        want = sorted(sel, key=lambda j: mz[j])  # stable
        assert order.tolist() == want


def test_fragment_introselect_invariants():
    """More than 16 kept fragments and 0 < k < kept - 1: the restated Rust
    1.88 select_nth_unstable_by (introselect) keeps k fragments whose
    intensities are the k largest (every kept one >= every dropped one),
    each fragment at most once, in ascending m/z.  Which of several tied
    fragments at the cut is kept follows the restated algorithm; no
    reference vector fixes it."""
    rng = np.random.default_rng(17)
    for trial in range(400):
        n = int(rng.integers(17, 160))
        it = rng.integers(1, 6 if trial % 2 else 40, n).astype(np.float32)
        mz = rng.uniform(200, 1800, n).astype(np.float32)
        k = int(rng.integers(1, n - 1))
        order = oracle.filter_sort_fragments(mz, it, [2] * n, [98] * n, True, True, k)
        assert len(order) == k and len(set(order.tolist())) == k
        dropped = np.setdiff1d(np.arange(n), order)
        assert it[order].min() >= it[dropped].max()
        assert np.all(np.diff(mz[order]) >= 0)


def test_fragment_introselect_small_and_extreme_cases():
    """<= 16 kept fragments: insertion sort, i.e. the stable order by
    descending intensity; k == kept - 1: the first fragment of the lowest
    intensity is dropped and the last kept fragment takes its place, for
    any length."""
    for n in (17, 40, 100):
        it = np.array([5.0] * n, np.float32)
        it[3] = 1.0
        it[n - 5] = 1.0
        mz = np.arange(n, dtype=np.float32) + 200
        order = oracle.filter_sort_fragments(mz, it, [2] * n, [98] * n, True, True, n - 1)
        assert order.tolist() == [j for j in range(n) if j != 3]


# ----------------------------------------------------------------- library
def test_speclib_sorted_by_precursor_idx():
    case = GOLD["speclib_sort"]
    n = len(case["precursor_idx"])
    nf = len(case["fragment_mz"])
    arrays = (np.array(case["precursor_idx"], np.uint64), None,
              np.array(case["precursor_mz"], np.float32), None,
              np.array(case["precursor_rt"], np.float32), None,
              np.array(case["flat_frag_start_idx"], np.uint64),
              np.array(case["flat_frag_stop_idx"], np.uint64), None,
              np.array(case["fragment_mz"], np.float32),
              np.array(case["fragment_intensity"], np.float32), None, None, None,
              np.ones(nf, np.uint8), None, np.full(nf, 98, np.uint8))
    lib = oracle.Lib(arrays)
    h = C.cast(lib.h, C.POINTER(C.c_size_t * 2))
    assert h.contents[0] == n


def test_selection_end_to_end_invariants(run0, oracle_run0):
    """Planted elution apexes are found as rank-0 candidates near their RT."""
    od, ol = oracle_run0
    spec = run0["spec"]
    cand = oracle.select(od, ol, oracle.params(**spec.selection_config()))
    pidx, rank, score, _, _, _, center, start, stop = cand
    assert np.all(np.diff(pidx.astype(np.int64)) >= 0)
    assert np.all(start == center - spec.peak_length)
    assert np.all(stop <= spec.n_cycles)
    rt = run0["lib"][4]
    planted = set(run0["planted"].tolist())
    r0 = rank == 0
    hit = 0
    for p_, c in zip(pidx[r0], center[r0]):
        if int(p_) in planted and abs(float(c) - float(rt[int(p_)])) <= 2.5:
            hit += 1
    assert hit >= 0.8 * len(planted)
