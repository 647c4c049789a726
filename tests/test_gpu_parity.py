"""HIP path vs the CPU oracle, through the C ABI (libadsr.so).

Integer, index and candidate columns must match bit for bit; f32 XIC sums,
convolutions and scores are compared bit for bit as well (the device
evaluates the reference's operations in the reference's order, and ln is
bit-identical to glibc logf).
"""
import numpy as np
import pytest

import oracle
import synthetic
from conftest import assert_tables_equal, oracle_params_for, params_for, with_peak_everywhere

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ads():
    import alphadia_search_rs as m
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    return m


def _search(ads, gen, **extra):
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(gen["spec"], **extra)
    if gen.get("rt_tolerance_per_precursor") is not None and \
            "rt_tolerance_per_precursor" not in extra:
        p.update({"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]})
    return d, lib, p, ads.PeakGroupSelection(p).search(d, lib).to_arrays()


def _oracle(gen, **extra):
    od, ol = oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"])
    kw = dict(extra)
    if gen.get("rt_tolerance_per_precursor") is not None and \
            "rt_tolerance_per_precursor" not in kw:
        kw["rt_tolerance_per_precursor"] = gen["rt_tolerance_per_precursor"]
    return oracle.select(od, ol, oracle_params_for(gen["spec"], **kw))


def test_device_logf_matches_glibc_on_every_positive_float(ads):
    """The device ln against glibc logf on all 2^31 non-negative floats
    (zero, subnormals, normals, +inf and NaNs), in batches of 2^27.  The
    kernel behind adsr_device_logf returns logf_exact, or a NaN that glibc
    never returns where logf_normal (the fused kernels' path for positive
    normal inputs) differs from it, so both paths are checked.  The compare
    loop is the oracle's threaded orc_logf_check_range; NaN results compare
    equal whatever their payload."""
    import os
    import alphadia_search_rs._native as N
    L = oracle.lib()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    step = 1 << 27
    bad_total = 0
    first_bad = []
    y = np.empty(step, np.float32)
    for first in range(0, 1 << 31, step):
        x = np.arange(first, first + step, dtype=np.uint32).view(np.float32)
        N.check(N.lib.adsr_device_logf(N.ptr(x), step, N.ptr(y)))
        fb = oracle.C.c_size_t(0)
        bad = int(L.orc_logf_check_range(first, step, N.ptr(y), threads, oracle.C.byref(fb)))
        if bad:
            first_bad.append(hex(first + fb.value))
        bad_total += bad
    # negative inputs (NaN, -inf at -0) on a sample
    rng = np.random.default_rng(11)
    neg = np.concatenate([np.array([0x80000000, 0xff800000, 0x80000001], np.uint32),
                          rng.integers(0x80000000, 0xffffffff, 1 << 20, dtype=np.uint32)])
    yn = np.empty(len(neg), np.float32)
    N.check(N.lib.adsr_device_logf(N.ptr(neg.view(np.float32)), len(neg), N.ptr(yn)))
    want = np.log(neg.view(np.float32)).astype(np.float32)
    same = (yn.view(np.uint32) == want.view(np.uint32)) | (np.isnan(yn) & np.isnan(want))
    assert same.all(), neg[~same][:5]
    assert bad_total == 0, f"{bad_total} mismatches, first at {first_bad[:5]}"


def test_stage_convolution_and_score(ads, run0, oracle_run0):
    import ctypes
    import alphadia_search_rs._native as N
    od, ol = oracle_run0
    prm = oracle_params_for(run0["spec"])
    rows = [3, 17, 101, 555, 1999]
    k = oracle.gaussian_kernel(run0["spec"].fwhm_rt, 1.0, 20, 1.0)
    kd = np.zeros(32, np.float32)
    n = N.lib.adsr_gaussian_kernel(run0["spec"].fwhm_rt, 1.0, 20, 1.0, N.ptr(kd), 32)
    assert n == 20 and np.array_equal(kd[:n].view(np.uint32), k.view(np.uint32))
    for r in rows:
        one = oracle.select_one(od, ol, prm, r)
        if one is None:
            continue
        x = np.ascontiguousarray(one["xic"])
        F, L = x.shape
        conv = np.zeros_like(x)
        N.check(N.lib.adsr_convolution(N.ptr(k), len(k), N.ptr(x), F, L, N.ptr(conv)))
        assert np.array_equal(conv.view(np.uint32), one["conv"].view(np.uint32))
        sc = np.zeros(L, np.float32)
        w = np.ascontiguousarray(one["fragment_intensity"])
        c = np.ascontiguousarray(one["conv"])
        N.check(N.lib.adsr_axis_log_dot_product(N.ptr(c), F, L, N.ptr(w), N.ptr(sc)))
        assert np.array_equal(sc.view(np.uint32), one["score"].view(np.uint32))
        idx = np.zeros(L, np.uint64)
        val = np.zeros(L, np.float32)
        nm = N.c_size()
        N.check(N.lib.adsr_find_local_maxima(N.ptr(sc), L, one["cycle_start"], N.ptr(idx),
                                             N.ptr(val), ctypes.byref(nm)))
        oi, ov = oracle.find_local_maxima(one["score"], one["cycle_start"])
        assert nm.value == len(oi)
        assert np.array_equal(idx[: nm.value].astype(np.int64), oi)
        assert np.array_equal(val[: nm.value].view(np.uint32), ov.view(np.uint32))


def test_stage_dense_xic(ads, run0, oracle_run0):
    import alphadia_search_rs._native as N
    od, ol = oracle_run0
    prm = oracle_params_for(run0["spec"])
    d = ads.DIAData.from_arrays(*run0["dia"])
    lib = run0["lib"]
    for r in [0, 5, 77, 800, 1500]:
        one = oracle.select_one(od, ol, prm, r)
        if one is None:
            continue
        fm = np.ascontiguousarray(one["fragment_mz"])
        F = len(fm)
        L = one["cycle_stop"] - one["cycle_start"]
        out = np.zeros((F, L), np.float32)
        N.check(N.lib.adsr_dense_xic(d._h, float(lib[2][r]), one["cycle_start"],
                                     one["cycle_stop"], run0["spec"].mass_tolerance,
                                     N.ptr(fm), F, N.ptr(out)))
        assert np.array_equal(out.view(np.uint32), one["xic"].view(np.uint32))


def test_fused_scores_bitwise(ads, run0, oracle_run0):
    import ctypes
    import alphadia_search_rs._native as N
    od, ol = oracle_run0
    prm = oracle_params_for(run0["spec"])
    d = ads.DIAData.from_arrays(*run0["dia"])
    lib = ads.SpecLibFlat.from_arrays(*run0["lib"])
    p = params_for(run0["spec"])
    r0, r1, stride = 0, 256, 256
    scores = np.zeros((r1 - r0, stride), np.float32)
    cs = np.zeros(r1 - r0, np.uint64)
    ce = np.zeros(r1 - r0, np.uint64)
    c = p._c()
    N.check(N.lib.adsr_select_scores(d._h, lib._h, ctypes.byref(c), r0, r1, stride,
                                     N.ptr(scores), N.ptr(cs), N.ptr(ce)))
    for i, r in enumerate(range(r0, r1)):
        one = oracle.select_one(od, ol, prm, r)
        if one is None:
            assert ce[i] <= cs[i]
            continue
        assert cs[i] == one["cycle_start"] and ce[i] == one["cycle_stop"]
        L = one["cycle_stop"] - one["cycle_start"]
        got = scores[i, :L]
        bad = np.flatnonzero(got.view(np.uint32) != one["score"].view(np.uint32))
        assert bad.size == 0, f"row {r}: score differs at cycles {bad[:8]}"


def test_fused_plan_fragments(ads, run0, oracle_run0):
    import ctypes
    import alphadia_search_rs._native as N
    d = ads.DIAData.from_arrays(*run0["dia"])
    libo = ads.SpecLibFlat.from_arrays(*run0["lib"])
    p = params_for(run0["spec"])
    c = p._c()
    n = libo.num_precursors
    stride = 64
    cs = np.zeros(n, np.uint64)
    ce = np.zeros(n, np.uint64)
    ns = np.zeros(n, np.uint32)
    sel = np.zeros((n, stride), np.uint64)
    N.check(N.lib.adsr_select_plan(d._h, libo._h, ctypes.byref(c), 0, n, N.ptr(cs),
                                   N.ptr(ce), N.ptr(ns), N.ptr(sel), stride))
    lib = run0["lib"]
    for r in range(0, n, 7):
        a, b = int(lib[6][r]), int(lib[7][r])
        want = a + oracle.filter_sort_fragments(lib[9][a:b], lib[10][a:b], lib[14][a:b],
                                                lib[16][a:b], True, True, 12)
        assert ns[r] == len(want)
        assert np.array_equal(sel[r, : ns[r]].astype(np.int64), want)


@pytest.mark.parametrize("frag_min,frag_max,top_k", [(17, 32, 7), (33, 60, 20), (20, 80, 12)],
                         ids=["lanes32", "lanes64", "thread-per-row"])
def test_top_k_with_ties_follows_introselect(ads, frag_min, frag_max, top_k):
    """More than 16 kept fragments, top_k < kept - 1, intensities on a few
    levels (ties at the cut) and some equal fragment m/z (the m/z sort
    keeps the selection's order): the device plan (per-row selected
    fragments) and the candidate table equal the oracle's restated Rust
    select_nth_unstable_by.  The three libraries take the three plan paths:
    32-lane and 64-lane groups per row, and one thread per row (rows of at
    most 48 fragments in LDS, longer ones in global scratch)."""
    import ctypes
    import alphadia_search_rs._native as N
    spec = synthetic.config(2, n_cycles=300, n_windows=4, peaks_per_spectrum=600,
                            n_targets=300, frag_min=frag_min, frag_max=frag_max,
                            half_window_min=30, half_window_max=60)
    gen = synthetic.generate(spec)
    lib = [a.copy() for a in gen["lib"]]
    fs, fe = lib[6].astype(np.int64), lib[7].astype(np.int64)
    rng = np.random.default_rng(frag_max)
    lib[10] = rng.integers(1, 5, len(lib[10])).astype(np.float32)
    for r in range(0, len(fs), 3):
        a = fs[r]
        lib[9][a + 2] = lib[9][a + 1]   # equal m/z within the row
    gen = dict(gen, lib=tuple(lib))
    d = ads.DIAData.from_arrays(*gen["dia"])
    libo = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(gen["spec"], top_k_fragments=top_k)
    if gen.get("rt_tolerance_per_precursor") is not None:
        p.update({"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]})
    c = p._c()
    n = libo.num_precursors
    stride = 64
    cs = np.zeros(n, np.uint64)
    ce = np.zeros(n, np.uint64)
    ns = np.zeros(n, np.uint32)
    sel = np.zeros((n, stride), np.uint64)
    N.check(N.lib.adsr_select_plan(d._h, libo._h, ctypes.byref(c), 0, n, N.ptr(cs),
                                   N.ptr(ce), N.ptr(ns), N.ptr(sel), stride))
    order = np.argsort(lib[0], kind="stable")  # plan rows follow precursor_idx
    differs = 0
    for i in range(n):
        r = int(order[i])
        a, b = int(fs[r]), int(fe[r])
        want = a + oracle.filter_sort_fragments(lib[9][a:b], lib[10][a:b], lib[14][a:b],
                                                lib[16][a:b], True, True, top_k)
        assert ns[i] == len(want), i
        assert np.array_equal(sel[i, : ns[i]].astype(np.int64), want), i
        stable = a + np.array(sorted(sorted(
            [j for j in range(b - a) if lib[10][a + j] > 0
             and not (lib[16][a + j] == 121 and lib[14][a + j] == 1)],
            key=lambda j: -lib[10][a + j])[:top_k], key=lambda j: lib[9][a + j]), np.int64)
        differs += not np.array_equal(stable, want)
    assert differs > 0  # the stable order would differ from the reference's here
    got = _search(ads, gen, top_k_fragments=top_k)[3]
    want = _oracle(gen, top_k_fragments=top_k)
    assert_tables_equal(got, want, f"introselect {frag_min}-{frag_max} top {top_k}")


def test_workload0_candidates_match_oracle(ads, run0):
    got = _search(ads, run0)[3]
    want = _oracle(run0)
    assert len(got[0]) > 0
    assert_tables_equal(got, want, "workload 0")


@pytest.mark.parametrize("variant", ["decoys_uneven", "per_window_sigma", "ties", "skips",
                                     "cross_bin_sums", "top_k_drop_one", "long_fragment_lists",
                                     "uneven_rt", "repeated_rt"])
def test_small_variants_match_oracle(ads, variant):
    if variant in ("uneven_rt", "repeated_rt"):
        # RT axes other than one value per cycle: unevenly spaced but strictly
        # increasing (the device's guessed search walks further), and with
        # every value twice (the reference's binary search on equal values)
        spec = synthetic.config(1, n_cycles=500, n_windows=6, peaks_per_spectrum=600,
                                n_targets=800, half_window_min=40, half_window_max=40)
        gen = synthetic.generate(spec)
        dia = list(gen["dia"])
        c = dia[5].astype(np.float64)
        if variant == "uneven_rt":
            rt = c + 0.4 * np.sin(1.7 * c) + 3.0 * (c > 250)
        else:
            rt = np.floor(c / 2.0) * 2.0
        dia[6] = rt.astype(np.float32)
        gen = dict(gen, dia=tuple(dia))
        got = _search(ads, gen)[3]
        want = _oracle(gen)
    elif variant == "decoys_uneven":
        spec = synthetic.config(2, n_cycles=600, n_windows=12, peaks_per_spectrum=800,
                                n_targets=1500)
        gen = synthetic.generate(spec)
        got = _search(ads, gen)[3]
        want = _oracle(gen)
    elif variant == "per_window_sigma":
        spec = synthetic.config(1, n_cycles=500, n_windows=6, peaks_per_spectrum=600,
                                n_targets=800, half_window_min=80, half_window_max=80)
        gen = synthetic.generate(spec)
        sig = np.linspace(1.5, 3.0, 6).astype(np.float32)
        got = _search(ads, gen, sigma_per_window=sig)[3]
        want = _oracle(gen, sigma_per_window=sig)
    elif variant == "skips":
        # rows the kernels skip: empty RT windows (c1 <= c0), no isolation
        # window containing the precursor, no fragment with intensity > 0
        spec = synthetic.config(0, n_cycles=400, n_windows=4, peaks_per_spectrum=800,
                                n_targets=700, half_window_min=30, half_window_max=30)
        gen = synthetic.generate(spec)
        lib = [a.copy() for a in gen["lib"]]
        lib[3][::7] = -1.0e4
        lib[4][::7] = -1.0e4
        lib[1][5::11] = 50.0
        lib[2][5::11] = 50.0
        fs, fe = lib[6].astype(np.int64), lib[7].astype(np.int64)
        for r in range(3, len(fs), 13):
            lib[10][fs[r]:fe[r]] = 0.0
        gen = dict(gen, lib=tuple(lib))
        got = _search(ads, gen)[3]
        want = _oracle(gen)
        skipped = set(range(0, 700, 7)) | set(range(5, 700, 11)) | set(range(3, 700, 13))
        assert not skipped & set(got[0].astype(np.int64).tolist())
    elif variant == "top_k_drop_one":
        # top_k = kept - 1: the reference drops the first minimum-intensity
        # fragment and moves the last kept one into its place
        spec = synthetic.config(0, n_cycles=300, n_windows=4, peaks_per_spectrum=800,
                                n_targets=600, half_window_min=40, half_window_max=40)
        gen = synthetic.generate(spec)
        got = _search(ads, gen, top_k_fragments=11)[3]
        want = _oracle(gen, top_k_fragments=11)
    elif variant == "long_fragment_lists":
        # 40-60 raw fragments per precursor (one lane group of 64 per row),
        # top 16 by intensity
        spec = synthetic.config(2, n_cycles=300, n_windows=6, peaks_per_spectrum=600,
                                n_targets=400, frag_min=40, frag_max=60,
                                half_window_min=30, half_window_max=60)
        gen = synthetic.generate(spec)
        got = _search(ads, gen, top_k_fragments=16)[3]
        want = _oracle(gen, top_k_fragments=16)
    elif variant == "cross_bin_sums":
        # every peak gets two copies at -6 and +6 ppm (other bins, same
        # spectrum) with very different intensities: XIC cells sum three or
        # more entries from different bins, whose order changes the bits
        spec = synthetic.config(0, n_cycles=300, n_windows=4, peaks_per_spectrum=700,
                                n_targets=500, half_window_min=40, half_window_max=40)
        gen = synthetic.generate(spec)
        dia = list(gen["dia"])
        start, stop, mz, it = dia[3], dia[4], dia[7], dia[8]
        new_mz, new_it, new_start, new_stop = [], [], [], []
        pos = 0
        for s0, s1 in zip(start, stop):
            m = mz[s0:s1]
            v = it[s0:s1]
            mm = np.concatenate([m, (m * np.float32(1 - 6e-6)).astype(np.float32),
                                 (m * np.float32(1 + 6e-6)).astype(np.float32)])
            vv = np.concatenate([v, (v * np.float32(0.37)).astype(np.float32),
                                 (v * np.float32(1.3e-3)).astype(np.float32)])
            o = np.argsort(mm, kind="stable")
            new_mz.append(mm[o])
            new_it.append(vv[o])
            new_start.append(pos)
            pos += len(mm)
            new_stop.append(pos)
        dia[3] = np.array(new_start, np.int64)
        dia[4] = np.array(new_stop, np.int64)
        dia[7] = np.concatenate(new_mz).astype(np.float32)
        dia[8] = np.concatenate(new_it).astype(np.float32)
        gen = dict(gen, dia=tuple(dia))
        got = _search(ads, gen)[3]
        want = _oracle(gen)
    else:
        # coarse intensities and many duplicate (bin, cycle) entries
        spec = synthetic.config(0, n_cycles=300, n_windows=4, peaks_per_spectrum=3000,
                                n_targets=600, half_window_min=40, half_window_max=40)
        gen = synthetic.generate(spec)
        dia = list(gen["dia"])
        dia[8] = np.round(dia[8] / 200.0).astype(np.float32) * 200.0
        # snap every 3rd peak onto its neighbour: same bin, same cycle
        mz = dia[7].copy()
        mz[1::3] = mz[0:-1:3][: len(mz[1::3])]
        dia[7] = mz
        gen = dict(gen, dia=tuple(dia))
        got = _search(ads, gen, top_k_fragments=9)[3]
        want = _oracle(gen, top_k_fragments=9)
    assert_tables_equal(got, want, variant)


def test_hit_list_overflow_scores_bitwise(ads):
    """Rows with more in-window entries than the device keeps in its per-row
    hit list (256): dense spectra, 40 ppm, 20-24 fragments and long RT
    windows; a fragment m/z present in every spectrum, so that one fragment
    alone overflows the list; precursors on a window boundary (two windows).
    Every score row is compared bit for bit, then the candidate table."""
    import ctypes
    import alphadia_search_rs._native as N
    spec = synthetic.config(2, n_cycles=700, n_windows=4, peaks_per_spectrum=3000,
                            n_targets=300, frag_min=20, frag_max=24,
                            half_window_min=150, half_window_max=300)
    gen = with_peak_everywhere(synthetic.generate(spec), 777.7777, 350.0)
    lib = [a.copy() for a in gen["lib"]]
    fs = lib[6].astype(np.int64)
    lib[9][fs[::4]] = np.float32(777.7777)
    lib[9][fs[2::8] + 5] = np.float32(777.7777)
    lib[1][1::9] = np.float32(550.0)   # window boundary: 400 + 150
    lib[2][1::9] = np.float32(550.0)
    gen = dict(gen, lib=tuple(lib))
    extra = dict(mass_tolerance=40.0, top_k_fragments=24)
    d = ads.DIAData.from_arrays(*gen["dia"])
    slib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(gen["spec"], **extra)
    p.update({"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]})
    n = slib.num_precursors
    stride = 1024
    scores = np.zeros((n, stride), np.float32)
    cs = np.zeros(n, np.uint64)
    ce = np.zeros(n, np.uint64)
    c = p._c()
    N.check(N.lib.adsr_select_scores(d._h, slib._h, ctypes.byref(c), 0, n, stride,
                                     N.ptr(scores), N.ptr(cs), N.ptr(ce)))
    od, ol = oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"])
    prm = oracle_params_for(gen["spec"], rt_tolerance_per_precursor=gen["rt_tolerance_per_precursor"],
                            **extra)
    big = single = two = 0
    for r in range(n):
        one = oracle.select_one(od, ol, prm, r)
        if one is None:
            assert ce[r] <= cs[r]
            continue
        assert cs[r] == one["cycle_start"] and ce[r] == one["cycle_stop"]
        L = one["cycle_stop"] - one["cycle_start"]
        assert L <= stride
        nz = one["xic"] != 0
        big += int(nz.sum() > 256)
        single += int(nz.sum(axis=1).max(initial=0) > 256)
        two += int(np.float32(lib[2][r]) == np.float32(550.0) and nz.sum() > 256)
        got = scores[r, :L]
        bad = np.flatnonzero(got.view(np.uint32) != one["score"].view(np.uint32))
        assert bad.size == 0, f"row {r}: score differs at cycles {bad[:8]}"
    # the cases are present: list overflow, one fragment alone over the
    # list, and overflow with two windows
    assert big > 50 and single > 20 and two > 3, (big, single, two)
    got = ads.PeakGroupSelection(p).search(d, slib).to_arrays()
    want = oracle.select(od, ol, prm)
    assert_tables_equal(got, want, "hit-list overflow")


def _long_row_run():
    """Workload 4's RT windows (half-windows of 50-400 cycles) on a
    1,200-cycle run of 1,000 precursors."""
    spec = synthetic.config(4, n_cycles=1200, n_windows=4, peaks_per_spectrum=600,
                            n_targets=500)
    return synthetic.generate(spec)


def test_long_rows_second_launch_bitwise(ads):
    """Rows longer than 704 cycles need more than a wave's 10 KB LDS share
    even for one fragment, so the first select launch defers them to a
    second launch over a device-side row list.  Both kinds of row are
    present; every score row is compared bit for bit, then the candidate
    table."""
    import ctypes
    import alphadia_search_rs._native as N
    gen = _long_row_run()
    d = ads.DIAData.from_arrays(*gen["dia"])
    slib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(gen["spec"])
    p.update({"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]})
    n = slib.num_precursors
    stride = 1024
    scores = np.zeros((n, stride), np.float32)
    cs = np.zeros(n, np.uint64)
    ce = np.zeros(n, np.uint64)
    c = p._c()
    N.check(N.lib.adsr_select_scores(d._h, slib._h, ctypes.byref(c), 0, n, stride,
                                     N.ptr(scores), N.ptr(cs), N.ptr(ce)))
    od, ol = oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"])
    prm = oracle_params_for(gen["spec"],
                            rt_tolerance_per_precursor=gen["rt_tolerance_per_precursor"])
    long_rows = short_rows = 0
    for r in range(n):
        one = oracle.select_one(od, ol, prm, r)
        if one is None:
            assert ce[r] <= cs[r]
            continue
        assert cs[r] == one["cycle_start"] and ce[r] == one["cycle_stop"]
        L = one["cycle_stop"] - one["cycle_start"]
        assert L <= stride
        long_rows += int(L > 704)
        short_rows += int(L <= 704)
        got = scores[r, :L]
        bad = np.flatnonzero(got.view(np.uint32) != one["score"].view(np.uint32))
        assert bad.size == 0, f"row {r} (L={L}): score differs at cycles {bad[:8]}"
    assert long_rows > 50 and short_rows > 400, (long_rows, short_rows)
    got = ads.PeakGroupSelection(p).search(d, slib).to_arrays()
    want = oracle.select(od, ol, prm)
    assert_tables_equal(got, want, "long rows")


def test_graph_replays_with_long_rows_match_oracle(ads):
    """The timed search: the autotuner's sweep, eager steps, then the search
    (plan, both select launches, compaction) captured in a hipGraph and
    replayed three times; the table of the last replay equals the oracle's."""
    import ctypes
    import alphadia_search_rs._native as N
    gen = _long_row_run()
    d = ads.DIAData.from_arrays(*gen["dia"])
    slib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(gen["spec"])
    p.update({"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]})
    c = p._c()
    tpb, ppw = ctypes.c_int(0), ctypes.c_int(0)
    step_ms, kern_ms, graph_ms = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    out = ctypes.c_void_p()
    N.check(N.lib.adsr_select_timed(d._h, slib._h, ctypes.byref(c), 1, 3,
                                    ctypes.byref(tpb), ctypes.byref(ppw),
                                    ctypes.byref(step_ms), ctypes.byref(kern_ms),
                                    ctypes.byref(graph_ms), ctypes.byref(out)))
    got = ads.CandidateCollection._wrap(out.value).to_arrays()
    assert graph_ms.value > 0.0 and tpb.value in (64, 128, 256)
    want = _oracle(gen)
    assert_tables_equal(got, want, "graph replay, long rows")


def test_concurrent_searches_match_sequential(ads, run0):
    """Searches from four Python threads on the same DIAData and SpecLibFlat
    (ctypes releases the GIL; each search has its own stream and
    workspace) give the tables of the same searches run one after another."""
    import threading
    d = ads.DIAData.from_arrays(*run0["dia"])
    lib = ads.SpecLibFlat.from_arrays(*run0["lib"])
    variants = [dict(), dict(candidate_count=5), dict(top_k_fragments=9),
                dict(mass_tolerance=12.0)]
    params = [params_for(run0["spec"], **v) for v in variants]
    want = [ads.PeakGroupSelection(p).search(d, lib).to_arrays() for p in params]
    got = [None] * len(params)
    errors = []

    def run(i):
        try:
            for _ in range(3):
                got[i] = ads.PeakGroupSelection(params[i]).search(d, lib).to_arrays()
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(params))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(len(params)):
        assert got[i] is not None, f"thread {i} did not finish"
        assert_tables_equal(got[i], want[i], f"concurrent variant {variants[i]}")


def _dense_fragment_run():
    """40 selected fragments per precursor, 20 of them on masses with a peak
    in each of cycles 20-29 (one or two 8-entry CSR groups per fragment) and
    20 at 0.5 Th above them, where no peak lies (empty CSR ranges, or a
    stray background peak), interleaved in m/z order: many fragments start
    inside one 64-group extraction tile, with empty ones between them and
    at tile edges."""
    rng = np.random.default_rng(31)
    C, W, K = 60, 2, 30
    win_lo = np.array([400.0, 700.0], np.float32)
    win_hi = np.array([700.0, 1000.0], np.float32)
    planted = np.sort(rng.uniform(300.0, 1900.0, 20)).astype(np.float32)
    spec_mz, spec_in, starts, stops = [], [], [], []
    pos = 0
    for c in range(C):
        for w in range(W):
            mz = rng.uniform(150.0, 2000.0, K).astype(np.float32)
            it = rng.lognormal(6.0, 1.5, K).astype(np.float32)
            if 20 <= c < 30 and w == 0:
                prof = np.float32(np.exp(-0.5 * ((c - 24.5) / 2.0) ** 2))
                mz = np.concatenate([mz, planted])
                it = np.concatenate([it, (rng.uniform(500, 5000, 20) * prof).astype(np.float32)])
            o = np.argsort(mz, kind="stable")
            spec_mz.append(mz[o])
            spec_in.append(it[o])
            starts.append(pos)
            pos += len(mz)
            stops.append(pos)
    n_spec = C * W
    idx = np.arange(n_spec)
    dia = ((idx % W).astype(np.int64), win_lo[idx % W], win_hi[idx % W],
           np.array(starts, np.int64), np.array(stops, np.int64),
           (idx // W).astype(np.int64), (idx // W).astype(np.float32),
           np.concatenate(spec_mz).astype(np.float32),
           np.concatenate(spec_in).astype(np.float32),
           np.zeros((W, 1, 1, 1), np.float32))
    P, F = 64, 40
    fmz = np.empty((P, F), np.float32)
    fmz[:, 0::2] = planted
    fmz[:, 1::2] = planted + np.float32(0.5)  # between planted masses, no peak
    fint = rng.uniform(0.01, 1.0, (P, F)).astype(np.float32)
    NF = P * F
    fs = (np.arange(P) * F).astype(np.uint64)
    local = np.tile(np.arange(F), P)
    ones = np.ones(NF, np.uint8)
    lib = (np.arange(P, dtype=np.uint64), np.full(P, 550.0, np.float32),
           np.full(P, 550.0, np.float32), np.full(P, 25.0, np.float32),
           np.full(P, 25.0, np.float32), np.full(P, 10, np.uint8), fs,
           fs + np.uint64(F), fmz.reshape(-1), fmz.reshape(-1), fint.reshape(-1), ones,
           ones.copy(), np.zeros(NF, np.uint8), (2 + local // 2).astype(np.uint8),
           ones.copy(), np.where(local % 2 == 0, 98, 121).astype(np.uint8))
    return dia, lib


def test_many_fragments_per_tile_with_empty_ones_bitwise(ads):
    import ctypes
    import alphadia_search_rs._native as N
    dia, lib_arrays = _dense_fragment_run()
    cfg = dict(fwhm_rt=4.7096, kernel_size=20, peak_length=5, mass_tolerance=5.0,
               rt_tolerance=100.0, candidate_count=3, top_k_fragments=40)
    d = ads.DIAData.from_arrays(*dia)
    lib = ads.SpecLibFlat.from_arrays(*lib_arrays)
    p = ads.SelectionParameters()
    p.update(cfg)
    n = lib.num_precursors
    stride = 64
    scores = np.zeros((n, stride), np.float32)
    cs = np.zeros(n, np.uint64)
    ce = np.zeros(n, np.uint64)
    c = p._c()
    N.check(N.lib.adsr_select_scores(d._h, lib._h, ctypes.byref(c), 0, n, stride,
                                     N.ptr(scores), N.ptr(cs), N.ptr(ce)))
    od, ol = oracle.Dia(dia), oracle.Lib(lib_arrays)
    prm = oracle.params(**cfg)
    for r in range(n):
        one = oracle.select_one(od, ol, prm, r)
        assert one is not None and one["fragment_mz"].size == 40
        L = one["cycle_stop"] - one["cycle_start"]
        hit = one["xic"].any(axis=1)
        assert hit.sum() >= 20 and (~hit).sum() >= 15  # planted hit, the others empty
        got = scores[r, :L]
        bad = np.flatnonzero(got.view(np.uint32) != one["score"].view(np.uint32))
        assert bad.size == 0, f"row {r}: score differs at cycles {bad[:8]}"
    got = ads.PeakGroupSelection(p).search(d, lib).to_arrays()
    want = oracle.select(od, ol, prm)
    assert len(want[0]) > 0
    assert_tables_equal(got, want, "40 fragments, empty ones interleaved")


@pytest.mark.parametrize("kernel_size", [20, 16, 28, 30])
def test_few_cell_convolution_scores_bitwise(ads, kernel_size):
    """Sparse XICs (a few hit cells per fragment row, pairs of cells at every
    distance, including the midpoints where two cells form one pair term) with
    negative (some scores NaN), zero and subnormal intensities, and a few huge ones (1e30, for which
    the device keeps the reference's full chain), at half widths 10, 8, 14 and
    15 (where every block takes the full chain): whole score rows bit for bit
    against the oracle, then the candidate table."""
    import ctypes
    import alphadia_search_rs._native as N
    spec = synthetic.config(0, n_cycles=300, n_windows=4, peaks_per_spectrum=500,
                            n_targets=400, half_window_min=60, half_window_max=60,
                            kernel_size=kernel_size)
    gen = synthetic.generate(spec)
    dia = list(gen["dia"])
    it = dia[8].copy()
    it[::5] = -it[::5]
    it[1::7] = 0.0
    it[2::11] = np.float32(3e-39)
    it[3::4001] = np.float32(1e30)
    dia[8] = it
    gen = dict(gen, dia=tuple(dia))
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(spec)
    n, stride = 400, 128
    scores = np.zeros((n, stride), np.float32)
    cs = np.zeros(n, np.uint64)
    ce = np.zeros(n, np.uint64)
    c = p._c()
    N.check(N.lib.adsr_select_scores(d._h, lib._h, ctypes.byref(c), 0, n, stride,
                                     N.ptr(scores), N.ptr(cs), N.ptr(ce)))
    od, ol = oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"])
    prm = oracle_params_for(spec)
    checked = 0
    for r in range(n):
        one = oracle.select_one(od, ol, prm, r)
        if one is None:
            assert ce[r] <= cs[r]
            continue
        L = one["cycle_stop"] - one["cycle_start"]
        assert cs[r] == one["cycle_start"] and L <= stride
        got, want = scores[r, :L], one["score"]
        # ln of a negative argument is NaN, whose sign and payload the
        # candidate table never shows (NaN is never a strict maximum)
        same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
        bad = np.flatnonzero(~same)
        assert bad.size == 0, f"row {r}: score differs at cycles {bad[:8]}"
        checked += 1
    assert checked > 300
    got = ads.PeakGroupSelection(p).search(d, lib).to_arrays()
    assert_tables_equal(got, oracle.select(od, ol, prm), f"kernel_size {kernel_size}")
