"""PeakGroupScoring.score on the device vs the oracle, through the C ABI.

Every feature is compared bit for bit except the three log10 features
(log10_b_ion_intensity, log10_y_ion_intensity, idf_intensity_dot_product),
whose device log10f may differ from glibc's by an ulp: those within 2e-6
relative.  Cases: workload 0 candidates from the device selection (host
table and the table kept in HBM), uneven fragment counts with top-k
truncation, a candidate whose precursor is not in the library (skipped), a
candidate centred before peak_length (cycle_start wraps, as usize does in
the reference's release build) and a precursor without positive fragments.
"""
import numpy as np
import pytest

import oracle
import synthetic
from conftest import params_for, with_peak_everywhere

pytestmark = pytest.mark.gpu

LOG10 = {"log10_b_ion_intensity", "log10_y_ion_intensity", "idf_intensity_dot_product"}


@pytest.fixture(scope="module")
def ads():
    import alphadia_search_rs as m
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    return m


def _assert_features_equal(got, want, what):
    assert list(got) == list(want), what
    n = len(want["precursor_idx"])
    assert len(got["precursor_idx"]) == n, f"{what}: {len(got['precursor_idx'])} rows, oracle {n}"
    assert np.array_equal(got["precursor_idx"], want["precursor_idx"]), what
    assert np.array_equal(got["rank"], want["rank"]), what
    for name in oracle.FEATURES:
        g, w = got[name], want[name]
        if name in LOG10:
            bad = np.flatnonzero(np.abs(g - w) > 2e-6 * np.maximum(1.0, np.abs(w)))
        else:  # +0 and -0 compare equal (an all-zero sum's sign is moot)
            bad = np.flatnonzero(~((g == w) | (np.isnan(g) & np.isnan(w))))
        assert bad.size == 0, (f"{what}: {name} differs at {bad.size} rows, first "
                               f"{bad[:5]}: got {g[bad[:5]]} want {w[bad[:5]]}")


def _run(ads, gen, tol=10.0, top_k=100, extra_rows=None, **sel_extra):
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    sel = ads.PeakGroupSelection(params_for(gen["spec"], **sel_extra))
    cands = sel.search(d, lib)
    cols = list(cands.to_arrays())
    if extra_rows:
        for row in extra_rows:
            for k in range(9):
                cols[k] = np.append(cols[k], np.asarray(row[k], cols[k].dtype))
        cands = ads.CandidateCollection.from_arrays(*cols)
    sp = ads.ScoringParameters()
    sp.update({"mass_tolerance": tol, "top_k_fragments": top_k})
    got = ads.PeakGroupScoring(sp).score(d, lib, cands).to_dict_arrays()
    od, ol = oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"])
    want = oracle.score(od, ol, tuple(cols), mass_tolerance=tol, top_k_fragments=top_k)
    return d, lib, sel, got, want


def test_scoring_workload0_matches_oracle(ads, run0):
    d, lib, sel, got, want = _run(ads, run0)
    assert len(want["precursor_idx"]) > 1000
    _assert_features_equal(got, want, "workload 0")
    # the table kept in HBM scores the same
    sp = ads.ScoringParameters()
    sp.update({"mass_tolerance": 10.0})
    dev = ads.PeakGroupScoring(sp).score(d, lib, sel.search_device(d, lib)).to_dict_arrays()
    _assert_features_equal(dev, want, "workload 0, device table")


def test_scoring_uneven_fragments_and_top_k(ads):
    gen = synthetic.generate(synthetic.config(0, n_cycles=400, n_windows=6, n_targets=400,
                                              frag_min=6, frag_max=24))
    for top_k in (100, 12, 5):
        _, _, _, got, want = _run(ads, gen, tol=15.0, top_k=top_k, top_k_fragments=24)
        _assert_features_equal(got, want, f"uneven fragments, top_k {top_k}")


def test_scoring_edge_rows(ads, run0):
    gen = dict(run0)
    lib = list(gen["lib"])
    # precursor 0 keeps no positive fragment
    fs, fe = int(lib[6][0]), int(lib[7][0])
    lib[10] = lib[10].copy()
    lib[10][fs:fe] = 0.0
    gen["lib"] = tuple(lib)
    pidx0 = int(lib[0][0])
    extra = [
        # not in the library: skipped
        (10 ** 9, 0, 1.0, 0, 0, 0, 50, 45, 56),
        # centred before peak_length: cycle_start wraps
        (int(lib[0][1]), 0, 1.0, 0, 0, 0, 2, (2 - 5) % (1 << 64), 8),
        # no positive fragment
        (pidx0, 0, 1.0, 0, 0, 0, 100, 95, 106),
    ]
    _, _, _, got, want = _run(ads, gen, extra_rows=extra)
    assert 10 ** 9 not in set(got["precursor_idx"].tolist())
    _assert_features_equal(got, want, "edge rows")


def test_scoring_entry_pool_overflow(ads):
    """Candidates whose bins hold more in-window entries than the kernel
    stages in LDS per pass of 64 bins (136): 21 peaks 2 ppm apart around
    each of 4 m/z values in every spectrum, fragments placed on them, 25 ppm
    (every one of a fragment's dense bins has an entry in each of the 11
    cycles).  The entries past the pool are read again from HBM; features
    identical to the oracle's."""
    spec = synthetic.config(0, n_cycles=300, n_windows=4, peaks_per_spectrum=800,
                            n_targets=300, half_window_min=40, half_window_max=40)
    gen = synthetic.generate(spec)
    base = np.linspace(300.0, 1500.0, 4)
    for m in base:
        for d in range(-20, 21, 2):
            gen = with_peak_everywhere(gen, np.float32(m * (1 + d * 1e-6)), 500.0 + d)
    lib = [a.copy() for a in gen["lib"]]
    fs, fe = lib[6].astype(np.int64), lib[7].astype(np.int64)
    for r in range(0, len(fs), 3):
        k = min(4, int(fe[r] - fs[r]))
        lib[9][fs[r]:fs[r] + k] = base[:k].astype(np.float32)
    gen = dict(gen, lib=tuple(lib))
    _, _, _, got, want = _run(ads, gen, tol=25.0)
    assert len(want["precursor_idx"]) > 500
    _assert_features_equal(got, want, "entry pool overflow")


def test_scoring_too_many_windows_is_unsupported(ads):
    """A precursor m/z inside more than 16 isolation windows: the device
    reports the limit (the reference would extract from all of them)."""
    import alphadia_search_rs._native as N
    spec = synthetic.config(0, n_cycles=60, n_windows=20, peaks_per_spectrum=200,
                            n_targets=40, half_window_min=10, half_window_max=10)
    gen = synthetic.generate(spec)
    dia = list(gen["dia"])
    dia[1] = np.full_like(dia[1], 400.0)   # every window covers 400-1000
    dia[2] = np.full_like(dia[2], 1000.0)
    d = ads.DIAData.from_arrays(*dia)
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    pidx = int(gen["lib"][0][0])
    cands = ads.CandidateCollection.from_arrays([pidx], [0], [1.0], [0], [0], [0],
                                                [30], [25], [36])
    with pytest.raises(N.AdsrError, match="isolation windows"):
        ads.PeakGroupScoring(ads.ScoringParameters()).score(d, lib, cands)


def test_scoring_more_than_64_fragments(ads):
    """Precursors with 40-128 fragments: with top_k above 64 each lane owns
    two fragments (the kernel's two-fragments-per-lane instance); every
    feature still equals the oracle's.  top_k 65 puts the stride just past
    one fragment per lane; 100 is the reference's default."""
    gen = synthetic.generate(synthetic.config(0, n_cycles=400, n_windows=6, n_targets=200,
                                              frag_min=40, frag_max=128))
    for top_k in (128, 100, 65):
        d, lib, sel, got, want = _run(ads, gen, tol=15.0, top_k=top_k)
        assert int(np.max(want["num_fragments"])) > 64
        _assert_features_equal(got, want, f"40-128 fragments, top_k {top_k}")
    # the table kept in HBM scores the same
    sp = ads.ScoringParameters()
    sp.update({"mass_tolerance": 15.0, "top_k_fragments": 65})
    dev = ads.PeakGroupScoring(sp).score(d, lib, sel.search_device(d, lib)).to_dict_arrays()
    _assert_features_equal(dev, want, "40-128 fragments, top_k 65, device table")


def test_scoring_more_than_128_fragments(ads):
    """Precursors with 130-250 fragments at top_k 256 and 200: four
    fragments per lane, every feature equal to the oracle's."""
    gen = synthetic.generate(synthetic.config(0, n_cycles=300, n_windows=4, n_targets=100,
                                              frag_min=130, frag_max=250))
    for top_k in (256, 200):
        _, _, _, got, want = _run(ads, gen, tol=15.0, top_k=top_k)
        assert int(np.max(want["num_fragments"])) > 128
        _assert_features_equal(got, want, f"130-250 fragments, top_k {top_k}")


def test_scoring_more_than_256_fragments_is_unsupported(ads):
    """A candidate keeping more than 256 fragments exceeds the kernel's four
    fragments per lane: the device reports the limit."""
    import alphadia_search_rs._native as N
    gen = synthetic.generate(synthetic.config(0, n_cycles=200, n_windows=4, n_targets=20,
                                              frag_min=260, frag_max=270))
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    cands = ads.PeakGroupSelection(params_for(gen["spec"])).search(d, lib)
    sp = ads.ScoringParameters()
    sp.update({"top_k_fragments": 1000})
    with pytest.raises(N.AdsrError, match="more than 256 fragments"):
        ads.PeakGroupScoring(sp).score(d, lib, cands)
