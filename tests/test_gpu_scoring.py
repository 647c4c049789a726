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
from conftest import params_for

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
