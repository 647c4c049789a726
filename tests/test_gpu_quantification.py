"""PeakGroupQuantification.quantify on the device vs the oracle.

The quantified library (both to_dict_arrays dicts of
src/speclib_flat_quantified/mod.rs:126-211) must equal the oracle's bit for
bit: per-fragment observed intensity, m/z, correlation and mass error from
the device kernel, and the host-side layout (fragments with intensity > 0,
fragments in candidate order, precursor columns stably sorted by
precursor_idx).  Cases: workload 0's candidates, uneven fragment counts
with top-k, and rows the reference drops (precursor not in the library,
zero cycles, centre outside the cycle range).
"""
import numpy as np
import pytest

import oracle
import synthetic
from conftest import params_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ads():
    import alphadia_search_rs as m
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    return m


def _assert_same(got, want, what):
    for part, g, w in (("precursor", got[0], want[0]), ("fragment", got[1], want[1])):
        assert set(g) == set(w), (what, part)
        for k in w:
            assert g[k].dtype == w[k].dtype, (what, part, k, g[k].dtype, w[k].dtype)
            assert g[k].shape == w[k].shape, (what, part, k, g[k].shape, w[k].shape)
            bad = np.flatnonzero(g[k] != w[k])
            assert bad.size == 0, (f"{what}: {part}.{k} differs at {bad.size} rows, first "
                                   f"{bad[:5]}: got {g[k][bad[:5]]} want {w[k][bad[:5]]}")


def _quantify(ads, gen, tol=10.0, top_k=10000, extra_rows=None, **sel_extra):
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    cols = list(ads.PeakGroupSelection(params_for(gen["spec"], **sel_extra))
                .search(d, lib).to_arrays())
    for row in extra_rows or []:
        for k in range(9):
            cols[k] = np.append(cols[k], np.asarray(row[k], cols[k].dtype))
    cands = ads.CandidateCollection.from_arrays(*cols)
    qp = ads.QuantificationParameters()
    qp.update({"tolerance_ppm": tol, "top_k_fragments": top_k})
    got = ads.PeakGroupQuantification(qp).quantify(d, lib, cands)
    od, ol = oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"])
    want = oracle.quantify(od, ol, gen["lib"], tuple(cols), tolerance_ppm=tol,
                           top_k_fragments=top_k)
    return got, want


def test_quantify_workload0_matches_oracle(ads, run0):
    q, want = _quantify(ads, run0)
    assert q.num_precursors == len(want[0]["precursor_idx"]) > 1000
    assert q.num_fragments == len(want[1]["mz"])
    _assert_same(q.to_dict_arrays(), want, "workload 0")


def test_quantify_uneven_fragments_top_k_and_dropped_rows(ads):
    gen = synthetic.generate(synthetic.config(0, n_cycles=400, n_windows=6, n_targets=400,
                                              frag_min=6, frag_max=24))
    pidx = int(gen["lib"][0][3])
    extra = [
        (10 ** 9, 0, 1.0, 0, 0, 0, 50, 45, 56),   # not in the library
        (pidx, 1, 1.0, 0, 0, 0, 50, 50, 50),      # zero cycles
        (pidx, 2, 1.0, 0, 0, 0, 60, 45, 56),      # centre past cycle_stop
        (pidx, 0, 1.0, 0, 0, 0, 2, (2 - 5) % (1 << 64), 8),  # wrapped start
    ]
    for top_k in (10000, 7):
        q, want = _quantify(ads, gen, tol=15.0, top_k=top_k, extra_rows=extra,
                            top_k_fragments=24)
        _assert_same(q.to_dict_arrays(), want, f"uneven, top_k {top_k}")


def test_quantify_more_than_64_fragments(ads):
    """Precursors with 40-128 fragments, all of them quantified (top_k
    10000): the kernel's two-fragments-per-lane instance, identical to the
    oracle."""
    gen = synthetic.generate(synthetic.config(0, n_cycles=400, n_windows=6, n_targets=200,
                                              frag_min=40, frag_max=128))
    nf = (gen["lib"][7] - gen["lib"][6]).astype(np.int64)
    assert nf.max() > 64
    q, want = _quantify(ads, gen, tol=15.0)
    assert q.num_fragments == len(want[1]["mz"]) > 0
    _assert_same(q.to_dict_arrays(), want, "40-128 fragments")


def test_quantify_more_than_128_fragments(ads):
    """Precursors with 130-250 fragments, all quantified: four fragments per
    lane, identical to the oracle."""
    gen = synthetic.generate(synthetic.config(0, n_cycles=300, n_windows=4, n_targets=100,
                                              frag_min=130, frag_max=250))
    nf = (gen["lib"][7] - gen["lib"][6]).astype(np.int64)
    assert nf.max() > 128
    q, want = _quantify(ads, gen, tol=15.0)
    assert q.num_fragments == len(want[1]["mz"]) > 0
    _assert_same(q.to_dict_arrays(), want, "130-250 fragments")
