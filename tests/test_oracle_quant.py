"""The quantification restatement (orc_quantify in oracle/adsr_oracle_score.c
and the layout functions of oracle/__init__.py) pinned three ways.

1. The reference's own unit tests, transcribed into
   golden/reference_kats.json: src/precursor_quantified/tests.rs
   (filter_fragments_by_intensity), src/speclib_flat_quantified/tests.rs
   (from_precursor_quantified_vec, the per-fragment precursor columns) and
   src/peak_group_quantification/tests.rs (QuantificationParameters).
2. An independent float64 numpy implementation of
   PeakGroupQuantification::quantify_precursor
   (src/peak_group_quantification/mod.rs:86-166) on the rank-0 candidates of
   workload 0: observed intensity (sum over cycles), observed m/z and ppm
   error, correlation with the median of the non-zero normalised profiles,
   rt_observed, the intensity > 0 filter and the flat layout, compared with
   oracle.quantify fragment by fragment.
3. Planted mistakes: four in the C restatement (tests/test_oracle_mutants.py
   builds them) and two in the layout, each caught by these checks.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_kats.json")))


def _pq(precursor, fragments, constants=None):
    pq = dict(precursor)
    n = len(fragments["mz"])
    const = constants or {"cardinality": 1, "charge": 1, "loss_type": 0, "type": 1}
    types = {"mz_library": np.float32, "mz": np.float32, "intensity": np.float32,
             "mz_observed": np.float32, "correlation_observed": np.float32,
             "mass_error_observed": np.float32}
    for k in ("mz_library", "mz", "intensity", "cardinality", "charge", "loss_type",
              "number", "position", "type", "mz_observed", "correlation_observed",
              "mass_error_observed"):
        v = fragments[k] if k in fragments else [const[k]] * n
        pq["fragment_" + k] = np.asarray(v, dtype=types.get(k, np.uint8))
    return pq


# ---- 1. reference unit tests ------------------------------------------------
@pytest.mark.parametrize("case", GOLD["quant_filter_by_intensity"], ids=lambda c: c["source"])
def test_filter_fragments_by_intensity_kats(case):
    pq = _pq(case["precursor"], case["fragments"])
    got = oracle.filter_fragments_by_intensity(pq, case["threshold"])
    if case["expect"] is None:
        assert got is None
        return
    for k, v in case["expect"].items():
        assert np.array_equal(got["fragment_" + k],
                              np.asarray(v, dtype=got["fragment_" + k].dtype)), k
    n = len(got["fragment_mz"])
    assert all(len(got["fragment_" + k]) == n for k in
               ("mz_library", "intensity", "cardinality", "charge", "loss_type", "number",
                "position", "type", "mz_observed", "correlation_observed",
                "mass_error_observed"))
    for k in ("precursor_idx", "mz", "rt", "naa", "rank", "rt_observed"):
        assert got[k] == pq[k]


def _speclib_case(case):
    spec = GOLD["quant_speclib_flat"]
    pqs = []
    for over in case["precursors"]:
        pre = dict(spec["base"]["precursor"])
        pre.update({k: v for k, v in over.items() if k != "fragments"})
        pqs.append(_pq(pre, over.get("fragments", spec["base"]["fragments"]),
                       spec["fragment_constants"]))
    return pqs


@pytest.mark.parametrize("case", GOLD["quant_speclib_flat"]["cases"],
                         ids=lambda c: c["source"])
def test_speclib_flat_quantified_kats(case):
    pqs = _speclib_case(case)
    prec, frag = oracle.speclib_flat_quantified(pqs)
    assert len(prec["mz"]) == case["num_precursors"]
    assert len(frag["mz"]) == case["num_fragments"]
    if "fragment_precursor_idx" in case:
        assert frag["precursor_idx"].tolist() == case["fragment_precursor_idx"]
        assert frag["rank"].tolist() == case["fragment_precursor_rank"]
    # precursor columns stably sorted by precursor_idx, each keeping its own
    # fragment range (src/speclib_flat_quantified/mod.rs:283-330)
    assert np.all(np.diff(prec["precursor_idx"].astype(np.int64)) >= 0)
    for i in range(len(prec["mz"])):
        a, b = int(prec["flat_frag_start_idx"][i]), int(prec["flat_frag_stop_idx"][i])
        assert np.all(frag["precursor_idx"][a:b] == prec["precursor_idx"][i])
        assert np.all(frag["rank"][a:b] == prec["rank"][i])


def test_speclib_keeps_fragments_in_input_order():
    """Precursors given as (10, 5): the precursor columns are sorted, the
    fragments stay in input order with the ranges following their precursor
    (src/speclib_flat_quantified/mod.rs:283-330 reorders only precursors)."""
    case = GOLD["quant_speclib_flat"]["cases"][3]
    pqs = _speclib_case(case)
    pqs[0]["fragment_intensity"] = np.array([1, 2, 3], np.float32)
    pqs[1]["fragment_intensity"] = np.array([4, 5, 6], np.float32)
    prec, frag = oracle.speclib_flat_quantified(pqs)
    assert prec["precursor_idx"].tolist() == [5, 10]
    assert frag["intensity"].tolist() == [1, 2, 3, 4, 5, 6]
    assert frag["precursor_idx"].tolist() == [10, 10, 10, 5, 5, 5]
    assert prec["flat_frag_start_idx"].tolist() == [3, 0]
    assert prec["flat_frag_stop_idx"].tolist() == [6, 3]


def test_quantification_parameters_kats():
    import alphadia_search_rs as ads
    kat = GOLD["quant_params"]
    p = ads.QuantificationParameters()
    assert p.tolerance_ppm == kat["defaults"]["tolerance_ppm"]
    assert p.top_k_fragments == kat["defaults"]["top_k_fragments"]
    for case in kat["updates"]:
        p = ads.QuantificationParameters()
        p.update(case["update"])
        assert p.tolerance_ppm == case["tolerance_ppm"]
        assert p.top_k_fragments == case["top_k_fragments"]


# ---- 2. float64 restatement -------------------------------------------------
def _pearson(a, b):
    if len(a) == 0 or a.sum() == 0 or b.sum() == 0:
        return 0.0
    da, db = a - a.mean(), b - b.mean()
    va, vb = (da * da).sum(), (db * db).sum()
    return 0.0 if va == 0 or vb == 0 else float(np.clip((da * db).sum() / math.sqrt(va * vb), -1, 1))


def _quantified_f64(dia, lib_arrays, cands, tol, top_k):
    """The PrecursorQuantified of every candidate after the intensity > 0
    filter, in candidate order, in float64 where the reference sums or
    divides; the observed m/z and ppm error in float32 scalar arithmetic (a
    difference of nearby f32 values)."""
    (pidx, pmzl, pmz, prtl, prt, naa, fs, fe, fmzl, fmz, fint, card, chg, loss,
     fnum, fpos, ftype) = lib_arrays
    order = np.argsort(pidx, kind="stable")
    spidx = pidx[order]
    rt = dia.rt()
    out = []
    for i in range(len(cands[0])):
        key = cands[0][i]
        j = int(np.searchsorted(spidx, key))
        if j >= len(spidx) or spidx[j] != key:
            continue
        r = int(order[j])
        c0, c1, cc = int(cands[7][i]), int(cands[8][i]), int(cands[6][i])
        if c1 - c0 == 0 or cc < c0 or cc >= c1:
            continue
        a, b = int(fs[r]), int(fe[r])
        sel = a + oracle.filter_sort_fragments(fmz[a:b], fint[a:b], fnum[a:b], ftype[a:b],
                                               top_k=top_k)
        fm = fmz[sel]
        xic32, mz32 = dia.dense_xic_mz(pmz[r], c0, c1, tol, fm)
        xic = xic32.astype(np.float64)
        F, L = xic.shape
        c = L // 2
        s_, e_ = max(c - 1, 0), min(c + 2, L)
        mean = xic[:, s_:e_].mean(axis=1)
        norm = np.where(mean[:, None] > 0, xic / np.where(mean > 0, mean, 1)[:, None], 0.0)
        nz = (norm != 0).any(axis=1)
        medf = np.median(norm[nz], axis=0) if nz.any() else np.zeros(L)
        corr = np.array([_pearson(medf, norm[f]) for f in range(F)])
        inten = xic.sum(axis=1)
        s32 = np.float32
        mzo = np.zeros(F, np.float32)
        err = np.zeros(F, np.float32)
        for f in range(F):
            ws, wt = s32(0), s32(0)
            for t in range(L):
                w = xic32[f, t]
                if w > 0:
                    ws = s32(ws + s32(mz32[f, t] * w))
                    wt = s32(wt + w)
            m = s32(ws / wt) if wt > 0 else s32(0)
            mzo[f] = m
            err[f] = s32(s32(s32(m - fm[f]) / fm[f]) * s32(1e6)) if m > 0 else s32(0)
        keep = inten > 0
        if not keep.any():
            continue
        out.append({"precursor_idx": int(key), "rank": int(cands[1][i]),
                    "rt_observed": float(rt[cc]), "mz": float(pmz[r]), "naa": int(naa[r]),
                    "L": L,
                    "fragment_mz": fm[keep], "fragment_number": fnum[sel][keep],
                    "fragment_intensity": inten[keep], "fragment_mz_observed": mzo[keep],
                    "fragment_mass_error_observed": err[keep],
                    "fragment_correlation_observed": corr[keep]})
    return out


def _check_quantify(run0, od, ol):
    gen = run0
    spec = gen["spec"]
    cands = oracle.select(od, ol, oracle.params(**spec.selection_config()))
    r0 = cands[1] == 0
    sample = np.flatnonzero(r0)[:300]
    cand = tuple(c[sample] for c in cands)
    tol, top_k = 10.0, 10000
    prec, frag = oracle.quantify(od, ol, gen["lib"], cand, tolerance_ppm=tol,
                                 top_k_fragments=top_k)
    want = _quantified_f64(od, gen["lib"], cand, tol, top_k)
    assert len(want) > 100
    # precursor columns: stable sort of the quantified precursors by index
    wp = sorted(range(len(want)), key=lambda k: want[k]["precursor_idx"])
    assert prec["precursor_idx"].tolist() == [want[k]["precursor_idx"] for k in wp]
    assert prec["rank"].tolist() == [want[k]["rank"] for k in wp]
    assert np.array_equal(prec["rt_observed"],
                          np.array([want[k]["rt_observed"] for k in wp], np.float32))
    # fragments in candidate order
    pos = 0
    for q in want:
        n = len(q["fragment_mz"])
        sl = slice(pos, pos + n)
        assert np.all(frag["precursor_idx"][sl] == q["precursor_idx"])
        assert np.array_equal(frag["mz"][sl], q["fragment_mz"])
        assert np.array_equal(frag["number"][sl], q["fragment_number"])
        # an f32 sum of L non-negative terms lies within L * 2^-24 of the
        # exact sum, relative
        got_i = frag["intensity"][sl].astype(np.float64)
        assert np.all(np.abs(got_i - q["fragment_intensity"]) <=
                      q["L"] * 2.0 ** -24 * q["fragment_intensity"]), (q["precursor_idx"],)
        assert np.array_equal(frag["mz_observed"][sl], q["fragment_mz_observed"])
        assert np.array_equal(frag["mass_error_observed"][sl],
                              q["fragment_mass_error_observed"])
        assert np.all(np.abs(frag["correlation_observed"][sl] -
                             q["fragment_correlation_observed"]) <= 1e-4), (q["precursor_idx"],)
        pos += n
    assert pos == len(frag["mz"])


def test_quantify_against_float64_implementation(run0, oracle_run0):
    _check_quantify(run0, *oracle_run0)


# ---- 3. planted layout mistakes ----------------------------------------------
LAYOUT_MUTANTS = {
    1: ("intensity filter keeps the threshold itself",
        [lambda: test_filter_fragments_by_intensity_kats(GOLD["quant_filter_by_intensity"][1])]),
    2: ("fragments reordered with their sorted precursors",
        [test_speclib_keeps_fragments_in_input_order]),
}


@pytest.mark.parametrize("m", sorted(LAYOUT_MUTANTS), ids=lambda m: LAYOUT_MUTANTS[m][0])
def test_planted_layout_mistake_is_caught(m):
    prev = oracle.LAYOUT_MUTANT
    oracle.LAYOUT_MUTANT = m
    try:
        for check in LAYOUT_MUTANTS[m][1]:
            with pytest.raises(AssertionError):
                check()
    finally:
        oracle.LAYOUT_MUTANT = prev
