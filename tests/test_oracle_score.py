"""The scoring restatement (oracle/adsr_oracle_score.c) pinned three ways.

1. The reference's own unit tests, transcribed: src/peak_group_scoring/
   tests.rs, src/utils/tests.rs, src/idf/tests.rs and
   src/dense_xic_observation/tests.rs (file:line at each case).
2. Closed forms and textbook routines: Pearson correlation, medians and
   sample standard deviations against numpy in float64, Stirling's formula
   against math.lgamma within its error bound.
3. An independent float64 numpy implementation of every feature of
   PeakGroupScoring.score (src/peak_group_scoring/mod.rs:105-341), run on
   the candidates of workload 0 and compared column by column; this pins
   the glue: which inputs feed which feature and the column order.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle

u8 = np.uint8
B, Y = 98, 121


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def P(a):
    return oracle.p(a)


def corr(x, y):
    x, y = f32(x), f32(y)
    return oracle.lib().orc_correlation_safe(P(x), P(y), len(x))


def median_axis_0(a):
    a = f32(a)
    out = np.zeros(a.shape[1], np.float32)
    oracle.lib().orc_median_axis_0(P(a), a.shape[0], a.shape[1], P(out))
    return out


def normalize(a, dil):
    a = f32(a)
    out = np.zeros_like(a)
    oracle.lib().orc_normalize_profiles(P(a), a.shape[0], a.shape[1], dil, P(out))
    return out


def ranked_mask(v, r1, r2):
    v = f32(v)
    m = np.zeros(max(len(v), 1), u8)
    oracle.lib().orc_ranked_mask(P(v), len(v), r1, r2, P(m))
    return [bool(x) for x in m[:len(v)]]


def hyperscore(types, inten, matched, w=None):
    t, i, m = np.array(types, u8), f32(inten), np.array(matched, u8)
    ww = None if w is None else f32(w)
    return oracle.lib().orc_hyperscore_weighted(P(t), P(i), P(m), P(ww), len(t))


def longest(types, nums, matched):
    b, y = C.c_uint8(), C.c_uint8()
    t, n, m = np.array(types, u8), np.array(nums, u8), np.array(matched, u8)
    oracle.lib().orc_longest_ion_series(P(t), P(n), P(m), len(t), C.byref(b), C.byref(y))
    return b.value, y.value


# ---- 1. reference KATs -------------------------------------------------------
def test_median_axis_0_kats():
    # src/peak_group_scoring/tests.rs:13-34
    assert median_axis_0([[1, 2, 3], [4, 5, 6], [7, 8, 9]]).tolist() == [4, 5, 6]
    assert median_axis_0([[1, 2], [3, 4], [5, 6], [7, 8]]).tolist() == [4, 5]
    assert median_axis_0([[1, 2], [3, 4], [5, 6]]).tolist() == [3, 4]


def test_normalize_profiles_kats():
    # src/peak_group_scoring/tests.rs:37-103
    r = normalize([[1, 2, 3, 4, 5], [2, 4, 6, 8, 10], [0, 0, 0, 0, 0]], 1)
    third = np.float32(1.0) / np.float32(3.0)
    assert r[0, 0] == third and r[0, 4] == np.float32(5.0) / np.float32(3.0)
    assert r[1, 0] == np.float32(2.0) / np.float32(6.0)
    assert (r[2] == 0).all()
    r = normalize([[1, 2], [3, 4]], 1)
    assert r[0, 0] == np.float32(1.0) / np.float32(1.5)
    assert r[1, 1] == np.float32(4.0) / np.float32(3.5)
    r = normalize([[1, 0, 3], [1, 2, 3]], 0)
    assert (r[0] == 0).all() and r[1].tolist() == [0.5, 1.0, 1.5]


def test_correlation_kats():
    # src/peak_group_scoring/tests.rs:106-190
    assert abs(corr([1, 2, 3], [1, 2, 3]) - 1) < 1e-6
    assert abs(corr([1, 2, 3], [2, 4, 6]) - 1) < 1e-6
    assert abs(corr([1, 2, 3], [3, 2, 1]) + 1) < 1e-6
    assert corr([1, 2, 3], [0, 0, 0]) == 0.0
    assert corr([1, 2], [1, 1]) == 0.0
    assert corr([0, 0, 0], [1, 2, 3]) == 0.0


def test_hyperscore_kats():
    # src/peak_group_scoring/tests.rs:213-345
    assert hyperscore([B, Y], [100, 200], [0, 0]) == 0.0
    h = hyperscore([B, B, Y, Y], [100, 200, 150, 250], [1, 1, 1, 1])
    assert 10.0 < h < 20.0
    h1 = hyperscore([B, Y], [50, 75], [1, 1])
    assert 7.0 < h1 < 10.0
    # 1! = Γ(2) is exact zero by the special case (utils.rs:302-304)
    assert h1 == np.float32(oracle.lib().orc_logf(50.0)) + np.float32(oracle.lib().orc_logf(75.0))
    assert hyperscore([B, B], [100, 200], [1, 1]) > 0
    # inverse mass error (tests.rs:614-748): smaller errors score higher
    small = hyperscore([B, Y], [1000, 2000], [1, 1], 1 / (np.abs([1.0, 1.0]) + 0.1))
    large = hyperscore([B, Y], [1000, 2000], [1, 1], 1 / (np.abs([50.0, 50.0]) + 0.1))
    assert small > large and math.isfinite(large)
    assert hyperscore([B, Y], [0, 0], [1, 1], [1, 1]) == 0.0


def test_gamma_ln_against_lgamma():
    # Stirling's series: ln Γ(x) - stirling(x) = 1/(12x) - ..., so the
    # restated approximation sits within 1/(12x) + f32 rounding of lgamma
    L = oracle.lib()
    assert L.orc_gamma_ln(1.0) == 0.0 and L.orc_gamma_ln(2.0) == 0.0
    assert L.orc_gamma_ln(0.0) == 0.0 and L.orc_gamma_ln(-3.0) == 0.0
    for n in range(2, 65):
        x = n + 1.0
        assert abs(L.orc_gamma_ln(x) - math.lgamma(x)) <= 1 / (12 * x) + 1e-5 * x


def test_longest_ion_series_kats():
    # src/peak_group_scoring/tests.rs:347-610
    assert longest([B, B, B, Y, Y, Y], [1, 2, 3, 1, 2, 3], [1] * 6) == (3, 3)
    assert longest([B, B, Y, Y], [3, 4, 1, 3], [1] * 4) == (2, 1)
    assert longest([B, B, B, Y, Y, Y], [2, 3, 4, 1, 2, 3], [0, 1, 1, 1, 1, 0]) == (2, 2)
    assert longest([B, Y], [1, 1], [0, 0]) == (0, 0)
    assert longest([B, B, B, B], [1, 2, 4, 5], [1] * 4) == (2, 0)
    assert longest([B, B, B, Y, Y], [3, 1, 2, 3, 2], [1] * 5) == (3, 2)
    assert longest([B, B, B, Y, Y, Y], [2, 2, 2, 3, 3, 3], [1] * 6) == (1, 1)
    assert longest([], [], []) == (0, 0)


def test_intensity_series_dot_fwhm_kats():
    L = oracle.lib()
    t = np.array([97, B, 99, 120, Y, 122], u8)
    i = f32([50, 100, 75, 80, 200, 60])
    m = np.ones(6, u8)
    # src/peak_group_scoring/tests.rs:834-892
    assert [L.orc_intensity_ion_series(P(t), P(i), P(m), 6, k)
            for k in (97, B, 99, 120, Y, 122)] == [50, 100, 75, 80, 200, 60]
    t2, i2, m2 = np.array([B, Y], u8), f32([100, 200]), np.array([0, 1], u8)
    assert L.orc_intensity_ion_series(P(t2), P(i2), P(m2), 2, B) == 0.0
    assert L.orc_intensity_ion_series(P(t2), P(i2), P(m2), 2, Y) == 200.0
    # :945-979
    a, b = f32([1, 2, 3]), f32([4, 5, 6])
    assert L.orc_dot(P(a), P(b), 3) == 32.0 and L.orc_dot(P(a), P(b), 0) == 0.0
    # :914-931 triangular peak, RT 1 per cycle: half maximum one cycle out
    prof, rt = f32([0, 50, 100, 50, 0]), f32(np.arange(17))
    assert 0.0 < L.orc_fwhm_rt(P(prof), 5, 10, P(rt)) < 5.0
    assert L.orc_fwhm_rt(P(prof), 0, 0, P(rt)) == 0.0


def test_crate_utils_kats():
    L = oracle.lib()
    # src/utils/tests.rs:275-291
    def med(v):
        v = f32(v)
        return L.orc_median(P(v), len(v))

    def std(v):
        v = f32(v)
        return L.orc_std(P(v), len(v))
    assert med([3, 1, 4, 1, 5]) == 3.0 and med([2, 4, 1, 3]) == 2.5
    assert med([42]) == 42.0 and med([]) == 0.0 and med([-3, 1, -1, 3, 0]) == 0.0
    assert abs(std([1, 2, 3, 4, 5]) - 1.5811) < 1e-4 and std([42]) == 0.0
    assert abs(std([1, 5]) - 2.8284) < 1e-4 and std([7, 7, 7]) == 0.0
    assert std([-2, 0, 2]) == 2.0
    # :295-412 create_ranked_mask
    assert ranked_mask([1, 3, 2], 0, 2) == [False, True, True]
    assert ranked_mask([5, 1, 3, 2], 0, 2) == [True, False, True, False]
    assert ranked_mask([1, 1, 1], 0, 2) == [True, True, False]
    assert ranked_mask([3, 1, 3, 2], 0, 2) == [True, False, True, False]
    assert ranked_mask([1, 2, 3], 2, 1) == [False] * 3
    assert ranked_mask([1, 2, 3], 1, 5) == [True, True, False]
    assert ranked_mask([1, 2, 3, 4, 5], 0, 3) == [False, False, True, True, True]
    assert ranked_mask([-1, 2, -3, 4], 0, 2) == [False, True, False, True]
    assert ranked_mask([5, 1, 3, 2, 4], 2, 4) == [False, False, True, True, False]
    assert ranked_mask([5, 1, 3, 2, 4], 4, 5) == [False, True, False, False, False]
    # :416-440 count_values_above
    v = f32([1, 2, 3, 4])
    assert L.orc_count_above(P(v), 4, 2.5, None) == 2
    assert L.orc_count_above(P(f32([2, 2, 2])), 3, 2.0, None) == 0
    assert L.orc_count_above(P(v), 4, 1.5, P(np.array([1, 0, 1, 0], u8))) == 1
    # :89-193 weighted mean absolute error
    def wmae(e, w):
        e, w = f32(e), f32(w)
        return L.orc_weighted_mae(P(e), P(w), len(e))
    assert abs(wmae([10, -20, 30, -40], [1, 2, 3, 4]) - 30) < 1e-5
    assert abs(wmae([10, -20, 30, -40], [1, 0, 3, 0]) - 25) < 1e-5
    assert abs(wmae([10, 0, 30, 0], [1, 2, 3, 4]) - 25) < 1e-5
    assert wmae([10, -20, 30], [0, 0, 0]) == 0.0 and wmae([], []) == 0.0
    assert wmae([25], [2]) == 25.0
    assert abs(wmae([10, 20, 30], [1, -2, 3]) - 25) < 1e-5
    # :5-85, 197-272 weighted mean m/z and mass errors
    def mzerr(mz, xic, lib_mz):
        mz, xic, lm = f32(mz), f32(xic), f32(lib_mz)
        mo, er = np.zeros(len(lm), np.float32), np.zeros(len(lm), np.float32)
        L.orc_fragment_mz_errors(P(mz), P(xic), mz.shape[0], mz.shape[1], P(lm), P(mo), P(er))
        return mo, er
    mo, _ = mzerr([[10, 20, 30], [5, 15, 25]], [[1, 2, 3], [4, 0, 2]], [1, 1])
    assert abs(mo[0] - 23.333334) < 1e-5 and abs(mo[1] - 11.666667) < 1e-5
    mo, _ = mzerr([[100, 200, 300, 400], [10, 20, 30, 40]],
                  [[1, 0, 3, 0], [0, 2, 0, 4]], [1, 1])
    assert mo[0] == 250.0 and abs(mo[1] - 33.333332) < 1e-5
    mo, er = mzerr([[200.01, 200.02, 200.03], [300.05, 300.10, 300.15], [0, 0, 0]],
                   [[1000, 2000, 3000], [500, 1000, 500], [0, 0, 0]], [200, 300, 400])
    assert abs(mo[0] - 200.02333) < 1e-5 and abs(er[0] - 116.67) < 0.1
    assert abs(mo[1] - 300.1) < 1e-5 and abs(er[1] - 333.33) < 0.1
    assert mo[2] == 0.0 and er[2] == 0.0
    mo, er = mzerr([[250.005]], [[1500.0]], [250.0])
    assert mo[0] == np.float32(250.005) and abs(er[0] - 20.0) < 0.1


def test_idf_kats():
    L = oracle.lib()
    # src/idf/tests.rs:3-41
    for lib_mz, query, want in [([100.0, 100.0], [100.0, 200.0], [0.0, math.log(2.0)]),
                                ([150.0], [150.0], [0.0])]:
        lm = f32(lib_mz)
        c = L.orc_idf_counts(P(lm), len(lm))
        got = [L.orc_idf(c, float(len(lm)), q) for q in query]
        assert all(abs(g - w) < 1e-6 for g, w in zip(got, want))
    c = L.orc_idf_counts(None, 0)
    assert [L.orc_idf(c, 0.0, q) for q in (100.0, 200.0, 300.0)] == [1.0, 1.0, 1.0]


def _raw(delta, lo, hi, ps, pe, cyc, rt, mz, it):
    return (np.array(delta, np.int64), f32(lo), f32(hi), np.array(ps, np.int64),
            np.array(pe, np.int64), np.array(cyc, np.int64), f32(rt), f32(mz), f32(it),
            np.zeros((1, 1, 1, 1), np.float32))


def test_dense_xic_mz_kats():
    # src/dense_xic_observation/tests.rs:127-228
    d = oracle.Dia(_raw([0], [199], [201], [0], [2], [10], [100], [200.0, 200.1], [1000, 2000]))
    xic, mz = d.dense_xic_mz(200.0, 10, 11, 1000.0, [200.05])
    assert xic[0, 0] == 3000.0
    assert abs(mz[0, 0] - (200.0 * 1000.0 + 200.1 * 2000.0) / 3000.0) < 1e-4
    d = oracle.Dia(_raw([0], [249], [251], [0], [3], [5], [50], [249.9, 250.0, 250.1],
                        [500, 1000, 1500]))
    xic, mz = d.dense_xic_mz(250.0, 5, 6, 2000.0, [250.0])
    assert xic[0, 0] == 3000.0
    assert abs(mz[0, 0] - (249.9 * 500 + 250.0 * 1000 + 250.1 * 1500) / 3000) < 1e-4
    d = oracle.Dia(_raw([0, 1], [299, 299], [301, 301], [0, 2], [2, 4], [0, 1], [10, 20],
                        [300.0, 300.1, 300.0, 300.1], [1000, 0, 0, 2000]))
    xic, mz = d.dense_xic_mz(300.0, 0, 2, 2000.0, [300.05])
    assert xic[0].tolist() == [1000.0, 2000.0]
    assert abs(mz[0, 0] - 300.0) < 2e-4 and abs(mz[0, 1] - 300.1) < 2e-4


# ---- 2. textbook routines -----------------------------------------------------
def test_correlation_median_std_against_numpy():
    rng = np.random.default_rng(3)
    L = oracle.lib()
    for n in (2, 3, 11, 24, 64):
        for _ in range(20):
            x = rng.lognormal(2, 1, n).astype(np.float32)
            y = (x * rng.uniform(0.5, 1.5, n)).astype(np.float32)
            want = np.corrcoef(x.astype(np.float64), y.astype(np.float64))[0, 1]
            assert abs(corr(x, y) - want) < 2e-5
            assert L.orc_median(P(x), n) == np.float32(np.median(x.astype(np.float64)))
            assert abs(L.orc_std(P(x), n) - np.std(x.astype(np.float64), ddof=1)) <= \
                1e-5 * np.std(x.astype(np.float64), ddof=1) + 1e-6


# ---- 3. every feature against an independent float64 implementation ---------
def _features_f64(dia, gen, cand, tol, top_k, counts, total, keep):
    """PeakGroupScoring.score features for one candidate, in float64 numpy
    from the dense XIC/MZ and the filtered fragments."""
    pmz, prt, naa, frags = keep
    pidx, _rank, score, cs, ce, cc = cand
    row = int(np.searchsorted(gen["sorted_pidx"], pidx))
    fmz, fint, ftype, fnum, fmzl = frags[row]
    F = len(fmz)
    c0, c1 = int(cs), int(ce)
    xic32, mz32 = dia.dense_xic_mz(pmz[row], c0, c1, tol, fmz)
    xic = xic32.astype(np.float64)
    L = xic.shape[1]
    c = L // 2
    s, e = max(c - 1, 0), min(c + 2, L)
    mean = xic[:, s:e].mean(axis=1) if F else np.zeros(0)
    norm = np.where(mean[:, None] > 0, xic / np.where(mean > 0, mean, 1)[:, None], 0.0)
    nz = (norm != 0).any(axis=1)
    med = np.median(norm, axis=0) if F else np.zeros(L)
    medf = np.median(norm[nz], axis=0) if nz.any() else np.zeros(L)

    def pearson(a, b):
        if len(a) == 0 or a.sum() == 0 or b.sum() == 0:
            return 0.0
        da, db = a - a.mean(), b - b.mean()
        va, vb = (da * da).sum(), (db * db).sum()
        return 0.0 if va == 0 or vb == 0 else float(np.clip((da * db).sum() / np.sqrt(va * vb), -1, 1))
    corrs = np.array([pearson(medf, norm[f]) for f in range(F)])
    oint = xic.sum(axis=1)
    matched = oint > 0
    rank_int = np.argsort(np.argsort(-fint, kind="stable"), kind="stable")

    def hs(inten, w=None):
        w = np.ones(F) if w is None else w
        sel = matched & (inten != 0)
        nb, ny = int((sel & (ftype == B)).sum()), int((sel & (ftype == Y)).sum())
        if nb == 0 and ny == 0:
            return 0.0
        sb = (inten * w)[sel & (ftype == B)].sum()
        sy = (inten * w)[sel & (ftype == Y)].sum()
        st = lambda n: 0.0 if n <= 1 else (n + 0.5) * math.log(n + 1) - (n + 1) + 0.5 * 1.837877
        return st(nb) + st(ny) + (math.log(sb) if sb > 0 else 0) + (math.log(sy) if sy > 0 else 0)

    def longest(t):
        v = sorted(fnum[matched & (ftype == t)].tolist())
        best = cur = 1 if v else 0
        for a, b in zip(v, v[1:]):
            cur = cur + 1 if b == a + 1 else 1
            best = max(best, cur)
        return best
    # observed m/z and ppm error in float32 scalar arithmetic: the error is a
    # difference of nearby f32 values, which float64 would not reproduce
    s32 = np.float32
    err = np.zeros(F)
    for f in range(F):
        ws, wt = s32(0), s32(0)
        for j in range(L):
            wj = xic32[f, j]
            if wj > 0:
                ws = s32(ws + s32(mz32[f, j] * wj))
                wt = s32(wt + wj)
        m = s32(ws / wt) if wt > 0 else s32(0)
        lm = s32(fmz[f])
        err[f] = float(s32(s32(s32(m - lm) / lm) * s32(1e6))) if m > 0 else 0.0
    ok = (fint > 0) & (err != 0)
    wmae = (np.abs(err) * fint)[ok].sum() / fint[ok].sum() if ok.any() else 0.0
    g = oracle.grid()
    bins = [oracle.lib().orc_find_closest_index(oracle.p(g), len(g), float(m)) for m in fmzl]
    idf = np.array([math.log(total / max(int(counts[b]), 1)) for b in bins])
    rank_idf = np.argsort(np.argsort(-idf, kind="stable"), kind="stable")
    rt = dia.rt()
    h = L // 2
    fwhm = 0.0
    for i in range(h):
        if (medf[h - i] + medf[h + i]) / 2 <= medf[h] / 2:
            # usize arithmetic wraps for candidates centred before peak_length
            fwhm = float(rt[(c0 + h + i) % (1 << 64)] - rt[(c0 + h - i) % (1 << 64)])
            break
    cnt = lambda t, m=None: float(((corrs > t) & (np.ones(F, bool) if m is None else m)).sum())
    rng = lambda r1, r2: (rank_int >= r1) & (rank_int < r2)
    ib = oint[matched & (ftype == B)].sum()
    iy = oint[matched & (ftype == Y)].sum()
    return {
        "score": score, "mean_correlation": corrs.mean() if F else 0.0,
        "median_correlation": float(np.median(corrs)) if F else 0.0,
        "correlation_std": float(np.std(corrs, ddof=1)) if F > 1 else 0.0,
        "intensity_correlation": pearson(oint, fint), "num_fragments": F,
        "num_scans": L, "num_over_95": cnt(0.95), "num_over_90": cnt(0.90),
        "num_over_80": cnt(0.80), "num_over_50": cnt(0.5), "num_over_0": cnt(0.0),
        "num_over_0_rank_0_5": cnt(0, rng(0, 6)), "num_over_0_rank_6_11": cnt(0, rng(6, 12)),
        "num_over_0_rank_12_17": cnt(0, rng(12, 18)), "num_over_0_rank_18_23": cnt(0, rng(18, 24)),
        "num_over_50_rank_0_5": cnt(0.5, rng(0, 6)), "num_over_50_rank_6_11": cnt(0.5, rng(6, 12)),
        "num_over_50_rank_12_17": cnt(0.5, rng(12, 18)),
        "num_over_50_rank_18_23": cnt(0.5, rng(18, 24)),
        "hyperscore_intensity_observation": hs(oint), "hyperscore_intensity_library": hs(fint),
        "hyperscore_inverse_mass_error": hs(oint, 1 / (np.abs(err) + 0.1)),
        "rt_observed": rt[int(cc)], "delta_rt": rt[int(cc)] - prt[row],
        "longest_b_series": longest(B), "longest_y_series": longest(Y), "naa": naa[row],
        "weighted_mass_error": wmae, "log10_b_ion_intensity": math.log10(ib + 1e-8),
        "log10_y_ion_intensity": math.log10(iy + 1e-8), "fwhm_rt": fwhm,
        "idf_hyperscore": hs(idf), "idf_xic_dot_product": float((idf * corrs).sum()),
        "idf_intensity_dot_product": math.log10((idf * oint).sum() + 1e-8),
        "median_profile_sum": med.sum(), "median_profile_sum_filtered": medf.sum(),
        "num_profiles": F, "num_profiles_filtered": float(nz.sum()),
        "num_over_0_top6_idf": cnt(0, rank_idf < 6), "num_over_50_top6_idf": cnt(0.5, rank_idf < 6),
    }, corrs


THRESHOLDS = (0.0, 0.5, 0.8, 0.9, 0.95)


def test_score_features_against_float64_implementation(run0, oracle_run0):
    od, ol = oracle_run0
    gen = run0
    spec = gen["spec"]
    cands = oracle.select(od, ol, oracle.params(**spec.selection_config()))
    tol, top_k = 10.0, 100
    feats = oracle.score(od, ol, cands, mass_tolerance=tol, top_k_fragments=top_k)
    assert len(feats["precursor_idx"]) == len(cands[0]) > 100
    assert list(feats)[2:] == oracle.FEATURES
    lib_arrays = gen["lib"]
    (pidx, _pmzl, pmz, _prtl, prt, naa, fs, fe, fmzl, fmz, fint, _c, _ch, _lo, fnum,
     _pos, ftype) = lib_arrays
    order = np.argsort(pidx, kind="stable")
    gen = dict(gen, sorted_pidx=pidx[order])
    frags = []
    L = oracle.lib()
    for r in order:
        a, b = int(fs[r]), int(fe[r])
        sel = np.zeros(max(b - a, 1), np.uintp)
        k = L.orc_filter_sort_fragments(b - a, P(f32(fmz[a:b])), P(f32(fint[a:b])),
                                        P(np.ascontiguousarray(fnum[a:b])),
                                        P(np.ascontiguousarray(ftype[a:b])), 1, 1, top_k,
                                        P(sel))
        idx = a + sel[:k].astype(np.int64)
        frags.append((fmz[idx].astype(np.float64), fint[idx].astype(np.float64),
                      ftype[idx], fnum[idx], fmzl[idx]))
    counts = np.ctypeslib.as_array(L.orc_idf_counts(P(f32(fmzl)), len(fmzl)),
                                   (len(oracle.grid()),)).copy()
    keep = (pmz[order], prt[order], naa[order], frags)
    rows = np.random.default_rng(0).choice(len(cands[0]), 150, replace=False)
    loose = {"log10_b_ion_intensity", "log10_y_ion_intensity", "idf_intensity_dot_product"}
    ppm = {}
    checked = 0
    for i in rows:
        # to_arrays order: precursor_idx, rank, score, scan_*, cycle_center,
        # cycle_start, cycle_stop
        cand = (cands[0][i], cands[1][i], cands[2][i], cands[7][i], cands[8][i], cands[6][i])
        want, corrs = _features_f64(od, gen, cand, tol, top_k, counts, float(len(fmzl)), keep)
        near = any(np.any(np.abs(corrs - t) < 1e-4) for t in THRESHOLDS)
        for name, w in want.items():
            g = float(feats[name][i])
            if name.startswith("num_over") and near:
                continue
            rel = ppm.get(name, 1e-4)
            assert abs(g - float(w)) <= rel * max(1.0, abs(float(w))), (i, name, g, w)
        checked += 1
    assert checked == 150
