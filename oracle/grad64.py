"""FP64 model of the differentiable selection's rank-0 score and gradient.

TEST INFRASTRUCTURE: the checker of the device backward pass, used by
tests/ and bench.py's cpu_baseline leg.  The product never imports it.

Model, per precursor p with rank-0 apex a (fixed by the forward pass),
selected fragments f, XIC X (exact f32 sums from the oracle's
DenseXICObservation, src/dense_xic_observation/mod.rs:30-64) and the kernel
of its first valid quadrupole window o(p):

    s*_p = sum_f w_f ln(1 + Y_f),  Y_f = sum_j k_j(sigma_o) X_f[a - half + j]

with k_j = exp(-x_j^2 / 2 sigma^2) / sum_m exp(-x_m^2 / 2 sigma^2) over
x = -half .. half-1 (src/kernel/mod.rs:32-56; the normalisation constant
cancels) and the symmetric tap use of src/convolution/scalar.rs:5-56 (taps
0..half on both sides).  The loss is the hinge over target/decoy pairs,
sum max(0, m - (s*_t - s*_d)).  The gradient is analytic:
d s*/d w_f = ln(1 + Y_f) and d s*/d sigma = sum_f w_f (dY_f/dsigma) / (1 + Y_f)
with dk_j/dsigma = k_j (x_j^2 - sum_m k_m x_m^2) / sigma^3; it is pinned by
central finite differences of the same model (check_gradient).
"""
from __future__ import annotations

import numpy as np

from . import Dia, Lib, filter_sort_fragments, params, select_one


def oracle_state(gen, weights, sigma):
    """Oracle DIAData, library with `weights` as fragment intensities, and
    parameters with one sigma per window, for a synthetic run."""
    lib = list(gen["lib"])
    lib[10] = weights
    od, ol = Dia(gen["dia"]), Lib(tuple(lib))
    prm = params(**gen["spec"].selection_config(),
                 rt_tolerance_per_precursor=gen["rt_tolerance_per_precursor"],
                 sigma_per_window=sigma)
    return od, ol, prm


class F64Model:
    """The FP64 restatement above, over the rows with apex >= 0."""

    def __init__(self, gen, weights, sigma, apex, dia=None, lib_o=None, prm=None):
        spec = gen["spec"]
        if dia is None:
            od, ol, prm = oracle_state(gen, weights, sigma)
        else:  # oracle handles already built with these weights and sigma
            od, ol = dia, lib_o
        lib = gen["lib"]
        self.klen = int(np.ceil(spec.kernel_size / 2.0) * 2)
        self.half = self.klen // 2
        self.rows = []  # (row, window, fragment indices, X window f64)
        fs, fe = lib[6].astype(np.int64), lib[7].astype(np.int64)
        for r in np.flatnonzero(apex >= 0):
            one = select_one(od, ol, prm, int(r), max_cells=64 * 4096)
            assert one is not None
            c0 = one["cycle_start"]
            a = int(apex[r]) - c0
            assert self.half <= a < one["xic"].shape[1] - self.half
            sel = filter_sort_fragments(
                lib[9][fs[r]:fe[r]], weights[fs[r]:fe[r]], lib[14][fs[r]:fe[r]],
                lib[16][fs[r]:fe[r]], top_k=spec.top_k_fragments)
            assert np.array_equal(lib[9][fs[r]:fe[r]][sel], one["fragment_mz"])
            obs = od.valid_observations(float(lib[2][r]))
            X = one["xic"][:, a - self.half:a + self.half + 1].astype(np.float64)
            self.rows.append((int(r), int(obs[0]), fs[r] + sel, X))
        self.n_prec = len(lib[0])
        self.n_frag = len(lib[10])

    def taps(self, s):
        x = np.arange(-self.half, self.half, dtype=np.float64)
        e = np.exp(-x * x / (2.0 * s * s))
        k = e / e.sum()
        dk = k * (x * x - (k * x * x).sum()) / (s ** 3)
        return k[: self.half + 1], dk[: self.half + 1]

    def _y(self, X, k):
        h = self.half
        # Y = X[a] k[h] + sum_d (X[a-d] + X[a+d]) k[h-d]; X column h is the apex
        sym = np.concatenate([X[:, :h] + X[:, h + 1:][:, ::-1], X[:, h:h + 1]], axis=1)
        return sym @ k

    def s_star(self, sigma, w):
        s = np.zeros(self.n_prec)
        for r, o, fi, X in self.rows:
            k, _ = self.taps(float(sigma[o]))
            s[r] = (w[fi] * np.log1p(self._y(X, k))).sum()
        return s

    def grads(self, sigma, w, g):
        gw = np.zeros(self.n_frag)
        gs = np.zeros(len(sigma))
        for r, o, fi, X in self.rows:
            k, dk = self.taps(float(sigma[o]))
            y = self._y(X, k)
            gw[fi] = g[r] * np.log1p(y)
            gs[o] += g[r] * (w[fi] * self._y(X, dk) / (1.0 + y)).sum()
        return gw, gs


def hinge64(s, pairs, m):
    return np.maximum(0.0, m - (s[pairs[:, 0]] - s[pairs[:, 1]])).sum()


def rel_errors(got, ref, rtol):
    """(max relative error, indices beyond rtol) with an absolute floor of
    rtol * 1e-6 * max|ref| for entries that are zero or nearly so."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    floor = rtol * 1e-6 * max(np.abs(ref).max(), 1e-30) if ref.size else 1e-30
    err = np.abs(got - ref)
    bad = np.flatnonzero(err > rtol * np.abs(ref) + floor)
    rel = float((err / np.maximum(np.abs(ref), floor)).max()) if ref.size else 0.0
    return rel, bad


def pick_margin(s64, pairs):
    """A hinge margin away from every kink of the pairs' score differences,
    with some pairs active and some not."""
    diff = s64[pairs[:, 0]] - s64[pairs[:, 1]]
    grid = np.linspace(np.quantile(diff, 0.3), np.quantile(diff, 0.7), 201)
    return float(grid[np.argmax([np.abs(diff - x).min() for x in grid])])


def check_gradient(model, sigma, weights, pairs, m, h=1e-4, n_weight_fd=40, seed=3):
    """The analytic FP64 gradient of the hinge over `pairs` at margin m and
    its check against FP64 central differences of the same model (every
    sigma, a sample of the non-zero weight gradients).  Returns
    (grad weights, grad sigma, max relative error of the sigma gradient
    against the differences, worst ratio |fd - g| / (1e-4 |g| + 1e-9) of the
    sampled weights: the check passes at <= 1)."""
    sig64, w64 = sigma.astype(np.float64), weights.astype(np.float64)
    s64 = model.s_star(sig64, w64)
    diff = s64[pairs[:, 0]] - s64[pairs[:, 1]]
    active = (m - diff) > 0
    g = np.zeros(model.n_prec)
    np.add.at(g, pairs[active, 0], -1.0)
    np.add.at(g, pairs[active, 1], 1.0)
    gw, gs = model.grads(sig64, w64, g)

    def L(sg, ww):
        return hinge64(model.s_star(sg, ww), pairs, m)

    fd_s = np.zeros_like(sig64)
    for o in range(len(sig64)):
        e = np.zeros_like(sig64)
        e[o] = h
        fd_s[o] = (L(sig64 + e, w64) - L(sig64 - e, w64)) / (2 * h)
    rel_s, _ = rel_errors(gs, fd_s, 1e-4)
    # weights: |fd - g| <= 1e-4 |g| + 1e-9 (a loss of order 1e3 carries about
    # 2e-13 of rounding, about 1e-9 once divided by 2h); the returned figure
    # is the largest |fd - g| / (1e-4 |g| + 1e-9), so <= 1 passes
    nz = np.flatnonzero(gw)
    worst_w = 0.0
    for j in np.random.default_rng(seed).choice(nz, size=min(n_weight_fd, nz.size),
                                                replace=False):
        e = np.zeros_like(w64)
        e[j] = h
        fd = (L(sig64, w64 + e) - L(sig64, w64 - e)) / (2 * h)
        worst_w = max(worst_w, abs(fd - gw[j]) / (1e-4 * abs(gw[j]) + 1e-9))
    return gw, gs, rel_s, worst_w
