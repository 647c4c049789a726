"""The four GPU workloads at their stated sizes, against the oracle.

  workload 1  6,000 cycles x 60 windows, 1,500 peaks per spectrum (554M
              peaks), 400k precursors (200k targets + 200k decoys), 12
              fragments, +-150 cycles, 10 ppm, top-3
  workload 2  the same run with 6-24 fragments, per-precursor half-windows
              of 30-200 cycles, clustered precursor m/z, top_k 24
  workload 3  workload 2's run: forward, hinge loss and backward over all
              400k precursors; the gradient of a loss over 500 target/decoy
              pairs (1,000 precursors) checked against the FP64 model and
              FP64 central differences
  workload 4  20,000 cycles x 100 windows, 1,000 peaks per spectrum (2.07B
              peaks), 2M precursors, 6-24 fragments, half-windows 50-400,
              5 ppm, top-5

Every candidate table is compared with oracle.select over the whole library,
bit for bit (src/peak_group_selection/mod.rs:94-128: the rayon collect keeps
library order).  The DIAData is the device build, which is the default of
DIAData.from_arrays when a device is present.
"""
import os
import sys
import time

import numpy as np
import pytest

import oracle
import synthetic
from conftest import assert_tables_equal, oracle_params_for, params_for

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _log(msg):
    print(f"[fullsize] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def ads():
    import alphadia_search_rs as m
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    return m


def _gen(cid):
    t = time.time()
    gen = synthetic.generate(synthetic.config(cid), threads=_threads())
    _log(f"workload {cid}: {gen['n_peaks']} peaks, "
         f"{gen['spec'].n_precursors} precursors generated in {time.time() - t:.1f}s")
    return gen


def _rt_extra(gen):
    r = gen.get("rt_tolerance_per_precursor")
    return {} if r is None else {"rt_tolerance_per_precursor": r}


def _device_and_oracle(ads, gen):
    t = time.time()
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(gen["spec"], **_rt_extra(gen))
    got = ads.PeakGroupSelection(p).search(d, lib).to_arrays()
    _log(f"device build + search {time.time() - t:.1f}s, {len(got[0])} candidates")
    t = time.time()
    od, ol = oracle.Dia(gen["dia"], nthreads=_threads()), oracle.Lib(gen["lib"])
    want = oracle.select(od, ol, oracle_params_for(gen["spec"], **_rt_extra(gen)),
                         nthreads=_threads())
    _log(f"oracle build + select {time.time() - t:.1f}s, {len(want[0])} candidates")
    return d, lib, p, got, want, od, ol


def test_workload1_full_size(ads):
    gen = _gen(1)
    _, _, _, got, want, _, _ = _device_and_oracle(ads, gen)
    assert len(want[0]) > 1_000_000
    assert_tables_equal(got, want, "workload 1, 400k precursors")


@pytest.fixture(scope="module")
def run2(ads):
    gen = _gen(2)
    d, lib, p, got, want, od, ol = _device_and_oracle(ads, gen)
    out = dict(gen=gen, d=d, lib=lib, p=p, got=got, want=want, od=od, ol=ol)
    yield out
    out.clear()


def test_workload2_full_size(run2):
    assert len(run2["want"][0]) > 1_000_000
    assert_tables_equal(run2["got"], run2["want"], "workload 2, 400k precursors")


def test_workload3_full_size_gradient_on_pair_subset(ads, run2):
    """Forward over all 400k precursors at init (sigma 2.0 per window, the
    library intensities as weights) equals workload 2's search bit for bit.
    Then with perturbed sigma and weights: the forward equals the oracle's
    rank-0 scores over all 400k precursors, and the hinge loss over the
    first 500 target/decoy pairs (1,000 precursors) is differentiated
    through the full-size state.  Only those 1,000 rows have a non-zero
    loss gradient, so the device's sigma gradient is their sum alone; it and
    the weight gradient must match the FP64 model within 1e-4 relative, and
    the FP64 gradient must match FP64 central differences (h = 1e-4)."""
    import torch
    from alphadia_search_rs.autograd import SelectionGradState, hinge_loss, rank_zero_scores
    from oracle.grad64 import F64Model, hinge64 as _hinge64
    from test_gpu_grad import _close, _rank0, H, RTOL

    gen, d, lib, p = run2["gen"], run2["d"], run2["lib"], run2["p"]
    P = lib.num_precursors
    W = d.num_observations
    st = SelectionGradState(d, lib, p)
    t = time.time()
    w0 = torch.tensor(gen["lib"][10], device="cuda")
    s0, _ = st.forward(torch.full((W,), 2.0), w0)
    want0, _ = _rank0(run2["want"], P)
    np.testing.assert_array_equal(s0.cpu().numpy().view(np.uint32), want0.view(np.uint32))
    _log(f"workload 3 forward at init equals the search ({time.time() - t:.1f}s)")

    rng = np.random.default_rng(7)
    sigma = np.linspace(1.6, 2.8, W).astype(np.float32)
    weights = (gen["lib"][10] * rng.uniform(0.5, 1.5, gen["lib"][10].shape)).astype(np.float32)
    lib_w = list(gen["lib"])
    lib_w[10] = weights
    ol_w = oracle.Lib(tuple(lib_w))
    prm = oracle.params(**gen["spec"].selection_config(),
                        rt_tolerance_per_precursor=gen["rt_tolerance_per_precursor"],
                        sigma_per_window=sigma)
    t = time.time()
    want_s, apex = _rank0(oracle.select(run2["od"], ol_w, prm, nthreads=_threads()), P)
    _log(f"oracle with sigma/weights {time.time() - t:.1f}s")

    sig_t = torch.tensor(sigma, requires_grad=True)
    w_t = torch.tensor(weights, device="cuda", requires_grad=True)
    s_t = rank_zero_scores(st, sig_t, w_t)
    s_gpu = s_t.detach().cpu().numpy()
    np.testing.assert_array_equal(s_gpu.view(np.uint32), want_s.view(np.uint32))

    pairs_all = gen["pairs"]
    sub = pairs_all[:500]
    rows = np.concatenate([sub[:, 0], sub[:, 1]])
    apex_sub = np.full(P, -1, np.int64)
    apex_sub[rows] = apex[rows]
    t = time.time()
    model = F64Model(gen, weights, sigma, apex_sub, dia=run2["od"], lib_o=ol_w, prm=prm)
    _log(f"FP64 model of {len(model.rows)} rows built in {time.time() - t:.1f}s")
    s64 = model.s_star(sigma.astype(np.float64), weights.astype(np.float64))
    _close(s64[rows][apex[rows] >= 0], s_gpu[rows][apex[rows] >= 0], "FP64 s* vs device s*")

    diff = s64[sub[:, 0]] - s64[sub[:, 1]]
    grid = np.linspace(np.quantile(diff, 0.3), np.quantile(diff, 0.7), 201)
    m = float(grid[np.argmax([np.abs(diff - x).min() for x in grid])])
    assert np.abs(diff - m).min() > 1e-3
    active = (m - diff) > 0
    assert 0 < active.sum() < len(sub)

    loss = hinge_loss(s_t, torch.tensor(sub[:, 0], device="cuda"),
                      torch.tensor(sub[:, 1], device="cuda"), m)
    loss.backward()
    gw_gpu = w_t.grad.cpu().numpy()
    gs_gpu = sig_t.grad.numpy()

    g = np.zeros(P)
    g[sub[active, 0]] = -1.0
    g[sub[active, 1]] = 1.0
    sig64, w64 = sigma.astype(np.float64), weights.astype(np.float64)
    gw_ref, gs_ref = model.grads(sig64, w64, g)

    def L(sg, ww):
        return _hinge64(model.s_star(sg, ww), sub, m)

    fd_s = np.zeros_like(sig64)
    for o in range(len(sig64)):
        e = np.zeros_like(sig64)
        e[o] = H
        fd_s[o] = (L(sig64 + e, w64) - L(sig64 - e, w64)) / (2 * H)
    _close(gs_ref, fd_s, "FP64 d/dsigma analytic vs central differences")
    nz = np.flatnonzero(gw_ref)
    for j in np.random.default_rng(3).choice(nz, size=min(40, nz.size), replace=False):
        e = np.zeros_like(w64)
        e[j] = H
        fd = (L(sig64, w64 + e) - L(sig64, w64 - e)) / (2 * H)
        assert abs(fd - gw_ref[j]) <= RTOL * abs(gw_ref[j]) + 1e-9, (j, fd, gw_ref[j])
    _close(gs_gpu, gs_ref, "device d/dsigma vs FP64 (400k-row state, 1k-row loss)")
    _close(gw_gpu, gw_ref, "device d/dweights vs FP64 (400k-row state, 1k-row loss)")
    assert np.count_nonzero(gw_gpu[gw_ref == 0]) == 0


def test_workload4_full_size(ads):
    gen = _gen(4)
    _, _, _, got, want, _, _ = _device_and_oracle(ads, gen)
    assert len(want[0]) > 5_000_000
    assert_tables_equal(got, want, "workload 4, 2M precursors")
