"""Differentiable selection (adsr_diff_* through alphadia_search_rs.autograd)
against an FP64 restatement of the rank-0 score (oracle/grad64.py).

Model, per precursor p with rank-0 apex a (fixed by the forward pass),
selected fragments f, XIC X (exact f32 sums from the oracle) and the kernel
of its first valid quadrupole window o(p):

    s*_p = sum_f w_f ln(1 + Y_f),  Y_f = sum_j k_j(sigma_o) X_f[a - half + j]

with k_j = exp(-x_j^2 / 2 sigma^2) / sum_m exp(-x_m^2 / 2 sigma^2) over
x = -half .. half-1 (src/kernel/mod.rs:32-56; the normalisation constant
cancels) and the symmetric tap use of src/convolution/scalar.rs:5-56
(taps 0..half on both sides).  The loss is the hinge over target/decoy
pairs, sum max(0, m - (s*_t - s*_d)).

Checks (tolerance 1e-4 relative, elementwise, with an absolute floor of
1e-4 * 1e-6 * max|reference| for entries that are zero or nearly so):
  * the device forward equals the oracle's rank-0 score bit for bit, at
    sigma = 2.0 equals plain search() bit for bit, and the FP64 model's s*
    agrees with it to 1e-5 (pins apex, XIC, window and fragment mapping);
  * the FP64 analytic gradient agrees with FP64 central finite differences
    (h = 1e-4) of the FP64 model, for every sigma and a sample of weights;
  * the device gradients agree with the FP64 gradient.
"""
import numpy as np
import pytest

import oracle
import synthetic
from oracle.grad64 import F64Model, hinge64 as _hinge64, oracle_state as _oracle_state

pytestmark = pytest.mark.gpu

SPEC = synthetic.config(3, n_cycles=1200, n_windows=20, peaks_per_spectrum=500,
                        n_targets=500)
H = 1e-4
RTOL = 1e-4


@pytest.fixture(scope="module")
def env():
    import torch
    import alphadia_search_rs as ads
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    gen = synthetic.generate(SPEC)
    lib = gen["lib"]
    assert np.all(np.diff(lib[0].astype(np.int64)) > 0)  # rows = input order
    rng = np.random.default_rng(7)
    W = SPEC.n_windows
    sigma = np.linspace(1.6, 2.8, W).astype(np.float32)
    weights = (lib[10] * rng.uniform(0.5, 1.5, lib[10].shape)).astype(np.float32)
    return dict(torch=torch, ads=ads, gen=gen, sigma=sigma, weights=weights)


def _params(gen, **extra):
    from conftest import params_for
    p = params_for(gen["spec"], **extra)
    p.update({"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]})
    return p


def _rank0(table, n):
    pid, rank, score, cc = table[0], table[1], table[2], table[6]
    s = np.zeros(n, np.float32)
    c = np.full(n, -1, np.int64)
    m = rank == 0
    s[pid[m].astype(np.int64)] = score[m]
    c[pid[m].astype(np.int64)] = cc[m].astype(np.int64)
    return s, c


def _close(got, ref, what):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    floor = RTOL * 1e-6 * max(np.abs(ref).max(), 1e-30)
    err = np.abs(got - ref)
    bad = np.flatnonzero(err > RTOL * np.abs(ref) + floor)
    rel = (err / np.maximum(np.abs(ref), floor)).max()
    print(f"{what}: max rel err {rel:.3e} over {len(ref)} values")
    assert bad.size == 0, (f"{what}: {bad.size} values off, first {bad[:5]}: "
                           f"got {got[bad[:5]]} ref {ref[bad[:5]]}")


def test_diff_forward_at_init_equals_search(env):
    """sigma = 2.0 everywhere and the library intensities as weights: s* is
    the rank-0 score of plain search() (fwhm 4.7096 gives sigma 2.0f)."""
    torch, ads, gen = env["torch"], env["ads"], env["gen"]
    from alphadia_search_rs.autograd import SelectionGradState
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = _params(gen)
    table = ads.PeakGroupSelection(p).search(d, lib).to_arrays()
    want, _ = _rank0(table, lib.num_precursors)
    st = SelectionGradState(d, lib, p)
    w = torch.tensor(gen["lib"][10], device="cuda")
    s, has = st.forward(torch.full((d.num_observations,), 2.0), w)
    s = s.cpu().numpy()
    assert has.cpu().numpy().sum() > 0.5 * lib.num_precursors
    np.testing.assert_array_equal(s.view(np.uint32), want.view(np.uint32))


def test_diff_gradients_match_fp64(env):
    torch, ads, gen = env["torch"], env["ads"], env["gen"]
    from alphadia_search_rs.autograd import (SelectionGradState, hinge_loss,
                                             rank_zero_scores)
    sigma, weights = env["sigma"], env["weights"]
    lib_arrays = list(gen["lib"])
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*lib_arrays)
    st = SelectionGradState(d, lib, _params(gen))
    P = lib.num_precursors

    # forward: bitwise against the oracle with the same sigma and weights
    od, ol, prm = _oracle_state(gen, weights, sigma)
    want_s, apex = _rank0(oracle.select(od, ol, prm), P)
    sig_t = torch.tensor(sigma, requires_grad=True)
    w_t = torch.tensor(weights, device="cuda", requires_grad=True)
    s_t = rank_zero_scores(st, sig_t, w_t)
    s_gpu = s_t.detach().cpu().numpy()
    np.testing.assert_array_equal(s_gpu.view(np.uint32), want_s.view(np.uint32))

    model = F64Model(gen, weights, sigma, apex)
    s64 = model.s_star(sigma.astype(np.float64), weights.astype(np.float64))
    _close(s64, s_gpu, "FP64 model s* vs device s*")

    # hinge margin away from every kink
    pairs = gen["pairs"]
    diff = s64[pairs[:, 0]] - s64[pairs[:, 1]]
    grid = np.linspace(np.quantile(diff, 0.3), np.quantile(diff, 0.7), 201)
    m = float(grid[np.argmax([np.abs(diff - x).min() for x in grid])])
    assert np.abs(diff - m).min() > 1e-3
    active = (m - diff) > 0
    assert 0 < active.sum() < len(pairs)

    loss = hinge_loss(s_t, torch.tensor(pairs[:, 0], device="cuda"),
                      torch.tensor(pairs[:, 1], device="cuda"), m)
    loss.backward()
    gw_gpu = w_t.grad.cpu().numpy()
    gs_gpu = sig_t.grad.numpy()

    g = np.zeros(P)
    g[pairs[active, 0]] = -1.0
    g[pairs[active, 1]] = 1.0
    sig64, w64 = sigma.astype(np.float64), weights.astype(np.float64)
    gw_ref, gs_ref = model.grads(sig64, w64, g)

    # FP64 finite differences of the FP64 model pin the analytic gradient
    def L(sg, ww):
        return _hinge64(model.s_star(sg, ww), pairs, m)

    fd_s = np.zeros_like(sig64)
    for o in range(len(sig64)):
        e = np.zeros_like(sig64)
        e[o] = H
        fd_s[o] = (L(sig64 + e, w64) - L(sig64 - e, w64)) / (2 * H)
    _close(gs_ref, fd_s, "FP64 d/dsigma analytic vs central differences")
    rng = np.random.default_rng(3)
    nz = np.flatnonzero(gw_ref)
    for j in rng.choice(nz, size=min(60, nz.size), replace=False):
        e = np.zeros_like(w64)
        e[j] = H
        fd = (L(sig64, w64 + e) - L(sig64, w64 - e)) / (2 * H)
        assert abs(fd - gw_ref[j]) <= RTOL * abs(gw_ref[j]) + 1e-9, (j, fd, gw_ref[j])

    # device gradients against FP64
    _close(gs_gpu, gs_ref, "device d/dsigma vs FP64")
    _close(gw_gpu, gw_ref, "device d/dweights vs FP64")
    assert np.count_nonzero(gw_gpu[gw_ref == 0]) == 0

    # deterministic: a second backward gives identical bits
    gw2, gs2 = st.backward(torch.tensor(g, dtype=torch.float32, device="cuda"))
    np.testing.assert_array_equal(gw2.cpu().numpy().view(np.uint32),
                                  gw_gpu.view(np.uint32))
    np.testing.assert_array_equal(gs2.numpy().astype(np.float32), gs_gpu)


def test_diff_backward_before_forward_raises(env):
    ads, gen, torch = env["ads"], env["gen"], env["torch"]
    from alphadia_search_rs.autograd import SelectionGradState
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    st = SelectionGradState(d, lib, _params(gen))
    with pytest.raises(RuntimeError):
        st.backward(torch.zeros(lib.num_precursors, device="cuda"))
    with pytest.raises(ValueError):
        st.forward(torch.full((3,), 2.0), torch.zeros(lib.num_fragments, device="cuda"))


def test_hinge_loss_kernel_matches_torch_ops(env):
    """adsr_hinge_loss (HingeLoss) against the same loss written as torch
    operations: index_select, clamp(min=0), sum.  The loss is bit-identical
    (the same per-pair terms, summed by the same torch reduction) and so is
    the gradient, including pairs exactly at the kink (margin - diff == 0,
    where torch.clamp passes the gradient) and precursors in several pairs."""
    torch = env["torch"]
    from alphadia_search_rs.autograd import hinge_loss
    rng = np.random.default_rng(11)
    n, npairs, margin = 5000, 3000, 1.0
    s = rng.normal(0.0, 1.0, n).astype(np.float32)
    t = rng.integers(0, n, npairs)
    d = rng.integers(0, n, npairs)
    s[t[:5]] = np.float32(1.5)  # margin - (1.5 - 0.5) == 0 exactly
    s[d[:5]] = np.float32(0.5)
    d[5:9] = d[0]               # repeated decoys
    s_dev = torch.tensor(s, device="cuda")
    t_dev = torch.tensor(t, device="cuda")
    d_dev = torch.tensor(d, device="cuda")

    a = s_dev.clone().requires_grad_(True)
    got = hinge_loss(a, t_dev, d_dev, margin)
    got.backward()
    b = s_dev.clone().requires_grad_(True)
    want = torch.clamp(margin - (b.index_select(0, t_dev) - b.index_select(0, d_dev)),
                       min=0.0).sum()
    want.backward()
    m = np.float32(margin) - (s[t] - s[d])
    assert (m == 0).sum() >= 5 and (m > 0).sum() > 100 and (m < 0).sum() > 100
    assert got.detach().cpu().numpy().view(np.uint32) == want.detach().cpu().numpy().view(np.uint32)
    ga, gb = a.grad.cpu().numpy(), b.grad.cpu().numpy()
    np.testing.assert_array_equal(ga.view(np.uint32), gb.view(np.uint32))

    # a second call with a scaled output gradient, and an empty pair list
    a.grad = None
    (2.5 * hinge_loss(a, t_dev, d_dev, margin)).backward()
    np.testing.assert_array_equal(a.grad.cpu().numpy(), 2.5 * gb)
    e = torch.empty(0, dtype=torch.int64, device="cuda")
    assert float(hinge_loss(s_dev, e, e, margin)) == 0.0

    with pytest.raises(ValueError):
        hinge_loss(s_dev, t_dev, torch.full_like(d_dev, n), margin)


def test_async_forward_matches_sync_and_defers_status(env):
    """adsr_diff_forward_async (the forward rank_zero_scores uses) queued on
    torch's stream gives the synchronous forward's s* and has-candidate
    bits; a device status error of an asynchronous forward (a precursor
    inside more than four isolation windows) is raised by the next call
    that waits for it, while the synchronous forward raises it at once."""
    torch, ads, gen = env["torch"], env["ads"], env["gen"]
    import alphadia_search_rs._native as N
    from alphadia_search_rs.autograd import SelectionGradState
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    st = SelectionGradState(d, lib, _params(gen))
    sig = torch.tensor(env["sigma"])
    w = torch.tensor(env["weights"], device="cuda")
    s_sync, h_sync = st.forward(sig, w)
    s_async, h_async = st.forward(sig, w, asynchronous=True)
    torch.cuda.synchronize()
    a, b = s_sync.cpu().numpy(), s_async.cpu().numpy()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert torch.equal(h_sync, h_async)
    assert st.timings()[0] > 0.0

    dia = list(gen["dia"])
    dia[1] = np.full_like(dia[1], 400.0)   # every window covers 400-1000
    dia[2] = np.full_like(dia[2], 1000.0)
    d2 = ads.DIAData.from_arrays(*dia)
    st2 = SelectionGradState(d2, lib, _params(gen))
    with pytest.raises(N.AdsrError, match="isolation windows"):
        st2.forward(sig, w)
    st2.forward(sig, w, asynchronous=True)  # queued; the status is read later
    with pytest.raises(N.AdsrError, match="isolation windows"):
        st2.backward(torch.zeros(lib.num_precursors, device="cuda"))


def test_backward_of_an_earlier_forward_raises(env):
    """The state holds the last forward's apexes and fragment selections: a
    backward through an earlier forward on the same state raises instead of
    returning the later forward's gradient; the last one's still works."""
    torch, ads, gen = env["torch"], env["ads"], env["gen"]
    from alphadia_search_rs.autograd import SelectionGradState, rank_zero_scores
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    st = SelectionGradState(d, lib, _params(gen))
    sig = torch.tensor(env["sigma"], requires_grad=True)
    w = torch.tensor(env["weights"], device="cuda", requires_grad=True)
    s1 = rank_zero_scores(st, sig, w)
    s2 = rank_zero_scores(st, sig * 1.01, w)
    with pytest.raises(RuntimeError, match="ran on this SelectionGradState since"):
        s1.sum().backward()
    s2.sum().backward()
    assert w.grad is not None and int((w.grad != 0).sum()) > 0


def test_forward_without_grad_raises_at_once(env):
    """A forward nothing differentiates waits for the device: its status
    error (a precursor in more than four isolation windows) is raised by
    rank_zero_scores itself, and settle() raises a deferred one."""
    torch, ads, gen = env["torch"], env["ads"], env["gen"]
    import alphadia_search_rs._native as N
    from alphadia_search_rs.autograd import SelectionGradState, rank_zero_scores
    dia = list(gen["dia"])
    dia[1] = np.full_like(dia[1], 400.0)
    dia[2] = np.full_like(dia[2], 1000.0)
    d2 = ads.DIAData.from_arrays(*dia)
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    st = SelectionGradState(d2, lib, _params(gen))
    sig = torch.tensor(env["sigma"])
    w = torch.tensor(env["weights"], device="cuda")
    with torch.no_grad():
        with pytest.raises(N.AdsrError, match="isolation windows"):
            rank_zero_scores(st, sig, w)
    st.forward(sig, w, asynchronous=True)
    with pytest.raises(N.AdsrError, match="isolation windows"):
        st.settle()


def test_backward_from_saved_windows_equals_reextraction(env):
    """The forward leaves the XIC window around each block-sparse row's
    apex for the backward, which re-extracts only the other rows' windows.
    A state created with ADSR_DIFF_REEXTRACT=1 re-extracts every window; the
    two give the same s* and gradients bit for bit (the windows hold the
    same f32 sums, added in the same entry order)."""
    import os
    torch, ads, gen = env["torch"], env["ads"], env["gen"]
    from alphadia_search_rs.autograd import SelectionGradState
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    P = lib.num_precursors
    rng = np.random.default_rng(11)
    g = torch.tensor(rng.choice([-1.0, 0.0, 0.5, 1.0], size=P).astype(np.float32),
                     device="cuda")
    sig = torch.tensor(env["sigma"])
    w = torch.tensor(env["weights"], device="cuda")
    out = []
    for reextract in (False, True):
        if reextract:
            os.environ["ADSR_DIFF_REEXTRACT"] = "1"
        try:
            st = SelectionGradState(d, lib, _params(gen))
        finally:
            os.environ.pop("ADSR_DIFF_REEXTRACT", None)
        s, has = st.forward(sig, w)
        gw, gs = st.backward(g)
        out.append((s.cpu().numpy(), has.cpu().numpy(), gw.cpu().numpy(),
                    np.asarray(gs, np.float64)))
        del st
    (s0, h0, gw0, gs0), (s1, h1, gw1, gs1) = out
    assert h0.sum() > P // 2
    np.testing.assert_array_equal(s0.view(np.uint32), s1.view(np.uint32))
    np.testing.assert_array_equal(h0, h1)
    assert np.count_nonzero(gw0) > 0
    np.testing.assert_array_equal(gw0.view(np.uint32), gw1.view(np.uint32))
    np.testing.assert_array_equal(gs0.view(np.uint64), gs1.view(np.uint64))
