"""Differentiable candidate selection: gradients of each precursor's rank-0
selection score with respect to a Gaussian sigma per quadrupole window and
the library fragment intensities.

The reference has no backward pass; this module extends
``PeakGroupSelection.search`` (src/peak_group_selection/mod.rs:99-123).  The
forward pass is that search with ``sigma_per_window`` and the weights in
place of ``fragment_intensity``; its output is s* per precursor, the score of
the rank-0 candidate (0 for a precursor without candidates).  The backward
holds each precursor's apex and its selected fragments fixed (both are
piecewise constant in the inputs), recomputes the XIC window around the
apex on the device and differentiates

    s* = sum_f w_f * ln(1 + Y_f),  Y_f = sum_d k_d(sigma) * X_f[apex + d]

with k the normalised Gaussian of src/kernel/mod.rs:32-56.  Both passes run
in libadsr.so (adsr_diff_forward / adsr_diff_backward); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np
import torch

from . import DIAData, SelectionParameters, SpecLibFlat
from . import _native as N
from ._native import check


def _dev_ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


class SelectionGradState:
    """Device state of one (DIAData, SpecLibFlat, SelectionParameters)
    triple for repeated forward/backward passes.

    Requirements: a gfx950 device, and fragment ranges that no two
    precursors share (per-fragment gradients are scattered without atomics).
    Rows of s* follow ascending ``precursor_idx`` (stable), the order of the
    candidate table.
    """

    def __init__(self, dia: DIAData, lib: SpecLibFlat,
                 params: SelectionParameters):
        if not N.device_available():
            raise RuntimeError("differentiable selection needs a gfx950 device")
        if not dia._h or not lib._h:
            raise ValueError("empty DIAData or SpecLibFlat")
        self._h = None
        self.num_observations = dia.num_observations
        self.num_precursors = lib.num_precursors
        self.num_fragments = lib.num_fragments
        self._dia, self._lib = dia, lib  # keep the handles alive
        c = params._c()
        out = C.c_void_p()
        check(N.lib.adsr_diff_create(dia._h, lib._h, C.byref(c), C.byref(out)))
        self._h = out.value
        self._fwd_done = False
        # forwards run so far: a backward belongs to the forward whose apexes,
        # fragment selections, weights and taps the state holds
        self.forward_count = 0

    def __del__(self, _free=N.lib.adsr_diff_free):
        # the free function is bound at definition: module globals may be
        # gone when the interpreter shuts down
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    def _check_inputs(self, sigma: torch.Tensor, weights: torch.Tensor):
        if sigma.dtype != torch.float32 or sigma.numel() != self.num_observations:
            raise ValueError(
                f"sigma must be float32 with {self.num_observations} values")
        if (weights.dtype != torch.float32 or not weights.is_cuda
                or weights.numel() != self.num_fragments):
            raise ValueError(
                f"weights must be a float32 device tensor with "
                f"{self.num_fragments} values")

    def forward(self, sigma: torch.Tensor, weights: torch.Tensor,
                with_has: bool = True, asynchronous: bool = False):
        """Return (s_star f32[P], has_candidate bool[P] or None when
        ``with_has`` is false) on the device.

        ``asynchronous``: the forward is queued on torch's current stream
        and the call returns without waiting for it; device status errors
        are then raised by the next backward (or forward, or timings)."""
        self._check_inputs(sigma, weights)
        w = weights.detach().contiguous()
        sig = np.ascontiguousarray(
            sigma.detach().to("cpu", torch.float32).numpy().reshape(-1))
        dev = w.device
        s_star = torch.empty(self.num_precursors, dtype=torch.float32, device=dev)
        has = (torch.empty(self.num_precursors, dtype=torch.uint8, device=dev)
               if with_has else None)
        has_p = _dev_ptr(has) if has is not None else None
        if asynchronous:
            stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
            check(N.lib.adsr_diff_forward_async(self._h, N.ptr(sig), _dev_ptr(w),
                                                _dev_ptr(s_star), has_p, stream))
        else:
            torch.cuda.synchronize(dev)  # the library runs on its own stream
            check(N.lib.adsr_diff_forward(self._h, N.ptr(sig), _dev_ptr(w),
                                          _dev_ptr(s_star), has_p))
        self._fwd_done = True
        self.forward_count += 1
        return s_star, (has.bool() if has is not None else None)

    def backward(self, grad_s: torch.Tensor, forward_id: int | None = None):
        """Return (grad_weights f32[n_fragments] on the device,
        grad_sigma f64[n_observations] on the host) for dLoss/ds* = grad_s,
        at the last forward.  ``forward_id``: the ``forward_count`` right
        after the forward the gradient is for; another forward since then
        raises RuntimeError (the state holds only the last forward)."""
        if not self._fwd_done:
            raise RuntimeError("backward before forward")
        if forward_id is not None and forward_id != self.forward_count:
            raise RuntimeError(
                f"backward of forward {forward_id}, but forward {self.forward_count} "
                "ran on this SelectionGradState since: its apexes and fragment "
                "selections replaced those of the earlier forward; use one state "
                "per forward that is differentiated")
        g = grad_s.detach().to(torch.float32).contiguous()
        if not g.is_cuda or g.numel() != self.num_precursors:
            raise ValueError("grad_s must be a device tensor with one value "
                             "per precursor")
        gw = torch.empty(self.num_fragments, dtype=torch.float32, device=g.device)
        gs = np.zeros(self.num_observations, dtype=np.float64)
        torch.cuda.synchronize(g.device)
        check(N.lib.adsr_diff_backward(self._h, _dev_ptr(g), _dev_ptr(gw),
                                       N.ptr(gs)))
        return gw, torch.from_numpy(gs)

    def settle(self):
        """Wait for the last forward and raise its device status error, if
        any (an asynchronous forward reports errors here or in the next
        call that waits for it)."""
        self.timings()

    def timings(self):
        """Device milliseconds of the last forward and backward pass."""
        f, b = C.c_double(), C.c_double()
        check(N.lib.adsr_diff_timings(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value


class RankZeroScore(torch.autograd.Function):
    """s* = RankZeroScore.apply(state, sigma, weights).

    ``sigma``: float32[n_observations] (host or device); ``weights``:
    float32[n_fragments] on the device.  Gradients flow to both; the sigma
    gradient is accumulated in float64 and returned in sigma's dtype."""

    @staticmethod
    def forward(ctx, state: SelectionGradState, sigma, weights):
        # queued without waiting only when a backward will follow: its
        # device status is then read by that backward; a forward nothing
        # differentiates waits and raises at once
        grad = bool(ctx.needs_input_grad[1] or ctx.needs_input_grad[2])
        s_star, _ = state.forward(sigma, weights, with_has=False, asynchronous=grad)
        ctx.state = state
        ctx.forward_id = state.forward_count
        ctx.sigma_meta = (sigma.dtype, sigma.device)
        return s_star

    @staticmethod
    def backward(ctx, grad_s):
        gw, gs = ctx.state.backward(grad_s, forward_id=ctx.forward_id)
        dt, dev = ctx.sigma_meta
        return None, gs.to(device=dev, dtype=dt), gw


def rank_zero_scores(state: SelectionGradState, sigma, weights):
    """s* per precursor (rows in candidate-table order), differentiable.

    When sigma or the weights require a gradient, the forward is queued on
    torch's current stream without waiting: a device status error (for
    example a precursor in more than four isolation windows) is raised by
    the backward, by ``state.settle()`` or by the next call on the state
    that waits.  Otherwise the forward waits and raises at once.  One
    backward per forward: a second forward on the same state before the
    first one's backward makes that backward raise."""
    return RankZeroScore.apply(state, sigma, weights)


# id(targets) -> (weakrefs to the targets and decoys tensors, versions, row
# count) of a checked pair list; only tensors used as passed in are kept, and
# an entry is dropped with its targets tensor, so a new tensor in a reused
# block is checked again
_checked_pairs: dict = {}


def _forget(key):
    def drop(_ref):
        _checked_pairs.pop(key, None)
    return drop


def _pair_indices(t, device) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t)
    return t.to(device=device, dtype=torch.int64).contiguous()


def _check_pairs(t: torch.Tensor, d: torch.Tensor, n_rows: int,
                 cacheable: bool) -> None:
    """Raise on an index outside [0, n_rows).  A pair of index tensors the
    caller passed as they are (no conversion) is checked once while both
    tensor objects live and their versions stay the same; converted
    temporaries are checked every time (the kernel itself skips pairs out
    of range, so it never reads out of bounds)."""
    if t.shape != d.shape or t.dim() != 1:
        raise ValueError("targets and decoys must be 1-D and of equal length")
    if cacheable:
        hit = _checked_pairs.get(id(t))
        if (hit is not None and hit[0]() is t and hit[1]() is d
                and hit[2:] == (t._version, d._version, n_rows)):
            return
    if t.numel():
        lo = torch.minimum(t.min(), d.min())
        hi = torch.maximum(t.max(), d.max())
        if bool(lo < 0) or bool(hi >= n_rows):
            raise ValueError(f"pair indices must lie in [0, {n_rows})")
    if cacheable:
        _checked_pairs[id(t)] = (weakref.ref(t, _forget(id(t))), weakref.ref(d),
                                 t._version, d._version, n_rows)


class HingeLoss(torch.autograd.Function):
    """sum over pairs of max(0, margin - (s*_target - s*_decoy)) in one
    device kernel (adsr_hinge_loss) on the current stream: the per-pair
    terms and the gradient, -1 to the target and +1 to the decoy of every
    pair with margin - (s*_t - s*_d) >= 0 (the subgradient torch.clamp
    uses).  The loss is torch's sum of the terms, bit-identical to
    ``torch.clamp(margin - (s[t] - s[d]), min=0).sum()``."""

    @staticmethod
    def forward(ctx, s_star, targets, decoys, margin):
        if s_star.dtype != torch.float32 or not s_star.is_cuda or s_star.dim() != 1:
            raise ValueError("s_star must be a 1-D float32 device tensor")
        s = s_star.detach().contiguous()
        n = s.numel()
        t = _pair_indices(targets, s.device)
        d = _pair_indices(decoys, s.device)
        _check_pairs(t, d, n, cacheable=t is targets and d is decoys)
        terms = torch.empty(t.numel(), dtype=torch.float32, device=s.device)
        g = torch.zeros(n, dtype=torch.float32, device=s.device)
        stream = C.c_void_p(torch.cuda.current_stream(s.device).cuda_stream)
        check(N.lib.adsr_hinge_loss(_dev_ptr(s), n, _dev_ptr(t), _dev_ptr(d),
                                    t.numel(), float(margin), _dev_ptr(terms),
                                    _dev_ptr(g), stream))
        ctx.save_for_backward(g)
        return terms.sum()

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return g * grad_out, None, None, None


def hinge_loss(s_star: torch.Tensor, targets: torch.Tensor,
               decoys: torch.Tensor, margin: float = 1.0) -> torch.Tensor:
    """sum over pairs of max(0, margin - (s*_target - s*_decoy)),
    differentiable with respect to s* (see HingeLoss)."""
    return HingeLoss.apply(s_star, targets, decoys, margin)
