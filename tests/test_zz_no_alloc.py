"""adsr_select_rows_noalloc (the search without device allocation) through
the C ABI, against adsr_select_device + adsr_device_candidates_export (the
existing entry point whose kernels it shares) and the oracle.

Two calls in a row: the whole library, then a smaller row range, whose
candidates must be the matching block of the whole table (library order,
src/peak_group_selection/mod.rs:101-119).  A column buffer that is too small
and a row range beyond the scratch array are rejected with an error.
"""
import ctypes

import numpy as np
import pytest

import oracle
import synthetic
from conftest import assert_tables_equal, oracle_params_for, params_for

pytestmark = pytest.mark.gpu


def _cols(torch, cap):
    return [torch.empty(cap, dtype=torch.float32 if i == 2 else torch.int64, device="cuda")
            for i in range(9)]


def _host(cols, n):
    out = []
    for i, c in enumerate(cols):
        a = c[:n].cpu().numpy()
        out.append(a if i == 2 else a.view(np.uint64))
    return tuple(out)


def _noalloc(N, d, lib, c, r0, r1, cols, cap, stream):
    n_out = ctypes.c_size_t(0)
    st = N.lib.adsr_select_rows_noalloc(
        d._h, lib._h, ctypes.byref(c), r0, r1, *[t.data_ptr() for t in cols], cap,
        ctypes.byref(n_out), stream)
    return st, n_out.value


def test_noalloc_search_matches_device_search_twice():
    import torch
    import alphadia_search_rs as ads
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    spec = synthetic.config(2, n_cycles=600, n_windows=12, peaks_per_spectrum=700,
                            n_targets=1500)
    gen = synthetic.generate(spec)
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    extra = {"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]}
    p = params_for(spec, **extra)
    n = lib.num_precursors
    # the existing entry point: the table in HBM, exported on torch's stream
    dc = ads.PeakGroupSelection(p).search_device(d, lib)
    full = _host(dc.to_torch(), len(dc))
    want = oracle.select(oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"]),
                         oracle_params_for(spec, **extra))
    assert_tables_equal(full, want, "device search")
    c = p._c()
    stream = torch.cuda.current_stream().cuda_stream
    cap = n * spec.candidate_count
    cols = _cols(torch, cap)
    # call 1: every row
    st, k = _noalloc(N, d, lib, c, 0, n, cols, cap, stream)
    N.check(st)
    assert k == len(full[0]) > 0
    assert_tables_equal(_host(cols, k), full, "no-allocation search, all rows")
    # call 2: a smaller row range into the same buffers
    r0, r1 = 217, 1904
    for t in cols:
        t.fill_(-1)
    st, k2 = _noalloc(N, d, lib, c, r0, r1, cols, cap, stream)
    N.check(st)
    m = (full[0] >= r0) & (full[0] < r1)  # precursor_idx == library row here
    part = tuple(col[m] for col in full)
    assert k2 == len(part[0]) > 0 and k2 < k
    assert_tables_equal(_host(cols, k2), part, "no-allocation search, rows 217-1904")
    # too small a column buffer: an error, the count still reported
    st, k3 = _noalloc(N, d, lib, c, 0, n, cols, 10, stream)
    assert st != 0 and k3 == k
    assert "columns hold 10" in N.lib.adsr_last_error().decode()
    # more rows than the scratch array holds: an error naming the limit
    big = synthetic.config(0, n_cycles=200, n_windows=4, peaks_per_spectrum=50,
                           n_targets=400_000, frag_min=24, frag_max=24,
                           planted_fraction=0.0)
    bg = synthetic.generate(big)
    d2 = ads.DIAData.from_arrays(*bg["dia"])
    lib2 = ads.SpecLibFlat.from_arrays(*bg["lib"])
    c2 = params_for(big, top_k_fragments=24, candidate_count=10)._c()
    st, _ = _noalloc(N, d2, lib2, c2, 0, lib2.num_precursors, cols, cap, stream)
    assert st != 0 and "rows per call" in N.lib.adsr_last_error().decode()
