"""Candidate table kept in HBM (adsr_select_device) vs the oracle and the
host table.

search_device().to_torch() must hold the same nine columns, row for row and
bit for bit, as the oracle's selection and as search().to_arrays()
(src/candidate/entry.rs:124-173), on torch's current stream and on a side
stream, and for an empty result.
"""
import numpy as np
import pytest

import oracle
from conftest import assert_tables_equal, oracle_params_for, params_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ads():
    import alphadia_search_rs as m
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    return m


def _host_view(cols):
    out = []
    for i, c in enumerate(cols):
        a = c.cpu().numpy()
        out.append(a if i == 2 else a.view(np.uint64))
    return tuple(out)


def test_device_table_equals_host_table(ads, run0):
    import torch
    d = ads.DIAData.from_arrays(*run0["dia"])
    lib = ads.SpecLibFlat.from_arrays(*run0["lib"])
    sel = ads.PeakGroupSelection(params_for(run0["spec"]))
    want = oracle.select(oracle.Dia(run0["dia"]), oracle.Lib(run0["lib"]),
                         oracle_params_for(run0["spec"]))
    assert_tables_equal(sel.search(d, lib).to_arrays(), want, "host table")
    dc = sel.search_device(d, lib)
    assert len(dc) == len(want[0]) > 0
    cols = dc.to_torch()
    assert [c.dtype for c in cols] == [torch.int64] * 2 + [torch.float32] + [torch.int64] * 6
    assert all(c.is_cuda for c in cols)
    torch.cuda.synchronize()
    assert_tables_equal(_host_view(cols), want, "device table")
    assert_tables_equal(dc.to_collection().to_arrays(), want, "device table to host")
    # DLPack capsules of the same columns, read back through torch
    from torch.utils.dlpack import from_dlpack
    caps = dc.to_dlpack()
    back = [from_dlpack(cp) for cp in caps]
    torch.cuda.synchronize()
    assert all(b.is_cuda for b in back)
    assert_tables_equal(_host_view(back), want, "device table through DLPack")


def test_device_table_on_side_stream(ads, run0):
    import torch
    d = ads.DIAData.from_arrays(*run0["dia"])
    lib = ads.SpecLibFlat.from_arrays(*run0["lib"])
    sel = ads.PeakGroupSelection(params_for(run0["spec"], candidate_count=5))
    want = oracle.select(oracle.Dia(run0["dia"]), oracle.Lib(run0["lib"]),
                         oracle_params_for(run0["spec"], candidate_count=5))
    assert_tables_equal(sel.search(d, lib).to_arrays(), want, "host table, top-5")
    dc = sel.search_device(d, lib)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        cols = dc.to_torch(stream=s)
        total = torch.stack([c.sum() for i, c in enumerate(cols) if i != 2]).sum()
    s.synchronize()
    assert_tables_equal(_host_view(cols), want, "device table, side stream")
    assert int(total.item()) == int(sum(int(w.astype(np.int64).sum())
                                        for i, w in enumerate(want) if i != 2))


def test_device_table_empty(ads, run0):
    import torch
    lib_arrays = list(run0["lib"])
    lib_arrays[2] = np.full_like(lib_arrays[2], 5000.0)  # no isolation window
    d = ads.DIAData.from_arrays(*run0["dia"])
    lib = ads.SpecLibFlat.from_arrays(*lib_arrays)
    dc = ads.PeakGroupSelection(params_for(run0["spec"])).search_device(d, lib)
    assert len(dc) == 0 and dc.is_empty()
    cols = dc.to_torch()
    assert len(cols) == 9 and all(c.numel() == 0 for c in cols)
    assert dc.to_collection().len() == 0
    torch.cuda.synchronize()
