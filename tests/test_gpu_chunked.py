"""Chunked searches: a library searched as device-resident chunks of rows
(each chunk planned and selected in turn, one long-row launch and one
compaction over all rows).  Libraries of
more than 1M rows are chunked by default (workload 4, covered at full size
by test_gpu_fullsize.py); here ADSR_CHUNK_ROWS=700 forces 1,000 rows into
two chunks in a child process, on workload 4's RT windows (rows too long for
the first launch in both chunks).  The table of search() and of the timed
search's last hipGraph replay must equal the oracle's bit for bit
(src/peak_group_selection/mod.rs:101-119: library order)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import ctypes, os, sys
sys.path[:0] = [ROOT, os.path.join(ROOT, "alphadia-search-rs_amd"), os.path.join(ROOT, "tests")]
import numpy as np
import alphadia_search_rs as ads
import alphadia_search_rs._native as N
import oracle, synthetic
from conftest import assert_tables_equal, oracle_params_for, params_for
spec = synthetic.config(4, n_cycles=1200, n_windows=4, peaks_per_spectrum=600, n_targets=500)
gen = synthetic.generate(spec)
rt = gen["rt_tolerance_per_precursor"]
d = ads.DIAData.from_arrays(*gen["dia"])
lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
p = params_for(spec)
p.update({"rt_tolerance_per_precursor": rt})
want = oracle.select(oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"]),
                     oracle_params_for(spec, rt_tolerance_per_precursor=rt))
assert len(want[0]) > 0
assert_tables_equal(ads.PeakGroupSelection(p).search(d, lib).to_arrays(), want, "chunked search")
c = p._c()
tpb, ppw = ctypes.c_int(0), ctypes.c_int(0)
a, b, g = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
out = ctypes.c_void_p()
N.check(N.lib.adsr_select_timed(d._h, lib._h, ctypes.byref(c), 1, 3, ctypes.byref(tpb),
                                ctypes.byref(ppw), ctypes.byref(a), ctypes.byref(b),
                                ctypes.byref(g), ctypes.byref(out)))
assert g.value > 0.0
got = ads.CandidateCollection._wrap(out.value).to_arrays()
assert_tables_equal(got, want, "chunked search, graph replay")
L = (lib.num_precursors, int((want[1] == 0).sum()))
print("chunked ok", L, flush=True)
'''


def test_chunked_search_matches_oracle():
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    env = dict(os.environ, ADSR_CHUNK_ROWS="700", ADSR_QUIET="1")
    r = subprocess.run([sys.executable, "-c", "ROOT = %r\n" % ROOT + CHILD], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "chunked ok" in r.stdout
