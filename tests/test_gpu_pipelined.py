"""Pipelined host tables: search() of libraries of 128k rows or more runs in
chunks of rows (ADSR_PIPE_CHUNKS overrides the count), each chunk
compacted after the ones before it and tabled, and its columns copied to
host memory while the next chunk is selected.  The table must equal the
oracle's bit for bit in library order (src/peak_group_selection/
mod.rs:101-119), whatever the chunk count.  Child processes force the
chunk count on small runs: 3 and 6 chunks of a 6,000-row library, a
library whose long RT windows need a second select launch (the search
then falls back to one compaction), a search without candidates, and a
device status error (every precursor in more than four isolation
windows), which must be raised.  Full-size workloads 1-3 take the
pipelined path by default (test_gpu_fullsize.py)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import os, sys
sys.path[:0] = [ROOT, os.path.join(ROOT, "alphadia-search-rs_amd"), os.path.join(ROOT, "tests")]
import numpy as np
import alphadia_search_rs as ads
import oracle, synthetic
from conftest import assert_tables_equal, oracle_params_for, params_for

def check(spec, what, mutate=None):
    gen = synthetic.generate(spec)
    if mutate:
        gen = mutate(gen)
    rt = gen.get("rt_tolerance_per_precursor")
    extra = {} if rt is None else {"rt_tolerance_per_precursor": rt}
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(spec)
    p.update(extra)
    want = oracle.select(oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"]),
                         oracle_params_for(spec, **extra))
    sel = ads.PeakGroupSelection(p)
    for rep in range(2):  # the second search reuses the pooled workspace
        got = sel.search(d, lib).to_arrays()
        assert_tables_equal(got, want, f"{what}, search {rep}")
    print(what, "ok", len(want[0]), flush=True)

case = sys.argv[1]
if case == "plain":
    check(synthetic.config(1, n_cycles=600, n_windows=6, peaks_per_spectrum=500,
                           n_targets=3000), "6000 rows")
elif case == "long_rows":
    check(synthetic.config(4, n_cycles=1200, n_windows=4, peaks_per_spectrum=600,
                           n_targets=1500), "long rows")
elif case == "empty":
    def no_fragments(gen):
        lib = list(gen["lib"])
        lib[10] = np.zeros_like(lib[10])
        return dict(gen, lib=tuple(lib))
    check(synthetic.config(1, n_cycles=600, n_windows=6, peaks_per_spectrum=500,
                           n_targets=3000), "no candidates", no_fragments)
elif case == "status":
    spec = synthetic.config(1, n_cycles=600, n_windows=6, peaks_per_spectrum=500,
                            n_targets=3000)
    gen = synthetic.generate(spec)
    dia = list(gen["dia"])
    dia[1] = np.full_like(dia[1], 100.0)   # every window spans every precursor
    dia[2] = np.full_like(dia[2], 2000.0)
    d = ads.DIAData.from_arrays(*dia)
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    try:
        ads.PeakGroupSelection(params_for(spec)).search(d, lib)
    except Exception as e:
        assert "isolation windows" in str(e), e
        print("status ok", flush=True)
'''


@pytest.mark.parametrize("case,chunks", [("plain", "3"), ("plain", "8"), ("long_rows", "2"),
                                         ("empty", "3"), ("status", "3")])
def test_pipelined_search_matches_oracle(case, chunks):
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    print(f"[pipelined] {case} ADSR_PIPE_CHUNKS={chunks}", file=sys.stderr, flush=True)
    env = dict(os.environ, ADSR_PIPE_CHUNKS=chunks, ADSR_QUIET="1")
    r = subprocess.run([sys.executable, "-c", "ROOT = %r\n" % ROOT + CHILD, case], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert " ok" in r.stdout, r.stdout
