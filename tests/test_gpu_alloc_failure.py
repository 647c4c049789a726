"""Device allocation failures, forced from the host.

ADSR_FAIL_ALLOC=n makes the n-th allocation of a process fail as out of
memory (no memory is exhausted and no kernel sees a missing buffer: the
call that allocates returns before launching).  Each case runs in a child
process: it builds a DIAData on the device, a library, and searches; the
step whose allocation fails must raise MemoryError naming the buffer, and
the same sequence run again in that process must give the oracle's table.
The allocation numbers cover the device builder, the library upload, the
search workspace, the device table and the host table.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import synthetic
from conftest import oracle_params_for

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys, numpy as np
import alphadia_search_rs as ads, synthetic
from conftest import params_for
gen = synthetic.generate(synthetic.config(0, n_cycles=200, n_windows=3, n_targets=60))
want = np.load(sys.argv[1])
def run():
    d = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    return ads.PeakGroupSelection(params_for(gen["spec"])).search(d, lib).to_arrays()
failed = ""
try:
    got = run()
except MemoryError as e:
    failed = str(e)
    got = run()
for i, (a, b) in enumerate(zip(got, (want[f"c{k}"] for k in range(9)))):
    assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), i
print("FAILED:" + failed if failed else "NO-FAILURE", flush=True)
"""


@pytest.fixture(scope="module")
def expected(tmp_path_factory):
    gen = synthetic.generate(synthetic.config(0, n_cycles=200, n_windows=3, n_targets=60))
    want = oracle.select(oracle.Dia(gen["dia"]), oracle.Lib(gen["lib"]),
                         oracle_params_for(gen["spec"]))
    assert len(want[0]) > 50
    path = str(tmp_path_factory.mktemp("alloc") / "want.npz")
    np.savez(path, **{f"c{k}": c for k, c in enumerate(want)})
    return path


def test_forced_device_allocation_failures_recover(expected):
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join(
        [os.path.join(ROOT, "alphadia-search-rs_amd"), ROOT, os.path.join(ROOT, "tests"),
         env.get("PYTHONPATH", "")])
    seen = []
    for n in (1, 3, 8, 13, 20, 26, 33, 40, 45):
        print(f"[alloc] ADSR_FAIL_ALLOC={n}", file=sys.stderr, flush=True)
        env["ADSR_FAIL_ALLOC"] = str(n)
        r = subprocess.run([sys.executable, "-c", CHILD, expected], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (n, r.stdout[-2000:], r.stderr[-4000:])
        line = r.stdout.strip().splitlines()[-1]
        print(f"[alloc] {n}: {line}", file=sys.stderr, flush=True)
        seen.append(line)
    failures = [s for s in seen if s.startswith("FAILED:")]
    assert len(failures) >= 6, seen
    assert all("failed" in s or "table" in s or "memory" in s.lower() for s in failures), seen
