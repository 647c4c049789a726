"""The CPU suite under AddressSanitizer.

libadsr.so's host code (alphadia-search-rs_amd/csrc/asan.mk) and the oracle
(`make -C oracle asan`) are built with clang's -fsanitize=address against
the shared ASan runtime; the rest of the CPU tests then run in a child
process with that runtime preloaded and ADSR_LIBRARY / ADSR_ORACLE_LIBRARY
pointing at the instrumented builds.  Any heap, stack or global overflow,
use after free or double free in the C ABI, the host builders, the
candidate tables or the oracle aborts the child with a report.  Leak
checking is off: the interpreter itself does not free everything at exit.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "alphadia-search-rs_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang"

pytestmark = pytest.mark.skipif(os.environ.get("ADSR_ASAN_CHILD") == "1",
                                reason="already inside the sanitizer run")


def test_cpu_suite_under_address_sanitizer(tmp_path):
    jobs = str(min(8, os.cpu_count() or 2))
    subprocess.run(["make", "-s", "-j", jobs, "-f", os.path.join(CSRC, "asan.mk")],
                   check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    rt = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"],
                        check=True, capture_output=True, text=True).stdout.strip()
    assert os.path.exists(rt), rt
    log = str(tmp_path / "asan")
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": rt,
        "ASAN_OPTIONS": f"detect_leaks=0:abort_on_error=1:log_path={log}",
        "ADSR_LIBRARY": os.path.join(CSRC, "_build", "asan", "libadsr.so"),
        "ADSR_ORACLE_LIBRARY": os.path.join(ROOT, "oracle", "_build_asan", "liboracle.so"),
        "ADSR_ASAN_CHILD": "1",
    })
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"),
                        "-m", "not gpu", "-x", "-q", "-p", "no:cacheprovider",
                        "--deselect", "tests/test_oracle_mutants.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    reports = sorted(glob.glob(log + "*"))
    text = "".join(open(f).read()[:4000] for f in reports)
    assert r.returncode == 0 and not reports, (r.stdout[-3000:], r.stderr[-2000:], text)
    assert " passed" in r.stdout.splitlines()[-1]
