import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "alphadia-search-rs_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

os.environ.setdefault("ADSR_QUIET", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long-running check")


def has_gpu() -> bool:
    import alphadia_search_rs._native as N
    return N.device_available()


@pytest.fixture(scope="session")
def run0():
    """Workload 0 at full size (1,000 cycles x 8 windows, 2k precursors)."""
    import synthetic
    return synthetic.generate(synthetic.config(0))


@pytest.fixture(scope="session")
def oracle_run0(run0):
    import oracle
    return oracle.Dia(run0["dia"]), oracle.Lib(run0["lib"])


def params_for(spec, **extra):
    import alphadia_search_rs as ads
    p = ads.SelectionParameters()
    cfg = spec.selection_config()
    cfg.update(extra)
    p.update(cfg)
    return p


def oracle_params_for(spec, **extra):
    import oracle
    cfg = spec.selection_config()
    cfg.update(extra)
    return oracle.params(**cfg)


def assert_tables_equal(got, want, what=""):
    names = ["precursor_idx", "rank", "score", "scan_center", "scan_start",
             "scan_stop", "cycle_center", "cycle_start", "cycle_stop"]
    assert len(got) == 9 and len(want) == 9
    assert len(got[0]) == len(want[0]), \
        f"{what}: {len(got[0])} candidates, oracle has {len(want[0])}"
    for n, a, b in zip(names, got, want):
        if n == "score":
            assert a.dtype == np.float32
            bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        else:
            assert a.dtype == np.uint64
            bad = np.flatnonzero(a != b)
        assert bad.size == 0, (
            f"{what}: column {n} differs at {bad.size} rows, first {bad[:5]}: "
            f"got {a[bad[:5]]} want {b[bad[:5]]}")


def with_peak_everywhere(gen, mz_add, intensity):
    """Adds one peak at mz_add to every spectrum (kept sorted by m/z)."""
    dia = list(gen["dia"])
    start, stop, mz, it = dia[3], dia[4], dia[7], dia[8]
    new_mz, new_it, new_start, new_stop = [], [], [], []
    pos = 0
    for s0, s1 in zip(start, stop):
        m = np.concatenate([mz[s0:s1], np.array([mz_add], np.float32)])
        v = np.concatenate([it[s0:s1], np.array([intensity], np.float32)])
        o = np.argsort(m, kind="stable")
        new_mz.append(m[o])
        new_it.append(v[o])
        new_start.append(pos)
        pos += len(m)
        new_stop.append(pos)
    dia[3] = np.array(new_start, np.int64)
    dia[4] = np.array(new_stop, np.int64)
    dia[7] = np.concatenate(new_mz).astype(np.float32)
    dia[8] = np.concatenate(new_it).astype(np.float32)
    return dict(gen, dia=tuple(dia))
