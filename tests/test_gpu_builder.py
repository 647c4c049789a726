"""DIAData CSR built on the device vs the host builder and the oracle.

adsr_dia_create_device must give the binned CSR of
src/dia_data_builder/mod.rs:52-139 bit for bit: slice_starts, the cycle
index and intensity of every entry in (bin, cycle, spectrum, peak) order,
window bounds and cycle counts.  The runs below stress each part of the
device pipeline: peaks on grid points, on exact f32 midpoints (ties go to
the lower bin), NaN / +-inf / out-of-range m/z, cycles that wrap at 2^16,
shuffled and duplicate cycles, spectra longer than one snap item, windows
of many sort tiles and of exactly one, bins whose runs cross tile borders,
empty spectra, empty windows and MS1 (negative delta_scan_idx) spectra.
"""
import numpy as np
import pytest

import oracle
import synthetic
from conftest import assert_tables_equal, params_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ads():
    import alphadia_search_rs as m
    import alphadia_search_rs._native as N
    if not N.device_available():
        pytest.fail("gfx950 device not available for a gpu-marked test")
    return m


def _assert_same_csr(dev, host, od=None):
    assert dev.num_observations == host.num_observations
    if od is not None:
        assert dev.num_observations == od.n_obs
    for o in range(dev.num_observations):
        a, b = dev._observation(o), host._observation(o)
        refs = [b] if od is None else [b, od.observation(o)]
        for ref in refs:
            for k in ("slice_starts", "cycle_indices", "intensities",
                      "isolation_window"):
                assert a[k].shape == ref[k].shape, (o, k)
                assert np.array_equal(a[k].view(np.uint8), ref[k].view(np.uint8)), (o, k)
            assert a["num_cycles"] == ref["num_cycles"], o
    ra, da = dev._rt_index()
    rb, db = host._rt_index()
    assert np.array_equal(ra, rb) and da == db


def _edge_run(seed, n_cycles=40, n_windows=4, peaks=(0, 900), hot_bin=False,
              special=False, wrap=False, dup=False, shuffle=False, ms1=False,
              empty_window=False):
    rng = np.random.default_rng(seed)
    grid = oracle.grid()
    spectra = []
    for c in range(n_cycles):
        if ms1:
            spectra.append((-1, 0.0, 0.0, c, rng.integers(0, 40)))
        for w in range(n_windows):
            if empty_window and w == n_windows - 2:
                continue
            cyc = c
            if wrap and rng.random() < 0.3:
                cyc = c + 65536 * int(rng.integers(1, 3))
            if dup and rng.random() < 0.2:
                cyc = max(c - 1, 0)
            spectra.append((w, 400.0 + 50 * w, 450.0 + 50 * w, cyc,
                            rng.integers(peaks[0], peaks[1] + 1)))
    if shuffle:
        spectra = [spectra[i] for i in rng.permutation(len(spectra))]
    mz_parts, in_parts, rows = [], [], []
    start = 0
    for (w, lo, hi, cyc, n) in spectra:
        mz = rng.uniform(140.0, 2010.0, int(n)).astype(np.float32)
        if special and n > 0:
            k = min(int(n), 40)
            i = rng.integers(0, len(grid) - 1, k)
            mid = ((grid[i].astype(np.float64) + grid[i + 1]) / 2).astype(np.float32)
            mz[:k // 2] = grid[i[:k // 2]]
            mz[k // 2:k] = mid[k // 2:]
            extra = np.array([np.nan, np.inf, -np.inf, 0.0, -5.0, 100.0, 149.99,
                              150.0, grid[0], grid[-1], 2000.0, 2000.5, 3000.0],
                             np.float32)
            m = min(len(extra), int(n) - k)
            mz[k:k + m] = extra[:m]
        if hot_bin and n > 0:
            mz[: min(5, int(n))] = np.float32(500.0)
        mz = np.sort(mz) if rng.random() < 0.5 else mz
        it = rng.lognormal(6, 1.5, int(n)).astype(np.float32)
        mz_parts.append(mz)
        in_parts.append(it)
        rows.append((w, lo, hi, start, start + int(n), cyc, 0.5 * (cyc % 65536)))
        start += int(n)
    r = np.array(rows, dtype=np.float64)
    dia = (r[:, 0].astype(np.int64), r[:, 1].astype(np.float32),
           r[:, 2].astype(np.float32), r[:, 3].astype(np.int64),
           r[:, 4].astype(np.int64), r[:, 5].astype(np.int64),
           r[:, 6].astype(np.float32),
           np.concatenate(mz_parts) if mz_parts else np.zeros(0, np.float32),
           np.concatenate(in_parts) if in_parts else np.zeros(0, np.float32),
           np.zeros((1, 1, 1, 1), np.float32))
    return dia


EDGE = {
    "plain": dict(),
    "special_mz": dict(special=True),
    "wrap_dup_shuffle": dict(wrap=True, dup=True, shuffle=True),
    "ms1_empty_window": dict(ms1=True, empty_window=True, peaks=(0, 300)),
    "long_spectra": dict(n_cycles=12, peaks=(2000, 5000), hot_bin=True),
    "hot_bin_many_tiles": dict(n_cycles=300, n_windows=2, peaks=(50, 120),
                               hot_bin=True, shuffle=True),
    "single_window": dict(n_cycles=30, n_windows=1, peaks=(1, 40)),
}


@pytest.mark.parametrize("name", sorted(EDGE))
def test_device_csr_equals_host_and_oracle(ads, name):
    dia = _edge_run(11, **EDGE[name])
    print(f"builder case {name}: {len(dia[0])} spectra, {len(dia[7])} peaks",
          flush=True)
    dev = ads.DIAData.from_arrays(*dia, device_build=True)
    host = ads.DIAData.from_arrays(*dia, device_build=False)
    _assert_same_csr(dev, host, oracle.Dia(dia))


def test_device_csr_tile_boundaries(ads):
    """Windows of exactly one, exactly two and one-past-two sort tiles."""
    n = np.array([4096, 8192, 8193, 1, 4095], np.int64)
    rng = np.random.default_rng(5)
    mz = rng.uniform(300, 310, int(n.sum())).astype(np.float32)
    it = np.arange(n.sum(), dtype=np.float32)
    stops = np.cumsum(n)
    dia = (np.arange(5, dtype=np.int64), np.full(5, 400, np.float32),
           np.full(5, 450, np.float32), stops - n, stops,
           np.zeros(5, np.int64), np.zeros(5, np.float32), mz, it,
           np.zeros((1, 1, 1, 1), np.float32))
    dev = ads.DIAData.from_arrays(*dia, device_build=True)
    host = ads.DIAData.from_arrays(*dia, device_build=False)
    _assert_same_csr(dev, host, oracle.Dia(dia))


def test_device_csr_golden_cases(ads):
    """The reference's builder KATs (src/dia_data_builder/tests.rs:31-269)
    through the device builder."""
    from test_host import GOLD, _dia_arrays
    case = GOLD["builder_three_windows"]
    dia = _dia_arrays(case["arrays"])
    d = ads.DIAData.from_arrays(*dia, device_build=True)
    assert d.num_observations == case["num_observations"]
    for mz, want in case["valid"]:
        assert d.get_valid_observations(mz) == want
    for o, w in enumerate(case["windows"]):
        assert d._observation(o)["isolation_window"].tolist() == w
    _assert_same_csr(d, ads.DIAData.from_arrays(*dia, device_build=False))
    case = GOLD["builder_cycle_order"]
    d = ads.DIAData.from_arrays(*_dia_arrays(case["arrays"]), device_build=True)
    g = oracle.grid()
    b = oracle.lib().orc_find_closest_index(oracle.p(g), len(g), case["mz"])
    o = d._observation(0)
    s, e = o["slice_starts"][b], o["slice_starts"][b + 1]
    assert o["cycle_indices"][s:e].tolist() == case["cycles"]
    assert o["intensities"][s:e].tolist() == case["intensities"]


def test_device_built_search_matches_host_built(ads):
    """Candidates from a device-built DIAData equal those from the host
    builder's (workload 0 with shuffled spectra)."""
    gen = synthetic.generate(synthetic.config(0, n_cycles=200, n_windows=6,
                                              n_targets=300))
    dia = list(gen["dia"])
    perm = np.random.default_rng(3).permutation(len(dia[0]))
    for i in range(7):
        dia[i] = dia[i][perm]
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    p = params_for(gen["spec"])
    got = ads.PeakGroupSelection(p).search(
        ads.DIAData.from_arrays(*dia, device_build=True), lib).to_arrays()
    want = ads.PeakGroupSelection(p).search(
        ads.DIAData.from_arrays(*dia, device_build=False), lib).to_arrays()
    assert_tables_equal(got, want, "device-built vs host-built DIAData")
    assert len(got[0]) > 0


def test_device_build_reports_timings(ads):
    dia = _edge_run(2, n_cycles=10)
    d = ads.DIAData.from_arrays(*dia, device_build=True)
    t = d.build_timings()
    assert set(t) == {"prepare_ms", "copy_ms", "kernels_ms"}
    assert all(v >= 0 for v in t.values()) and t["kernels_ms"] > 0
    h = ads.DIAData.from_arrays(*dia, device_build=False)
    assert all(v == 0 for v in h.build_timings().values())
