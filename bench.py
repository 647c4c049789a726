"""Throughput of PeakGroupSelection.search on one MI355X.

Default workload: 6,000 cycles x 60 windows, 1,500 peaks per spectrum
(~554M peaks with planted elution profiles), 400k precursors (200k targets +
200k shifted decoys), 12 fragments, +-150 cycles, 10 ppm, top-3.

A step is one full search over the library (plan, bins, select, compaction
kernels) with the run and the library already resident in HBM.  The value
is precursors per second of the device pipeline; the host copy of the
candidate table is reported separately.  The oracle (the C restatement of
the reference's scalar path) is timed on the host's cores on a sample of the
same library for cpu_baseline, and its candidates are compared with the
device's on that sample.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "alphadia-search-rs_amd"))

METRIC = ("precursors/s (XIC→conv→score→top-k) at 400k precs; "
          "fwd+bwd precursors/s")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="library rows timed on the host for cpu_baseline "
                         "(0 = the whole library)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threads-per-block", type=int, default=0)
    ap.add_argument("--precursors-per-wave", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU split is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import alphadia_search_rs as ads
    import alphadia_search_rs._native as N
    import synthetic

    if not N.device_available():
        print("bench.py: no gfx950 device visible", file=sys.stderr)
        sys.exit(1)
    os.environ.setdefault("ADSR_QUIET", "1")
    spec = synthetic.config(args.workload)
    t0 = time.time()
    last = [0.0]

    def progress(c, total):
        if time.time() - last[0] > 10:
            last[0] = time.time()
            log(f"generating cycles {c}/{total}")

    gen = synthetic.generate(spec, progress=progress)
    log(f"generated {gen['n_peaks']} peaks, {spec.n_precursors} precursors in "
        f"{time.time() - t0:.1f}s")
    t1 = time.time()
    dia = ads.DIAData.from_arrays(*gen["dia"])
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    log(f"DIAData/SpecLibFlat built and uploaded in {time.time() - t1:.1f}s")
    params = ads.SelectionParameters()
    params.update(spec.selection_config())
    if gen["rt_tolerance_per_precursor"] is not None:
        params.update({"rt_tolerance_per_precursor": gen["rt_tolerance_per_precursor"]})
    c = params._c()

    tpb = C.c_int(args.threads_per_block)
    ppw = C.c_int(args.precursors_per_wave)
    step_ms = C.c_double()
    kern_ms = C.c_double()
    graph_ms = C.c_double()
    out = C.c_void_p()
    log(f"timing {args.steps} steps after {args.warmup} warmup")
    N.check(N.lib.adsr_select_timed(dia._h, lib._h, C.byref(c), args.warmup,
                                    args.steps, C.byref(tpb), C.byref(ppw),
                                    C.byref(step_ms), C.byref(kern_ms),
                                    C.byref(graph_ms), C.byref(out)))
    cand = ads.CandidateCollection._wrap(out.value)
    n_prec = lib.num_precursors
    value = n_prec / (step_ms.value / 1e3)
    log(f"step {step_ms.value:.3f} ms, select kernel {kern_ms.value:.3f} ms, "
        f"graph step {graph_ms.value:.3f} ms, {cand.len()} candidates, "
        f"shape {tpb.value} threads x {ppw.value} per grab")

    # PCIe-inclusive: the Python API call including the table copy to host
    p2 = ads.PeakGroupSelection(params)
    p2.search(dia, lib)
    t2 = time.perf_counter()
    api_reps = 3
    for _ in range(api_reps):
        cc2 = p2.search(dia, lib)
    api_s = (time.perf_counter() - t2) / api_reps
    log(f"search() incl. candidate table on host: {api_s * 1e3:.1f} ms")

    abytes = C.c_double()
    ents = C.c_double()
    hits = C.c_double()
    N.check(N.lib.adsr_select_algorithmic_bytes(dia._h, lib._h, C.byref(c),
                                                C.byref(abytes), C.byref(ents),
                                                C.byref(hits)))
    achieved = abytes.value / (kern_ms.value / 1e3) / 1e9
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "precursors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms.value, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, planted elution profiles)",
        "config": {
            "workload": (f"plain selection, {spec.n_cycles} cycles x "
                         f"{spec.n_windows} windows, {gen['n_peaks']} peaks, "
                         f"{n_prec} precursors ({spec.n_targets} targets + "
                         f"decoys), {spec.frag_min}-{spec.frag_max} fragments, "
                         f"+-{spec.half_window_max:g} cycles, "
                         f"{spec.mass_tolerance:g} ppm, top-{spec.candidate_count}"),
            "candidates": cand.len(),
            "select_kernel_ms": round(kern_ms.value, 4),
            "graph_ms_per_step": round(graph_ms.value, 4),
            "threads_per_block": tpb.value,
            "precursors_per_wave": ppw.value,
            "search_api_ms_incl_host_copy": round(api_s * 1e3, 2),
            "algorithmic_bytes_per_search": abytes.value,
            "csr_entries_scanned": ents.value,
            "entries_in_rt_window": hits.value,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
    }

    if not args.no_cpu_baseline:
        import oracle
        nt = os.cpu_count() or 1
        t3 = time.time()
        odia = oracle.Dia(gen["dia"], nthreads=nt)
        olib = oracle.Lib(gen["lib"])
        log(f"oracle run built in {time.time() - t3:.1f}s")
        kw = spec.selection_config()
        if gen["rt_tolerance_per_precursor"] is not None:
            kw["rt_tolerance_per_precursor"] = gen["rt_tolerance_per_precursor"]
        oprm = oracle.params(**kw)
        rows = n_prec if args.cpu_sample <= 0 else min(args.cpu_sample, n_prec)
        if rows >= n_prec:
            r0, r1, r2, r3 = 0, n_prec, n_prec, n_prec
        else:
            # decoys are the second half of the library: sample both halves
            r0, r1 = 0, rows // 2
            r2 = n_prec // 2
            r3 = r2 + (rows - rows // 2)
        oracle.select(odia, olib, oprm, r0, min(r1, 1000), nthreads=nt)  # warm
        t4 = time.perf_counter()
        oa = oracle.select(odia, olib, oprm, r0, r1, nthreads=nt)
        ob = oracle.select(odia, olib, oprm, r2, r3, nthreads=nt)
        cpu_s = time.perf_counter() - t4
        cpu_rate = rows / cpu_s
        got = cand.to_arrays()
        pidx = got[0].astype(np.int64)
        m = ((pidx >= r0) & (pidx < r1)) | ((pidx >= r2) & (pidx < r3))
        identical = True
        for k in range(9):
            want = np.concatenate([oa[k], ob[k]])
            g = got[k][m]
            if len(g) != len(want) or not np.array_equal(g.view(np.uint8), want.view(np.uint8)):
                identical = False
        log(f"cpu oracle {cpu_rate:.0f} prec/s on {nt} threads; sample parity "
            f"{'identical' if identical else 'MISMATCH'}")
        result["cpu_baseline"] = {
            "value": round(cpu_rate, 1),
            "unit": "precursors/s",
            "cores": nt,
            "kind": "port",
            "sample": (f"library rows [{r0},{r1}) and [{r2},{r3}) ({rows} "
                       f"precursors, targets and decoys) of the same run; C "
                       f"restatement of the reference scalar path, dynamic "
                       f"chunks of 64 precursors over {nt} threads; "
                       f"{cpu_s:.1f}s"),
        }
        result["config"]["cpu_sample_candidates_identical"] = identical
        result["config"]["gpu_over_cpu"] = round(value / cpu_rate, 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
