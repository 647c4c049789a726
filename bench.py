"""Throughput of PeakGroupSelection.search on one MI355X.

Workloads (--workload, default 1):
  1  6,000 cycles x 60 windows, 1,500 peaks per spectrum (~554M peaks with
     planted elution profiles), 400k precursors (200k targets + 200k shifted
     decoys), 12 fragments, +-150 cycles, 10 ppm, top-3
  2  as 1 with 6-24 fragments, per-precursor half-windows of 30-200 cycles,
     clustered precursor m/z and top_k_fragments 24 (launch shape autotuned)
  3  the data of 2, forward + backward: rank-0 scores, a hinge loss over the
     target/decoy pairs and its gradient with respect to a Gaussian sigma
     per window and every fragment intensity (value = fwd+bwd precursors/s)
  4  20,000 cycles x 100 windows, 1,000 peaks per spectrum (~2B peaks),
     2M precursors (1M targets + 1M decoys), 6-24 fragments, half-windows
     of 50-400 cycles, 5 ppm, top-5 (--with-backward adds the fwd+bwd rate)

A step is one full search over the library (plan, bins, select, compaction
kernels) with the run and the library already resident in HBM; for workload
3 it is the forward, the loss and the backward.  The value is precursors per
second of the device pipeline; the host copy of the candidate table is
reported separately.  The oracle (the C restatement of the reference's
scalar path) is timed on the host's cores on a sample of the same library
for cpu_baseline, and its candidates are compared with the device's on that
sample.

The run's DIAData is built on the device (adsr_dia_create_device) and,
unless --no-csr-check, also by the host builder; the two CSRs must be equal
byte for byte.  Both build times are reported under config.dia_build.

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


def fwd_bwd(args, gen, dia, lib, params):
    """Workload 3: time forward (rank-0 scores) + hinge loss + backward.

    Returns (ms per step, forward device ms, backward device ms, extras)."""
    import numpy as np
    import torch

    from alphadia_search_rs.autograd import (SelectionGradState, hinge_loss,
                                             rank_zero_scores)
    st = SelectionGradState(dia, lib, params)
    sigma = torch.full((dia.num_observations,), 2.0, requires_grad=True)
    w = torch.tensor(gen["lib"][10], device="cuda", requires_grad=True)
    pairs = torch.tensor(np.asarray(gen["pairs"], np.int64), device="cuda")
    t_idx, d_idx = pairs[:, 0].contiguous(), pairs[:, 1].contiguous()
    stats = {}

    def step():
        sigma.grad = None
        w.grad = None
        s = rank_zero_scores(st, sigma, w)
        loss = hinge_loss(s, t_idx, d_idx, args.margin)
        loss.backward()
        return s, loss

    for _ in range(max(args.warmup, 1)):
        s, loss = step()
    torch.cuda.synchronize()
    fwd_ms = bwd_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s, loss = step()
        f, b = st.timings()
        fwd_ms += f
        bwd_ms += b
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stats["loss"] = float(loss.detach().cpu())
    stats["active_pairs"] = int(((args.margin - (s[t_idx] - s[d_idx])) > 0).sum().cpu())
    stats["grad_sigma_norm"] = float(sigma.grad.norm())
    stats["grad_weight_nonzero"] = int((w.grad != 0).sum().cpu())
    return step_ms, fwd_ms / args.steps, bwd_ms / args.steps, stats


def measured_traffic(workload):
    """HBM bytes per select_kernel launch from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy); None
    when no summary covers this workload."""
    path = os.path.join(ROOT, "profiles", f"select_workload{workload}_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if (len(parts) == 4 and parts[0] == "traffic_bytes_per_launch"
                        and parts[1] == f"workload={workload}"
                        and parts[2] == "kernel=select_kernel"):
                    return int(parts[3])
    except OSError:
        pass
    return None


def measured_valu(workload):
    """select_kernel VALU figures from the committed counter summary
    (VALU busy fraction, VALU instructions per precursor, lane utilization,
    VALU-issue floor in ms); None when no summary covers this workload."""
    path = os.path.join(ROOT, "profiles", f"select_workload{workload}_valu.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if (len(parts) == 7 and parts[0] == "valu_summary"
                        and parts[1] == f"workload={workload}"
                        and parts[2] == "kernel=select_kernel"):
                    kv = dict(p.split("=") for p in parts[3:])
                    return {k: float(v) for k, v in kv.items()}
    except (OSError, ValueError):
        pass
    return None


def copy_probe(mib):
    """A plain copy of known size, for scaling HBM byte counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    assert bool((dst[:: max(1, n // 1000)] == src[:: max(1, n // 1000)]).all())
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "ms_per_copy": round(ms, 4),
                      "GB_per_s": round(2 * n * 4 / (ms / 1e3) / 1e9, 1)}),
          flush=True)


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
    ap.add_argument("--no-scoring", action="store_true",
                    help="skip PeakGroupScoring.score of the run's candidates")
    ap.add_argument("--no-csr-check", action="store_true",
                    help="skip the host build of the run and the byte "
                         "comparison of its CSR with the device-built one")
    ap.add_argument("--threads-per-block", type=int, default=0)
    ap.add_argument("--precursors-per-wave", type=int, default=0)
    ap.add_argument("--margin", type=float, default=1.0,
                    help="hinge margin of the workload-3 loss")
    ap.add_argument("--with-backward", action="store_true",
                    help="workload 4: also time forward + loss + backward")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB "
                         "(5 times) and print its size and time: the "
                         "calibration for FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.workload not in (1, 2, 3, 4):
        print("bench.py: --workload is 1, 2, 3 or 4", file=sys.stderr)
        sys.exit(2)
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
    # workload 3 runs on the data of workload 2 (same generator seeds)
    spec = synthetic.config(2 if args.workload == 3 else args.workload)
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
    dia = ads.DIAData.from_arrays(*gen["dia"], device_build=True)
    dia_wall = time.time() - t1
    bt = dia.build_timings()
    log(f"DIAData built on the device in {dia_wall:.2f}s (prepare "
        f"{bt['prepare_ms']:.0f} ms, peak copy {bt['copy_ms']:.0f} ms, kernels "
        f"{bt['kernels_ms']:.1f} ms)")
    t1 = time.time()
    lib = ads.SpecLibFlat.from_arrays(*gen["lib"])
    log(f"SpecLibFlat built and uploaded in {time.time() - t1:.1f}s")
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

    fb = None
    if args.workload == 3 or (args.workload == 4 and args.with_backward):
        log("timing forward + hinge loss + backward")
        fb = fwd_bwd(args, gen, dia, lib, params)
        log(f"fwd+bwd step {fb[0]:.3f} ms (forward {fb[1]:.3f} ms, backward "
            f"{fb[2]:.3f} ms on the device), {fb[3]}")

    abytes = C.c_double()
    ents = C.c_double()
    hits = C.c_double()
    N.check(N.lib.adsr_select_algorithmic_bytes(dia._h, lib._h, C.byref(c),
                                                C.byref(abytes), C.byref(ents),
                                                C.byref(hits)))
    achieved = abytes.value / (kern_ms.value / 1e3) / 1e9
    fwd_rate = value
    ms = step_ms.value
    kind = "plain selection"
    if args.workload == 2:
        kind = "uneven batched selection (per-precursor RT windows)"
    if args.workload == 4:
        kind = "large uneven batched selection (per-precursor RT windows)"
    if fb is not None and args.workload == 3:
        ms = fb[0]
        value = n_prec / (ms / 1e3)
        kind = ("forward + hinge loss + backward (d/dsigma per window, "
                "d/dintensity per fragment) on the uneven batched selection")
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "precursors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, planted elution profiles)",
        "config": {
            "workload": (f"{kind}, {spec.n_cycles} cycles x "
                         f"{spec.n_windows} windows, {gen['n_peaks']} peaks, "
                         f"{n_prec} precursors ({spec.n_targets} targets + "
                         f"decoys), {spec.frag_min}-{spec.frag_max} fragments, "
                         f"+-{spec.half_window_min:g}..{spec.half_window_max:g} cycles, "
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
            "traffic": measured_traffic(args.workload),
        },
    }
    valu = measured_valu(args.workload)
    if valu is not None:
        # the kernel's binding limit: VALU issue under the reference's exact
        # f32 operation order (counter summary under profiles/)
        result["config"]["select_kernel_valu"] = valu

    build = {"device_wall_s": round(dia_wall, 3),
             **{k: round(v, 3) for k, v in bt.items()}}
    if not args.no_csr_check:
        # the same run through the host builder (src/dia_data_builder), and
        # every observation's CSR compared byte for byte
        t1 = time.time()
        hdia = ads.DIAData.from_arrays(*gen["dia"], device_build=False)
        build["host_builder_wall_s"] = round(time.time() - t1, 3)
        same = hdia.num_observations == dia.num_observations
        for o in range(dia.num_observations):
            if not same:
                break
            a, b = dia._observation(o), hdia._observation(o)
            same = all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))
                       for k in ("slice_starts", "cycle_indices", "intensities",
                                 "isolation_window")) and \
                a["num_cycles"] == b["num_cycles"]
        del hdia
        build["csr_identical_to_host_builder"] = bool(same)
        log(f"host builder {build['host_builder_wall_s']:.2f}s; device CSR "
            f"{'identical' if same else 'DIFFERENT'}")
    result["config"]["dia_build"] = build

    if not args.no_scoring:
        # PeakGroupScoring.score of this search's candidate table, handed
        # over in HBM (the next step of the reference pipeline)
        sp = ads.ScoringParameters()
        sp.update({"mass_tolerance": spec.mass_tolerance})
        scorer = ads.PeakGroupScoring(sp)
        dcands = ads.PeakGroupSelection(params).search_device(dia, lib)
        scorer.score(dia, lib, dcands)  # warm: fragment plan, IDF table
        t1 = time.perf_counter()
        feats = scorer.score(dia, lib, dcands)
        score_s = time.perf_counter() - t1
        tm = feats.device_timings()
        n_c = len(dcands)
        scoring = {
            "candidates": n_c,
            "scored": len(feats),
            "plan_ms": round(tm["plan_ms"], 3),
            "kernel_ms": round(tm["kernel_ms"], 3),
            "candidates_per_s_device": round(n_c / (tm["kernel_ms"] / 1e3), 1)
            if tm["kernel_ms"] > 0 else None,
            "call_s_incl_feature_copy": round(score_s, 4),
        }
        log(f"scoring: {n_c} candidates, kernel {tm['kernel_ms']:.2f} ms, plan "
            f"{tm['plan_ms']:.2f} ms, call {score_s * 1e3:.1f} ms")
        result["config"]["scoring"] = scoring
        fd = feats.to_dict_arrays()
        del feats
        # PeakGroupQuantification.quantify of the rank-0 candidates (the
        # reference quantifies the post-FDR subset, about one per precursor)
        tab = dcands.to_collection().to_arrays()
        r0m = tab[1] == 0
        qcols = tuple(c[r0m] for c in tab)
        qcands = ads.CandidateCollection.from_arrays(*qcols)
        qp = ads.QuantificationParameters()
        qp.update({"tolerance_ppm": spec.mass_tolerance})
        quant = ads.PeakGroupQuantification(qp)
        quant.quantify(dia, lib, qcands)  # warm
        t1 = time.perf_counter()
        qlib = quant.quantify(dia, lib, qcands)
        q_s = time.perf_counter() - t1
        qt = qlib.device_timings()
        result["config"]["quantification"] = {
            "candidates": int(r0m.sum()),
            "precursors": qlib.num_precursors,
            "fragments": qlib.num_fragments,
            "kernel_ms": round(qt["kernel_ms"], 3),
            "candidates_per_s_device": round(int(r0m.sum()) / (qt["kernel_ms"] / 1e3), 1)
            if qt["kernel_ms"] > 0 else None,
            "call_s_incl_host_layout": round(q_s, 4),
        }
        log(f"quantification: {int(r0m.sum())} candidates, kernel {qt['kernel_ms']:.2f} ms, "
            f"call {q_s * 1e3:.1f} ms")
        del qlib

    if fb is not None:
        fb_rate = n_prec / (fb[0] / 1e3)
        result["config"].update({
            "forward_only_precursors_per_s": round(fwd_rate, 1),
            "fwd_bwd_precursors_per_s": round(fb_rate, 1),
            "fwd_bwd_over_forward": round(fb_rate / fwd_rate, 3),
            "diff_forward_device_ms": round(fb[1], 4),
            "diff_backward_device_ms": round(fb[2], 4),
            "hinge_margin": args.margin,
            **{k: v for k, v in fb[3].items()},
        })

    if not args.no_cpu_baseline:
        import oracle
        # the host share of one GPU is 16 cores, whatever os.cpu_count() says
        nt = max(1, min(os.cpu_count() or 1, 16))
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
        if not args.no_scoring:
            # the oracle scores a sample of the same table; its features must
            # equal the device's (log10 columns within 2e-6)
            tab = dcands.to_collection().to_arrays()
            k = min(len(tab[0]), 100_000)
            samp = tuple(c[:k] for c in tab)
            # the reference builds the IDF table with the library: warm it
            oracle.score(odia, olib, tuple(c[:1] for c in tab),
                         mass_tolerance=spec.mass_tolerance, nthreads=nt)
            t5 = time.perf_counter()
            ofeat = oracle.score(odia, olib, samp, mass_tolerance=spec.mass_tolerance,
                                 nthreads=nt)
            o_s = time.perf_counter() - t5
            same = len(ofeat["precursor_idx"]) == k
            for name in ofeat:
                if not same:
                    break
                g, w = fd[name][:k], ofeat[name]
                if name in ("log10_b_ion_intensity", "log10_y_ion_intensity",
                            "idf_intensity_dot_product"):
                    same = bool(np.all(np.abs(g - w) <= 2e-6 * np.maximum(1.0, np.abs(w))))
                else:
                    same = bool(np.all(g == w))
            result["config"]["scoring"].update({
                "cpu_oracle_candidates_per_s": round(k / o_s, 1),
                "cpu_sample_candidates": k,
                "cpu_sample_features_identical": same,
            })
            log(f"cpu oracle scoring {k / o_s:.0f} candidates/s on {nt} threads; "
                f"sample features {'identical' if same else 'DIFFERENT'}")
            # the oracle (one thread) quantifies a 30k sample of the rank-0
            # table; the device quantifies the same sample for the comparison
            qs = tuple(c[:30_000] for c in qcols)
            qd = quant.quantify(dia, lib, ads.CandidateCollection.from_arrays(*qs)).to_dict_arrays()
            t6 = time.perf_counter()
            oq = oracle.quantify(odia, olib, gen["lib"], qs, tolerance_ppm=spec.mass_tolerance)
            oq_s = time.perf_counter() - t6
            qsame = all(np.array_equal(qd[i][k], oq[i][k]) for i in (0, 1) for k in oq[i])
            result["config"]["quantification"].update({
                "cpu_oracle_candidates_per_s": round(len(qs[0]) / oq_s, 1),
                "cpu_oracle_threads": 1,
                "cpu_sample_candidates": len(qs[0]),
                "cpu_sample_identical": qsame,
            })
            log(f"cpu oracle quantification {len(qs[0]) / oq_s:.0f} candidates/s "
                f"(1 thread); sample {'identical' if qsame else 'DIFFERENT'}")
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
                       f"{cpu_s:.1f}s"
                       + ("; forward only (the reference has no backward "
                          "pass), so an upper bound on a CPU fwd+bwd rate"
                          if fb is not None else "")),
        }
        result["config"]["cpu_sample_candidates_identical"] = identical
        result["config"]["gpu_over_cpu"] = round(value / cpu_rate, 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
