"""Throughput of PeakGroupSelection.search on one MI355X, and of its
forward + backward pass.

The default run (--workload 1) measures both halves of the metric:
  value      workload 1: 6,000 cycles x 60 windows, 1,500 peaks per
             spectrum (554M peaks with planted elution profiles), 400k
             precursors (200k targets + 200k shifted decoys), 12 fragments,
             +-150 cycles, 10 ppm, top-3; precursors/s of the search
  fwd_bwd    workload 3 on workload 2's data (6-24 fragments, per-precursor
             half-windows of 30-200 cycles, clustered precursor m/z, top_k
             24): rank-0 scores, a hinge loss over the 200k target/decoy
             pairs and its gradient with respect to a Gaussian sigma per
             window and every fragment intensity; fwd+bwd precursors/s, the
             same run's forward-only rate, and a check of the gradient of a
             1,000-precursor loss against the FP64 model (1e-4 relative)
Other workloads (--workload 2, 3, 4) time one workload alone; workload 4 is
20,000 cycles x 100 windows (2B peaks), 2M precursors, 6-24 fragments,
half-windows of 50-400, 5 ppm, top-5 (--with-backward adds fwd+bwd).

A step is one full search over the library (plan, select and compaction
kernels) with the run and the library already resident in HBM: value is
device-resident throughput.  search() including the candidate table's copy
to host memory (the end of the reference's timed region,
src/peak_group_selection/mod.rs:99-123) is reported beside it.  The oracle
(the C restatement of the reference's scalar path) is timed on the host's
cores for cpu_baseline: one warm-up, then the median of 5 runs over the
whole library; its candidates must equal the device's.

The run's DIAData is built on the device and, unless --no-csr-check, also by
the host builder; the two CSRs must be equal byte for byte.

Prints one JSON line.
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "alphadia-search-rs_amd"))

METRIC = ("precursors/s (XIC→conv→score→top-k) at 400k precs; "
          "fwd+bwd precursors/s")
HBM_PEAK_GBS = 8000.0
KERNEL_SOURCE = os.path.join(ROOT, "alphadia-search-rs_amd", "csrc", "adsr_device.hip")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    """Threads for host work: the process's CPU affinity, at most 16 (the
    host share of one GPU on the measurement machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16)), n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def source_sha16():
    with open(KERNEL_SOURCE, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def profile_line(name, key, workload):
    """The machine-readable line `key workload=W kernel=select_kernel
    source=S k=v...` of a committed counter summary, as a dict, when S is
    the sha256 prefix of the current kernel source (a summary measured on
    another build is not used).  None otherwise."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if (len(parts) >= 4 and parts[0] == key
                        and parts[1] == f"workload={workload}"
                        and parts[2] == "kernel=select_kernel"):
                    kv = dict(p.split("=", 1) for p in parts[3:])
                    if kv.pop("source", None) != source_sha16():
                        return None
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


def workload_text(kind, spec, gen):
    return (f"{kind}, {spec.n_cycles} cycles x {spec.n_windows} windows, "
            f"{gen['n_peaks']} peaks, {spec.n_precursors} precursors "
            f"({spec.n_targets} targets + decoys), {spec.frag_min}-{spec.frag_max} "
            f"fragments, +-{spec.half_window_min:g}..{spec.half_window_max:g} cycles, "
            f"{spec.mass_tolerance:g} ppm, top-{spec.candidate_count}")


class Run:
    """A generated workload with its DIAData (device build), SpecLibFlat and
    SelectionParameters."""

    def __init__(self, workload, threads):
        import alphadia_search_rs as ads
        import synthetic
        self.workload = workload
        self.spec = synthetic.config(2 if workload == 3 else workload)
        t0 = time.time()
        self.gen = synthetic.generate(self.spec, threads=threads)
        log(f"workload {workload}: generated {self.gen['n_peaks']} peaks, "
            f"{self.spec.n_precursors} precursors in {time.time() - t0:.1f}s")
        t1 = time.time()
        self.dia = ads.DIAData.from_arrays(*self.gen["dia"], device_build=True)
        self.dia_wall = time.time() - t1
        self.build_timings = self.dia.build_timings()
        bt = self.build_timings
        log(f"DIAData built on the device in {self.dia_wall:.2f}s (prepare "
            f"{bt['prepare_ms']:.0f} ms, peak copy {bt['copy_ms']:.0f} ms, kernels "
            f"{bt['kernels_ms']:.1f} ms)")
        self.lib = ads.SpecLibFlat.from_arrays(*self.gen["lib"])
        self.params = ads.SelectionParameters()
        self.params.update(self.spec.selection_config())
        if self.gen["rt_tolerance_per_precursor"] is not None:
            self.params.update({"rt_tolerance_per_precursor":
                                self.gen["rt_tolerance_per_precursor"]})
        self.n_prec = self.lib.num_precursors

    def oracle_params(self):
        import oracle
        kw = self.spec.selection_config()
        if self.gen["rt_tolerance_per_precursor"] is not None:
            kw["rt_tolerance_per_precursor"] = self.gen["rt_tolerance_per_precursor"]
        return oracle.params(**kw)


def time_search(run, args):
    """adsr_select_timed (autotuned shape, HIP events around each step and
    around the select launches, hipGraph replay), the algorithmic bytes, and
    search() with the candidate table copied to host memory."""
    import alphadia_search_rs as ads
    import alphadia_search_rs._native as N
    c = run.params._c()
    tpb = C.c_int(args.threads_per_block)
    ppw = C.c_int(args.precursors_per_wave)
    step_ms, kern_ms, graph_ms = C.c_double(), C.c_double(), C.c_double()
    out = C.c_void_p()
    log(f"timing {args.steps} steps after {args.warmup} warmup")
    N.check(N.lib.adsr_select_timed(run.dia._h, run.lib._h, C.byref(c), args.warmup,
                                    args.steps, C.byref(tpb), C.byref(ppw),
                                    C.byref(step_ms), C.byref(kern_ms),
                                    C.byref(graph_ms), C.byref(out)))
    cand = ads.CandidateCollection._wrap(out.value)
    log(f"step {step_ms.value:.3f} ms, select kernel {kern_ms.value:.3f} ms, "
        f"graph step {graph_ms.value:.3f} ms, {cand.len()} candidates, "
        f"shape {tpb.value} threads x {ppw.value} per grab")
    # search() as a caller makes it: the table ends in host memory
    sel = ads.PeakGroupSelection(run.params)
    sel.search(run.dia, run.lib)
    api = []
    for _ in range(max(5, args.steps // 2)):
        t = time.perf_counter()
        sel.search(run.dia, run.lib)
        api.append((time.perf_counter() - t) * 1e3)
    api_ms = statistics.median(api)
    log(f"search() incl. candidate table in host memory: {api_ms:.2f} ms (median)")
    abytes, ents, hits = C.c_double(), C.c_double(), C.c_double()
    N.check(N.lib.adsr_select_algorithmic_bytes(run.dia._h, run.lib._h, C.byref(c),
                                                C.byref(abytes), C.byref(ents),
                                                C.byref(hits)))
    return dict(step_ms=step_ms.value, kern_ms=kern_ms.value, graph_ms=graph_ms.value,
                tpb=tpb.value, ppw=ppw.value, cand=cand, api_ms=api_ms,
                abytes=abytes.value, entries=ents.value, hits=hits.value)


def fwd_bwd(args, run):
    """Forward (rank-0 scores) + hinge loss over all target/decoy pairs +
    backward, timed over `steps` steps.  Returns the state too, for the
    gradient check."""
    import numpy as np
    import torch

    from alphadia_search_rs.autograd import (SelectionGradState, hinge_loss,
                                             rank_zero_scores)
    st = SelectionGradState(run.dia, run.lib, run.params)
    sigma = torch.full((run.dia.num_observations,), 2.0, requires_grad=True)
    w = torch.tensor(run.gen["lib"][10], device="cuda", requires_grad=True)
    pairs = torch.tensor(np.asarray(run.gen["pairs"], np.int64), device="cuda")
    t_idx, d_idx = pairs[:, 0].contiguous(), pairs[:, 1].contiguous()

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
    stats = {
        "loss": float(loss.detach().cpu()),
        "active_pairs": int(((args.margin - (s[t_idx] - s[d_idx])) > 0).sum().cpu()),
        "grad_sigma_norm": float(sigma.grad.norm()),
        "grad_weight_nonzero": int((w.grad != 0).sum().cpu()),
    }
    return dict(step_ms=step_ms, fwd_ms=fwd_ms / args.steps, bwd_ms=bwd_ms / args.steps,
                stats=stats, state=st)


def grad_check(run, st, threads):
    """The device gradient of a hinge loss over the first 500 target/decoy
    pairs (1,000 precursors) of the full-size state, against the oracle's
    FP64 model (oracle/grad64.py), with perturbed sigma and weights."""
    import numpy as np
    import torch

    import oracle
    from oracle.grad64 import F64Model, check_gradient, pick_margin, rel_errors
    from alphadia_search_rs.autograd import hinge_loss, rank_zero_scores
    gen = run.gen
    W = run.dia.num_observations
    rng = np.random.default_rng(7)
    sigma = np.linspace(1.6, 2.8, W).astype(np.float32)
    weights = (gen["lib"][10] * rng.uniform(0.5, 1.5, gen["lib"][10].shape)).astype(np.float32)
    sub = np.asarray(gen["pairs"][:500])
    rows = np.concatenate([sub[:, 0], sub[:, 1]])
    sig_t = torch.tensor(sigma, requires_grad=True)
    w_t = torch.tensor(weights, device="cuda", requires_grad=True)
    s_t = rank_zero_scores(st, sig_t, w_t)
    s_dev = s_t.detach().cpu().numpy()
    # apex of each subset row: the rank-0 centre of the oracle's search with
    # the same sigma and weights (equal to the device's, as the tests show)
    lib_w = list(gen["lib"])
    lib_w[10] = weights
    od, ol = oracle.Dia(gen["dia"], nthreads=threads), oracle.Lib(tuple(lib_w))
    kw = gen["spec"].selection_config()
    prm = oracle.params(**kw, rt_tolerance_per_precursor=gen["rt_tolerance_per_precursor"],
                        sigma_per_window=sigma)
    apex = np.full(run.n_prec, -1, np.int64)
    for lo, hi in ((0, 500), (run.n_prec // 2, run.n_prec // 2 + 500)):
        t = oracle.select(od, ol, prm, lo, hi, nthreads=threads)
        m0 = t[1] == 0
        apex[t[0][m0].astype(np.int64)] = t[6][m0].astype(np.int64)
    model = F64Model(gen, weights, sigma, apex, dia=od, lib_o=ol, prm=prm)
    s64 = model.s_star(sigma.astype(np.float64), weights.astype(np.float64))
    has = apex[rows] >= 0
    rel_fwd, _ = rel_errors(s_dev[rows][has], s64[rows][has], 1e-4)
    m = pick_margin(s64, sub)
    loss = hinge_loss(s_t, torch.tensor(sub[:, 0], device="cuda"),
                      torch.tensor(sub[:, 1], device="cuda"), m)
    loss.backward()
    gw_ref, gs_ref, fd_rel_s, fd_w = check_gradient(model, sigma, weights, sub, m)
    rel_s, bad_s = rel_errors(sig_t.grad.numpy(), gs_ref, 1e-4)
    rel_w, bad_w = rel_errors(w_t.grad.cpu().numpy(), gw_ref, 1e-4)
    ok = (bad_s.size == 0 and bad_w.size == 0 and fd_rel_s <= 1e-4 and fd_w <= 1.0
          and rel_fwd <= 1e-4)
    return {"precursors": int(rows.size), "pairs": int(len(sub)), "margin": round(m, 6),
            "forward_rel_err_vs_fp64": float(rel_fwd),
            "fp64_sigma_grad_vs_central_differences_rel_err": float(fd_rel_s),
            "fp64_weight_grad_vs_central_differences_worst_over_bound": float(fd_w),
            "device_sigma_grad_rel_err": float(rel_s),
            "device_weight_grad_rel_err": float(rel_w),
            "within_1e-4_of_fp64": bool(ok)}


def cpu_baseline(run, cand, threads, affinity, value):
    """The oracle over the whole library: one warm-up, then the median of 5
    timed runs; its table must equal the device's."""
    import numpy as np

    import oracle
    t0 = time.time()
    odia = oracle.Dia(run.gen["dia"], nthreads=threads)
    olib = oracle.Lib(run.gen["lib"])
    log(f"oracle run built in {time.time() - t0:.1f}s")
    oprm = run.oracle_params()
    oracle.select(odia, olib, oprm, nthreads=threads)  # warm-up
    times = []
    for _ in range(5):
        t = time.perf_counter()
        want = oracle.select(odia, olib, oprm, nthreads=threads)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    got = cand.to_arrays()
    identical = all(len(g) == len(w) and np.array_equal(g.view(np.uint8), w.view(np.uint8))
                    for g, w in zip(got, want))
    rate = run.n_prec / med
    log(f"cpu oracle {rate:.0f} prec/s on {threads} threads (median of 5); table "
        f"{'identical' if identical else 'MISMATCH'}")
    return {
        "value": round(rate, 1),
        "unit": "precursors/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"all {run.n_prec} precursors of the same run, one warm-up then "
                   f"the median of 5 runs ({', '.join(f'{t:.3f}' for t in times)} s); "
                   f"C restatement of the reference scalar path, dynamic chunks of 64 "
                   f"precursors over {threads} threads (CPU affinity {affinity}, capped "
                   f"at the 16-core host share of one GPU); {cpu_model()}"),
    }, identical, odia, olib


def scoring_and_quant(run, params, dcands, odia, olib, threads, check):
    """PeakGroupScoring.score of the search's table handed over in HBM, and
    PeakGroupQuantification.quantify of its rank-0 rows; with `check`, the
    oracle's features and quantified dicts on samples of the same tables."""
    import numpy as np

    import alphadia_search_rs as ads
    spec = run.spec
    sp = ads.ScoringParameters()
    sp.update({"mass_tolerance": spec.mass_tolerance})
    scorer = ads.PeakGroupScoring(sp)
    scorer.score(run.dia, run.lib, dcands)  # warm: fragment plan, IDF table
    t1 = time.perf_counter()
    feats = scorer.score(run.dia, run.lib, dcands)
    score_s = time.perf_counter() - t1
    tm = feats.device_timings()
    n_c = len(dcands)
    scoring = {
        "candidates": n_c, "scored": len(feats),
        "plan_ms": round(tm["plan_ms"], 3), "kernel_ms": round(tm["kernel_ms"], 3),
        "candidates_per_s_device": round(n_c / (tm["kernel_ms"] / 1e3), 1)
        if tm["kernel_ms"] > 0 else None,
        "call_s_incl_feature_copy": round(score_s, 4),
    }
    log(f"scoring: {n_c} candidates, kernel {tm['kernel_ms']:.2f} ms, call "
        f"{score_s * 1e3:.1f} ms")
    fd = feats.to_dict_arrays()
    del feats
    tab = dcands.to_collection().to_arrays()
    r0m = tab[1] == 0
    qcols = tuple(c[r0m] for c in tab)
    qcands = ads.CandidateCollection.from_arrays(*qcols)
    qp = ads.QuantificationParameters()
    qp.update({"tolerance_ppm": spec.mass_tolerance})
    quant = ads.PeakGroupQuantification(qp)
    quant.quantify(run.dia, run.lib, qcands)  # warm
    t1 = time.perf_counter()
    qlib = quant.quantify(run.dia, run.lib, qcands)
    q_s = time.perf_counter() - t1
    qt = qlib.device_timings()
    quantification = {
        "candidates": int(r0m.sum()), "precursors": qlib.num_precursors,
        "fragments": qlib.num_fragments, "kernel_ms": round(qt["kernel_ms"], 3),
        "candidates_per_s_device": round(int(r0m.sum()) / (qt["kernel_ms"] / 1e3), 1)
        if qt["kernel_ms"] > 0 else None,
        "call_s_incl_host_layout": round(q_s, 4),
    }
    log(f"quantification: {int(r0m.sum())} candidates, kernel {qt['kernel_ms']:.2f} ms")
    del qlib
    if check:
        import oracle
        k = min(len(tab[0]), 100_000)
        samp = tuple(c[:k] for c in tab)
        oracle.score(odia, olib, tuple(c[:1] for c in tab),
                     mass_tolerance=spec.mass_tolerance, nthreads=threads)
        t5 = time.perf_counter()
        ofeat = oracle.score(odia, olib, samp, mass_tolerance=spec.mass_tolerance,
                             nthreads=threads)
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
        scoring.update({"cpu_oracle_candidates_per_s": round(k / o_s, 1),
                        "cpu_oracle_threads": threads,
                        "cpu_sample_candidates": k,
                        "cpu_sample_features_identical": same})
        log(f"cpu oracle scoring {k / o_s:.0f} candidates/s on {threads} threads; "
            f"sample {'identical' if same else 'DIFFERENT'}")
        qs = tuple(c[:30_000] for c in qcols)
        qd = quant.quantify(run.dia, run.lib,
                            ads.CandidateCollection.from_arrays(*qs)).to_dict_arrays()
        t6 = time.perf_counter()
        oq = oracle.quantify(odia, olib, run.gen["lib"], qs, tolerance_ppm=spec.mass_tolerance,
                             nthreads=threads)
        oq_s = time.perf_counter() - t6
        qsame = all(np.array_equal(qd[i][kk], oq[i][kk]) for i in (0, 1) for kk in oq[i])
        quantification.update({"cpu_oracle_candidates_per_s": round(len(qs[0]) / oq_s, 1),
                               "cpu_oracle_threads": threads,
                               "cpu_sample_candidates": len(qs[0]),
                               "cpu_sample_identical": qsame})
        log(f"cpu oracle quantification {len(qs[0]) / oq_s:.0f} candidates/s on "
            f"{threads} threads; sample {'identical' if qsame else 'DIFFERENT'}")
    return scoring, quantification


def csr_check(run):
    """The same run through the host builder (src/dia_data_builder), every
    observation's CSR compared byte for byte with the device build."""
    import numpy as np

    import alphadia_search_rs as ads
    t1 = time.time()
    hdia = ads.DIAData.from_arrays(*run.gen["dia"], device_build=False)
    wall = time.time() - t1
    same = hdia.num_observations == run.dia.num_observations
    for o in range(run.dia.num_observations):
        if not same:
            break
        a, b = run.dia._observation(o), hdia._observation(o)
        same = all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))
                   for k in ("slice_starts", "cycle_indices", "intensities",
                             "isolation_window")) and a["num_cycles"] == b["num_cycles"]
    log(f"host builder {wall:.2f}s; device CSR {'identical' if same else 'DIFFERENT'}")
    return wall, bool(same)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--no-fwd-bwd", action="store_true",
                    help="workload 1: skip the workload-3 forward + backward leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scoring", action="store_true",
                    help="skip PeakGroupScoring.score and quantify of the candidates")
    ap.add_argument("--no-csr-check", action="store_true",
                    help="skip the host build of the run and the byte "
                         "comparison of its CSR with the device-built one")
    ap.add_argument("--threads-per-block", type=int, default=0)
    ap.add_argument("--precursors-per-wave", type=int, default=0)
    ap.add_argument("--margin", type=float, default=1.0,
                    help="hinge margin of the fwd+bwd loss")
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
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this "
              "build measures one GPU's share; the multi-GPU split is a design only",
              file=sys.stderr)
        sys.exit(2)

    import alphadia_search_rs._native as N
    if not N.device_available():
        print("bench.py: no gfx950 device visible", file=sys.stderr)
        sys.exit(1)
    os.environ.setdefault("ADSR_QUIET", "1")
    threads, affinity = host_threads()
    import alphadia_search_rs as ads

    run = Run(args.workload, threads)
    ts = time_search(run, args)
    value = run.n_prec / (ts["step_ms"] / 1e3)
    achieved = ts["abytes"] / (ts["kern_ms"] / 1e3) / 1e9
    kind = {1: "plain selection",
            2: "uneven batched selection (per-precursor RT windows)",
            3: "uneven batched selection (per-precursor RT windows)",
            4: "large uneven batched selection (per-precursor RT windows)"}[args.workload]
    traffic = profile_line(f"select_workload{args.workload}_hbm_traffic.md",
                           "traffic_bytes_per_launch", args.workload)
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "precursors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ts["step_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, planted elution profiles)",
        "config": {
            "workload": workload_text(kind, run.spec, run.gen),
            "candidates": ts["cand"].len(),
            "select_kernel_ms": round(ts["kern_ms"], 4),
            "graph_ms_per_step": round(ts["graph_ms"], 4),
            "threads_per_block": ts["tpb"],
            "precursors_per_grab": ts["ppw"],
            "search_api_ms_incl_host_copy": round(ts["api_ms"], 3),
            "search_api_precursors_per_s_incl_host_copy":
                round(run.n_prec / (ts["api_ms"] / 1e3), 1),
            "search_api_over_step": round(ts["api_ms"] / ts["step_ms"], 3),
            "algorithmic_bytes_per_search": ts["abytes"],
            "csr_entries_scanned": ts["entries"],
            "entries_in_rt_window": ts["hits"],
            "host_threads": threads,
            "host_cpu": cpu_model(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": int(traffic["bytes"]) if traffic else None,
        },
    }
    valu = profile_line(f"select_workload{args.workload}_valu.md", "valu_summary",
                        args.workload)
    if valu is not None:
        result["config"]["select_kernel_valu"] = valu
    build = {"device_wall_s": round(run.dia_wall, 3),
             **{k: round(v, 3) for k, v in run.build_timings.items()}}
    if not args.no_csr_check:
        build["host_builder_wall_s"], build["csr_identical_to_host_builder"] = csr_check(run)
    result["config"]["dia_build"] = build

    odia = olib = None
    if not args.no_cpu_baseline:
        cb, identical, odia, olib = cpu_baseline(run, ts["cand"], threads, affinity, value)
        result["cpu_baseline"] = cb
        result["config"]["cpu_candidates_identical"] = identical
        result["config"]["gpu_over_cpu"] = round(value / cb["value"], 1)
    if not args.no_scoring:
        dcands = ads.PeakGroupSelection(run.params).search_device(run.dia, run.lib)
        sc, qu = scoring_and_quant(run, run.params, dcands, odia, olib, threads,
                                   check=odia is not None)
        result["config"]["scoring"] = sc
        result["config"]["quantification"] = qu
        del dcands
    del odia, olib

    # forward + backward: workload 3 (alone), workload 4 --with-backward, or
    # the default run's second leg on workload 2's data
    fb_run = None
    own_run = False
    if args.workload == 3 or (args.workload == 4 and args.with_backward):
        fb_run, fwd_ts, own_run = run, ts, True
    elif args.workload == 1 and not args.no_fwd_bwd:
        del run, ts
        fb_run = Run(3, threads)
        fwd_ts = time_search(fb_run, args)
    if fb_run is not None:
        log("timing forward + hinge loss + backward")
        fb = fwd_bwd(args, fb_run)
        fwd_rate = fb_run.n_prec / (fwd_ts["step_ms"] / 1e3)
        fb_rate = fb_run.n_prec / (fb["step_ms"] / 1e3)
        log(f"fwd+bwd step {fb['step_ms']:.3f} ms (forward {fb['fwd_ms']:.3f} ms, backward "
            f"{fb['bwd_ms']:.3f} ms on the device), {fb['stats']}")
        leg = {
            "workload": workload_text(
                "forward + hinge loss + backward (d/dsigma per window, d/dintensity per "
                "fragment) on the uneven batched selection", fb_run.spec, fb_run.gen),
            "fwd_bwd_precursors_per_s": round(fb_rate, 1),
            "ms_per_fwd_bwd_step": round(fb["step_ms"], 4),
            "forward_only_precursors_per_s": round(fwd_rate, 1),
            "forward_only_ms_per_step": round(fwd_ts["step_ms"], 4),
            "fwd_bwd_over_forward": round(fb_rate / fwd_rate, 3),
            "diff_forward_device_ms": round(fb["fwd_ms"], 4),
            "diff_backward_device_ms": round(fb["bwd_ms"], 4),
            "hinge_margin": args.margin,
            **fb["stats"],
        }
        if not args.no_cpu_baseline:
            leg["gradient_check"] = grad_check(fb_run, fb["state"], threads)
            log(f"gradient check: {leg['gradient_check']}")
        if own_run:
            result["value"] = round(fb_rate, 1)
            result["ms_per_step"] = round(fb["step_ms"], 4)
        result["config"]["fwd_bwd"] = leg
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
