#!/usr/bin/env python
"""Placement throughput on one MI355X.

Workload: the 16S-scale shape -- GTR+Gamma4, 2,000-taxon Yule tree (3,997
edges, branch lengths ~Exp(0.03)), 1,600 columns, 100,000 queries with
uniform 100-500-column masked lengths and 2% substitutions -- synthesised
from a fixed seed.  One step = one full pass of pplacer's per-query placement
(stage-1 ranking of every (query, edge) pair, baseball + Brent, keep, refine,
like-weight ratios) over the whole batch, with the queries and the tree's
GLVs already resident in HBM.

Prints one JSON line.  value = query x edge evaluations per second over the
whole placement pass (Q * E / step time).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (= FP64 matrix) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed) and FP32 matrix (dense MFMA) peak


WORKLOAD_TEXT = {
    "gtr2000": "16S-scale GTR+G4: 2000-taxon Yule tree (3997 edges), 1600 columns, %d queries, "
               "masked length 100-500, 2%% substitutions",
    "prot500": "protein 20-state empirical matrix +G4: 500-taxon Yule tree (997 edges), 400 columns, "
               "%d queries, masked length 60-400, 2%% substitutions",
    "gtr10000": "GTR+G4 stream shape: 10000-taxon Yule tree (19997 edges), 1500 columns, %d queries, "
                "masked length 80-600",
    "gtr50": "GTR+G4: 50-taxon Yule tree (97 edges), 1000 columns, %d queries, fragments 150-300",
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def pitch_flops(stats, R=4, S=4):
    """Algorithmic flops of the baseball kernel, priced with the reference's
    per-probe work (U2): a pendant probe evolves the pendant GLV (2*R*S*S per
    masked site) and runs log_like3 (4*R*S); a distal probe evolves the
    proximal and distal GLVs (2 * 2*R*S*S) and runs log_like3."""
    return ((2 * R * S * S + 4 * R * S) * stats.n_site_probes_pend_pitch
            + (4 * R * S * S + 4 * R * S) * stats.n_site_probes_dist_pitch)


def mfma_flops(stats, R=4, S=20):
    """Flops the FP32 20-state engine issues on the matrix cores: per masked
    site of a distal probe, the proximal and distal evolves of every rate
    (2 * R*S*S multiply-adds).  Pendant probes use no MFMA."""
    return 4 * R * S * S * stats.n_site_probes_dist_pitch


def _per(x, ms):
    """x per second over ms milliseconds (None without a time)."""
    return x / (ms * 1e-3) if ms > 0 else None


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size)."""
    for fn in ("pmc_traffic.txt", "pmc_mfma_prot500_20k.txt"):
        path = os.path.join(ROOT, "profiles", fn)
        if not os.path.exists(path):
            continue
        for line in open(path):
            f = line.split()
            if len(f) >= 2 and not line.startswith("#") and f[0] == kernel:
                return float(f[1])
    return None


def kernel_work(stats, R=4, S=4, w=8):
    """Algorithmic work of the stage-1 and stage-2 kernels.
    Stage 1 per (query, edge, masked site): the reference's masked_logdot
    does R*S products and sums (2*R*S flops) on a supernode row (R*S*8 bytes
    + exponent); the table form gathers one w-byte log-util entry and the
    4-byte supernode exponent, and does one FP64 add.
    Stage 2 per probe and masked site: pendant probe 2*R*S flops (util);
    distal probe adds two R*S*S evolves (4*R*S*S) and the pi*x*y products
    (2*R*S)."""
    s1_ref_flops = 2 * R * S * stats.n_site_pairs
    s1_gather_bytes = (w + 4) * stats.n_site_pairs
    s2_flops = (2 * R * S) * stats.n_site_probes_pend + (4 * R * S * S + 4 * R * S) * stats.n_site_probes_dist
    return s1_ref_flops, s1_gather_bytes, s2_flops


def copy_calibration(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's blit kernel): a
    kernel of known traffic for scaling FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)):
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) or hip.hipDeviceSynchronize():
        raise RuntimeError("hipMemcpy failed")
    hip.hipFree(src)
    hip.hipFree(dst)


def cpu_baseline(wl, queries, n_sample, threads):
    """The oracle (FP64 restatement of pplacer, one query per worker like a
    pplacer -j child) on a bounded sample of the same workload."""
    from oracle import pyoracle as po
    S = wl.n_states
    ref_seqs = [s for _, s in wl.ref_aln]
    model = po.Model(wl.exch, po.emper_freq(S, ref_seqs), po.discrete_gamma(wl.n_cats, wl.alpha))
    st, mask, keep, n = po.premask(S, ref_seqs, [s for _, s in wl.queries])
    import numpy as np
    m = np.asarray(mask, dtype=bool)
    leaf = {i: np.frombuffer(s.encode(), np.uint8)[m].tobytes().decode()
            for i, s in wl.leaf_seqs().items()}
    pre = po.Precompute(model, wl.tree, leaf, n)
    pre.pull_and_supernodes()
    core = po.Core(pre)
    sample = queries[:n_sample]
    E = wl.tree.n_nodes - 1
    t0 = time.perf_counter()
    st, res, status, stats = core.place_many(sample, n_threads=threads)
    dt = time.perf_counter() - t0
    # the same oracle on one thread (pplacer -j 1), over the first eighth of the sample
    one = sample[:max(1, len(sample) // 8)]
    t1 = time.perf_counter()
    core.place_many(one, n_threads=1)
    dt1 = time.perf_counter() - t1
    return {"value": len(sample) * E / dt, "unit": "query x edge evals/s", "cores": threads,
            "kind": "port",
            "sample": "%d queries of the same workload (FP64 oracle, one query per worker as "
                      "pplacer -j N, N = %d threads of the %d cores in this process's affinity "
                      "mask, %.1f s wall, %.1f ms/query/thread)"
                      % (len(sample), threads, len(os.sched_getaffinity(0)), dt,
                         1000 * dt * threads / len(sample)),
            "queries_per_s": len(sample) / dt,
            "one_thread": {"value": len(one) * E / dt1, "queries_per_s": len(one) / dt1,
                           "sample": "%d queries, 1 thread, %.1f s" % (len(one), dt1)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=100_000)
    ap.add_argument("--workload", default="gtr2000")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--other-form", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the other precision form (reported under 'forms')")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--calc-pp", action="store_true",
                    help="-p: posterior probabilities of the kept placements (keep_at_most 20, "
                         "keep_factor 0.001, max_strikes 20)")
    ap.add_argument("--stream", action="store_true",
                    help="place the queries as a stream of uneven batches of 1k-50k (one pass per "
                         "form, the tree GLVs resident throughout)")
    ap.add_argument("--tune", action="store_true",
                    help="tune each form's launch shape (bucket streams per stage-2 phase, chunk "
                         "size) on the resident queries before its warmup (untimed)")
    ap.add_argument("--graph", action="store_true",
                    help="capture each form's placement pass into a HIP graph and replay it "
                         "(no per-phase or per-launch times are recorded then)")
    ap.add_argument("--fp32-graph", action=argparse.BooleanOptionalAction, default=True,
                    help="when the FP32 form runs beside the FP64 one (4 states), run it as the "
                         "batched form: its pass captured into a HIP graph in the warmup and "
                         "replayed in the timed steps (the FP64 form stays eager, so that its "
                         "per-launch times feed the roofline)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a known-size "
                         "kernel for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(queries would shard across GPUs with replicated tree GLVs).", file=sys.stderr)
        sys.exit(2)

    from pplacer_amd import placer, synth
    from pplacer_amd._lib import PP_FP32, PP_FP64

    if args.stream:
        run_stream(args)
        return

    log("synthesising workload %s (%d queries, seed %d)" % (args.workload, args.queries, args.seed))
    t0 = time.perf_counter()
    wl = synth.make_workload(args.workload, seed=args.seed, n_queries=args.queries)
    log("synthesised in %.1f s; building model, pre-mask and tree GLVs" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    E = wl.tree.n_nodes - 1
    N = len(queries[0])
    S = wl.n_states
    log("tree ready in %.1f s (G=%d, N=%d); uploading queries" % (time.perf_counter() - t0,
                                                                  wl.tree.n_nodes, N))
    pl.upload_queries(queries)
    prefs = placer.default_prefs(calc_pp=args.calc_pp)
    if args.graph:
        pl.set_launch_config(graph=1)
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb << 20)
        log("copy calibration: %d MiB read + %d MiB written" % (args.copy_calibration_mb,
                                                               args.copy_calibration_mb))

    # the FP32 form as the other form of a 4-state run: replayed from a graph
    fp32_graph = (args.fp32_graph and not args.graph and not args.tune and S == 4
                  and args.precision == "fp64" and not args.calc_pp)

    def run_form(prec_name):
        prec = PP_FP64 if prec_name == "fp64" else PP_FP32
        pl.set_launch_config(graph=1 if (args.graph or (fp32_graph and prec_name == "fp32")) else 0)
        if args.tune:
            from pplacer_amd import tune
            best, trials = tune.tune(pl, prefs, prec, log=log)
            log("%s launch shape tuned over %d trials: %s" % (prec_name, len(trials), best))
        for i in range(args.warmup):
            log("%s warmup %d" % (prec_name, i))
            pl.place_resident(prefs, prec)
        pl.synchronize()
        times, stats_l = [], []
        for i in range(args.steps):
            pl.synchronize()
            t0 = time.perf_counter()
            st = pl.place_resident(prefs, prec)
            pl.synchronize()
            times.append(time.perf_counter() - t0)
            stats_l.append(st)
            log("%s step %d: %.3f s (table %.1f ms, stage1 %.1f ms, pitch %.1f ms, refine %.1f ms%s)" % (
                prec_name, i, times[-1], st.t_table_ms, st.t_stage1_ms, st.t_pitch_ms, st.t_refine_ms,
                ", -p %.1f ms" % st.t_pp_ms if args.calc_pp else ""))
        return times, stats_l

    forms = {}
    order = [args.precision] + ([("fp32" if args.precision == "fp64" else "fp64")] if args.other_form else [])
    for f in order:
        times, stats_l = run_form(f)
        st = stats_l[-1]
        ms = 1000.0 * sum(times) / len(times)
        w = 8 if f == "fp64" else 4
        s1_ref_flops, s1_gather_bytes, s2_flops = kernel_work(st, S=S, w=w)
        forms[f] = {
            "ms_per_step": ms,
            "value": st.n_logdots / (ms / 1000.0),
            "queries_per_s": args.queries / (ms / 1000.0),
            "t_table_ms": st.t_table_ms, "t_stage1_ms": st.t_stage1_ms, "t_pitch_ms": st.t_pitch_ms,
            "t_refine_ms": st.t_refine_ms, "t_pp_ms": st.t_pp_ms,
            # (graph mode records no per-phase times: these read null)
            "stage1_ref_equiv_tflops": _per(s1_ref_flops / 1e12, st.t_stage1_ms),
            "stage1_gather_GBps": _per(s1_gather_bytes / 1e9, st.t_stage1_ms),
            "stage2_tflops": _per(s2_flops / 1e12, st.t_pitch_ms + st.t_refine_ms),
            "pitch_u2_tflops": _per(pitch_flops(st, S=S) / 1e12, st.t_pitch_ms),
            "pitch_u2_tflops_per_launch": _per(pitch_flops(st, S=S) / 1e12, st.t_pitch_kernel_ms),
            "pitch_mfma_tflops_per_launch": (_per(mfma_flops(st) / 1e12, st.t_pitch_kernel_ms)
                                             if S == 20 and f == "fp32" else None),
            "t_graph_ms": st.t_ml_ms if pl.launch_config()["graph"] else None,
            "pitch_launches": st.n_pitch_launches,
            "pitch_avg_launch_ms": st.t_pitch_kernel_ms / max(1, st.n_pitch_launches),
            "pitch_phase_ms_per_launch": st.t_pitch_ms / max(1, st.n_pitch_launches),
            "n_logdots": st.n_logdots, "n_like_calls": st.n_like_calls,
            "n_pitched": st.n_pitched, "n_launches": st.n_launches,
            # FP32 form: queries placed again in FP64 after the prox_bl assertion
            "n_fp64_replaced": st.n_fp64_replaced, "t_fp64_replace_ms": st.t_fp64_replace_ms,
            "launch": pl.launch_config(), "launch_tuned": bool(args.tune),
        }
    main_f = forms[args.precision]
    st_main = main_f
    # dominant kernel: the one with the largest device time
    kern = max([("stage1_kernel", st_main["t_stage1_ms"]), ("pitch_kernel", st_main["t_pitch_ms"]),
                ("refine_kernels", st_main["t_refine_ms"])], key=lambda kv: kv[1])[0]
    if args.graph:  # no phase times in graph mode
        kern = "pitch_kernel"
    peak = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
    real = "double" if args.precision == "fp64" else "float"
    if kern == "stage1_kernel":
        name = "stage1_kernel<%s,16>" % real
        roofline = {"bound": "hbm", "kernel": name,
                    "achieved": st_main["stage1_gather_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": (st_main["stage1_gather_GBps"] or 0) / HBM_PEAK_GBS,
                    "traffic": read_traffic(name),
                    "note": "gathered table bytes; most are served by L1/L2"}
    elif S == 20 and real == "float":
        # the FP32 20-state baseball kernel evolves on the matrix cores
        # (v_mfma_f32_4x4x1f32); achieved = its MFMA flops per launch
        name = "pitch_prot_kernel<float,4>"
        achieved = st_main["pitch_mfma_tflops_per_launch"]
        roofline = {"bound": "mfma", "pipe": "MFMA fp32 (state-basis evolves, 4x4x1 16-block tiles)",
                    "kernel": name, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak if achieved else None, "traffic": read_traffic(name),
                    "avg_launch_ms": st_main["pitch_avg_launch_ms"],
                    "launches": st_main["pitch_launches"],
                    "u2_achieved": st_main["pitch_u2_tflops_per_launch"],
                    "work": "MFMA flops: 4*R*S^2 per masked site of a distal probe (both evolves, every rate)"}
    else:
        # 4 states: one FP64 site or two FP32 sites per worker lane; the FP64
        # kernel runs in the 512- and 768-lane launch classes (masked lengths
        # up to 448 and beyond), named together as MAXBD = *
        # (template arguments: real, rates, launch class, sites per lane,
        # Price-CAT, FP64 site-order log sums); 20 states in FP64: the
        # eigenbasis engine
        name = ("pitch_mw_kernel<double,4,*,1,false,true>" if S == 4 and real == "double"
                else "pitch_mw_kernel<float,4,512,2,false,false>" if S == 4
                else "pitch_prot_kernel<double,4>")
        # per launch: the flops of the phase's launches over the sum of their
        # own durations (HIP events around each launch, on its stream); the
        # launches of one phase run concurrently on several streams, so the
        # phase as a whole reaches phase_achieved
        achieved = st_main["pitch_u2_tflops_per_launch"]
        # bound by the vector ALU: the 4-state contractions are 4x4, below any
        # MFMA tile, and the rows stay in registers (no HBM stream per probe)
        roofline = {"bound": "valu",
                    "pipe": ("vector ALU %s (4-state contractions do not use MFMA)" % args.precision
                             if S == 4 else "vector ALU fp64 (eigenbasis evolves)"),
                    "kernel": name, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak if achieved else None, "traffic": read_traffic(name),
                    "avg_launch_ms": st_main["pitch_avg_launch_ms"],
                    "launches": st_main["pitch_launches"],
                    "phase_achieved": st_main["pitch_u2_tflops"],
                    "phase_frac": st_main["pitch_u2_tflops"] / peak if st_main["pitch_u2_tflops"] else None,
                    "work": "reference U2 flops of the pitched probes: (2RS^2+4RS) per masked site "
                            "of a pendant probe, (4RS^2+4RS) of a distal probe"}
        if S == 4 and real == "double" and not args.tune:
            roofline["launch_note"] = ("the FP64 baseball launches run one after another on the main "
                                       "stream (launch shape pitch_streams[0] = 0), so each launch's "
                                       "duration is its own; --tune may spread them over streams "
                                       "(4 streams measured within 1% per pass at 50,000-query "
                                       "chunks, launches overlapping)")
        if kern != "pitch_kernel":
            roofline["note"] = "the refinement kernels took longer than the baseball kernel"
    cpu = None
    if not args.no_cpu:
        # pplacer -j N with N = the host cores this process may use, capped
        # at the 16 cores a one-GPU share of the host holds
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        log("CPU baseline: %d queries on %d threads" % (args.cpu_sample, threads))
        cpu = cpu_baseline(wl, queries, args.cpu_sample, threads)
    out = {
        "metric": "placements/sec (query x edge evals/s), full placement pass",
        "value": main_f["value"],
        "unit": "query x edge evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": main_f["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded simulation)",
        "config": {"workload": WORKLOAD_TEXT.get(args.workload, args.workload + ", %d queries") % args.queries,
                   "queries": args.queries, "edges": E, "columns": N,
                   "form": "plain FP64" if args.precision == "fp64" else "batched FP32",
                   "calc_pp": bool(args.calc_pp)},
        "queries_per_s": main_f["queries_per_s"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    if "fp32" in forms and "fp64" in forms:
        # the batched FP32 form against the plain FP64 form, same workload
        out["fp32_over_fp64"] = forms["fp32"]["value"] / forms["fp64"]["value"]
    if cpu:
        out["gpu_vs_cpu"] = main_f["value"] / cpu["value"]
    print(json.dumps(out), flush=True)
    pl.close()


def run_stream(args):
    """Queries arriving in uneven batches of 1,000-50,000 (seeded sizes), each
    uploaded, placed and downloaded while the tree GLVs stay resident.  The
    timed region is the sum of the batches' placement passes (inputs resident
    in HBM); uploads and downloads are reported beside it."""
    import numpy as np

    from pplacer_amd import placer, synth
    from pplacer_amd._lib import PP_FP32, PP_FP64, Placement

    log("synthesising stream %s (%d queries, seed %d)" % (args.workload, args.queries, args.seed))
    t0 = time.perf_counter()
    wl = synth.make_workload(args.workload, seed=args.seed, n_queries=args.queries,
                             batched_queries=True)
    log("synthesised in %.1f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    pl, queries, statd, rates = placer.setup_from_workload(wl)
    E = wl.tree.n_nodes - 1
    log("tree ready in %.1f s (G=%d, N=%d)" % (time.perf_counter() - t0, wl.tree.n_nodes,
                                                 len(queries[0])))
    rng = np.random.default_rng(args.seed + 7)
    sizes = []
    while sum(sizes) < len(queries):
        sizes.append(int(min(rng.integers(1000, 50001), len(queries) - sum(sizes))))
    prefs = placer.default_prefs()
    forms = {}
    for f in [args.precision] + ([("fp32" if args.precision == "fp64" else "fp64")]
                                 if args.other_form else []):
        prec = PP_FP64 if f == "fp64" else PP_FP32
        pl.upload_queries(queries[: sizes[0]])
        pl.place_resident(prefs, prec)  # warmup on the first batch
        pl.synchronize()
        t_place = t_io = 0.0
        n_logdots = n_queries = 0
        q0 = 0
        for k, n in enumerate(sizes):
            t0 = time.perf_counter()
            pl.upload_queries(queries[q0: q0 + n])
            pl.synchronize()
            t1 = time.perf_counter()
            st = pl.place_resident(prefs, prec)
            pl.synchronize()
            t2 = time.perf_counter()
            out = (Placement * (n * prefs.keep_at_most))()
            n_out = np.zeros(n, np.int32)
            status = np.zeros(n, np.int32)
            placer.check(placer.lib().pp_results_download(
                pl._ctx, out, n_out.ctypes.data_as(placer.C.POINTER(placer.C.c_int)),
                status.ctypes.data_as(placer.C.POINTER(placer.C.c_int))), "results_download")
            t3 = time.perf_counter()
            t_place += t2 - t1
            t_io += (t1 - t0) + (t3 - t2)
            n_logdots += st.n_logdots
            n_queries += n
            q0 += n
            if k % 8 == 0 or k == len(sizes) - 1:
                log("%s batch %d/%d: %d queries, %.1f ms placing" % (f, k + 1, len(sizes), n,
                                                                     1000 * (t2 - t1)))
        forms[f] = {"ms_total": 1000 * t_place, "value": n_logdots / t_place,
                    "queries_per_s": n_queries / t_place, "io_ms_total": 1000 * t_io,
                    "value_with_io": n_logdots / (t_place + t_io), "batches": len(sizes)}
    main_f = forms[args.precision]
    out = {
        "metric": "placements/sec (query x edge evals/s), full placement pass",
        "value": main_f["value"], "unit": "query x edge evals/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": main_f["ms_total"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded simulation)",
        "config": {"workload": (WORKLOAD_TEXT.get(args.workload, args.workload + ", %d queries")
                                % args.queries) + ", uneven batches of 1k-50k",
                   "queries": args.queries, "edges": E, "batches": len(sizes),
                   "form": "plain FP64" if args.precision == "fp64" else "batched FP32"},
        "forms": forms,
    }
    print(json.dumps(out), flush=True)
    pl.close()


if __name__ == "__main__":
    main()
