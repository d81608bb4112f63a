#!/usr/bin/env python3
"""Benchmark: GCUPS of the batched persistent path vs the plain per-batch launch.

Workload (default): the affine-gap DNA trimmer configuration — 20,000 reads,
lengths lognormal(median 250, sigma 0.5) clipped 50-1000, uniform ACGT with
1% N, against 64 adapters of 20-80 nt plus their 64 reverse complements;
match +1, mismatch -3, gap open -5, extension -1, FP32.  One step = score +
traceback gate + direction recompute + start selection + walk over all
2,560,000 pairs, results copied back; sequences and the work list are
resident on the device before the timed region (pw_prepare).

Prints one JSON line.  Also runs, untimed for the headline value:
  * the plain per-batch launch (every pair padded to the batch's longest
    read and adapter, static schedule) for the batched/plain ratio;
  * the CPU baseline: the oracle's restatement of the reference OpenCL-CPU
    schedule (C + OpenMP) on a bounded sample of the same workload, whose hits
    are also compared with the GPU's hits on the same pairs.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCUPS batched-persistent vs plain launch, bit-exact hits; p50/p99 latency"
# Published FP32 peak of MI355X: 157.3 TFLOP/s, an FMA counted as 2 ops.  The score pass issues FP32 adds and
# maxima (one op each; v_max3 does two), never an FMA.  Its instruction-issue ceiling, 78.6 T lane-instructions/s
# (256 CUs x 128 FP32 lanes/clk x 2.4 GHz, the rate behind the 157.3 figure), is reported beside it as issue_frac:
# VALU instructions issued per second (counted per cell by rocprofv3) over that ceiling.
FP32_PEAK_TFLOPS = 157.3
VALU_ISSUE_PEAK_TOPS = 157.3 / 2
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "score_config2_traffic.txt")
OPS_PER_CELL = {"linear": 7, "affine": 17}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_job(cfg):
    from pypaswas_amd import workloads as W
    from pypaswas_amd.Core.SmithWatermanHIP import job_arrays
    from pypaswas_amd.Core.Scores import DnaRnaScore, Blosum62Score
    import logging
    gen = {"config2": W.config2, "config1": W.config1, "config4": W.config4}[cfg]
    queries, targets, overrides = gen()
    settings = W.settings_dict(overrides)
    logger = logging.getLogger("bench")
    score = (Blosum62Score if settings.matrix_name.upper() == "BLOSUM62" else DnaRnaScore)(logger, settings)
    dna = score.score_type == "DNA_RNA"
    q, t = W.reference_order(queries, targets, dna)
    return q, t, settings, score


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="config2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-batch launch")
    ap.add_argument("--cpu-reads", type=int, default=1600, help="reads in the CPU-baseline sample")
    ap.add_argument("--cpu-one-core-reads", type=int, default=60, help="reads of the one-thread CPU run")
    ap.add_argument("--cpu-device-gib", type=float, default=8.0,
                    help="MAX_MEM_ALLOC_SIZE of the CPU device the reference's batching is sized for")
    ap.add_argument("--no-protein", action="store_true", help="skip the uneven-length protein comparison")
    ap.add_argument("--protein", default="2000x20000",
                    help="queries x targets of the uneven-length protein comparison (default: the full set)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving run")
    ap.add_argument("--no-engine", action="store_true", help="skip the one-shot engine call and Aligner.process")
    ap.add_argument("--no-autotune", action="store_true", help="use the built-in shape cost model")
    ap.add_argument("--serving-jobs", type=int, default=3000, help="requests in the serving run")
    ap.add_argument("--serving-qps", type=float, default=2000.0, help="Poisson arrival rate of the serving run")
    ap.add_argument("--serving-sweep", default="8000,20000",
                    help="further Poisson rates of the serving sweep (comma-separated; saturation is added)")
    ap.add_argument("--serving-lanes", type=int, default=4, help="stream lanes of the server")
    ap.add_argument("--serving-arena-mib", type=int, default=16384, help="direction arena of the server")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving measurement")
    ap.add_argument("--protein-only", action="store_true", help="run only the protein comparison")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device three times (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build targets a single MI355X and has no multi-GPU form; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from pypaswas_amd._native import Device
    from pypaswas_amd.Core.SmithWatermanHIP import job_arrays

    if args.protein_only:
        nq, nt = (int(v) for v in args.protein.split("x"))
        print(json.dumps(protein_comparison(Device(0), nq, nt, not args.no_autotune)), flush=True)
        return

    if args.serving_only:
        dev = Device(0)
        runs = [serving_run(dev, args.serving_jobs, qps, args.serving_lanes, args.serving_arena_mib)
                for qps in [args.serving_qps] + [float(v) for v in args.serving_sweep.split(",") if v] + [None]]
        print(json.dumps(runs), flush=True)
        return

    t0 = time.time()
    q, t, settings, score = build_job(args.workload)
    log("workload built: %d queries x %d targets in %.1fs" % (len(q), len(t), time.time() - t0))
    qp, tp, pq, pt, mss, same = job_arrays(q, t, score.highest_score, float(settings.filter_factor))
    if args.workload == "config4":
        # 64 one-query jobs (query i vs target i and its reverse complement), run as one batch
        from pypaswas_amd.workloads import config4_pair_mask
        keep = config4_pair_mask(q, t, pq, pt)
        pq, pt, mss, same = pq[keep], pt[keep], mss[keep], same[keep]
    dev = Device(0)
    dev.set_scoring(score.to_float32(), ord(score.char_offset), score.gap_score,
                    score.gap_extension if score.gap_extension is not None else 0.0, score.lower_limit_score,
                    score.minimum_score)
    minimum = float(settings.minimum_score)
    kind = "affine" if score.gap_extension is not None else "linear"
    tuning = None
    if not args.no_autotune:
        # measured kernel shapes for this gap model, before the plans are built (untimed)
        sec = dev.autotune()
        tab = dev.tuning(kind == "affine")
        best = tab["rates"].argmax(axis=2)
        tuning = {"seconds": round(sec, 2), "group_chunk_width": tab["group_nb_long"],
                  "group_rates_cells_per_us": {w: round(float(r), 1) for w, r in
                                               zip((32, 48, 64, 80, 96), tab["group_rates"])},
                  "residues_per_lane": [[int([1, 2, 4, 8][k]) for k in row] for row in best],
                  "buckets": "rows: lane-side length <=64,128,256,512,1024,2048,more; "
                             "columns: swept length <=32,64,128,256,512,1024,more"}
        log("autotune: %.2fs, multi-chunk width %d" % (sec, tab["group_nb_long"]))

    t0 = time.time()
    plan = dev.prepare(qp, tp, pq, pt, mss, same, mode=1)
    log("batched plan prepared in %.1fs" % (time.time() - t0))
    plan_tuning = None
    if not args.no_autotune:
        plan_tuning = dev.tune_plan(plan)
        log("plan tuned in %.1fs: %s" % (plan_tuning["seconds"], plan_tuning))
    for i in range(args.warmup):
        r = dev.execute(plan, minimum)
        log("warmup %d: %.1f ms (score %.1f ms, dirs %.1f, tb %.1f, walk %.1f), %d traceback pairs, %d hits" %
            (i, r.ms_total, r.ms_score, r.ms_directions, r.ms_traceback, r.ms_walk, r.n_traceback_pairs,
             int(r.hits["shown"].sum()) if len(r.hits) else 0))
    score_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = dev.execute(plan, minimum)
        score_ms.append(r.ms_score)
    elapsed = time.perf_counter() - t0
    cells = r.cells
    gcups = cells * args.steps / elapsed / 1e9
    ms_step = elapsed * 1000 / args.steps
    log("batched: %.1f ms/step, %.1f GCUPS, score kernel %.2f ms" % (ms_step, gcups, np.mean(score_ms)))
    last = r
    # the same plan through its captured HIP graph (pw_execute_graph): one
    # capture (untimed, with a warmup replay), then as many timed replays
    r = dev.execute(plan, minimum, graph=True)
    r = dev.execute(plan, minimum, graph=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = dev.execute(plan, minimum, graph=True)
    el_g = time.perf_counter() - t0
    graph = {"ms_per_step": round(el_g * 1000 / args.steps, 3), "gcups": round(cells * args.steps / el_g / 1e9, 2),
             "identical_results": bool(np.array_equal(r.maxima, last.maxima) and np.array_equal(r.hits, last.hits)
                                       and r.ops == last.ops)}
    log("graph replays: %.2f ms/step, %.1f GCUPS, identical=%s" % (graph["ms_per_step"], graph["gcups"],
                                                                   graph["identical_results"]))
    del r
    plan.release()

    plain = None
    if not args.no_plain:
        plan0 = dev.prepare(qp, tp, pq, pt, mss, same, mode=0)
        r0 = dev.execute(plan0, minimum)
        t1 = time.perf_counter()
        r0 = dev.execute(plan0, minimum)
        el0 = time.perf_counter() - t1
        plan0.release()
        same_hits = (np.array_equal(r0.maxima, last.maxima) and np.array_equal(r0.hits, last.hits)
                     and r0.ops == last.ops)
        plain = {"gcups": cells / el0 / 1e9, "ms_per_step": el0 * 1000, "score_kernel_ms": r0.ms_score,
                 "identical_results": bool(same_hits)}
        log("plain: %.1f ms/step, %.2f GCUPS, identical=%s" % (el0 * 1000, plain["gcups"], same_hits))

    engine = None
    if not args.no_engine:
        engine = engine_window(dev, q, t, settings, score, qp, tp, pq, pt, mss, same, minimum, cells, ms_step)

    avg_score_s = float(np.mean(score_ms)) / 1000.0
    achieved = OPS_PER_CELL[kind] * cells / avg_score_s / 1e12
    traffic = None
    issue_frac = None
    if args.workload == "config2" and os.path.exists(TRAFFIC_FILE):
        # HBM bytes per score pass from the committed FETCH_SIZE / WRITE_SIZE runs (see the file for the command),
        # and the VALU instructions per cell counted there (SQ_INSTS_VALU x 64 lanes / cells)
        prof = json.load(open(TRAFFIC_FILE))
        traffic = prof.get("bytes_per_pass")
        if prof.get("valu_instructions") and prof.get("cells"):
            instr_per_cell = prof["valu_instructions"] * 64 / prof["cells"]
            issue_frac = round(instr_per_cell * cells / avg_score_s / 1e12 / VALU_ISSUE_PEAK_TOPS, 4)
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "issue_peak": round(VALU_ISSUE_PEAK_TOPS, 2), "issue_frac": issue_frac,
                "kernel": "score pass (group_kernel<%s,NB> + score_kernel<%s>)" % (
                    ("true",) * 2 if kind == "affine" else ("false",) * 2),
                "ops_per_cell": OPS_PER_CELL[kind], "cells_per_launch": cells,
                "kernel_ms": round(float(np.mean(score_ms)), 3)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(q, t, settings, score, dev, args.cpu_reads, args.cpu_one_core_reads,
                           args.cpu_device_gib)

    serving = None
    sweep = None
    if not args.no_serving:
        serving = serving_run(dev, args.serving_jobs, args.serving_qps, args.serving_lanes, args.serving_arena_mib)
        # offered load from the headline rate up to saturation (every request at once)
        sweep = []
        for qps in [float(v) for v in args.serving_sweep.split(",") if v] + [None]:
            r = serving_run(dev, args.serving_jobs, qps, args.serving_lanes, args.serving_arena_mib)
            sweep.append({k: r[k] for k in r if k not in ("cells", "hits")})

    protein = None
    if not args.no_protein:  # last: it changes the device's scoring constants
        nq, nt = (int(v) for v in args.protein.split("x"))
        protein = protein_comparison(dev, nq, nt, not args.no_autotune)

    out = {
        "metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "affine-gap DNA trimmer: 20,000 reads lognormal(median 250, sigma 0.5) clipped "
                               "50-1000 with 1% N x 64 adapters 20-80 nt + 64 reverse complements; +1/-3, "
                               "gap -5/-1; batched persistent kernel" if args.workload == "config2" else
                               args.workload,
                   "pairs": int(len(pq)), "cells": cells, "mode": "batched-persistent"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_launch": plain,
        "graph_replay": graph,
        "end_to_end": engine,
        "batched_over_plain": round(gcups / plain["gcups"], 2) if plain else None,
        "protein_uneven": protein,
        "latency_ms": ({"p50": serving["p50_ms"], "p99": serving["p99_ms"], "workload": "serving mix"}
                       if serving else None),
        "serving": serving,
        "serving_sweep": sweep,
        "traceback_pairs": last.n_traceback_pairs,
        "hits": int(last.hits["shown"].sum()) if len(last.hits) else 0,
        "autotune": tuning,
        "plan_tuning": plan_tuning,
        "stage_ms": {"score": round(last.ms_score, 3), "directions": round(last.ms_directions, 3),
                     "traceback": round(last.ms_traceback, 3), "walk": round(last.ms_walk, 3),
                     "host_readback": round(last.ms_readback, 3), "host_tail": round(last.ms_tail, 3),
                     "host_assemble": round(last.ms_assemble, 3), "call_wall": round(last.ms_total, 3)},
    }
    print(json.dumps(out), flush=True)


def engine_window(dev, q, t, settings, score, qp, tp, pq, pt, mss, same, minimum, cells, resident_ms):
    """The reference's timed window (SmithWaterman.py:525-542: copy, score,
    traceback, walk, hits) for the drop-in, on the same workload:
      * engine_call: one pw_align_pairs call on the packed inputs (upload,
        planning, score, gate, traceback, walk, records back), median of 3;
      * aligner_process: Aligner.process on the sequence records with
        --framework HIP (pair arrays, the engine call, Hit objects, filter and
        HitList), one call after an untimed one."""
    import logging
    from pypaswas_amd.Core.Programs import Aligner
    from pypaswas_amd.Core.SWSeqRecord import SWSeqRecord
    runs = []
    for _ in range(4):
        t0 = time.perf_counter()
        r = dev.align_pairs(qp, tp, pq, pt, mss, same, minimum, mode=1)
        runs.append((time.perf_counter() - t0, r.ms_plan, r.ms_upload, r.ms_total))
    runs = sorted(runs[1:])  # the first call sizes the context's work areas
    wall, ms_plan, ms_upload, ms_lib = runs[len(runs) // 2]
    out = {"engine_call_ms": round(wall * 1e3, 3), "engine_call_gcups": round(cells / wall / 1e9, 2),
           "ms_plan": round(ms_plan, 3), "ms_upload": round(ms_upload, 3),
           "ms_execute": round(ms_lib - ms_plan - ms_upload, 3),
           "over_resident": round(wall * 1e3 / resident_ms, 2)}
    log("engine call: %.1f ms (plan %.1f, upload %.1f), %.1f GCUPS, %.1fx the resident execution" %
        (wall * 1e3, ms_plan, ms_upload, cells / wall / 1e9, wall * 1e3 / resident_ms))
    st = settings
    st.framework = "HIP"
    recs_q = [SWSeqRecord(s, i) for i, s in q]
    recs_t = [SWSeqRecord(s, i) for i, s in t]
    aligner = Aligner(logging.getLogger("bench"), score, st)
    aligner.process(recs_q, recs_t, None)
    aligner.hitlist.hits.clear()
    aligner.hitlist.real_hits.clear()
    t0 = time.perf_counter()
    hl = aligner.process(recs_q, recs_t, None)
    el = time.perf_counter() - t0
    out.update({"aligner_process_ms": round(el * 1e3, 1), "aligner_process_gcups": round(cells / el / 1e9, 2),
                "aligner_hits": len(hl.real_hits)})
    log("Aligner.process: %.1f ms, %.1f GCUPS, %d hits" % (el * 1e3, cells / el / 1e9, len(hl.real_hits)))
    aligner.smith_waterman.dev.close()
    return out


_SERVING_JOBS = {}


def serving_run(dev, n_jobs, qps, lanes, arena_mib, seed=5):
    """Serving mix through the continuous-batching server at a fixed Poisson
    rate (qps None: every request at time 0, the saturation point): 70% reads
    (100-300 nt) x 8-adapter panels (+RC), 25% protein pairs (100-2,000,
    BLOSUM62 -11/-1), 5% long DNA pairs (5,000-20,000 nt, +RC).  Latency =
    completion poll - arrival.  A short untimed burst first sizes the lanes'
    work areas."""
    from pypaswas_amd import workloads as W
    from pypaswas_amd.serving import Job, Server, poisson_arrivals, replay
    t0 = time.time()
    schemes = W.serving_schemes()
    key = (seed, n_jobs)
    if key not in _SERVING_JOBS:  # the same requests for every rate of a sweep
        raw = W.serving_jobs(seed, n_jobs)
        jobs = [Job.all_vs_all(W.SERVING_KINDS.index(k), schemes[W.SERVING_KINDS.index(k)], q, t)
                for k, q, t in raw]
        warm = [Job.all_vs_all(W.SERVING_KINDS.index(k), schemes[W.SERVING_KINDS.index(k)], q, t)
                for k, q, t in W.serving_jobs(seed + 1000, 400)]
        _SERVING_JOBS[key] = (raw, jobs, warm)
        log("serving: %d jobs built in %.1fs" % (n_jobs, time.time() - t0))
    raw, jobs, warm = _SERVING_JOBS[key]
    srv = Server(dev, lanes=lanes, direction_bytes=int(arena_mib) << 20)
    try:
        for i, sc in enumerate(schemes):
            srv.set_scheme(i, sc)
        replay(srv, warm)
        st0 = srv.stats()
        arrivals = poisson_arrivals(n_jobs, qps, seed) if qps else None
        watch = {}
        results, lat, wall = replay(srv, jobs, arrivals, watch)
        st = srv.stats()
    finally:
        srv.close()
    cells = float(sum(j.cells for j in jobs))
    kinds = np.array([k for k, _, _ in raw])
    out = {"jobs": n_jobs, "qps": qps if qps else "all at once", "lanes": lanes, "arena_mib": arena_mib,
           "wall_s": round(wall, 3), "jobs_per_s": round(n_jobs / wall, 1),
           "peak_queued": watch.get("peak_queued", 0), "peak_running": watch.get("peak_running", 0),
           "gcups": round(cells / wall / 1e9, 2), "cells": cells,
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "max_ms": round(float(lat.max()) * 1e3, 3),
           "batches": st["batches"] - st0["batches"],
           "jobs_per_batch": round(n_jobs / max(1, st["batches"] - st0["batches"]), 2),
           "peak_arena_mib": round(st["peak_arena_bytes"] / 2 ** 20, 1),
           "hits": int(sum(int(r.hits["shown"].sum()) if len(r.hits) else 0 for r in results)),
           "host_ms_launch": round((st["host_us_launch"] - st0["host_us_launch"]) / 1e3, 1),
           "host_ms_finish": round((st["host_us_finish"] - st0["host_us_finish"]) / 1e3, 1),
           "device_ms_batches": round((st["device_us"] - st0["device_us"]) / 1e3, 1),
           "reallocs": st["reallocs"] - st0["reallocs"]}
    for k in W.SERVING_KINDS:
        m = kinds == k
        if m.any():
            out["p50_ms_" + k] = round(float(np.percentile(lat[m], 50)) * 1e3, 3)
            out["p99_ms_" + k] = round(float(np.percentile(lat[m], 99)) * 1e3, 3)
    log("serving: %d jobs at %s/s in %.2fs, %.1f GCUPS, p50 %.2f ms, p99 %.2f ms, %d batches, %.2f jobs per "
        "batch, peak queue %d" % (n_jobs, qps if qps else "inf", wall, out["gcups"], out["p50_ms"], out["p99_ms"],
                                  out["batches"], out["jobs_per_batch"], out["peak_queued"]))
    return out


def copy_calibration(mib):
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copied %d MiB x 3" % mib)


def protein_comparison(dev, nq, nt, tune=True):
    """Uneven-length protein all-vs-all sample (BLOSUM62, -11/-1; lengths lognormal
    median 350, sigma 0.6, clipped 30-5000): batched persistent vs plain launch,
    same pairs, results compared."""
    import logging
    from pypaswas_amd import workloads as W
    from pypaswas_amd.Core.Scores import Blosum62Score
    from pypaswas_amd.Core.SmithWatermanHIP import job_arrays
    queries, targets, overrides = W.config3(n_queries=nq, n_targets=nt)
    settings = W.settings_dict(overrides)
    score = Blosum62Score(logging.getLogger("bench"), settings)
    q, t = W.reference_order(queries, targets, False)
    qp, tp, pq, pt, mss, same = job_arrays(q, t, score.highest_score, float(settings.filter_factor))
    dev.set_scoring(score.to_float32(), ord(score.char_offset), score.gap_score, score.gap_extension,
                    score.lower_limit_score, score.minimum_score)
    minimum = float(settings.minimum_score)
    out = {"pairs": int(len(pq)), "workload": "%d x %d proteins, lognormal(350, 0.6) clipped 30-5000, BLOSUM62 "
                                              "-11/-1, 100 families at ~40%% pairwise identity" % (nq, nt)}
    res = {}
    # at the full 2,000 x 20,000 size (~7e12 cells) one timed execution per mode
    # after two untimed ones (the first sizes the work areas, the second still
    # allocates the pinned result arrays and grows the areas the first call's
    # later rounds asked for), and no plan tuning (it times the scoring pass
    # 4 x 12 times)
    big = float(sum(len(s) for _, s in q)) * float(sum(len(s) for _, s in t)) > 1e12
    warm, reps = (2, 1) if big else (2, 3)
    for mode, name in ((1, "batched"), (0, "plain")):
        t0 = time.perf_counter()
        plan = dev.prepare(qp, tp, pq, pt, mss, same, mode=mode)
        out[name + "_prepare_s"] = round(time.perf_counter() - t0, 3)
        if mode == 1 and tune and not big:
            out["plan_tuning"] = dev.tune_plan(plan)
        for k in range(warm):
            t0 = time.perf_counter()
            dev.execute(plan, minimum)
            if k == 0:
                # the first execution sizes the work areas (a single-call user's time)
                out[name + "_first_call_s"] = round(time.perf_counter() - t0, 3)
            log("protein %s: warm execution done" % name)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = dev.execute(plan, minimum)
            times.append(time.perf_counter() - t0)
        el = float(np.median(times))
        plan.release()
        res[name] = r
        out[name + "_gcups"] = round(r.cells / el / 1e9, 2)
        out[name + "_score_kernel_ms"] = round(r.ms_score, 3)
        out[name + "_stage_ms"] = {"score": round(r.ms_score, 3), "directions": round(r.ms_directions, 3),
                                   "traceback": round(r.ms_traceback, 3), "walk": round(r.ms_walk, 3),
                                   "host_readback": round(r.ms_readback, 3), "host_tail": round(r.ms_tail, 3),
                                   "host_assemble": round(r.ms_assemble, 3), "call_wall": round(r.ms_total, 3)}
        out["traceback_pairs"] = r.n_traceback_pairs
        log("protein %s: %.1f ms, %.2f GCUPS (score kernels %.2f ms)" % (name, el * 1000, r.cells / el / 1e9,
                                                                        r.ms_score))
    out["cells"] = res["batched"].cells
    out["timed_executions"] = reps
    out["untimed_executions"] = warm
    out["batched_over_plain"] = round(out["batched_gcups"] / out["plain_gcups"], 2)
    out["identical_results"] = bool(np.array_equal(res["batched"].maxima, res["plain"].maxima)
                                    and np.array_equal(res["batched"].hits, res["plain"].hits)
                                    and res["batched"].ops == res["plain"].ops)
    out["hits"] = int(res["batched"].hits["shown"].sum()) if len(res["batched"].hits) else 0
    return out


# The reference's own cpu-OpenCL run on its benchmark (8 ankyrin queries of 4,624 residues x the 340-target file of
# 281,798 residues, BLOSUM80, i7-2600K, 8 threads): data/desktop/timings/canisLupusAnkyrinPRED.fasta_340.fasta.CPU.log.txt
# logs "Processing 8- vs 340-sequences" and 45 "Time spent on Smith-Waterman" lines summing to 106.42 s.  Its published
# 0.82 GCUPS divides data/cellUpdates.csv:35 (80,113,762,410 cells) by that time, but calcCellUpdates.py multiplies the
# target residues by the 284,295 residues of the 346-sequence canisLupusAnkyrinPRED.fasta, not by the 4,624 query
# residues the run aligned: the cells of that run are 4,624 x 281,798 = 1.303e9.
REFERENCE_CPU_LOGGED = {"cells": 4624 * 281798, "seconds": 106.42, "gcups": round(4624 * 281798 / 106.42 / 1e9, 4),
                        "published_gcups": 0.82, "threads": 8, "cpu": "i7-2600K (4C/8T, 4.5 GHz)"}


def reference_batches(lengths, target_len, n_targets, mem_bytes, fill, affine):
    """Read batches of SmithWaterman.align_sequences (SmithWaterman.py:441-486,
    recompile=T): from read `index`, padded length = that read's length rounded up
    to 8 (reads are sorted longest first), and as many reads as
    _get_max_number_sequences (:202-221) lets the score matrices of a batch fill
    mem_bytes * fill: (len x tlen x (1 + 4 or 12 B) + len x tlen x 4 / 64 B) x targets
    per read.  Returns (first read, count) per batch."""
    basic = 1 + (3 if affine else 1) * 4
    tlen = -(-target_len // 8) * 8
    out, index = [], 0
    while index < len(lengths):
        length = -(-lengths[index] // 8) * 8
        per_read = (length * tlen * basic + length * tlen * 4.0 / 64) * n_targets
        count = max(1, int(mem_bytes * fill // per_read))
        count = min(count, len(lengths) - index)
        out.append((index, count))
        index += count
    return out


def cpu_identity():
    """nproc (the CPUs this process may run on) and the CPU model string."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return len(os.sched_getaffinity(0)), model


def cpu_baseline(q, t, settings, score, dev, n_reads, one_core_reads, device_gib):
    """The oracle's restatement of the reference OpenCL-CPU schedule (full FP32
    matrices per pair, 8x8 blocks, one parallel sweep per block anti-diagonal,
    the reverse traceback sweep and the host walk) on a stride sample of the
    reads x all targets, batched as the reference batches for a CPU device whose
    MAX_MEM_ALLOC_SIZE is device_gib (SmithWatermanOcl.py:131-134), on every CPU
    this process may use; then the first one_core_reads reads of the sample as
    one batch on one thread.  The hits of the full sample are compared with the
    GPU's on the same pairs."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from pypaswas_amd.Core.SmithWatermanHIP import job_arrays
    p = O.Params(matrix=score.matrix, char_offset=ord(score.char_offset), gap=score.gap_score,
                 gap_ext=score.gap_extension, lower_limit=score.lower_limit_score, minimum=score.minimum_score,
                 filter_factor=float(settings.filter_factor), highest=float(score.highest_score))
    stride = max(1, len(q) // n_reads)
    sample = q[::stride][:n_reads]
    nproc, model = cpu_identity()
    omp_env = os.environ.get("OMP_NUM_THREADS")
    threads = nproc
    targets = [s for _, s in t]
    tids = [i for i, _ in t]
    tcells = float(sum(len(s) for s in targets))
    batches = reference_batches([len(s) for _, s in sample], max(len(s) for s in targets), len(targets),
                                device_gib * 2.0 ** 30, float(settings.maximum_memory_usage),
                                score.gap_extension is not None)

    def run(reads, plan, nthreads, collect):
        hits = {}
        t0 = time.perf_counter()
        for b, n in plan:
            chunk = reads[b:b + n]
            qs = [s for _, s in chunk]
            br = O.sched_batch(qs, targets, p, nthreads=nthreads)
            ws = O.walk_c(br.directions, br.starting_points, qs, targets, [i for i, _ in chunk], tids, p.minimum)
            del br
            if collect:
                for w in ws:
                    hits.setdefault((b + w.qi, w.ti), []).append(
                        (bool(w.shown), w.s_start, w.s_end, w.t_start, w.t_end, w.matches, w.mismatches, w.score))
        return time.perf_counter() - t0, hits

    cells = float(sum(len(s) for _, s in sample)) * tcells
    el, o_hits = run(sample, batches, threads, True)
    one = sample[:one_core_reads]
    cells1 = float(sum(len(s) for _, s in one)) * tcells
    el1, _ = run(one, [(0, len(one))], 1, False)
    # the same sample on the GPU
    qp, tp, pq, pt, mss, same = job_arrays(sample, t, score.highest_score, float(settings.filter_factor))
    r = dev.align_pairs(qp, tp, pq, pt, mss, same, float(settings.minimum_score), mode=1)
    g_hits = {}
    for h in r.hits.tolist():
        rec = dict(zip(r.hits.dtype.names, h))
        g_hits.setdefault((rec["query"], rec["target"]), []).append(
            (bool(rec["shown"]), rec["s_start"], rec["s_end"], rec["t_start"], rec["t_end"], rec["matches"],
             rec["mismatches"], float(rec["score"])))
    parity = "bit-exact" if o_hits == g_hits else "MISMATCH"
    log("cpu baseline: %.2e cells in %.1fs on %d threads (%d batches) = %.3f GCUPS; 1 thread: %.2e cells in %.1fs "
        "= %.4f GCUPS; %s; GPU parity on the sample: %s (%d walks)"
        % (cells, el, threads, len(batches), cells / el / 1e9, cells1, el1, cells1 / el1 / 1e9, model, parity,
           sum(len(v) for v in o_hits.values())))
    return {"value": round(cells / el / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": "%d reads (every %d-th in reader order) x all %d targets of the same workload, %.3g cells, in "
                      "%d batches sized as the reference sizes them for a %g GiB CPU device at "
                      "maximum_memory_usage %s; oracle restatement of the reference OpenCL-CPU schedule (8x8 "
                      "blocks, 2x2 work-items of 4x4 cells, full FP32 matrices) + host walk, OpenMP"
                      % (len(sample), stride, len(t), cells, len(batches), device_gib,
                         settings.maximum_memory_usage),
            "seconds": round(el, 2), "nproc": nproc, "cpu_model": model, "omp_num_threads_env": omp_env,
            "one_core": {"value": round(cells1 / el1 / 1e9, 4), "cells": cells1, "seconds": round(el1, 2),
                         "sample": "the first %d reads of the sample x all targets, one batch" % len(one)},
            "gpu_parity_on_sample": parity,
            "reference_logged": REFERENCE_CPU_LOGGED}


if __name__ == "__main__":
    main()
