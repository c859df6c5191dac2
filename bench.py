"""Benchmark: rw throughput (Mnodes/s) on the 10M-AND deep random AIG, one MI355X.

Workload (one "step"): one `rw` pass (gls_rewrite) over the strashed 10M-AND
AIG, on a device-resident copy of the strashed input (the copy is part of the
step).  The input is generated, uploaded and strashed before timing.
value = ANDs entering rw / seconds per step.

Also printed in the same JSON line:
  roofline      the dominant kernel's algorithmic bytes per launch / its average
                launch time (HIP events on the context stream), vs 8 TB/s HBM;
  cpu_baseline  the CPU oracle's sequential rewriter that follows the
                reference's replacement method ("port", the ABC-proxy), built
                -O3 -march=native on the host, pinned to one core, median of 5,
                with the oracle of this package's pass beside it;
  qor           AND counts, depth, and equivalence: the library's device check
                (gls_cec, 4,096 patterns) and the oracle's host simulation;
  e2e           `strash; rw` end to end on the input AND count;
  hash          strash-table probes and probes per second (device counters).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nands N] [--cpu-sample N] [--cut-mode M]
       [--stage rw|cuts|strash] [--workload rw10m|batch|serve]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per AND per rw pass: 3*K*s_c + 64 with K = 8 cuts of s_c = 20 B
K_CUTS, CUT_BYTES = 8, 20
B_RW_PER_AND = 3 * K_CUTS * CUT_BYTES + 64
# cut enumeration per AND: two fanin cut sets read, its own written (3*K*s_c),
# plus its two fanin literals and its reference count (12 B)
B_CUT_PER_AND = 3 * K_CUTS * CUT_BYTES + 12
# HBM bytes per k_cuts_triple launch on the default 10M workload and schedule:
# FETCH_SIZE of a rocprofv3 --pmc run over all 12,913 launches, 430.7 KiB per
# launch, x 2.0 (the counter reads half of a plain 256 MiB copy's bytes) =
# 882,074 B; WRITE_SIZE, 160.0 B per node on a 200k-AND AIG of the same level
# shape (the same counter reads exactly a plain copy's bytes), x 617.3 nodes
# per launch = 98,770 B.  Reads are about 2.9x the algorithmic bytes because a
# node is recomputed in the cones of its consumers within the launch.
TRAFFIC_BYTES_PER_LAUNCH = 980_844
TRAFFIC_NOTE = ("profiled constant of the three-level cut kernel (profiles/k_cuts_triple_rw10m_traffic.md): "
                "FETCH_SIZE measured over the 10M pass; the WRITE_SIZE half (98,770 B) is extrapolated from a "
                "200k-AND AIG of the same level shape at 160.0 B per node")


def cut_kernel_name(mode):
    """The cut-enumeration kernel that a rewrite with this schedule launches."""
    import culs_amd
    mode &= ~culs_amd.GLS_RW_GRAPH_CUTS
    return {culs_amd.GLS_RW_XCD_CUTS: "k_cuts_xcd", culs_amd.GLS_RW_LEVEL_CUTS: "k_cuts_level",
            culs_amd.GLS_RW_COOP_CUTS: "k_cuts_levels", culs_amd.GLS_RW_DATAFLOW_CUTS: "k_cuts_dataflow",
            culs_amd.GLS_RW_PAIR_CUTS: "k_cuts_pair", culs_amd.GLS_RW_PLAN_CUTS: "k_cuts_plan",
            culs_amd.GLS_RW_QUAD_CUTS: "k_cuts_quad"}.get(mode, "k_cuts_triple")


def cut_launches(mode, nlev, levels=3):
    """Cut-enumeration launches per rw pass over nlev AND levels (`levels` per
    launch of the planned schedule)."""
    import culs_amd
    mode &= ~culs_amd.GLS_RW_GRAPH_CUTS
    if mode in (culs_amd.GLS_RW_XCD_CUTS, culs_amd.GLS_RW_COOP_CUTS, culs_amd.GLS_RW_DATAFLOW_CUTS):
        return 1
    per = {culs_amd.GLS_RW_LEVEL_CUTS: 1, culs_amd.GLS_RW_PAIR_CUTS: 2, culs_amd.GLS_RW_PLAN_CUTS: levels,
           culs_amd.GLS_RW_QUAD_CUTS: 4}.get(mode, 3)
    return (nlev + per - 1) // per


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nands", type=int, default=10_000_000)
    ap.add_argument("--pis", type=int, default=1024)
    ap.add_argument("--pos", type=int, default=4096)
    ap.add_argument("--window", type=int, default=50_000)
    ap.add_argument("--p", type=float, default=0.001)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="ANDs of the CPU oracle sample (same generator); 0 = the whole workload")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (torch) 5 times, to scale HBM counters")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="rw10m: skip the batched and served legs that the default run appends")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU baseline runs (median)")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--cut-mode", type=int, default=0,
                    help="cut-enumeration schedule (gls_rewrite flags): 0 three levels per launch, 2 "
                         "cooperative, 4 dataflow, 8 one XCD, 16 per level, 32 two levels per launch; +128 "
                         "launches the schedule from one captured HIP graph")
    ap.add_argument("--cut-levels", type=int, default=0,
                    help="levels per launch of the planned schedule (gls_tuning.cut_levels; 0: the default)")
    ap.add_argument("--stage", choices=["rw", "cuts", "strash"], default="rw",
                    help="cuts: time only reorder + cut enumeration (gls_enumerate_cuts); strash: only the "
                         "strash of the generated input; both for counter runs")
    ap.add_argument("--workload", choices=["rw10m", "batch", "serve"], default="rw10m",
                    help="rw10m: the headline line; batch: 4096 mixed AIGs packed into one arena vs one at a "
                         "time; serve: a Poisson stream of `b; rw; rwz; b` jobs, served vs a per-command loop")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="arrivals per second (Poisson)")
    ap.add_argument("--serve-max", type=int, default=2_000_000, help="largest AND count of a random job")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-max", type=int, default=200_000, help="largest AND count of the batch mix")
    ap.add_argument("--no-loop", action="store_true",
                    help="batch: time only the batched pass (no one-at-a-time loop, no identity check); "
                         "for counter runs")
    ap.add_argument("--no-tune", action="store_true", help="batch: time the pass with the default settings")
    ap.add_argument("--batch-check", type=int, default=256,
                    help="members compared array by array with their one-at-a-time result")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the path is a single-device kernel pipeline "
              "(sharding across GPUs is by independent AIGs, not built here); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        import torch
        x = torch.empty(a.copy_calibration << 20, dtype=torch.uint8, device="cuda").fill_(1)
        y = torch.empty_like(x)
        for _ in range(5):
            torch.bitwise_not(x, out=y)     # one elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel(), "copy_bytes_written": y.numel(), "copies": 5}))
        return

    import culs_amd

    if a.workload == "batch":
        print(json.dumps(bench_batch(a)), flush=True)
        return
    if a.workload == "serve":
        print(json.dumps(bench_serve(a)), flush=True)
        return
    ctx = culs_amd.Context(0)
    if not ctx.has_device:
        raise SystemExit("no GPU")
    if a.cut_levels:
        ctx.set_tuning(-1, cut_levels=a.cut_levels)
    cut_levels = ctx.get_tuning(1)["cut_levels"]
    t0 = time.time()
    raw = ctx.gen_geometric(a.pis, a.nands, a.pos, a.window, a.p, a.seed)
    g_in = raw.arrays() if not a.no_check else None
    log("generated %d ANDs in %.1fs" % (a.nands, time.time() - t0))
    src = raw.copy()
    src.strash()
    st_info = src.last_op()
    s = src.stats()
    n_in = s.nAnds
    log("strash: %d -> %d ANDs, depth %d, %.1f ms" % (a.nands, n_in, s.levels, st_info["total_ms"]))
    if a.stage == "strash":
        return
    # the input strash again (device time of gls_strash, the upload excluded), for `strash; rw` end to end
    st_runs = [st_info]
    for _ in range(2):
        c = raw.copy()
        c.strash()
        st_runs.append(c.last_op())
        c.free()
    strash_ms = sorted(r["total_ms"] for r in st_runs)[len(st_runs) // 2]
    strash_probes = st_runs[-1]["hash_probes"]

    def step():
        c = src.copy()
        if a.stage == "cuts":
            c.reorder()
            c.enumerate_cuts(a.cut_mode)
            return c
        c.rewrite(cut_mode=a.cut_mode)
        return c

    if a.stage == "cuts":
        for i in range(a.steps):
            t = time.perf_counter()
            step().free()
            log("cuts step %d: %.1f ms" % (i, (time.perf_counter() - t) * 1e3))
        return

    for _ in range(a.warmup):
        step().free()
    ctx_sync = culs_amd.lib()
    times, infos, out = [], [], None
    t_all = time.perf_counter()
    for i in range(a.steps):
        t = time.perf_counter()
        c = step()
        st = c.stats()          # host-side counters; the pass itself synchronises
        times.append(time.perf_counter() - t)
        infos.append(c.last_op())
        if out is not None:
            out.free()
        out = c
        log("step %d: %.1f ms -> %d ANDs" % (i, times[-1] * 1e3, st.nAnds))
    total = time.perf_counter() - t_all
    ms_per_step = 1e3 * total / a.steps
    value = n_in / (total / a.steps) / 1e6
    info = infos[-1]
    n_out = out.stats().nAnds
    depth_out = out.stats().levels

    # dominant kernel: the cut enumeration kernel (k_cuts_triple: three levels
    # per launch).  Every 64 levels a window of 8 consecutive launches is timed with
    # HIP events on the first and last dispatch (the context stream), so the
    # average is the per-launch kernel time as rocprofv3 reports it.
    ks = sum(i["kernel_samples"] for i in infos)
    kms = sum(i["kernel_sample_ms"] for i in infos)
    knodes = sum(i["kernel_sample_nodes"] for i in infos)
    nlev = max(1, info["levels_enumerated"])
    if ks > 0 and kms > 0:
        avg_launch_s = kms / 1e3 / ks
        bytes_per_launch = knodes * B_CUT_PER_AND / ks
        achieved = knodes * B_CUT_PER_AND / (kms / 1e3) / 1e9
    else:
        avg_launch_s, bytes_per_launch, achieved = 0.0, n_in * B_CUT_PER_AND / nlev, 0.0
    # the measured traffic holds for the default workload and schedule only
    default_workload = (a.cut_mode == 0 and a.nands == 10_000_000 and a.pis == 1024 and a.pos == 4096 and
                        a.window == 50_000 and a.p == 0.001 and a.seed == 3)
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6),
                "achieved_basis": "algorithmic bytes (%d B per node) / measured launch time" % B_CUT_PER_AND,
                "limiter": "instruction issue along each node's dependent chain and the launch latency of "
                           "the depth-long launch sequence, not HBM bandwidth (counter profile: "
                           "profiles/k_cuts_triple_geom200k_counters.md)",
                "traffic": TRAFFIC_BYTES_PER_LAUNCH if default_workload else None,
                "traffic_note": TRAFFIC_NOTE if default_workload else None,
                "kernel": cut_kernel_name(a.cut_mode), "launches_per_step": cut_launches(a.cut_mode, nlev, cut_levels),
                "timed_launches": ks,
                "avg_launch_us": round(avg_launch_s * 1e6, 3), "bytes_per_launch": int(bytes_per_launch),
                "bytes_per_node": B_CUT_PER_AND,
                "pass_alg_GBs": round(n_in * B_RW_PER_AND / (total / a.steps) / 1e9, 3)}

    qor = {"and_in": a.nands, "and_strashed": n_in, "and_out": n_out, "depth_in": s.levels, "depth_out": depth_out}
    if g_in is not None:
        # the library's device check (gls_cec) on 4,096 patterns, and the
        # oracle's host simulation on 128 patterns as an independent check
        t = time.time()
        ci = raw.cec(out, words=64, seed=11)
        qor["cec"] = {"equivalent": bool(ci["equivalent"]), "patterns": ci["patterns"],
                      "method": "gls_cec: device bit-parallel simulation, input vs rw output",
                      "s": round(time.time() - t, 3)}
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ffi as O
        t = time.time()
        qor["sim_equivalent"] = bool(O.sim_cmp(g_in, out.arrays(), words=2, seed=11) == -1)
        log("equivalence checks: %s, oracle %.1fs" % (qor["cec"], time.time() - t))

    cpu = None
    if not a.no_cpu:
        ns = a.cpu_sample if a.cpu_sample > 0 else a.nands
        sample = g_in if (g_in is not None and ns == a.nands) else \
            ctx.gen_geometric(a.pis, ns, a.pos, a.window, a.p, a.seed).arrays()
        cpu = cpu_baseline(sample, ns, a.cpu_runs)
        log("cpu baseline: %s" % cpu)

    e2e_ms = strash_ms + ms_per_step
    hash_info = {"strash_probes": int(strash_probes),
                 "strash_probes_per_s": round(strash_probes / (strash_ms / 1e3), 1) if strash_ms > 0 else None,
                 "rw_probes": int(info["hash_probes"]),
                 "rw_probes_per_s": round(info["hash_probes"] / (info["total_ms"] / 1e3), 1)
                 if info["total_ms"] > 0 else None,
                 "probe": "one compare-and-swap on a strash table key slot, or one value update through a "
                          "pre-inserted key's stored location (device counters in the sweep kernels)"}

    res = {
        "metric": "rw Mnodes/s on 10M-AND AIG",
        "value": round(value, 4),
        "unit": "Mnodes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "rw (one pass) on the strashed deep random AIG: %d PIs, %d ANDs, %d POs, "
                               "fanin distance ~ Geometric(p=%g) within a %d window, seed %d" %
                               (a.pis, a.nands, a.pos, a.p, a.window, a.seed),
                   "nands": a.nands, "levels": s.levels},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "qor": qor,
        "stages_ms": {k: round(info[k], 3) for k in ("reorder_ms", "cuts_ms", "table_ms", "eval_ms", "select_ms",
                                                      "build_ms", "strash_ms", "total_ms")},
        "rw_counts": {k: info[k] for k in ("candidates", "selected", "new_nodes", "rounds", "est_gain")},
        "strash_ms": round(strash_ms, 3),
        "e2e": {"metric": "strash; rw Mnodes/s on the input ANDs", "value": round(a.nands / e2e_ms / 1e3, 4),
                "unit": "Mnodes/s", "strash_ms": round(strash_ms, 3), "rw_ms": round(ms_per_step, 3)},
        "hash": hash_info,
    }
    if info["cut_graph_nodes"]:
        res["cut_graph"] = {"capture_instantiate_ms": round(info["cut_graph_ms"], 3),
                            "kernel_nodes": info["cut_graph_nodes"],
                            "note": "cuts_ms holds capture, instantiation and the graph's execution"}
    if cpu is not None and cpu.get("e2e_value"):
        res["e2e"]["vs_cpu"] = round(res["e2e"]["value"] / cpu["e2e_value"], 3)
    if cpu is not None:
        res["vs_cpu"] = round(value / cpu["value"], 3)
    if not a.no_batch:
        # the batched and served workloads, in the same run (their own keys;
        # `value` stays the single-AIG rw throughput)
        for x in (out, src, raw):
            x.free()
        b = bench_batch(a)
        res["batch"] = {k: b[k] for k in ("metric", "value", "unit", "ms_per_step", "loop", "speedup_vs_loop",
                                          "identical_to_alone", "config", "stages_ms", "tuning")}
        log("batch: %s" % res["batch"])
        v = bench_serve(a)
        res["serve"] = {k: v[k] for k in ("metric", "value", "unit", "served", "loop", "throughput_vs_loop",
                                          "capacity", "identical_to_loop", "config")}
        log("serve: %s" % res["serve"])
    print(json.dumps(res), flush=True)


def build_native_oracle():
    """The oracle compiled for this host (-O3 -march=native) into a temporary
    directory; None if gcc is not usable here."""
    import subprocess
    import tempfile
    d = tempfile.mkdtemp(prefix="gls_cpu_")
    path = os.path.join(d, "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-fPIC", "-shared", "-std=c11", "-o", path,
           os.path.join(ROOT, "oracle", "oracle.c")]
    try:
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=120)
    except (OSError, subprocess.SubprocessError):
        return None, None
    return path, " ".join(cmd[:5])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sample, n_raw, runs):
    """CPU rewriting on the host's own cores: the oracle's sequential
    rewriter that follows the reference's own replacement method
    (orc_rw_culs, rewrite.cu:811-938 — the ABC-proxy of the CPU baseline) is
    `value`; the oracle's restatement of this package's pass (orc_rw) runs
    beside it.  One thread pinned to one core, median of `runs` runs, the
    oracle built -O3 -march=native for this host; the input strash is timed
    the same way for `strash; rw` end to end."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import statistics
    import oracle_ffi as O
    path, flags = build_native_oracle()
    core = None
    old_aff = None
    if hasattr(os, "sched_getaffinity"):
        old_aff = os.sched_getaffinity(0)
        core = min(old_aff)
        os.sched_setaffinity(0, {core})
    try:
        import contextlib
        with (O.using(path) if path else contextlib.nullcontext()):
            t_st, t_culs, t_rw = [], [], []
            ss = None
            for _ in range(runs):
                t = time.perf_counter()
                ss, _ = O.strash(*sample)
                t_st.append(time.perf_counter() - t)
            for _ in range(runs):
                t = time.perf_counter()
                rc, _ = O.rw_culs(*ss)
                t_culs.append(time.perf_counter() - t)
            for _ in range(max(1, (runs + 1) // 2)):
                t = time.perf_counter()
                rp, _ = O.rw(*ss)
                t_rw.append(time.perf_counter() - t)
    finally:
        if old_aff is not None:
            os.sched_setaffinity(0, old_aff)
    n = O.n_ands(ss)
    m_st, m_culs, m_rw = statistics.median(t_st), statistics.median(t_culs), statistics.median(t_rw)
    return {"value": round(n / m_culs / 1e6, 4), "unit": "Mnodes/s", "cores": 1, "kind": "port",
            "sample": "the workload's generator and seed at %d ANDs (%d after strash); oracle orc_rw_culs (the reference's "
                      "sequential replacement method, the ABC-proxy), median of %d runs of %.2f s"
                      % (n_raw, n, runs, m_culs),
            "build": flags or "shipped oracle/liboracle.so (gcc unavailable here)",
            "pinned_core": core, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "runs_s": [round(x, 3) for x in t_culs],
            "this_method": {"value": round(n / m_rw / 1e6, 4), "median_s": round(m_rw, 3),
                            "runs_s": [round(x, 3) for x in t_rw], "what": "orc_rw: the oracle of this package's pass"},
            "strash_median_s": round(m_st, 3),
            "e2e_value": round(n_raw / (m_st + m_culs) / 1e6, 4),
            "qor_ands": {"rw_culs": O.n_ands(rc), "rw": O.n_ands(rp)}}


def bench_batch(a):
    """Batched rw: --batch-count seeded random AIGs (ANDs log-uniform in
    [1000, --batch-max], PIs = max(16, ANDs/200), POs = PIs/2, fanins uniform over
    the PIs and the previous 1000 ANDs, seeds 1000+i), strashed, then one rw
    pass packed into one device arena vs the same pass one AIG at a time.
    value = total ANDs entering rw / batched rw time."""
    import numpy as np
    import culs_amd
    ctx = culs_amd.Context(0)
    rng = np.random.default_rng(12345)
    t0 = time.time()
    aigs = []
    for i in range(a.batch_count):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(a.batch_max))))
        pis = max(16, n // 200)
        aigs.append(ctx.gen_random(pis, n, max(1, pis // 2), 1000, 1000 + i))
    log("generated %d AIGs in %.1fs" % (len(aigs), time.time() - t0))
    ctx.batch_strash(aigs)
    n_in = sum(x.n_ands for x in aigs)
    depth = max(x.stats().levels for x in aigs)
    log("strashed: %d ANDs in total, deepest %d levels" % (n_in, depth))
    tuning = None
    if not a.no_tune:
        # the batched pass's settings for this arena size, picked by timing it
        t_tune = time.time()
        settings, tuned_ms, default_ms = ctx.autotune(aigs, culs_amd.GLS_TUNE_RW, reps=1)
        tuning = {"settings": settings, "start_ms": round(default_ms, 3), "tuned_ms": round(tuned_ms, 3),
                  "search_s": round(time.time() - t_tune, 1),
                  "method": "gls_autotune: coordinate descent over the pass settings, one timed batched rw "
                            "on device copies per candidate; the timed steps below run with the result"}
        log("autotune: %s, %.1f ms (started from %.1f ms) in %.1fs" % (settings, tuned_ms, default_ms,
                                                                       time.time() - t_tune))

    def copies():
        return [x.copy() for x in aigs]

    for _ in range(a.warmup):
        w = copies()
        ctx.batch_rewrite(w)
        for x in w:
            x.free()
    times, out = [], None
    for i in range(a.steps):
        c = copies()
        t = time.perf_counter()
        ctx.batch_rewrite(c)
        times.append(time.perf_counter() - t)
        if out is not None:
            for x in out:
                x.free()
        out = c
        log("batch step %d: %.1f ms" % (i, times[-1] * 1e3))
    t_batch = sum(times) / len(times)
    info = out[0].last_op()
    if a.no_loop:
        return {"metric": "batched rw Mnodes/s, %d mixed AIGs" % len(aigs),
                "value": round(n_in / t_batch / 1e6, 4), "unit": "Mnodes/s",
                "ms_per_step": round(t_batch * 1e3, 3), "loop": None, "tuning": tuning,
                "stages_ms": {k: round(info[k], 3) for k in ("cuts_ms", "eval_ms", "strash_ms", "total_ms")}}
    # the plain loop: the same pass, one AIG at a time (each already on the device)
    c = copies()
    t = time.perf_counter()
    for x in c:
        x.rewrite()
    t_loop = time.perf_counter() - t
    log("one-at-a-time loop: %.1f ms" % (t_loop * 1e3))
    same_count = all(x.n_ands == y.n_ands for x, y in zip(out, c))
    idx = np.linspace(0, len(c) - 1, min(a.batch_check, len(c))).astype(int)
    same_arrays = all(all(np.array_equal(p, q) for p, q in zip(out[k].arrays(), c[k].arrays())) for k in idx)
    n_out = sum(x.n_ands for x in out)
    res = {
        "metric": "batched rw Mnodes/s, %d mixed AIGs" % len(aigs),
        "value": round(n_in / t_batch / 1e6, 4),
        "unit": "Mnodes/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(t_batch * 1e3, 3),
        "higher_is_better": True, "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "batched rw over %d seeded random AIGs, ANDs log-uniform in [1000, %d], "
                               "PIs = max(16, ANDs/200), POs = PIs/2, fanin window 1000, seeds 1000+i"
                               % (len(aigs), a.batch_max),
                   "ands_total": n_in, "deepest": depth},
        "loop": {"value": round(n_in / t_loop / 1e6, 4), "ms": round(t_loop * 1e3, 3)},
        "speedup_vs_loop": round(t_loop / t_batch, 3),
        "identical_to_alone": {"and_counts_all": bool(same_count), "arrays_checked": int(len(idx)),
                               "arrays_equal": bool(same_arrays)},
        "qor": {"and_in": n_in, "and_out": n_out},
        "stages_ms": {k: round(info[k], 3) for k in ("reorder_ms", "cuts_ms", "table_ms", "eval_ms", "select_ms",
                                                      "build_ms", "strash_ms", "total_ms")},
        "tuning": tuning,
    }
    for x in out + c + aigs:
        x.free()
    return res


def bench_serve(a):
    """Serving: --serve-requests jobs arriving as a Poisson stream (--serve-rate
    per second), each `b; rw; rwz; b`; half seeded random AIGs with ANDs
    log-uniform in [1000, --serve-max] (PIs = max(16, ANDs/200), POs = PIs/2,
    fanin window 1000), half n-bit array multipliers, n uniform in 8..64.
    Served: the continuous-batching server (jobs stay on the device for their
    whole script).  Baseline: one job at a time in arrival order, the AIG
    copied to the host and back around every command, as the reference does
    around rw.  Latency = completion - arrival."""
    import threading
    import numpy as np
    import culs_amd
    script = "b; rw; rwz; b"
    rng = np.random.default_rng(777)
    ctx = culs_amd.Context(0)
    t0 = time.time()
    jobs, arrivals, t = [], [], 0.0
    for i in range(a.serve_requests):
        t += rng.exponential(1.0 / a.serve_rate)
        arrivals.append(t)
        if rng.random() < 0.5:
            n = int(np.exp(rng.uniform(np.log(1000), np.log(a.serve_max))))
            pis = max(16, n // 200)
            jobs.append(ctx.gen_random(pis, n, max(1, pis // 2), 1000, 2000 + i).arrays())
        else:
            jobs.append(ctx.gen_multiplier(int(rng.integers(8, 65)), 0.0, 2000 + i).arrays())
    n_in = sum(len(g[1]) - g[0] - 1 for g in jobs)
    log("generated %d jobs (%d ANDs) in %.1fs" % (len(jobs), n_in, time.time() - t0))

    def pct(x, q):
        return float(np.percentile(np.asarray(x), q))

    # served
    srv = culs_amd.Server(0, hbm_budget=96 << 30, max_active=4096)
    warm = srv.submit(jobs[0], script)
    srv.wait(warm)
    srv.release(warm)
    ids = [None] * len(jobs)
    start = time.perf_counter()

    def submitter():
        for k, g in enumerate(jobs):
            d = arrivals[k] - (time.perf_counter() - start)
            if d > 0:
                time.sleep(d)
            ids[k] = srv.submit(g, script)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    infos = [srv.wait(j) for j in ids]
    served_span = time.perf_counter() - start
    lat_s = [(i["done_ms"] - i["submit_ms"]) for i in infos]
    served_out = [srv.result(j) for j in ids]
    log("served: makespan %.2fs p50 %.1f ms p99 %.1f ms" % (served_span, pct(lat_s, 50), pct(lat_s, 99)))
    # capacity: every job submitted at once
    start = time.perf_counter()
    ids2 = [srv.submit(g, script) for g in jobs]
    for j in ids2:
        srv.wait(j)
    sat_span = time.perf_counter() - start
    srv.close()
    log("served, all at once: %.2fs" % sat_span)

    # plain loop with a host round trip per command
    lat_l, loop_out, busy = [], [], 0.0
    start = time.perf_counter()
    for k, g in enumerate(jobs):
        d = arrivals[k] - (time.perf_counter() - start)
        if d > 0:
            time.sleep(d)
        t_job = time.perf_counter()
        cur = g
        for cmd in ("b", "rw", "rwz", "b"):
            x = ctx.from_arrays(*cur)
            if cmd == "b":
                x.balance()
            else:
                x.rewrite(use_zeros=(cmd == "rwz"))
            cur = x.arrays()
            x.free()
        loop_out.append(cur)
        busy += time.perf_counter() - t_job
        lat_l.append((time.perf_counter() - start - arrivals[k]) * 1e3)
    loop_span = time.perf_counter() - start
    log("loop: makespan %.2fs p50 %.1f ms p99 %.1f ms" % (loop_span, pct(lat_l, 50), pct(lat_l, 99)))
    identical = all(all(np.array_equal(p, q) for p, q in zip(x, y)) for x, y in zip(served_out, loop_out))
    res = {
        "metric": "served b;rw;rwz;b Mnodes/s, Poisson %g req/s" % a.serve_rate,
        "value": round(n_in / served_span / 1e6, 4),
        "unit": "Mnodes/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(served_span * 1e3, 3),
        "higher_is_better": True, "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "%d jobs `%s`, Poisson arrivals %g/s; half random AIGs (ANDs log-uniform "
                               "[1000, %d], window 1000), half n-bit multipliers (n in 8..64); seeds 2000+i"
                               % (len(jobs), script, a.serve_rate, a.serve_max),
                   "ands_total": n_in},
        "served": {"p50_ms": round(pct(lat_s, 50), 3), "p99_ms": round(pct(lat_s, 99), 3),
                   "makespan_s": round(served_span, 3), "jobs_per_s": round(len(jobs) / served_span, 3)},
        "loop": {"p50_ms": round(pct(lat_l, 50), 3), "p99_ms": round(pct(lat_l, 99), 3),
                 "makespan_s": round(loop_span, 3), "jobs_per_s": round(len(jobs) / loop_span, 3),
                 "value": round(n_in / loop_span / 1e6, 4)},
        "throughput_vs_loop": round(loop_span / served_span, 3),
        "capacity": {"served_all_at_once_s": round(sat_span, 3), "loop_busy_s": round(busy, 3),
                     "served_jobs_per_s": round(len(jobs) / sat_span, 3),
                     "loop_jobs_per_s": round(len(jobs) / busy, 3), "ratio": round(busy / sat_span, 3)},
        "identical_to_loop": bool(identical),
    }
    return res


if __name__ == "__main__":
    main()
