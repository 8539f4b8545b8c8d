"""Benchmark: RelaxMap's hot path on one MI355X.

A step is one pass of the hot path over the resident synthetic network:
steady-state flow (Network::initiate) and the full optimisation to a
converged codelength (main()'s super-step loop), i.e. the time to converged
codelength.  ``value`` is the move-sweep throughput of the mixed-precision
path in GTEPS (arc visits of all sweeps / sweep wall time); the FP64 path is
run on the same network beside it.  The workload is the 500M-arc
configuration: a directed Chung-Lu graph with planted communities (32M nodes,
degree exponent 2.1 in [4, 100000], community sizes exponent 1.2 in
[50, 50000], mu 0.4, FP32 lognormal weights, 1% dangling).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nodes N] [--no-cpu]
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "move-sweep GTEPS + time-to-converged codelength, mixed vs FP64, 500M edges"
HBM_PEAK_GBS = 8000.0
DEFAULT_NODES = 32 * 1024 * 1024
DEFAULT_MEAN_DEG = 18.5   # calibrated so the generated network has ~529M arcs


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def nmi(a, b):
    a = np.unique(a, return_inverse=True)[1].astype(np.int64)
    b = np.unique(b, return_inverse=True)[1].astype(np.int64)
    n = len(a)
    pair = a * (b.max() + 1) + b
    _, cnt = np.unique(pair, return_counts=True)
    ua, ca = np.unique(a, return_counts=True)
    ub, cb = np.unique(b, return_counts=True)
    h = lambda c: -(c / n * np.log(c / n)).sum()
    mi = h(ca) + h(cb) - h(cnt)
    return float(2 * mi / (h(ca) + h(cb)))


def run_path(graph, precision, steps, warmup, args, pull=None):
    from relaxmap_amd import Network
    net = Network(precision=precision, batches=args.batches)
    if args.fp64_pull if pull is None else pull:
        net.set_fp64_flow(False)
    t = time.perf_counter()
    graph.upload_to(net)
    log(f"{precision}: uploaded {graph.n} nodes / {graph.m} arcs in {time.perf_counter() - t:.1f} s")
    rec = []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        it, L0 = net.initiate()
        t1 = time.perf_counter()
        r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
        t2 = time.perf_counter()
        tm = net.timing()
        rec.append({"step_s": t2 - t0, "initiate_s": t1 - t0, "partition_s": t2 - t1, "iterations": it,
                    "sweep_s": r.time_sweeps_s, "contract_s": r.time_contract_s, "edge_visits": r.edge_visits,
                    "sweeps": r.sweeps, "codelength": r.computed_codelength, "n_module": r.n_module,
                    "flow_ms": tm["flow_ms"], "super_steps": r.super_steps})
        log(f"{precision} {'warmup' if i < warmup else 'step'} {i}: {t2 - t0:.3f} s (flow {t1 - t0:.3f} s, "
            f"{it} it; partition {t2 - t1:.3f} s: sweeps {r.time_sweeps_s:.3f} s, contraction {r.time_contract_s:.3f} s, "
            f"{r.sweeps} sweeps, {r.edge_visits / max(r.time_sweeps_s, 1e-9) / 1e9:.2f} GTEPS); "
            f"L={r.computed_codelength / math.log(2):.5f} bits, {r.n_module} modules")
    return net, rec[warmup:]


def profile_pass(graph, precision, args):
    """One extra, untimed pass with per-launch HIP events (profiling mode)."""
    from relaxmap_amd import Network
    net = Network(precision=precision, batches=args.batches)
    if args.fp64_pull:
        net.set_fp64_flow(False)
    graph.upload_to(net)
    net.set_profiling(True)
    net.initiate()
    net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
    stats = net.kernel_stats()
    net.close()
    return stats


REF_EXE = os.path.join(ROOT, "oracle", "_ref", "ompRelaxmap")


def host_info():
    """Threads this process may run on (what nproc reports), CPU model, RAM."""
    threads = len(os.sched_getaffinity(0))
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    mem = None
    try:
        for ln in open("/proc/meminfo"):
            if ln.startswith("MemTotal"):
                mem = round(int(ln.split()[1]) / 2 ** 20, 1)
                break
    except OSError:
        pass
    return {"threads": threads, "cpu_model": model, "ram_gib": mem}


def run_reference(path, outdir, threads, timeout):
    """ompRelaxmap seed 1, prior, threshold 1e-3, vThresh 0, maxIter 10, one
    attempt; the timings and codelength it prints (OmpRelaxmap.cpp:101, 222,
    229, 233, 334-335)."""
    cmd = [REF_EXE, "1", path, str(threads), "1", "1e-3", "0", "10", outdir, "prior"]
    t = time.perf_counter()
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout).stdout
        timed_out = False
    except subprocess.TimeoutExpired as e:
        out = (e.stdout or b"").decode() if isinstance(e.stdout, bytes) else (e.stdout or "")
        timed_out = True
    wall = time.perf_counter() - t
    num = r"([0-9.eE+-]+)"
    it1 = re.search(r"Iteration 1: code length =\s*" + num + r"\s*, elapsed time:\s*" + num, out)
    total = re.search(r"Overall Elapsed Time for Module Detection \(w/o file IO\): " + num, out)
    read = re.search(r"Time for reading input data : " + num, out)
    final = re.search(r"Computed Code Length = " + num, out)
    return {"first_sweep_s": float(it1.group(2)) if it1 else None,
            "time_to_converged_s": float(total.group(1)) if total else None,
            "read_s": float(read.group(1)) if read else None,
            "codelength_bits": float(final.group(1)) if final else None,
            "threads": threads, "wall_s": wall, "timed_out": timed_out}


def reference_partition(outdir, name, n):
    """The module of every node from the reference's .clu file (one raw
    module id per line after the header, OmpRelaxmap.cpp:240-249)."""
    path = os.path.join(outdir, name + ".clu")
    try:
        lines = open(path).read().split()
    except OSError:
        return None
    vals = np.array([int(x) for x in lines[2:]], dtype=np.int64)   # "*Vertices n" header
    return vals if len(vals) == n else None


def cpu_baseline(args, host):
    """The reference OpenMP binary (built from its sources into oracle/_ref)
    on a bounded sample: the 100k-node LFR-style network, all its sweeps, on
    every host thread this process may use."""
    if not os.path.exists(REF_EXE):
        log("reference binary oracle/_ref/ompRelaxmap not built; no CPU baseline")
        return None, None
    from relaxmap_amd import synth
    g = synth.config2()
    threads = host["threads"]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "lfr100k.net")
        g.write(path)
        info = run_reference(path, d, threads, args.cpu_timeout)
        part = reference_partition(d, "lfr100k", g.n)
    if info["first_sweep_s"] is None:
        log("reference run produced no sweep timing")
        return None, None
    t_it1 = info["first_sweep_s"]
    gteps = 2.0 * g.m / t_it1 / 1e9
    info.update({"graph_nodes": g.n, "graph_arcs": g.m})
    if part is not None and g.community[0] >= 0:
        info["nmi_planted"] = nmi(part, g.community)
    log(f"reference OpenMP ({threads} threads) on {g.n} nodes / {g.m} arcs: first sweep {t_it1:.3f} s "
        f"({gteps:.4f} GTEPS), converged in {info['time_to_converged_s']} s")
    base = {"value": gteps, "unit": "GTEPS", "cores": threads, "kind": "reference",
            "sample": f"reference ompRelaxmap (unmodified sources, -O3 -fopenmp, {threads} threads = nproc, prior, "
                      f"threshold 1e-3, maxIter 10) on the 100k-node LFR-style network ({g.m} arcs); "
                      f"value = 2 x arcs / time of its first full sweep"}
    return base, info


def gpu_same_input(g, args):
    """The GPU path on a network the reference also ran: the time from the
    host's parsed arrays to the converged codelength (upload, in-CSR and
    buckets, steady state, partition), the span of the reference's
    "w/o file IO" timer (OmpRelaxmap.cpp:103-229), and the kernel part alone
    (initiate + partition on the resident network)."""
    from relaxmap_amd import Network
    out = {}
    for prec in ("mixed", "fp64"):
        net = Network(precision=prec, batches=args.batches)   # warm-up run of the whole path
        g.upload_to(net)
        net.initiate()
        net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
        net.close()
        t0 = time.perf_counter()
        net = Network(precision=prec, batches=args.batches)
        g.upload_to(net)
        t1 = time.perf_counter()
        net.initiate()
        r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
        t2 = time.perf_counter()
        rec = {"time_to_converged_s": t2 - t0, "upload_s": t1 - t0, "initiate_and_partition_s": t2 - t1,
               "codelength_bits": r.computed_codelength / math.log(2), "modules": r.n_module}
        if g.community[0] >= 0:
            rec["nmi_planted"] = nmi(net.partition(), g.community)
        out[prec] = rec
        net.close()
    return out


def reference_large(args, host):
    """A second same-input point: the Chung-Lu generator of the 500M-arc
    workload at 2M nodes (about 31M arcs), the reference and the GPU."""
    from relaxmap_amd import synth
    g = synth.config5(n=args.ref_nodes, mean_deg=DEFAULT_MEAN_DEG)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "chunglu.net")
        t = time.perf_counter()
        g.write(path)
        log(f"second reference point: {g.n} nodes / {g.m} arcs written in {time.perf_counter() - t:.1f} s")
        ref = run_reference(path, d, host["threads"], args.ref_timeout)
        part = reference_partition(d, "chunglu", g.n)
    ref.update({"graph_nodes": g.n, "graph_arcs": g.m})
    if part is not None:
        ref["nmi_planted"] = nmi(part, g.community)
    log(f"reference on {g.n} nodes: converged in {ref['time_to_converged_s']} s (read {ref['read_s']} s), "
        f"L = {ref['codelength_bits']} bits")
    gpu = gpu_same_input(g, args)
    out = {"workload": f"directed Chung-Lu, the 500M-arc configuration's generator at {g.n} nodes",
           "reference": ref, "gpu": gpu}
    if ref["time_to_converged_s"]:
        out["speedup_vs_reference"] = {k: ref["time_to_converged_s"] / v["time_to_converged_s"]
                                       for k, v in gpu.items()}
    return out


def config3(args):
    """R-MAT scale 22 (edge factor 16, a, b, c, d = .57, .19, .19, .05, FP32
    U(0,1] weights, dangling nodes kept): the precision pattern's error on a
    skewed, weakly modular network.  Flows of both precisions, the
    codelength of one fixed partition (FP64's converged one) in both, and
    each precision's own converged codelength and time."""
    from relaxmap_amd import Network, synth
    t = time.perf_counter()
    g = synth.config3()
    log(f"config 3: R-MAT scale 22, {g.n} nodes / {g.m} arcs generated in {time.perf_counter() - t:.1f} s")
    res = {}
    part = None
    for prec in ("fp64", "mixed"):
        net = Network(precision=prec, batches=args.batches)
        g.upload_to(net)
        net.initiate()   # warm-up
        t0 = time.perf_counter()
        it, _ = net.initiate()
        t1 = time.perf_counter()
        r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
        t2 = time.perf_counter()
        rec = {"flows": net.flows(), "iterations": it, "initiate_s": t1 - t0, "time_to_converged_s": t2 - t0,
               "codelength_bits": r.computed_codelength / math.log(2), "modules": r.n_module,
               "move_sweep_gteps": r.edge_visits / max(r.time_sweeps_s, 1e-9) / 1e9}
        if part is None:
            part = net.partition()
        net.set_partition(part)
        rec["fixed_partition_codelength"] = net.calculate_codelength()
        res[prec] = rec
        net.close()
    d, m = res["fp64"], res["mixed"]
    out = {"workload": "R-MAT scale 22, edge factor 16, (0.57, 0.19, 0.19, 0.05), FP32 U(0,1] weights",
           "nodes": g.n, "arcs": g.m,
           "flow_l1_mixed_vs_fp64": float(np.abs(m["flows"] - d["flows"]).sum()),
           "codelength_rel": (m["fixed_partition_codelength"] - d["fixed_partition_codelength"])
           / d["fixed_partition_codelength"],
           "final_codelength_rel": (m["codelength_bits"] - d["codelength_bits"]) / d["codelength_bits"]}
    for prec in ("fp64", "mixed"):
        out[prec] = {k: v for k, v in res[prec].items() if k != "flows"}
    log(f"config 3: flow L1 mixed vs FP64 {out['flow_l1_mixed_vs_fp64']:.2e}, fixed-partition codelength rel "
        f"{out['codelength_rel']:.2e}")
    return out


TRAFFIC = {"mixed": os.path.join(ROOT, "profiles", "step_kernels_chunglu32m_mixed_traffic.txt"),
           "fp64": os.path.join(ROOT, "profiles", "step_kernels_chunglu32m_fp64_traffic.txt")}


def measured_traffic(kernel, precision="mixed"):
    """HBM bytes per dispatch of `kernel` on the 32M-node workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of that precision
    (calibrated by a plain copy of known size), or None."""
    try:
        lines = open(TRAFFIC[precision]).read().splitlines()
    except OSError:
        return None
    ff = wf = None
    for ln in lines:
        if ln.startswith("fetch_factor"):
            ff = float(ln.split()[1])
        elif ln.startswith("write_factor"):
            wf = float(ln.split()[1])
        elif not ln.startswith("#"):
            parts = ln.split("\t")
            if len(parts) >= 4 and parts[0] == kernel and ff and wf:
                return float(parts[2]) * 1024 * ff + float(parts[3]) * 1024 * wf
    return None


def config4(args):
    """64 directed SBMs (500-50k nodes, 10-200 blocks, Poisson(12), mu in
    [0.1, 0.5]) x 8 seeds = 512 independent instances: instances/s and the
    per-network codelength spread over seeds.  Primary: one batch of all 512
    instances (rmx_batch_upload); beside it, concurrent contexts taking the
    networks one by one."""
    from relaxmap_amd import synth
    from relaxmap_amd.batch import run_batched, run_instances
    graphs = synth.config4_graphs()
    seeds = list(range(1, 9))

    def spread_of(res):
        spread = []
        for gi in range(len(graphs)):
            L = np.array([res[(gi, s)][0] for s in seeds])
            spread.append(float((L.max() - L.min()) / L.min()))
        return {"median": float(np.median(spread)), "max": float(np.max(spread))}

    run_batched(graphs[:4], seeds[:2])   # warm-up
    res, dt = run_batched(graphs, seeds)
    out = {"instances": len(res), "networks": len(graphs), "seeds": len(seeds), "mode": "one batch",
           "nodes": int(sum(g.n for g in graphs)), "arcs": int(sum(g.m for g in graphs)), "seconds": dt,
           "instances_per_s": len(res) / dt, "codelength_rel_spread_over_seeds": spread_of(res)}
    log(f"config 4: {len(res)} instances as one batch in {dt:.2f} s ({len(res) / dt:.1f} instances/s)")
    run_instances(graphs[:4], seeds[:1], workers=args.workers)   # warm-up
    res_c, dt_c = run_instances(graphs, seeds, workers=args.workers)
    out["contexts"] = {"workers": args.workers, "seconds": dt_c, "instances_per_s": len(res_c) / dt_c,
                       "codelength_rel_spread_over_seeds": spread_of(res_c)}
    rel = [abs(res[k][0] - res_c[k][0]) / res_c[k][0] for k in res]
    out["codelength_rel_batch_vs_contexts"] = {"median": float(np.median(rel)), "max": float(np.max(rel))}
    log(f"config 4: {len(res_c)} instances by {args.workers} contexts in {dt_c:.2f} s "
        f"({len(res_c) / dt_c:.1f} instances/s)")
    if not args.no_cpu and os.path.exists(REF_EXE):
        out["reference"] = config4_reference(graphs, res, args)
    return out


def config4_reference(graphs, res, args):
    """The reference on the 64 networks with seed 1 (one process per network
    on all host threads, as it runs), scaled to the 512 instances by x 8.  A
    network on which the multi-threaded reference crashes (its parallel
    super-node sweep reads the module array past its end, Module.cpp:2642)
    is run again on one thread, the reference's sequential path."""
    threads = host_info()["threads"]
    no_io = wall = 0.0
    rel = []
    crashed = []
    with tempfile.TemporaryDirectory() as d:
        for gi, g in enumerate(graphs):
            path = os.path.join(d, f"sbm{gi}.net")
            g.write(path)
            r = run_reference(path, d, threads, args.cpu_timeout)
            if r["time_to_converged_s"] is None:
                crashed.append(gi)
                r = run_reference(path, d, 1, args.cpu_timeout)
                if r["time_to_converged_s"] is None:
                    log(f"config 4 reference: network {gi} gave no timing on {threads} threads or on one")
                    return None
            no_io += r["time_to_converged_s"]
            wall += r["wall_s"]
            if r["codelength_bits"]:
                gpu_bits = res[(gi, 1)][0]   # bits
                rel.append((gpu_bits - r["codelength_bits"]) / r["codelength_bits"])
    out = {"networks": len(graphs), "seed": 1, "threads": threads, "seconds_without_io": no_io,
           "seconds_wall": wall, "instances_per_s": len(graphs) / no_io,
           "scaled_512_instances_s": 8 * no_io,
           "crashed_multithreaded_rerun_on_one_thread": crashed,
           "gpu_batch_codelength_rel_vs_reference": {"median": float(np.median(rel)), "max": float(np.max(rel)),
                                                     "min": float(np.min(rel))} if rel else None}
    log(f"config 4 reference: 64 networks in {no_io:.2f} s without file IO ({wall:.1f} s wall), "
        f"{out['instances_per_s']:.1f} instances/s; 512 instances ~ {8 * no_io:.1f} s")
    return out


def roofline_of(stats, default_workload, precision="mixed"):
    """The roofline of the kernel with the most time in the profiled step
    (every entry of the profiling mode is one HIP kernel), and the table of
    every kernel with at least 1 % of the profiled kernel time: launches,
    time, algorithmic GB/s and, on the default workload, the measured HBM
    traffic per launch."""
    if not stats:
        return None, {}
    single = {k: v for k, v in stats.items() if "sequence" not in k}
    total = sum(ms for _, ms, _ in single.values())
    table = {}
    for k, (n, ms, b) in single.items():
        if ms < 0.01 * total:
            continue
        row = {"launches": n, "avg_ms": ms / max(n, 1), "total_ms": ms, "share": ms / total,
               "bytes_per_launch": b / max(n, 1),
               "achieved_GBs": (b / max(n, 1)) / (ms / max(n, 1) * 1e-3) / 1e9 if ms > 0 else 0.0}
        tr = measured_traffic(k, precision) if default_workload else None
        if tr:
            row["traffic_per_launch"] = tr
            row["traffic_over_algorithmic"] = tr / max(row["bytes_per_launch"], 1.0)
        table[k] = row
    name = max(single, key=lambda k: single[k][1])
    n, ms, b = single[name]
    achieved = (b / n) / (ms / n * 1e-3) / 1e9
    traffic = measured_traffic(name, precision) if default_workload else None
    roof = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": b / n, "avg_launch_ms": ms / n, "launches": n,
            "share_of_kernel_time": round(ms / total, 4)}
    if traffic:
        roof["traffic_over_algorithmic"] = round(traffic / (b / n), 2)
    return roof, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nodes", type=int, default=DEFAULT_NODES)
    ap.add_argument("--mean-deg", type=float, default=DEFAULT_MEAN_DEG)
    ap.add_argument("--batches", type=int, default=0,
                    help="sub-sweeps per sweep; 0: one per 32768 units of a level, 4 to 8")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="only the FP64 path (for its kernel trace); value is then the FP64 path's")
    ap.add_argument("--no-fp64-pull", action="store_true", help="skip the short FP64 run with the in-CSR pull steady state")
    ap.add_argument("--fp64-pull", action="store_true",
                    help="the FP64 steady state by the in-CSR pull instead of the blocked push / pull")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--cpu-timeout", type=float, default=120.0)
    ap.add_argument("--no-config4", action="store_true", help="skip the 512-instance batch measurement")
    ap.add_argument("--workers", type=int, default=8, help="concurrent contexts of the batch measurement")
    ap.add_argument("--no-config3", action="store_true", help="skip the R-MAT scale-22 precision measurement")
    ap.add_argument("--no-ref-large", action="store_true", help="skip the second same-input reference point")
    ap.add_argument("--ref-nodes", type=int, default=2_000_000)
    ap.add_argument("--ref-timeout", type=float, default=300.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the hot path here is a single-GPU build "
              "(a row partition of the steady state and an all-gather of module labels per batch "
              "would spread it over several); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from relaxmap_amd import device_count, synth
    if device_count() < 1:
        raise SystemExit("no GPU visible")
    t = time.perf_counter()
    g = synth.config5(n=args.nodes, mean_deg=args.mean_deg)
    log(f"generated Chung-Lu network: {g.n} nodes, {g.m} arcs in {time.perf_counter() - t:.1f} s")

    if args.no_mixed:   # the FP64 path alone, reported in the mixed path's place
        args.no_fp64 = True
        net_m, rec_m = run_path(g, "fp64", args.steps, args.warmup, args)
    else:
        net_m, rec_m = run_path(g, "mixed", args.steps, args.warmup, args)
    p_mixed = net_m.flows()
    part_m = net_m.partition()
    net_m.close()
    rec_d, p_fp64, part_d = None, None, None
    if not args.no_fp64:
        net_d, rec_d = run_path(g, "fp64", args.steps, args.warmup, args)
        p_fp64 = net_d.flows()
        part_d = net_d.partition()
        net_d.close()
    rec_p = None
    if not args.no_fp64 and not args.fp64_pull and not args.no_fp64_pull:
        # the FP64 path with the in-CSR pull steady state (the plain CSR pull SpMV): a short run beside it
        net_p, rec_p = run_path(g, "fp64", 2, 1, args, pull=True)
        net_p.close()

    sweep_s = sum(r["sweep_s"] for r in rec_m)
    visits = sum(r["edge_visits"] for r in rec_m)
    gteps = visits / sweep_s / 1e9
    step_ms = 1e3 * sum(r["step_s"] for r in rec_m) / len(rec_m)
    out = {"metric": METRIC, "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step_ms, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None,
           "dtype": "f64" if args.no_mixed else "f32 arcs + f64/int64 accumulation (mixed)", "data": "synthetic",
           "config": {"workload": "directed Chung-Lu with planted communities (500M-arc configuration)",
                      "nodes": g.n, "arcs": g.m, "degree_exponent": 2.1, "degree_range": [4, 100000],
                      "community_exponent": 1.2, "community_range": [50, 50000], "mu": 0.4,
                      "weights": "fp32 lognormal(0,1)", "dangling": 0.01, "seed": 5,
                      "threshold": 1e-3, "vthresh": 0.0, "max_iter": 10, "mode": "prior",
                      "batches": args.batches or "per level (8 at every level of this network)"}}
    mixed = {"time_to_converged_s": step_ms / 1e3, "move_sweep_gteps": gteps,
             "codelength_bits": rec_m[-1]["codelength"] / math.log(2), "modules": rec_m[-1]["n_module"],
             "flow_iterations": rec_m[-1]["iterations"], "sweeps": rec_m[-1]["sweeps"],
             "flow_s": rec_m[-1]["initiate_s"], "partition_s": rec_m[-1]["partition_s"]}
    out["mixed"] = mixed
    if rec_d:
        sd = sum(r["sweep_s"] for r in rec_d)
        vd = sum(r["edge_visits"] for r in rec_d)
        fp64 = {"time_to_converged_s": sum(r["step_s"] for r in rec_d) / len(rec_d), "move_sweep_gteps": vd / sd / 1e9,
                "codelength_bits": rec_d[-1]["codelength"] / math.log(2), "modules": rec_d[-1]["n_module"],
                "flow_iterations": rec_d[-1]["iterations"], "sweeps": rec_d[-1]["sweeps"],
                "flow_s": rec_d[-1]["initiate_s"], "partition_s": rec_d[-1]["partition_s"]}
        out["fp64"] = fp64
        out["speedup_mixed_vs_fp64"] = fp64["time_to_converged_s"] / mixed["time_to_converged_s"]
        if rec_p:
            tp = sum(r["step_s"] for r in rec_p) / len(rec_p)
            out["fp64_pull"] = {"time_to_converged_s": tp, "flow_s": rec_p[-1]["initiate_s"],
                                "flow_iterations": rec_p[-1]["iterations"], "steps": len(rec_p),
                                "codelength_bits": rec_p[-1]["codelength"] / math.log(2)}
            out["speedup_mixed_vs_fp64_pull"] = tp / mixed["time_to_converged_s"]
        out["flow_l1_mixed_vs_fp64"] = float(np.abs(p_mixed - p_fp64).sum())
        out["codelength_rel_mixed_vs_fp64"] = (mixed["codelength_bits"] - fp64["codelength_bits"]) / fp64["codelength_bits"]
        if g.community[0] >= 0:
            out["nmi_planted"] = {"mixed": nmi(part_m, g.community), "fp64": nmi(part_d, g.community)}
    if not args.no_profile:
        default_workload = args.nodes == DEFAULT_NODES and args.mean_deg == DEFAULT_MEAN_DEG
        precs = ["fp64"] if args.no_mixed else (["mixed"] + ([] if args.no_fp64 else ["fp64"]))
        for prec in precs:
            stats = profile_pass(g, prec, args)
            roof, table = roofline_of(stats, default_workload, prec)
            kern = {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()} for k, v in
                    sorted(table.items(), key=lambda kv: -kv[1]["total_ms"])}
            if prec == precs[0]:
                out["roofline"] = roof
                out["kernels"] = kern
            else:
                out["roofline_" + prec] = roof
                out["kernels_" + prec] = kern
    del g
    if not args.no_config3:
        out["config3"] = config3(args)
    if not args.no_config4:
        out["config4"] = config4(args)
    if not args.no_cpu:
        host = host_info()
        out["host"] = host
        base, info = cpu_baseline(args, host)
        out["cpu_baseline"] = base
        if info:
            out["cpu_baseline_detail"] = info
            from relaxmap_amd import synth
            gs = gpu_same_input(synth.config2(), args)
            out["gpu_on_cpu_sample"] = gs
            if info.get("time_to_converged_s"):
                out["speedup_vs_reference_on_sample"] = {
                    k: info["time_to_converged_s"] / v["time_to_converged_s"] for k, v in gs.items()}
            if not args.no_ref_large:
                out["reference_large"] = reference_large(args, host)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
