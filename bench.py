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


def run_path(graph, precision, steps, warmup, args):
    from relaxmap_amd import Network
    net = Network(precision=precision, batches=args.batches)
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
    graph.upload_to(net)
    net.set_profiling(True)
    net.initiate()
    net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
    stats = net.kernel_stats()
    net.close()
    return stats


def cpu_baseline(args):
    """The reference OpenMP binary (built from its sources into oracle/_ref)
    on a bounded sample: the 100k-node LFR-style network, all its sweeps."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ompRelaxmap")
    if not os.path.exists(exe):
        log("reference binary oracle/_ref/ompRelaxmap not built; no CPU baseline")
        return None, None
    from relaxmap_amd import synth
    g = synth.config2()
    threads = min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "lfr100k.net")
        g.write(path)
        cmd = [exe, "1", path, str(threads), "1", "1e-3", "0", "10", d, "prior"]
        t = time.perf_counter()
        try:
            out = subprocess.run(cmd, capture_output=True, text=True, timeout=args.cpu_timeout).stdout
        except subprocess.TimeoutExpired as e:
            out = (e.stdout or b"").decode() if isinstance(e.stdout, bytes) else (e.stdout or "")
        wall = time.perf_counter() - t
    it1 = re.search(r"Iteration 1: code length =\s*([0-9.eE+-]+)\s*, elapsed time:\s*([0-9.eE+-]+)", out)
    total = re.search(r"Overall Elapsed Time for Module Detection \(w/o file IO\): ([0-9.eE+-]+)", out)
    final = re.search(r"Computed Code Length = ([0-9.eE+-]+)", out)
    if not it1:
        log("reference run produced no sweep timing")
        return None, None
    t_it1 = float(it1.group(2))
    gteps = 2.0 * g.m / t_it1 / 1e9
    info = {"graph_nodes": g.n, "graph_arcs": g.m, "first_sweep_s": t_it1,
            "time_to_converged_s": float(total.group(1)) if total else None,
            "codelength_bits": float(final.group(1)) if final else None, "threads": threads, "wall_s": wall}
    log(f"reference OpenMP ({threads} threads) on {g.n} nodes / {g.m} arcs: first sweep {t_it1:.3f} s "
        f"({gteps:.4f} GTEPS), converged in {info['time_to_converged_s']} s")
    base = {"value": gteps, "unit": "GTEPS", "cores": threads, "kind": "reference",
            "sample": f"reference ompRelaxmap (unmodified sources, -O3 -fopenmp, {threads} threads, prior, "
                      f"threshold 1e-3, maxIter 10) on the 100k-node LFR-style network ({g.m} arcs); "
                      f"value = 2 x arcs / time of its first full sweep"}
    return base, info


def gpu_on_sample(args):
    """The GPU path on the CPU baseline's sample, for a same-input wall-time ratio."""
    from relaxmap_amd import Network, synth
    g = synth.config2()
    out = {}
    for prec in ("mixed", "fp64"):
        net = Network(precision=prec, batches=args.batches)
        g.upload_to(net)
        net.initiate()   # warm-up
        t = time.perf_counter()
        net.initiate()
        r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
        out[prec] = {"time_to_converged_s": time.perf_counter() - t,
                     "codelength_bits": r.computed_codelength / math.log(2)}
        net.close()
    return out


TRAFFIC = os.path.join(ROOT, "profiles", "sweep_flow_kernels_chunglu32m_traffic.txt")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` on the 32M-node workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated by a plain
    copy of known size), or None."""
    try:
        lines = open(TRAFFIC).read().splitlines()
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
    return out


def roofline_of(stats, default_workload):
    if not stats:
        return None, {}
    table = {k: {"launches": n, "avg_ms": ms / max(n, 1), "total_ms": ms,
                 "achieved_GBs": (b / max(n, 1)) / (ms / max(n, 1) * 1e-3) / 1e9 if ms > 0 else 0.0}
             for k, (n, ms, b) in stats.items()}
    single = [k for k in stats if not k.startswith("hub sequence")]   # one HIP kernel per entry
    name = max(single, key=lambda k: stats[k][1])
    n, ms, b = stats[name]
    achieved = (b / n) / (ms / n * 1e-3) / 1e9
    traffic = measured_traffic(name) if default_workload else None
    roof = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": b / n, "avg_launch_ms": ms / n, "launches": n}
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
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--cpu-timeout", type=float, default=120.0)
    ap.add_argument("--no-config4", action="store_true", help="skip the 512-instance batch measurement")
    ap.add_argument("--workers", type=int, default=8, help="concurrent contexts of the batch measurement")
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

    sweep_s = sum(r["sweep_s"] for r in rec_m)
    visits = sum(r["edge_visits"] for r in rec_m)
    gteps = visits / sweep_s / 1e9
    step_ms = 1e3 * sum(r["step_s"] for r in rec_m) / len(rec_m)
    out = {"metric": METRIC, "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step_ms, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 arcs + f64/int64 accumulation (mixed)", "data": "synthetic",
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
        out["flow_l1_mixed_vs_fp64"] = float(np.abs(p_mixed - p_fp64).sum())
        out["codelength_rel_mixed_vs_fp64"] = (mixed["codelength_bits"] - fp64["codelength_bits"]) / fp64["codelength_bits"]
        if g.community[0] >= 0:
            out["nmi_planted"] = {"mixed": nmi(part_m, g.community), "fp64": nmi(part_d, g.community)}
    if not args.no_profile:
        stats = profile_pass(g, "mixed", args)
        roof, table = roofline_of(stats, args.nodes == DEFAULT_NODES and args.mean_deg == DEFAULT_MEAN_DEG)
        out["roofline"] = roof
        out["kernels"] = {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in
                          sorted(table.items(), key=lambda kv: -kv[1]["total_ms"])[:12]}
    del g
    if not args.no_config4:
        out["config4"] = config4(args)
    if not args.no_cpu:
        base, info = cpu_baseline(args)
        out["cpu_baseline"] = base
        if info:
            out["cpu_baseline_detail"] = info
            gs = gpu_on_sample(args)
            out["gpu_on_cpu_sample"] = gs
            if info.get("time_to_converged_s"):
                out["speedup_vs_reference_on_sample"] = {
                    k: info["time_to_converged_s"] / v["time_to_converged_s"] for k, v in gs.items()}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
