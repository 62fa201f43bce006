#!/usr/bin/env python3
"""Benchmark: FDTD Mcell-updates/s of the PTask Yee graph on one MI355X.

Workload: a 384^3 binary32 Yee grid (six field arrays, 1.36 GB) advanced 200 time steps, fields
seeded U[-1,1), Courant number 0.5, tangential E faces copied through.  The graph is the looped
PTask graph (general-iteration meta ports, predicated back-channels) built against ptaskapi.h
(ptask_amd/csrc/apps/ptk_fdtd.cpp) and dispatched task by task by the runtime.

One benchmark "step" is one 200-time-step run of the graph.  Runs are chained: each run starts
from the previous run's output datablocks, which never leave HBM, so the timed region moves
nothing over PCIe (the first, untimed warmup run uploads; a final untimed run downloads).

Graph forms measured (all bit-identical to the six reference updates); `value` is the fastest:
  overlap3p -- overlap3 on a persistent grid of one block per CU: each block streams one
             whole-depth tile and an equal run of the remaining tiles' planes
  overlap3 -- 1 task per 3 steps, FdtdThreeStepOverlap: three whole Yee steps per pass on
             overlapped 58 x 10 column tiles, each step two planes behind the previous (16 B per
             cell-step + tile and slab overlap); 200 = 66 x 3 + 2: the last two steps run in one
             FdtdTwoStepOverlap task after the loop
  overlap2 -- 1 task per 2 steps, FdtdTwoStepOverlap: two whole Yee steps per pass on overlapped
             60 x 12 column tiles, the second step two planes behind the first (24 B per
             cell-step + tile and slab overlap)
  overlap -- 1 task/step, FdtdStepOverlap8: whole Yee step on overlapped 62 x 6 column tiles,
             one LDS barrier per plane (48 B per cell-step + tile overlap)
  halves  -- 2 tasks/step: FdtdHFields, FdtdEFields (72 B per cell-step)
  six     -- the reference's six tasks/step (96 B per cell-step)
  fused   -- 1 task/step, FdtdStepFused: whole step, halo recomputed by edge lanes (48 B)

The same line also carries, under "workloads", the other graphs of the hot path:
  captured -- the same 384^3 graph as one instance of the batched executor: plain per-task
             dispatch of the six reference tasks vs one hipGraph-captured iteration per step
  batched -- 2048 FDTD graphs (edge U{16..96}, 50 steps) + 512 groupby graphs (1M-8M rows, Zipf 1.1
             over 65,536 keys) through the batched executor, plain per-task dispatch vs captured
             batched iterations (instances/s, both on the same instances, result checksums compared)
  served  -- Poisson arrivals of FDTD / GMM / groupby graph instances at a few rates: p50/p99
             latency and graphs/s, continuous batching vs plain dispatch, alone-vs-mix identity
  gmm     -- EM-GMM N=2M, D=32, K=64, 25 EM steps (points x steps per second)
  pca     -- Gram-Schmidt PCA 65536 x 2048, 16 components, 10 power iterations

Prints one JSON line.  `--gpus N` with N > 1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0   # MI355X published HBM3E bandwidth
METRIC = "FDTD Mcell-updates/s; served graphs/s + p50/p99 latency vs plain dispatch"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed 200-time-step graph runs")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--edge", type=int, default=384)
    ap.add_argument("--fdtd-steps", type=int, default=200)
    ap.add_argument("--forms", default="overlap3p,overlap3,overlap2,overlap,halves,six,fused")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--suite", default="captured,batched,served,gmm,pca",
                    help="extra workloads after the FDTD headline (comma list; empty for none)")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-qps", default="20,100,300")
    return ap.parse_args()


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def field_elems(n: int) -> list[int]:
    from ptask_amd import fdtd as FD
    return [FD.elems(f, n, n, n) for f in range(6)]


def algorithmic_bytes(task: str, n: int) -> int:
    """HBM bytes one launch of `task` must move: each input array read once, each output array
    written once (binary32)."""
    E = field_elems(n)
    whole = ("FdtdStepFused", "FdtdStepOverlap", "FdtdStepOverlap8", "FdtdTwoStepOverlap",
             "FdtdThreeStepOverlap", "FdtdThreeStepPersistent")
    reads = {"FdtdHFields": range(6), "FdtdEFields": range(6), **{k: range(6) for k in whole}}
    writes = {"FdtdHFields": (0, 1, 2), "FdtdEFields": (3, 4, 5), **{k: range(6) for k in whole}}
    six = {"HxComputation": (0, (0, 4, 5)), "HyComputation": (1, (1, 3, 5)),
           "HzComputation": (2, (2, 3, 4)), "ExComputation": (3, (3, 1, 2)),
           "EyComputation": (4, (4, 0, 2)), "EzComputation": (5, (5, 0, 1))}
    base = task.split("_")[0]
    if base in six:
        own, ins = six[base]
        return 4 * (sum(E[f] for f in ins) + E[own])
    return 4 * (sum(E[f] for f in reads[base]) + sum(E[f] for f in writes[base]))


# Time steps one launch of a task kernel advances the fields by (1 unless listed).
STEPS_PER_LAUNCH = {"FdtdTwoStepOverlap": 2, "FdtdThreeStepOverlap": 3, "FdtdThreeStepPersistent": 3}


def run_graph(form: str, n: int, steps: int, repeat: int, warmup: int) -> dict:
    from ptask_amd import paths
    kernels = paths.kernel_path("fdtd.hsaco" if form == "six" else "fdtd_fused.hsaco")
    cmd = [paths.bin_path("ptk_fdtd"), "--kernels", kernels, "--form", form, "--n", str(n),
           "--steps", str(steps), "--repeat", str(repeat), "--warmup", str(warmup), "--chain",
           "--profile"]
    progress("run: " + " ".join(cmd))
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        raise RuntimeError(f"ptk_fdtd failed ({p.returncode}): {p.stderr[-2000:]}")
    r = json.loads(p.stdout.strip().splitlines()[-1])
    progress(f"{form}: {r['mcups']:.1f} Mcell-updates/s, {r['ms_per_run']:.2f} ms per "
             f"{steps}-step run ({time.time() - t0:.1f} s wall incl. setup)")
    return r


def kernel_roofline(r: dict, n: int) -> dict:
    """The dominant task kernel of a run: average launch duration from the HIP events the
    runtime recorded on the task's stream during the timed runs."""
    tasks = [t for t in r["tasks"] if t["dispatches"] > 0]
    dom = max(tasks, key=lambda t: t["device_ms"])
    avg_ms = dom["device_ms"] / dom["dispatches"]
    alg = algorithmic_bytes(dom["task"], n)
    achieved = alg / (avg_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom["task"],
            "kernel_avg_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": alg,
            "time_steps_per_launch": STEPS_PER_LAUNCH.get(dom["task"].split("_")[0], 1),
            "launches_timed": int(dom["dispatches"])}


def form_roofline(r: dict, n: int, form: str) -> dict:
    roof = kernel_roofline(r, n)
    roof["traffic"] = measured_traffic(roof["kernel"], n, form)
    return roof


def run_app(name: str, args: list, timeout: int = 600) -> dict:
    from ptask_amd import paths
    cmd = [paths.bin_path(name), *map(str, args)]
    progress("run: " + " ".join(cmd))
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=REPO)
    if p.returncode != 0:
        raise RuntimeError(f"{name} failed ({p.returncode}): {p.stderr[-2000:]}{p.stdout[-2000:]}")
    r = json.loads(p.stdout.strip().splitlines()[-1])
    progress(f"{name}: done in {time.time() - t0:.1f} s")
    return r


def measured_traffic(kernel: str, n: int, form: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary of its form (rocprofv3 --pmc
    runs of FETCH_SIZE and WRITE_SIZE scaled by a copy of known size, cross-checked with the
    read/write request counts by size), when it covers this workload."""
    path = os.path.join(REPO, "profiles", f"fdtd_384_{form}_hbm_traffic.md")
    if n != 384 or not os.path.exists(path):
        return None
    text = open(path).read()
    block = text[text.index("```json") + 7:]
    data = json.loads(block[:block.index("```")])
    return data.get(kernel.split("_")[0], {}).get("traffic_bytes")


def leg_batched() -> dict:
    from ptask_amd import paths
    r = run_app("ptk_batch", ["--kernels", paths.KERNEL_DIR])
    pl, ba = r["modes"]["plain"], r["modes"]["batched"]
    return {"workload": "2048 FDTD graphs (cubic edge U{16..96}, 50 steps) + 512 groupby graphs "
                        "(rows U[1M,8M], Zipf 1.1 over 65,536 keys), one executor, all resident",
            "unit": "graph instances/s",
            "plain_exec": round(pl["exec_instances_per_s"], 1),
            "batched_exec": round(ba["exec_instances_per_s"], 1),
            "speedup_exec": round(ba["exec_instances_per_s"] / pl["exec_instances_per_s"], 2),
            "plain_end_to_end": round(pl["instances_per_s"], 1),
            "batched_end_to_end": round(ba["instances_per_s"], 1),
            "speedup_end_to_end": round(ba["instances_per_s"] / pl["instances_per_s"], 2),
            "plain_exec_ms": pl["exec_ms"], "batched_exec_ms": ba["exec_ms"],
            "plain_launches": pl["launches"], "batched_graph_launches": ba["graph_launches"],
            "tiles": ba["tiles"], "checksums_equal": r.get("checksums_equal"),
            "footprint_bytes": r["footprint_bytes"]}


def leg_served(seconds: float, rates: list) -> dict:
    from ptask_amd import paths
    runs = []
    for q in rates:
        r = run_app("ptk_serve", ["--kernels", paths.KERNEL_DIR, "--qps", q, "--duration", seconds,
                                  "--identity", 4])
        runs.append({k: r[k] for k in ("mode", "qps", "requests", "graphs_per_s", "p50_ms", "p99_ms",
                                       "p50_ms_by_kind", "p99_ms_by_kind", "mean_active",
                                       "identity_checked", "identity_equal")})
    q = rates[len(rates) // 2]
    r = run_app("ptk_serve", ["--kernels", paths.KERNEL_DIR, "--qps", q, "--duration", seconds,
                              "--identity", 0, "--mode", "plain"])
    runs.append({k: r[k] for k in ("mode", "qps", "requests", "graphs_per_s", "p50_ms", "p99_ms",
                                   "p50_ms_by_kind", "p99_ms_by_kind", "mean_active")})
    return {"workload": "Poisson arrivals; FDTD edge U{32..192} x 20..100 steps, GMM N U[10K,1M] D=16 "
                        "K U{8..32} 20 EM steps, groupby rows U[1M,16M]; mix 0.5:0.25:0.25",
            "latency_kinds": ["fdtd", "gmm", "groupby"], "runs": runs}


def leg_captured(n: int, steps: int) -> dict:
    """The 384^3 FDTD graph as one instance of the batched executor: plain per-task dispatch
    (the six reference Yee tasks launched one by one, every step) against hipGraph-captured
    iterations (one replayed graph per three steps, the three-step kernel on overlapped tiles,
    then one-step graphs for the remaining steps % 3).  Device
    time of the iterations from the executor's events; wall time includes admission (field
    generation on the device) and retirement (result checksums)."""
    from ptask_amd import batch
    out = {"workload": f"FDTD {n}^3 binary32, {steps} time steps, one executor instance",
           "unit": "Mcell-updates/s"}
    sums = {}
    for name, mode in (("plain", batch.PLAIN), ("captured", batch.BATCHED)):
        with batch.Executor(8 << 30, mode) as ex:
            t0 = time.perf_counter()
            i = ex.submit_fdtd(n, n, n, steps, 1, keep=True)
            ex.drain()
            wall = time.perf_counter() - t0
            st = ex.stats()
            sums[name] = ex.checksums(i)
            ex_s = st["exec_us"] / 1e6
            out[name] = {"exec_ms": round(ex_s * 1e3, 3), "value": round(n ** 3 * steps / ex_s / 1e6, 1),
                         "wall_ms": round(wall * 1e3, 3), "launches": st["launches"],
                         "graph_launches": st["graph_launches"],
                         "fdtd_steps_per_launch": ex.fdtd_steps()[0] if mode == batch.BATCHED else 1}
        progress(f"captured leg {name}: {out[name]['value']} Mcell-updates/s")
    out["speedup"] = round(out["captured"]["value"] / out["plain"]["value"], 2)
    out["checksums_equal"] = sums["plain"] == sums["captured"]
    return out


def leg_gmm() -> dict:
    from ptask_amd import paths
    r = run_app("ptk_gmm", ["--kernels", paths.kernel_path("gmm.hsaco"), "--n", 2000000, "--d", 32,
                            "--k", 64, "--steps", 25, "--repeat", 3, "--warmup", 1, "--profile",
                            "--blocks", 512])
    acc = next(t for t in r["tasks"] if t["task"] == "GmmAccumulate")
    avg = acc["device_ms"] / acc["dispatches"]
    flops = 8 * 2000000 * 32 * 64
    return {"workload": "EM-GMM N=2,000,000 D=32 K=64 diagonal, 25 EM steps, looped graph",
            "value": r["point_steps_per_s"], "unit": "point-steps/s", "ms_per_run": r["ms_per_run"],
            "loglike": r["loglike"], "nonpositive_var": r["nonpositive_var"],
            "roofline": {"bound": "valu", "kernel": "GmmAccumulate", "kernel_avg_ms": round(avg, 4),
                         "achieved": round(flops / (avg / 1e3) / 1e12, 2), "peak": 157.3,
                         "unit": "TFLOP/s", "frac": round(flops / (avg / 1e3) / 1e12 / 157.3, 4),
                         "flops_per_launch": flops}}


def leg_pca() -> dict:
    from ptask_amd import paths
    r = run_app("ptk_pca", ["--kernels", paths.kernel_path("pca.hsaco"), "--m", 65536, "--n", 2048,
                            "--k", 16, "--j", 10, "--repeat", 3, "--warmup", 1, "--profile"])
    avg = r["gram_ms"] / max(1, r["gram_dispatches"])
    gbs = r["gram_bytes"] / (avg / 1e3) / 1e9
    return {"workload": "Gram-Schmidt PCA 65536 x 2048 binary32, 16 components, 10 power iterations",
            "value": r["runs_per_s"], "unit": "PCA runs/s", "ms_per_run": r["ms_per_run"],
            "singular_values": r["L"][:4],
            "roofline": {"bound": "hbm", "kernel": "PcaGram", "kernel_avg_ms": round(avg, 5),
                         "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": r["gram_bytes"]}}


def cpu_baseline(n: int, seconds: float) -> dict:
    """The oracle (oracle/fdtd.c, binary32, one core) timed on the same 384^3 grid for as many
    whole time steps as fit in about `seconds`."""
    import oracle
    p = oracle.make_params(n, n, n, 0.5)
    F = oracle.init_fields(n, n, n, seed=1)
    steps, t0 = 0, time.perf_counter()
    cur = F
    while True:
        cur = oracle.fdtd_run(cur, p, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    rate = n ** 3 * steps / el / 1e6
    del cur, F
    return {"value": round(rate, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{n}^3 grid, {steps} time step(s) of the binary32 oracle on one core "
                      f"({el:.1f} s)",
            "host_graph_64": host_graph_baseline()}


def host_graph_baseline(runs: int = 5) -> dict:
    """The plain CPU method: the looped 64^3 FDTD graph, 10 time steps, on the HostAccelerator
    with the oracle's host-task forms of the six Yee tasks; median of `runs` runs with one host
    task at a time and with as many as there are cores (at most 16, the box's CPU share)."""
    import statistics

    import oracle
    from ptask_amd import paths
    lib = os.path.join(oracle.ORACLE_DIR, "_build", "libptk_oracle.so")
    cores = max(1, min(os.cpu_count() or 1, 16))
    out = {"workload": "looped FDTD graph 64^3, 10 steps, six host tasks per step", "runs": runs,
           "unit": "Mcell-updates/s"}
    for label, conc in (("one_core", 1), ("all_cores", cores)):
        cmd = [paths.bin_path("ptk_fdtd"), "--kernels", lib, "--n", "64", "--steps", "10",
               "--repeat", str(runs), "--host-concurrency", str(conc)]
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if p.returncode != 0:
            out[label] = {"error": p.stderr[-300:]}
            continue
        r = json.loads(p.stdout.strip().splitlines()[-1])
        med = statistics.median(r["run_ms"])
        out[label] = {"host_concurrency": conc, "median_ms": round(med, 3),
                      "value": round(64 ** 3 * 10 / (med / 1e3) / 1e6, 2)}
    return out


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the path runs on a "
              "single MI355X (how it would shard across GPUs is a design, not a measured mode)", file=sys.stderr)
        return 2
    from ptask_amd import hip as H
    if H.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    props = H.device_props(0)
    progress(f"device {props.name.decode()} ({props.arch.decode()}), {props.compute_units} CUs")
    n, T = args.edge, args.fdtd_steps
    forms = [f for f in args.forms.split(",") if f]
    results = {f: run_graph(f, n, T, args.steps, args.warmup) for f in forms}
    head_form = max(results, key=lambda f: results[f]["mcups"])
    head = results[head_form]
    roof = form_roofline(head, n, head_form)
    line = {
        "metric": METRIC,
        "value": round(head["mcups"], 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_run"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded U[-1,1) fields from the shared counter-based generator)",
        "config": {"workload": f"FDTD Yee {n}^3 binary32, {T} time steps per step, looped PTask "
                               f"graph, {head_form} form, per-task dispatch by the runtime, "
                               "device-resident datablocks",
                   "grid": [n, n, n], "time_steps": T, "courant": 0.5, "form": head_form,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "forms": {f: {"mcups": round(r["mcups"], 1), "ms_per_run": round(r["ms_per_run"], 3),
                      "tasks_per_step": round(r["dispatches_last_run"] / T, 3),
                      "dispatches_per_run": r["dispatches_last_run"],
                      "roofline": form_roofline(r, n, f)} for f, r in results.items()},
        "checksums_equal_across_forms": len({tuple(c["fnv1a"] for c in r["checksums"])
                                             for r in results.values()}) == 1,
    }
    legs = {"batched": leg_batched, "gmm": leg_gmm, "pca": leg_pca,
            "captured": lambda: leg_captured(n, T),
            "served": lambda: leg_served(args.serve_seconds, args.serve_qps.split(","))}
    line["workloads"] = {}
    for name in [x for x in args.suite.split(",") if x]:
        try:
            line["workloads"][name] = legs[name]()
        except Exception as e:  # a failed leg is reported, the headline still stands
            progress(f"{name} failed: {e}")
            line["workloads"][name] = {"error": str(e)[-500:]}
    if not args.no_cpu:
        progress("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
