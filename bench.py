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
             dispatch of the six reference tasks, the batched kernels launched directly, and the
             same kernels as one hipGraph replay per iteration (capture alone vs plain dispatch)
  batched -- 2048 FDTD graphs (edge U{16..96}, 50 steps) + 512 groupby graphs (1M-8M rows, Zipf 1.1
             over 65,536 keys) through the batched executor, plain per-task dispatch vs captured
             batched iterations (instances/s, both on the same instances, result checksums
             compared; tile shapes from an untimed autotuning pass)
  served  -- Poisson arrivals of FDTD / GMM / groupby graph instances, both modes swept up to
             saturation: p50/p99 latency, graphs/s, capacity and rate within a p99 bound,
             continuous batching vs plain dispatch, every request's result identical alone
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
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-qps", default="25,50,100,200,400,800,1600,3200")
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


# Algorithmic HBM bytes per cell-step of each graph form: every array its tasks read once and
# every array they write once, binary32, per time step.
FORM_BYTES_PER_CELL_STEP = {"six": 96, "halves": 72, "overlap": 48, "fused": 48, "overlap2": 24,
                            "overlap3": 16, "overlap3p": 16}


def form_roofline(r: dict, n: int, form: str) -> dict:
    """The dominant kernel's roofline, plus the form's own rate: cell-steps per second of the whole
    graph times the form's algorithmic bytes per cell-step (for the six-task form, whose three H
    and three E tasks run concurrently on their own streams, this is the meaningful figure: each
    kernel's duration is stretched by its concurrent siblings)."""
    roof = kernel_roofline(r, n)
    roof["traffic"] = measured_traffic(roof["kernel"], n, form)
    b = FORM_BYTES_PER_CELL_STEP.get(form)
    if b:
        gbs = r["mcups"] * 1e6 * b / 1e9
        roof["graph"] = {"bytes_per_cell_step": b, "achieved": round(gbs, 1),
                         "frac": round(gbs / HBM_PEAK_GBS, 4)}
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


def hbm_ceiling() -> dict:
    """The measured streaming ceiling of this GPU beside the 8 TB/s peak: PtkStreamMix (six arrays
    read and three written with 16-byte accesses, the access mix of a half-step FDTD kernel
    without its stencil), PtkReadWords (reads only) and PtkCopyWords (copy), HIP-event timed."""
    from ptask_amd import paths
    out = {"unit": "GB/s", "peak": HBM_PEAK_GBS}
    for mode, mib, grid in (("mix", 217, 1024), ("read", 1024, 4096), ("copy", 1024, 1024)):
        r = run_app("ptk_copy", ["--kernels", paths.KERNEL_DIR, "--mode", mode, "--mib", mib,
                                 "--repeat", 10, "--grid", grid])
        out[mode] = r.get("gbs", r.get("read_plus_write_gbs"))
    return out


def leg_batched() -> dict:
    from ptask_amd import paths
    r = run_app("ptk_batch", ["--kernels", paths.KERNEL_DIR])
    pl, ba = r["modes"]["plain"], r["modes"]["batched"]
    return {"workload": "2048 FDTD graphs (cubic edge U{16..96}, 50 steps) + 512 groupby graphs "
                        "(rows U[1M,8M], Zipf 1.1 over 65,536 keys), one executor, all resident",
            "unit": "graph instances/s", "target_speedup": 5.0,
            "autotune": "separate untimed pass over the same mix; its choice is applied to the timed run",
            "autotune_pass_ms": r.get("autotune_pass_ms"), "tuning": r.get("tuning"),
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


SERVE_P99_BOUND_MS = 1000.0


def serve_run(mode: str, qps: float, seconds: float, identity: int = 0) -> dict:
    from ptask_amd import paths
    r = run_app("ptk_serve", ["--kernels", paths.KERNEL_DIR, "--qps", qps, "--duration", seconds,
                              "--identity", identity, "--mode", mode])
    keep = ("mode", "qps", "requests", "graphs_per_s", "offered_per_s", "p50_ms", "p99_ms",
            "p50_ms_by_kind", "p99_ms_by_kind", "mean_active", "identity_checked", "identity_equal")
    return {k: r[k] for k in keep if k in r}


def leg_served(seconds: float, rates: list) -> dict:
    """Both modes swept upwards in offered load until they saturate (graphs/s below 85% of the
    offered rate, or p99 above 10 s).  Per mode: the service capacity (largest graphs/s any run
    sustained) and the largest offered rate served within the p99 bound with at least 95% of the
    offered rate completed; their ratio batched / plain is the served speedup."""
    out = {"workload": "Poisson arrivals; FDTD edge U{32..192} x 20..100 steps, GMM N U[10K,1M] D=16 "
                       "K U{8..32} 20 EM steps, groupby rows U[1M,16M]; mix 0.5:0.25:0.25",
           "latency_kinds": ["fdtd", "gmm", "groupby"], "p99_bound_ms": SERVE_P99_BOUND_MS,
           "target_speedup": 5.0, "modes": {}}
    for mode in ("batched", "plain"):
        runs = []
        for q in rates:
            r = serve_run(mode, float(q), seconds)
            runs.append(r)
            progress(f"served {mode} {q} QPS: {r['graphs_per_s']:.1f} graphs/s, p99 {r['p99_ms']:.1f} ms")
            if r["graphs_per_s"] < 0.85 * r["offered_per_s"] or r["p99_ms"] > 10000:
                break
        ok = [r for r in runs if r["p99_ms"] <= SERVE_P99_BOUND_MS and r["graphs_per_s"] >= 0.95 * r["offered_per_s"]]
        out["modes"][mode] = {"runs": runs,
                              "capacity_graphs_per_s": round(max(r["graphs_per_s"] for r in runs), 2),
                              "max_qps_within_p99": max((r["qps"] for r in ok), default=0.0)}
    b, p = out["modes"]["batched"], out["modes"]["plain"]
    out["capacity_speedup"] = round(b["capacity_graphs_per_s"] / max(p["capacity_graphs_per_s"], 1e-9), 2)
    out["qps_within_p99_speedup"] = (round(b["max_qps_within_p99"] / p["max_qps_within_p99"], 2)
                                     if p["max_qps_within_p99"] else None)
    ident = serve_run("batched", float(rates[min(2, len(rates) - 1)]), seconds, identity=-1)
    out["identity"] = {"qps": ident["qps"], "requests": ident["requests"],
                       "checked": ident["identity_checked"], "equal": ident["identity_equal"]}
    return out


def leg_captured(n: int, steps: int) -> dict:
    """The 384^3 FDTD graph as one instance of the batched executor, three ways:
      plain    -- per-task dispatch of the six reference Yee tasks, one launch each, every step;
      direct   -- the batched iteration's kernels (three steps per launch on overlapped tiles, then
                  one-step launches for steps % 3) launched directly, one launch per kernel;
      captured -- the same kernels as one hipGraph replay per iteration.
    direct vs captured is the effect of capture alone; plain vs captured combines it with the
    fused kernels.  Device time of the iterations from the executor's events; wall time includes
    admission (field generation on the device) and retirement (result checksums)."""
    from ptask_amd import batch
    out = {"workload": f"FDTD {n}^3 binary32, {steps} time steps, one executor instance",
           "unit": "Mcell-updates/s"}
    sums = {}
    for name, mode, graphs in (("plain", batch.PLAIN, False), ("direct", batch.BATCHED, False),
                               ("captured", batch.BATCHED, True)):
        with batch.Executor(8 << 30, mode) as ex:
            ex.set_launch(graphs)
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
    out["speedup_capture_alone"] = round(out["captured"]["value"] / out["direct"]["value"], 3)
    out["speedup_vs_plain"] = round(out["captured"]["value"] / out["plain"]["value"], 2)
    out["checksums_equal"] = sums["plain"] == sums["captured"] == sums["direct"]
    out["ptaskapi"] = leg_captured_ptaskapi(n)
    return out


def leg_captured_ptaskapi(n: int) -> dict:
    """The reference's six-task graph built through ptaskapi (unrolled, graphfdtd.cpp:735-739), run
    three times: task-by-task dispatch by the runtime vs the runtime's launch capture (the first
    run recorded, the later runs replayed as one hipGraph launch each; Graph::EnableLaunchCapture).
    Same kernels both ways, so the ratio is the effect of capture alone.  Chained runs: each run
    starts from the previous run's device-resident outputs, and the timed region covers three runs
    after one warm-up run (recorded, under capture) from a drained device to completion."""
    from ptask_amd import paths
    res = {}
    for edge, steps in ((n, 20), (32, 50)):
        base = ["--kernels", paths.kernel_path("fdtd.hsaco"), "--form", "six", "--mode", "unrolled",
                "--n", edge, "--steps", steps, "--repeat", 3, "--warmup", 1, "--chain", "--seed", 3]
        plain, cap = run_app("ptk_fdtd", base), run_app("ptk_fdtd", base + ["--capture"])
        res[f"{edge}^3x{steps}"] = {
            "plain_ms_per_run": plain["ms_per_run"], "captured_ms_per_run": cap["ms_per_run"],
            "plain_mcups": round(edge ** 3 * steps / (plain["ms_per_run"] / 1e3) / 1e6, 1),
            "captured_mcups": round(edge ** 3 * steps / (cap["ms_per_run"] / 1e3) / 1e6, 1),
            "speedup": round(plain["ms_per_run"] / cap["ms_per_run"], 3),
            "tasks_per_run": plain["dispatches_last_run"], "captured_dispatches_per_run": cap["dispatches_last_run"],
            "replays": cap["capture"]["replays"],
            "checksums_equal": [c["fnv1a"] for c in plain["checksums"]] == [c["fnv1a"] for c in cap["checksums"]]}
        progress(f"ptaskapi capture {edge}^3 x {steps}: plain {plain['ms_per_run']} ms, captured {cap['ms_per_run']} ms")
    return res


def leg_gmm() -> dict:
    from ptask_amd import paths
    base = ["--kernels", paths.kernel_path("gmm.hsaco"), "--n", 2000000, "--d", 32, "--k", 64, "--steps", 25,
            "--repeat", 3, "--warmup", 1, "--profile", "--blocks", 512]
    flops = 8 * 2000000 * 32 * 64  # E-step 4 and M-step 4 flops per (point, component, dimension)
    forms = {}
    for acc in ("mfma", "valu"):
        r = run_app("ptk_gmm", base + ["--accumulate", acc])
        t = next(x for x in r["tasks"] if x["task"] == "GmmAccumulate")
        avg = t["device_ms"] / t["dispatches"]
        forms[acc] = {"run": r, "avg_ms": avg}
    r, avg = forms["mfma"]["run"], forms["mfma"]["avg_ms"]
    # fp32 matrix peak of MI355X; the VALU form without packed FP32 has half of it
    return {"workload": "EM-GMM N=2,000,000 D=32 K=64 diagonal, 25 EM steps, looped graph",
            "value": r["point_steps_per_s"], "unit": "point-steps/s", "ms_per_run": r["ms_per_run"],
            "loglike": r["loglike"], "nonpositive_var": r["nonpositive_var"],
            "valu_form": {"kernel": "GmmAccumulate", "kernel_avg_ms": round(forms["valu"]["avg_ms"], 4),
                          "ms_per_run": forms["valu"]["run"]["ms_per_run"],
                          "loglike": forms["valu"]["run"]["loglike"]},
            "roofline": {"bound": "mfma", "kernel": "GmmAccumulateMfma", "kernel_avg_ms": round(avg, 4),
                         "achieved": round(flops / (avg / 1e3) / 1e12, 2), "peak": 157.3,
                         "unit": "TFLOP/s", "frac": round(flops / (avg / 1e3) / 1e12 / 157.3, 4),
                         "flops_per_launch": flops}}


def leg_pca() -> dict:
    """The unrolled PCA graph (498 tasks per run) dispatched task by task and with launch capture,
    each timed without per-task profiling; a third, profiled run gives PcaGram's kernel time."""
    from ptask_amd import paths
    base = ["--kernels", paths.kernel_path("pca.hsaco"), "--m", 65536, "--n", 2048, "--k", 16, "--j", 10,
            "--repeat", 3, "--warmup", 1]
    plain = run_app("ptk_pca", base)
    cap = run_app("ptk_pca", base + ["--capture"])
    prof = run_app("ptk_pca", base + ["--profile"])
    if (plain["T_fnv1a"], plain["S_fnv1a"]) != (cap["T_fnv1a"], cap["S_fnv1a"]):
        raise RuntimeError("captured PCA runs differ from per-task dispatch")
    best = cap if cap["capture"]["armed"] and cap["ms_per_run"] < plain["ms_per_run"] else plain
    avg = prof["gram_ms"] / max(1, prof["gram_dispatches"])
    gbs = prof["gram_bytes"] / (avg / 1e3) / 1e9
    return {"workload": "Gram-Schmidt PCA 65536 x 2048 binary32, 16 components, 10 power iterations",
            "value": best["runs_per_s"], "unit": "PCA runs/s", "ms_per_run": best["ms_per_run"],
            "dispatch": "launch capture" if best is cap else "per task",
            "per_task_ms_per_run": plain["ms_per_run"], "captured_ms_per_run": cap["ms_per_run"],
            "capture_replays": cap["capture"]["replays"], "results_equal": True,
            "singular_values": best["L"][:4],
            "roofline": {"bound": "hbm", "kernel": "PcaGram", "kernel_avg_ms": round(avg, 5),
                         "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": prof["gram_bytes"]}}


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
            "cpu_model": cpu_model(), "cpus_available": len(os.sched_getaffinity(0)),
            "sample": f"{n}^3 grid, {steps} time step(s) of the binary32 oracle on one core "
                      f"({el:.1f} s)",
            "host_graph_64": host_graph_baseline()}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
    try:
        ceiling = hbm_ceiling()
        roof["measured_ceiling"] = ceiling
    except Exception as e:  # reported, the headline still stands
        progress(f"hbm ceiling probe failed: {e}")
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
