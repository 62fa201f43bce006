#!/usr/bin/env python3
"""Benchmark: FDTD Mcell-updates/s of the PTask Yee graph on one MI355X.

Workload: a 384^3 binary32 Yee grid (six field arrays, 1.36 GB) advanced 200 time steps, fields
seeded U[-1,1) on the device, Courant number 0.5, faces copied through.  One benchmark "step"
is one full 200-time-step run (6 task kernels per time step).  Inputs are resident in HBM
before timing starts; nothing crosses PCIe inside the timed region.

Prints one JSON line (see the repository's benchmark contract).  `--gpus N` with N > 1 exits:
this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0   # MI355X published HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed 200-time-step runs")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--edge", type=int, default=384)
    ap.add_argument("--fdtd-steps", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class FdtdDevice:
    """Six resident field arrays plus their out-of-place partners, and the six task kernels."""

    def __init__(self, H, FD, n: int, seed: int):
        from ptask_amd import paths
        self.H, self.FD, self.n = H, FD, n
        self.mod = H.Module(paths.kernel_path("fdtd.hsaco"))
        self.stream = H.Stream()
        self.params = FD.make_params(n, n, n, 0.5)
        self.cur, self.nxt = [], []
        init = self.mod.function("FdtdInitField")
        for f in range(6):
            cnt = FD.elems(f, n, n, n)
            a, b = H.DeviceBuffer(4 * cnt), H.DeviceBuffer(4 * cnt)
            args = H.ArgPack().ptr(a.ptr).u64(cnt).i32(f).u64(seed)
            H.launch(init, (2048, 1, 1), (256, 1, 1), args, self.stream)
            self.cur.append(a)
            self.nxt.append(b)
        self.stream.synchronize()
        self.fns = [self.mod.function(t[0]) for t in FD.TASKS]
        self.geo = [FD.default_geometry(*FD.extent(t[1], n, n, n)) for t in FD.TASKS]
        # pre-packed argument buffers for both ping-pong parities
        self.packs = {}
        for parity in (0, 1):
            cur, nxt = (self.cur, self.nxt) if parity == 0 else (self.nxt, self.cur)
            for t, (_, own, a, b) in enumerate(FD.TASKS):
                # H tasks read the old E from `cur`; E tasks read the new H from `nxt`
                ins = nxt if t >= 3 else cur
                args = FD.pack_task_args(H, cur[own].ptr, ins[a].ptr, ins[b].ptr, self.params,
                                         nxt[own].ptr)
                self.packs[(parity, t)] = args.finish()
        self.parity = 0

    def launch_task(self, t: int) -> None:
        import ctypes as C
        cbuf, nbytes = self.packs[(self.parity, t)]
        (gx, gy, gz), (bx, by, bz) = self.geo[t]
        self.H.check(self.H.lib().ptk_hip_launch(self.fns[t], gx, gy, gz, bx, by, bz, 0,
                                                 self.stream.handle, C.cast(cbuf, C.c_void_p),
                                                 nbytes), "launch")

    def time_step(self) -> None:
        for t in range(6):
            self.launch_task(t)
        self.parity ^= 1

    def run(self, steps: int) -> None:
        for _ in range(steps):
            self.time_step()


def measure_kernels(dev: FdtdDevice, H, steps: int) -> dict:
    """Average duration of each task kernel, from HIP events on the launch stream."""
    evs = [[(H.Event(), H.Event()) for _ in range(6)] for _ in range(steps)]
    for s in range(steps):
        for t in range(6):
            evs[s][t][0].record(dev.stream)
            dev.launch_task(t)
            evs[s][t][1].record(dev.stream)
        dev.parity ^= 1
    dev.stream.synchronize()
    out = {}
    for t, (name, own, _, _) in enumerate(dev.FD.TASKS):
        ms = [evs[s][t][0].elapsed_ms(evs[s][t][1]) for s in range(steps)]
        out[name] = sum(ms) / len(ms)
    for row in evs:
        for a, b in row:
            a.destroy()
            b.destroy()
    return out


def cpu_baseline(n: int, seconds: float) -> dict:
    """The oracle (oracle/fdtd.c, binary32, one core) timed on the same 384^3 grid for as many
    whole time steps as fit in about `seconds`."""
    import numpy as np
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
                      f"({el:.1f} s)"}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(every GPU box here has a single MI355X)", file=sys.stderr)
        return 2
    from ptask_amd import hip as H
    from ptask_amd import fdtd as FD
    if H.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    H.set_device(0)
    props = H.device_props(0)
    n, T = args.edge, args.fdtd_steps
    progress(f"device {props.name.decode()} ({props.arch.decode()}), {props.compute_units} CUs")
    dev = FdtdDevice(H, FD, n, args.seed)
    progress(f"fields resident: {n}^3, {sum(b.nbytes for b in dev.cur) / 1e9:.3f} GB x2")

    for w in range(args.warmup):
        dev.run(T)
        dev.stream.synchronize()
        progress(f"warmup {w + 1}/{args.warmup} done")

    e0, e1 = H.Event(), H.Event()
    H.synchronize()
    t0 = time.perf_counter()
    e0.record(dev.stream)
    for s in range(args.steps):
        dev.run(T)
    e1.record(dev.stream)
    e1.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1)
    H.synchronize()
    progress(f"timed {args.steps} x {T} time steps: {gpu_ms:.1f} ms (wall {wall * 1e3:.1f} ms)")

    cells = n ** 3
    ms_per_step = gpu_ms / args.steps
    mcups = cells * T * args.steps / (gpu_ms / 1e3) / 1e6

    kt = measure_kernels(dev, H, 10)
    dom = max(kt, key=kt.get)
    t_dom = [t for t in FD.TASKS if t[0] == dom][0]
    alg_bytes = 16 * FD.elems(t_dom[1], n, n, n)   # 3 reads + 1 write of binary32 per cell
    achieved = alg_bytes / (kt[dom] / 1e3) / 1e9
    six_task_gbs = 96 * cells * T * args.steps / (gpu_ms / 1e3) / 1e9
    progress("kernel ms: " + ", ".join(f"{k}={v:.4f}" for k, v in kt.items()))

    line = {
        "metric": "FDTD Mcell-updates/s; served graphs/s + p50/p99 latency vs plain dispatch",
        "value": round(mcups, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded U[-1,1) fields generated on the device)",
        "config": {"workload": f"FDTD Yee {n}^3 binary32, {T} time steps, six-task form, "
                               "plain per-task dispatch through the HIP C ABI",
                   "grid": [n, n, n], "time_steps": T, "courant": 0.5,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": dom,
                     "kernel_avg_ms": round(kt[dom], 5),
                     "algorithmic_bytes_per_launch": alg_bytes},
        "six_task_effective_GBs": round(six_task_gbs, 1),
        "kernel_avg_ms": {k: round(v, 5) for k, v in kt.items()},
    }
    if not args.no_cpu:
        progress("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
