#!/usr/bin/env python3
"""Throughput of the 2D per-cell phase-field step on one MI355X.

Workload: periodic 2048x2048 lattice, 1024 cells R=20 on a seeded jittered
32x32 grid (jitter +-4, Philox seed 7), fp32 phi, Q5.27 aggregate, default
coefficients. One step = one persistent launch that advances all 1024 cells
(tumble, interaction integral, velocity, stencil update, shifted store,
Q5.27 scatter, moments). Inputs are resident in HBM before timing; steps are
replayed from six-step HIP graphs.

Prints one JSON line: whole-job cell-steps/s, the HBM roofline of the step
kernel from its algorithmic bytes (per cell-step 12*W + 8*Q + 4*P*L/N: read
the phi window, read the S window, write the phi window, atomic
read-modify-write of S where q != 0, and the amortized clear of one S slot),
and the CPU oracle's throughput on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
for _p in (REPO, REPO / "phase-field-gh200_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0          # MI355X published HBM3E bandwidth
TRAFFIC_SUMMARY = REPO / "profiles" / "pf2d_step_L2048_N1024_traffic.md"


def q_nonzero_count(sim) -> int:
    """Pixels with q(phi) != 0 over every cell's active window (GPU count)."""
    import torch
    from pfh import abi
    phi = sim.phi[sim.steps_done % 2].view(sim.N, abi.TILE_PITCH, abi.TILE_PITCH)
    cls = sim.cls.to(torch.int64).cpu().numpy()
    total = 0
    for c in set(int(v) for v in cls):
        wx, wy, tx0, ty0 = abi.CLASSES[c]
        idx = torch.nonzero(sim.cls == c).flatten()
        win = phi[idx, ty0:ty0 + wy, tx0:tx0 + wx]
        q = torch.round(torch.clamp(win * win, max=4.0) * 134217728.0)
        total += int((q != 0).sum().item())
    return total


def window_area(sim) -> int:
    from pfh import abi
    cls = sim.classes()
    return int(sum(abi.CLASSES[int(c)][0] * abi.CLASSES[int(c)][1] for c in cls))


def read_traffic() -> float | None:
    """HBM bytes per step launch measured with rocprofv3 counters (FETCH_SIZE
    + WRITE_SIZE, calibrated on a copy kernel), when a summary is committed."""
    if not TRAFFIC_SUMMARY.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", TRAFFIC_SUMMARY.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(wl, budget_s: float = 12.0) -> dict:
    """The CPU oracle (binary32, the kernel's operation order) on the same
    workload, timed on this host's cores after two warm-up steps."""
    from oracle import pfo
    threads = min(16, os.cpu_count() or 1)
    orc = pfo.OracleSim2D(wl, threads=threads)
    orc.step(2)
    t0 = time.perf_counter()
    steps = 0
    while True:
        orc.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    return {"value": wl.params.num_cells * steps / el, "unit": "cell-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} steps x {wl.params.num_cells} cells of the same workload "
                      f"after 2 warm-up steps ({el:.1f} s wall)"}


def calibrate_copy(mib: int) -> int:
    import ctypes as C

    import torch
    from pfh import abi
    lib = abi.load()
    n = mib * 1024 * 1024
    src = torch.ones(n // 4, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(8):
        abi.check(lib.pfh_stream_copy(dst.data_ptr(), src.data_ptr(), n, s), "pfh_stream_copy")
    torch.cuda.synchronize()
    assert bool((dst == 1).all())
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 8}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200)
    ap.add_argument("--warmup", type=int, default=120)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run 8 plain copies of MIB MiB "
                         "(the known-size kernel that scales hardware byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the per-cell step does not shard across GPUs in this build)",
              file=sys.stderr)
        return 2

    import torch
    from pfh import params, sim2d

    wl = params.workload_c2()
    sim = sim2d.Sim2D(wl, device="cuda:0")
    sim.capture_graph()
    warm = ((args.warmup + 5) // 6) * 6
    steps = max(6, (args.steps // 6) * 6)
    sim.run(warm)
    sim.sync()
    q0 = q_nonzero_count(sim)
    W = window_area(sim)

    stream = sim.stream
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    sim.run(steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    q1 = q_nonzero_count(sim)
    if sim.fatal():
        print(json.dumps({"error": "fatal integrity flags", "flags": sim.flag_counts().tolist()}))
        return 1

    N = sim.N
    ms_per_step = ms / steps
    value = N * steps / (ms / 1e3)
    Qbar = 0.5 * (q0 + q1)
    bytes_per_launch = 12.0 * W + 8.0 * Qbar + 4.0 * sim.P * sim.L
    achieved = bytes_per_launch / (ms_per_step * 1e-3) / 1e9
    traffic = read_traffic()
    out = {
        "metric": "cell-steps/s, 2D 1024 cells: plain vs surrogate at <=1e-3 L2; HBM roofline %",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "lattice": [sim.L, sim.L], "cells": N,
                   "radius": wl.params.target_radius, "placement": "32x32 grid, jitter +-4",
                   "philox_seed": wl.params.seed, "method": "plain",
                   "graph_steps": sim2d.GRAPH_BODY},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "window_pixels_per_launch": W, "q_nonzero_per_launch": Qbar},
        "surrogate": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
