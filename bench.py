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

The "pf3d" object beside it measures the 3D periodic step on 256^3 with 128
cells (R=12, lambda=3.5, brick 56) the same way, with algorithmic bytes per
cell-step 16*B^3 + 8*Q3 + 4*P*L^2/N (two phi reads, one phi write and one S
read over the brick, the Q5.27 atomics, the amortized S clear).
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
BF16_PEAK_TFLOPS = 2500.0       # MI355X published dense BF16 MFMA rate
TRAFFIC_SUMMARY = REPO / "profiles" / "pf2d_step_L2048_N1024_traffic.md"
TRAFFIC_SUMMARY_3D = REPO / "profiles" / "pf3d_step_L256_N128_B56_traffic.md"
TRAFFIC_SUMMARY_BATCHED = REPO / "profiles" / "pf2d_batched_64replicas_kernel_trace.md"


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


def read_traffic(path: Path = TRAFFIC_SUMMARY, key: str = "traffic_bytes_per_launch"
                 ) -> float | None:
    """HBM bytes per step measured with rocprofv3 counters (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy kernel), when a summary is committed."""
    if not path.exists():
        return None
    m = re.search(key + r":\s*([0-9.eE+]+)", path.read_text())
    return float(m.group(1)) if m else None


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(wl, threads: int, budget_s: float, max_steps: int) -> tuple[float, int, float]:
    from oracle import pfo
    orc = pfo.OracleSim2D(wl, threads=threads)
    orc.step(2)
    t0 = time.perf_counter()
    steps = 0
    while True:
        orc.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= max_steps:
            break
    return wl.params.num_cells * steps / el, steps, el


def cpu_baseline(wl, budget_s: float = 12.0) -> dict:
    """The CPU oracle (binary32, the kernel's operation order) on the same
    workload, timed on this host's cores after two warm-up steps; beside it
    the same on one thread and the 384^2 / 16-cell workload's 1000 steps."""
    from pfh import params
    threads = min(16, os.cpu_count() or 1)
    rate, steps, el = _oracle_rate(wl, threads, budget_s, 200)
    rate1, steps1, el1 = _oracle_rate(wl, 1, 0.4 * budget_s, 20)
    c1 = params.workload_c1()
    from oracle import pfo
    orc = pfo.OracleSim2D(c1, threads=threads)
    t0 = time.perf_counter()
    orc.step(1000)
    el_c1 = time.perf_counter() - t0
    return {"value": rate, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{steps} steps x {wl.params.num_cells} cells of the same workload "
                      f"after 2 warm-up steps ({el:.1f} s wall)",
            "one_thread": {"value": rate1, "unit": "cell-steps/s",
                           "sample": f"{steps1} steps x {wl.params.num_cells} cells, 1 thread "
                                     f"({el1:.1f} s wall)"},
            "c1_full": {"workload": c1.name, "steps": 1000, "wall_s": el_c1,
                        "value": c1.params.num_cells * 1000 / el_c1, "unit": "cell-steps/s",
                        "cores": threads}}


def q_nonzero_3d(sim) -> int:
    """Voxels with q(phi) != 0 over every brick (GPU count)."""
    import torch
    phi = sim.phi[sim.steps_done % 2]
    q = torch.round(torch.clamp(phi * phi, max=4.0) * 134217728.0)
    return int((q != 0).sum().item())


def cpu_baseline_3d(wl, shards: int, budget_s: float = 10.0) -> dict:
    """The 3D CPU oracle (binary32, the kernels' operation order) on the same
    workload, timed on this host's cores after one warm-up step."""
    from oracle import pfo3
    threads = min(16, os.cpu_count() or 1)
    p = wl.params
    orc = pfo3.OracleSim3D(p, wl.cx, wl.cy, wl.cz, shards=shards, threads=threads)
    orc.step(1)
    t0 = time.perf_counter()
    steps = 0
    while True:
        orc.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 100:
            break
    return {"value": p.num_cells * steps / el, "unit": "cell-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} steps x {p.num_cells} cells of the same workload after "
                      f"1 warm-up step ({el:.1f} s wall)"}


def bench_3d(steps: int, warmup: int, with_cpu: bool) -> dict:
    """Throughput of the 3D periodic step (measure + update + finish launches,
    replayed from two-step HIP graphs) on the 256^3 / 128-cell workload."""
    import torch
    from pfh import params3d, sim3d
    wl = params3d.workload_c3()
    sim = sim3d.Sim3D(wl, device="cuda:0")
    sim.capture_graph()
    warm = max(2, (warmup // 2) * 2)
    steps = max(2, (steps // 2) * 2)
    sim.run(warm)
    sim.sync()
    q0 = q_nonzero_3d(sim)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(sim.stream)
    sim.run(steps)
    e1.record(sim.stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    q1 = q_nonzero_3d(sim)
    if sim.fatal():
        return {"error": "fatal integrity flags", "flags": sim.flag_counts().tolist()}
    N, B, L, P = sim.N, sim.B, sim.L, sim.P
    ms_per_step = ms / steps
    Qbar = 0.5 * (q0 + q1)
    bytes_per_step = 16.0 * N * B ** 3 + 8.0 * Qbar + 4.0 * P * L * L
    achieved = bytes_per_step / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": "cell-steps/s, 3D periodic 128 cells",
        "value": N * steps / (ms / 1e3),
        "unit": "cell-steps/s",
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms_per_step,
        "dtype": "f32",
        "config": {"workload": wl.name, "lattice": [L, L, L], "cells": N,
                   "radius": wl.params.target_radius, "lambda": wl.params.lambda_,
                   "brick_edge": B, "automatic_brick_edge": wl.params.automatic_brick_edge(),
                   "placement": "sequential, min separation R, seed 1234",
                   "measure_shards": sim.shards, "update_slabs": sim.slabs,
                   "graph_steps": sim3d.GRAPH_BODY},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(TRAFFIC_SUMMARY_3D, "traffic_bytes_per_step"),
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, "
                                       "calibrated on a copy kernel, committed in profiles/"
                                       + TRAFFIC_SUMMARY_3D.name + " (not measured in this run)",
                     "algorithmic_bytes_per_step": bytes_per_step,
                     "q_nonzero_per_step": Qbar},
    }
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_3d(wl, sim.shards)
    return out


def bench_surrogate(blocks: int, path=None) -> dict | None:
    """Surrogate + fallback against the plain method on the same 2048^2 /
    1024-cell state (blocks of the weights' K steps, check rule and
    tolerance of pfh.surrogate.SurrogateRunner), with committed weights."""
    from pfh import surrogate as su
    from pfh import surrogate_train as st
    path = su.WEIGHTS_PATH if path is None else path
    if not path.exists():
        return {"error": f"no weights at {path.name}"}
    net = su.Net.load(path)
    r = st.evaluate(net, blocks=blocks, K=net.K)
    inf = r["infer"]
    r["roofline"] = {"bound": "mfma", "kernel": inf["kernel"], "achieved": inf["tflops"],
                     "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": inf["tflops"] / BF16_PEAK_TFLOPS,
                     "flop_per_launch": inf["flop_per_launch"], "us_per_launch": inf["us"]}
    return r


def q_nonzero_ensemble(ens) -> tuple[int, int]:
    """Window pixels and pixels with q(phi) != 0 over every replica's cells."""
    import torch
    from pfh import abi
    phi = ens.phi[ens.steps_done % 2].view(ens.N, abi.TILE_PITCH, abi.TILE_PITCH)
    W = Q = 0
    for c in set(int(v) for v in ens.cls.cpu().numpy()):
        wx, wy, tx0, ty0 = abi.CLASSES[c]
        idx = torch.nonzero(ens.cls == c).flatten()
        W += int(idx.numel()) * wx * wy
        for chunk in torch.split(idx, 1024):
            win = phi[chunk, ty0:ty0 + wy, tx0:tx0 + wx]
            q = torch.round(torch.clamp(win * win, max=4.0) * 134217728.0)
            Q += int((q != 0).sum().item())
    return W, Q


def bench_batched(steps: int, warmup: int) -> dict:
    """Throughput of the batched ensemble: 64 independent replicas (sides
    256/512/1024, 16 to 750 cells at R=20, seeds 100..163) advanced by one
    persistent launch per step, replayed from six-step HIP graphs. Bytes per
    step: sum over replicas of N_r (12 W + 8 Q) + 4 P_r L_r."""
    import torch
    from pfh import ensemble, params
    wls = params.ensemble_c4()
    ens = ensemble.Ensemble2D(wls)
    ens.capture_graph()
    ens.run(warmup)
    ens.sync()
    W0, Q0 = q_nonzero_ensemble(ens)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(ens.stream)
    ens.run(steps)
    e1.record(ens.stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    W1, Q1 = q_nonzero_ensemble(ens)
    if ens.fatal():
        return {"error": "fatal integrity flags", "flags": ens.flag_counts().tolist()}
    ms_per_step = ms / steps
    bytes_per_step = 12.0 * 0.5 * (W0 + W1) + 8.0 * 0.5 * (Q0 + Q1) + 4.0 * ens.words
    achieved = bytes_per_step / (ms_per_step * 1e-3) / 1e9
    from collections import Counter
    return {
        "metric": "cell-steps/s, batched 2D ensemble of 64 replicas",
        "value": ens.N * steps / (ms / 1e3), "unit": "cell-steps/s",
        "steps": steps, "warmup": warmup, "ms_per_step": ms_per_step, "dtype": "f32",
        "config": {"workload": "ensemble_c4_64", "replicas": ens.R, "cells": ens.N,
                   "sides": dict(Counter(ens.Ls)), "radius": 20.0, "seeds": "100..163",
                   "placement": "grid + jitter per replica", "graph_steps": 6},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(TRAFFIC_SUMMARY_BATCHED, "traffic_bytes_per_step"),
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, "
                                       "calibrated on a copy kernel, committed in profiles/"
                                       + TRAFFIC_SUMMARY_BATCHED.name
                                       + " (not measured in this run)",
                     "algorithmic_bytes_per_step": bytes_per_step,
                     "window_pixels_per_step": 0.5 * (W0 + W1),
                     "q_nonzero_per_step": 0.5 * (Q0 + Q1)},
    }


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
    ap.add_argument("--workload", choices=["all", "2d", "3d", "batched"], default="all",
                    help="2d: the headline 2048^2 step only; 3d: the 256^3 step only; "
                         "batched: the 64-replica ensemble only")
    ap.add_argument("--steps-batched", type=int, default=120)
    ap.add_argument("--steps3d", type=int, default=600)
    ap.add_argument("--warmup3d", type=int, default=40)
    ap.add_argument("--sur-blocks", type=int, default=125,
                    help="blocks of the surrogate A/B run (125 x 8 = 1000 steps; 0 skips it)")
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

    if args.workload == "3d":
        print(json.dumps(bench_3d(args.steps3d, args.warmup3d, not args.no_cpu_baseline)),
              flush=True)
        return 0
    if args.workload == "batched":
        print(json.dumps(bench_batched(args.steps_batched, 12)), flush=True)
        return 0

    import torch
    from pfh import params, sim2d

    wl = params.workload_c2()
    sim = sim2d.Sim2D(wl, device="cuda:0")
    sim.capture_graph()
    # exactly the requested counts: whole six-step graph replays, single
    # launches of the same kernel for the rest
    warm = max(0, args.warmup)
    steps = max(1, args.steps)
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
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, "
                                       "calibrated on a copy kernel, committed in "
                                       "profiles/" + TRAFFIC_SUMMARY.name + " (not "
                                       "measured in this run)",
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "window_pixels_per_launch": W, "q_nonzero_per_launch": Qbar},
        "surrogate": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl)
    del sim
    torch.cuda.empty_cache()
    if args.sur_blocks > 0:
        from pfh import surrogate as su
        out["surrogate"] = bench_surrogate(args.sur_blocks)
        torch.cuda.empty_cache()
        # the 16-step variant over about the same number of steps
        out["surrogate_k16"] = bench_surrogate(max(1, args.sur_blocks // 2),
                                               su.WEIGHTS_PATH_K16)
        torch.cuda.empty_cache()
    if args.workload == "all":
        out["pf3d"] = bench_3d(args.steps3d, args.warmup3d, not args.no_cpu_baseline)
        torch.cuda.empty_cache()
        out["batched"] = bench_batched(args.steps_batched, 12)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
