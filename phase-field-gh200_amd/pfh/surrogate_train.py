"""Train the surrogate on the GPU and evaluate it against the plain method.

    python -m pfh.surrogate_train train --seeds 1000 1099 --out weights.npz
    python -m pfh.surrogate_train eval --weights weights.npz --blocks 25

`train` generates crops from 2048^2 / 1024-cell monolayers (two samples per
seed, after 300 and 420 steps: 204,800 crops for seeds 1000..1099), trains with Adam and reports the
held-out increment error of the network beside the no-change prediction.
`eval` advances the 2048^2 / 1024-cell workload (seed 7) with the surrogate
and, from the same state, with the plain method, and reports the relative
L2 distance of the fields, the acceptance rate and both throughputs.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch

from . import params, sim2d
from . import surrogate as su


def log(msg: str) -> None:
    print(f"[surrogate] {msg}", flush=True)
    sys.stdout.flush()


def cmd_train(a) -> int:
    seeds = list(range(a.seeds[0], a.seeds[1] + 1))
    t0 = time.perf_counter()
    X, D = su.make_dataset(seeds, samples_per_seed=a.samples, K=a.K, warm=a.warm, gap=a.gap,
                           log=log)
    n = D.numel() // su.PIX
    log(f"{n} crops in {time.perf_counter() - t0:.1f} s")
    held = min(2048, n // 10)
    ntr = n - held
    net = su.Net(seed=a.init_seed, scale=a.scale)
    net.K = a.K
    t0 = time.perf_counter()
    hist = su.train(net, X[:ntr * su.PIX * su.C0], D[:ntr * su.PIX], epochs=a.epochs,
                    batch=a.batch, lr=a.lr, log=log)
    dt = time.perf_counter() - t0
    ws = su.Workspace(256, net.dev)
    e_net = e_id = 0.0
    for o in range(ntr, n, 256):
        m = min(256, n - o)
        ws.X0[:m * su.PIX * su.C0].copy_(X[o * su.PIX * su.C0:(o + m) * su.PIX * su.C0])
        su.forward(net, ws, m, pred=False)
        d = D[o * su.PIX:(o + m) * su.PIX].double()
        e_net += float(((ws.dphi[:m * su.PIX].double() * net.scale - d) ** 2).sum())
        e_id += float((d ** 2).sum())
    net.save(Path(a.out))
    rep = {"K": a.K, "crops": n, "train_crops": ntr, "held_out": held, "epochs": a.epochs,
           "batch": a.batch, "train_seconds": dt, "loss_history": hist,
           "held_out_increment_rel_error_network": (e_net / e_id) ** 0.5,
           "held_out_increment_rel_error_no_change": 1.0, "weights": a.out}
    print(json.dumps(rep), flush=True)
    return 0


def _rel_l2_fields(a, b, origin) -> np.ndarray:
    """Per-cell relative L2 of two Sim2D states on the same crop windows."""
    n = a.N
    fa = torch.zeros(n * su.PIX, dtype=torch.float32, device=a.dev)
    fb = torch.zeros_like(fa)
    su.recrop_sim(a, origin, fa)
    su.recrop_sim(b, origin, fb)
    fa = fa.view(n, -1).double()
    fb = fb.view(n, -1).double()
    return (torch.linalg.norm(fa - fb, dim=1) / torch.linalg.norm(fb, dim=1)).cpu().numpy()


def evaluate(net: su.Net, blocks: int = 25, K: int | None = None, tol: float = 1e-3,
             warm: int = 300, wl=None, **runner_kw) -> dict:
    wl = wl or params.workload_c2()
    plain = sim2d.Sim2D(wl)
    plain.capture_graph()
    plain.run(warm)
    sur = sim2d.Sim2D(wl)
    sur.capture_graph()
    sur.run(warm)
    runner = su.SurrogateRunner(sur, net, K=K, tol=tol, **runner_kw)
    K = runner.K
    runner.prepare()
    # timing: plain K-step blocks from graphs, surrogate blocks
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(plain.stream)
    plain.run(blocks * K)
    e1.record(plain.stream)
    torch.cuda.synchronize()
    ms_plain = e0.elapsed_time(e1)
    t0 = time.perf_counter()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record(sur.stream)
    runner.run(blocks * K)
    s1.record(sur.stream)
    torch.cuda.synchronize()
    st = runner.stats()
    ms_sur = s0.elapsed_time(s1)
    wall_sur = time.perf_counter() - t0
    infer_t = time_infer(runner)
    ws = su.Workspace(plain.N, plain.dev)
    su.crop_sim(plain, ws)
    rel = _rel_l2_fields(sur, plain, ws.origin)
    N = plain.N
    return {
        "K": K, "tol": tol, "blocks": blocks, "steps": blocks * K,
        "plain_cell_steps_per_s": N * blocks * K / (ms_plain / 1e3),
        "surrogate_cell_steps_per_s": N * blocks * K / (ms_sur / 1e3),
        "speedup": ms_plain / ms_sur,
        "surrogate_wall_s": wall_sur,
        "accepted_blocks": st["accepted_blocks"],
        "fallback_blocks": st["fallback_blocks"],
        "fallback": st["fallback"],
        "fallback_cells": st.get("fallback_cells"),
        "accepted_cell_fraction": st.get("accepted_cell_fraction"),
        "max_accumulated_error": st.get("max_accumulated_error"),
        "tol_inc": runner.tol_inc, "budget": runner.budget,
        "max_check_error": st["max_check_error"],
        "rel_l2_vs_plain_max": float(rel.max()),
        "rel_l2_vs_plain_mean": float(rel.mean()),
        "fatal": bool(sur.fatal() or plain.fatal()),
        "infer": infer_t,
    }


# algorithmic MACs of one crop: conv1 (4 used input channels), conv2, conv3,
# output layer
INFER_FLOP_PER_CROP = 2 * su.PIX * 9 * (4 * su.CH + su.CH * su.CH * 2 + su.CH)


def time_infer(runner: su.SurrogateRunner, reps: int = 20) -> dict:
    """Average duration of the fused inference kernel (crops built in the
    kernel, as in a block) on the runner's stream, on the current state, and
    its MFMA rate."""
    sim = runner.sim
    sp = sim._sp()
    stream = sim.stream
    ws = runner.ws
    su.infer_cells(runner.net, sim, ws, W4=runner.W4d, b4=runner.b4d, dphi=False, stream=sp)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        su.infer_cells(runner.net, sim, ws, W4=runner.W4d, b4=runner.b4d, dphi=False, stream=sp)
    e1.record(stream)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    flop = INFER_FLOP_PER_CROP * sim.N
    return {"kernel": "k_sur_infer", "us": us, "flop_per_launch": flop,
            "tflops": flop / (us * 1e-6) / 1e12}


def cmd_eval(a) -> int:
    net = su.Net.load(Path(a.weights))
    out = evaluate(net, blocks=a.blocks, K=a.K or net.K, tol=a.tol)
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--seeds", type=int, nargs=2, default=[1000, 1099])
    t.add_argument("--samples", type=int, default=2)
    t.add_argument("--warm", type=int, default=300)
    t.add_argument("--gap", type=int, default=120)
    t.add_argument("--K", type=int, default=8, help="steps per surrogate block")
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--batch", type=int, default=128)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--scale", type=float, default=1e-3)
    t.add_argument("--init-seed", type=int, default=0)
    t.add_argument("--out", default=str(su.WEIGHTS_PATH))
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=str(su.WEIGHTS_PATH))
    e.add_argument("--blocks", type=int, default=25)
    e.add_argument("--tol", type=float, default=1e-3)
    e.add_argument("--K", type=int, default=0, help="steps per block (default: the weights')")
    a = ap.parse_args(argv)
    return cmd_train(a) if a.cmd == "train" else cmd_eval(a)


if __name__ == "__main__":
    sys.exit(main())
