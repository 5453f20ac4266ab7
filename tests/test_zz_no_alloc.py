"""The allocation-free 3D step (pfh3d_step_noalloc) against pfh3d_step with
caller-owned shard partials and arrival words, through the C ABI: two
calls in a row, the second on a smaller population, each bit-identical in
fields, aggregates, cell states, flags and counters."""
import ctypes as C
import sys

import numpy as np
import pytest

from pfh import abi, params3d

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-no-alloc] {msg}", flush=True)
    sys.stdout.flush()


def _pair(wl, shards):
    from pfh import sim3d
    a = sim3d.Sim3D(wl, device="cuda:0", shards=shards)
    b = sim3d.Sim3D(wl, device="cuda:0", shards=shards)
    return a, b


def _noalloc_step(sim, n):
    lib = abi.load()
    for _ in range(n):
        args = sim._args[sim.steps_done % len(sim._args)]
        saved = (args.partials, args.arrivals)
        args.partials, args.arrivals = None, None      # ignored by the entry point
        try:
            abi.check(lib.pfh3d_step_noalloc(C.byref(args), sim.slabs, sim._sp()),
                      "pfh3d_step_noalloc")
        finally:
            args.partials, args.arrivals = saved
        sim.steps_done += 1


def _same(a, b):
    a.sync()
    b.sync()
    assert np.array_equal(a.phi_now().view(np.uint32), b.phi_now().view(np.uint32))
    assert np.array_equal(a.S_now(), b.S_now())
    assert a.cells().tobytes() == b.cells().tobytes()
    assert np.array_equal(a.flag_counts(), b.flag_counts())
    assert a.counter_values() == b.counter_values()


def test_noalloc_step_equals_step_twice_second_smaller():
    import torch
    assert torch.cuda.is_available()
    big = params3d.workload_small()                                  # 8 cells, 96^3
    small = params3d.workload("pf3d_L96_N3_R8", 3, 8.0, 3.5, 96, seed=5)
    for wl, shards, steps in [(big, 5, 3), (small, 2, 2)]:
        say(f"{wl.name}: shards {shards}, {steps} steps")
        ref, nal = _pair(wl, shards)
        ref.step(steps)
        _noalloc_step(nal, steps)
        _same(ref, nal)


def test_noalloc_rejects_oversized_population():
    lib = abi.load()
    a = abi.StepArgs3()
    a.N, a.B, a.L, a.P, a.shards = 4096, 48, 512, 512, 3      # 12288 partials > 8192
    assert lib.pfh3d_step_noalloc(C.byref(a), 0, None) == -1
    assert b"8192" in lib.pfh_last_error()
