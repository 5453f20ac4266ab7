"""Launch-shape autotuner: times candidate shapes of the 3D step (measurement
shards per cell x update slabs per cell, i.e. the z extent of each
workgroup's LDS plane ring) and the queue order of the batched 2D ensemble
on the actual workload, with HIP events on the launch stream, and keeps the
fastest. Results are cached per (device, workload signature) in memory and,
when a cache path is given (or PFH_AUTOTUNE_CACHE is set), in a JSON file.

Tuning runs on a throwaway copy of the workload: the production object is
built afterwards with the chosen shape. Update slabs do not change any
result; measurement shards change the order of the fp64 moment sums (the
oracle takes the same shard count), so a tuned 3D run is compared with an
oracle run of the chosen shard count.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import torch

_MEMO: dict = {}


def _device_key(dev: torch.device) -> str:
    p = torch.cuda.get_device_properties(dev)
    return f"{p.name}:{getattr(p, 'gcnArchName', '')}:{p.multi_processor_count}"


class Autotuner:
    def __init__(self, cache_path: str | os.PathLike | None = None, device: str = "cuda:0"):
        self.dev = torch.device(device)
        path = cache_path or os.environ.get("PFH_AUTOTUNE_CACHE")
        self.path = Path(path) if path else None
        self.table: dict = {}
        if self.path and self.path.exists():
            self.table = json.loads(self.path.read_text())
        self.log: list = []          # every timing of the last tune call

    def _lookup(self, key: str):
        hit = _MEMO.get(key) or self.table.get(key)
        if hit is not None:
            _MEMO[key] = hit
        return hit

    def _store(self, key: str, val: dict) -> None:
        _MEMO[key] = val
        self.table[key] = val
        if self.path:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            tmp = self.path.with_suffix(".tmp")
            tmp.write_text(json.dumps(self.table, indent=1, sort_keys=True))
            os.replace(tmp, self.path)

    @staticmethod
    def _time(stream, fn, reps: int) -> float:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # ---- 3D: measurement shards x update slabs ----
    def candidates_3d(self, N: int, B: int) -> list[tuple[int, int]]:
        shards = [s for s in (1, 2, 3, 4, 6, 8) if s <= B // 4]
        slabs = [k for k in (1, 2, 3, 4, 6, 8, 12) if k <= B // 4]
        return [(s, k) for s in shards for k in slabs]

    def tune_3d(self, wl, steps: int = 6, warm: int = 2, force: bool = False) -> dict:
        """The fastest (shards, slabs) for a 3D workload: each candidate runs
        `warm` untimed and `steps` timed steps of the same population."""
        from . import sim3d
        p = wl.params
        key = f"3d|{_device_key(self.dev)}|L{p.Nx}|N{p.num_cells}|B{p.brick_edge()}"
        if not force:
            hit = self._lookup(key)
            if hit is not None:
                return hit
        best = None
        self.log = []
        cands = self.candidates_3d(int(p.num_cells), p.brick_edge())
        smax = max(s for s, _ in cands)
        sim = sim3d.Sim3D(wl, device=str(self.dev), shards=smax, slabs=1)
        # every candidate starts from the same state: the population is not
        # advanced hundreds of steps by the sweep itself
        start = (sim.phi_now().copy(), sim.S_now().copy(), sim.cells().copy(),
                 sim.counter_values())
        for shards, slabs in cands:
            sim.load_state(*start)
            sim.set_launch_shape(shards, slabs)
            sim.step(warm)
            ms = self._time(sim.stream, lambda: sim.step(1), steps)
            self.log.append({"shards": shards, "slabs": slabs, "ms": ms})
            if best is None or ms < best["ms"]:
                best = {"shards": shards, "slabs": slabs, "ms": ms}
        if sim.fatal():
            raise RuntimeError(f"autotune: fatal flags while tuning {key}")
        del sim
        torch.cuda.empty_cache()
        best["candidates"] = len(cands)
        self._store(key, best)
        return best

    # ---- batched 2D ensemble: queue order ----
    ORDERS = ("identity", "morton", "interleaved")

    def tune_ensemble(self, wls, steps: int = 6, warm: int = 2, force: bool = False) -> dict:
        """The fastest queue order of a batched ensemble: replicas in turn
        (identity), Morton order within each replica, or replicas interleaved
        cell by cell (neighbouring tickets in different aggregates)."""
        from . import ensemble
        sig = ",".join(f"{w.params.Nx}x{w.params.num_cells}" for w in wls)
        key = f"ens|{_device_key(self.dev)}|{len(wls)}|{hash(sig) & 0xFFFFFFFF:08x}"
        if not force:
            hit = self._lookup(key)
            if hit is not None:
                return hit
        self.log = []
        best = None
        ens = ensemble.Ensemble2D(wls, device=str(self.dev))
        for order in self.ORDERS:
            ens.set_order(order)
            ens.step(warm)
            ms = self._time(ens.stream, lambda: ens.step(1), steps)
            self.log.append({"order": order, "ms": ms})
            if best is None or ms < best["ms"]:
                best = {"order": order, "ms": ms}
        if ens.fatal():
            raise RuntimeError(f"autotune: fatal flags while tuning {key}")
        del ens
        torch.cuda.empty_cache()
        self._store(key, best)
        return best


def tuned_sim3d(wl, tuner: Autotuner | None = None, device: str = "cuda:0"):
    """A Sim3D with the autotuned launch shape."""
    from . import sim3d
    t = tuner or Autotuner(device=device)
    best = t.tune_3d(wl)
    return sim3d.Sim3D(wl, device=device, shards=best["shards"], slabs=best["slabs"]), best


def tuned_ensemble(wls, tuner: Autotuner | None = None, device: str = "cuda:0"):
    """An Ensemble2D with the autotuned queue order."""
    from . import ensemble
    t = tuner or Autotuner(device=device)
    best = t.tune_ensemble(wls)
    ens = ensemble.Ensemble2D(wls, device=device)
    ens.set_order(best["order"])
    return ens, best


__all__ = ["Autotuner", "tuned_sim3d", "tuned_ensemble"]
