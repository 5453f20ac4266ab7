"""Host-side 3D simulation over the C ABI: buffers, the fresh initial
condition and the per-step sequence of the reference's Sim3D for the periodic
base-brick path (src/pf3d/sim.cu:1053-1177, 1280-1368, 1460-1651).

Device memory and streams come from PyTorch; every kernel is the gfx950
library's (lib/libpfh.so) -- there is no other implementation to fall back to.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi
from .params3d import SUPPORT_MARGIN, SimParams3D, Workload3D, initial_polarity, \
    s_pitch_for, seed_radius

GRAPH_BODY = 2   # fields and aggregates alternate every begun step


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def default_shards(N: int, B: int, device_index: int) -> int:
    """Measurement workgroups per cell: fill one pass over the compute units
    (the reference's one-wave policy, include/pf3d/measure_shards.hpp:36-51),
    at most 4 and with at least four planes per shard."""
    cus = torch.cuda.get_device_properties(device_index).multi_processor_count
    s = max(1, min(4, cus // max(1, N), B // 4))
    return s


class Sim3D:
    """One periodic 3D population on one GPU (mirror of pf3d::Sim3D)."""

    def __init__(self, wl: Workload3D, device: str = "cuda:0", shards: int = 0,
                 slabs: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("Sim3D needs a GPU; the kernels have no CPU path")
        errs = wl.params.validate()
        if errs:
            raise ValueError("; ".join(errs))
        self.lib = abi.load()
        self.p: SimParams3D = wl.params
        self.dev = torch.device(device)
        torch.cuda.set_device(self.dev)
        self.stream = torch.cuda.current_stream(self.dev)
        p = self.p
        self.N = int(p.num_cells)
        self.L = int(p.Nx)
        self.P = s_pitch_for(self.L)
        self.B = p.brick_edge()
        self.words = self.L * self.L * self.P
        self.voxels = self.B ** 3
        dix = self.dev.index or 0
        self.shards = shards if shards > 0 else default_shards(self.N, self.B, dix)
        self.slabs = slabs if slabs > 0 else int(self.lib.pfh3d_update_slabs(self.N, self.B, dix))
        N = self.N
        z = dict(device=self.dev)
        self.phi = [torch.zeros(N * self.voxels, dtype=torch.float32, **z) for _ in range(2)]
        self.S = [torch.zeros(self.words, dtype=torch.int32, **z) for _ in range(2)]
        self.cell = torch.zeros(N * 256, dtype=torch.uint8, **z)
        self.accepted = torch.zeros(N * 256, dtype=torch.uint8, **z)
        self.flags = torch.zeros(abi.FLAG3_COUNT, dtype=torch.int32, **z)
        self.counters = torch.zeros(2, dtype=torch.int64, **z)
        self.partials = torch.zeros(N * self.shards * abi.PARTIAL3_BYTES, dtype=torch.uint8, **z)
        self.arrivals = torch.zeros(N, dtype=torch.int32, **z)
        self.shards_alloc = self.shards
        self.seed_radius = seed_radius(p)
        self._args = [self._bake(s) for s in range(GRAPH_BODY)]
        self._graph = None
        self._init_fresh(wl)

    # ---- arguments (src/pf3d/sim.cu:1301-1325, 1496-1519) ----
    def _bake(self, slot: int) -> abi.StepArgs3:
        p = self.p
        a = abi.StepArgs3()
        a.phi_in = _ptr(self.phi[slot])
        a.phi_out = _ptr(self.phi[1 - slot])
        a.S_in = _ptr(self.S[slot])
        a.S_out = _ptr(self.S[1 - slot])
        a.cells = _ptr(self.cell)
        a.accepted = _ptr(self.accepted)
        a.flags = _ptr(self.flags)
        a.counters = _ptr(self.counters)
        a.partials = _ptr(self.partials)
        a.arrivals = _ptr(self.arrivals)
        a.N, a.B, a.L, a.P = self.N, self.B, self.L, self.P
        a.shards = self.shards
        a.max_shift = abi.MAX_SHIFT_3D
        a.support_margin = SUPPORT_MARGIN
        a.full_moment_every = p.full_moment_every
        a.dt = float(np.float32(p.dt))
        a.bulk_scale = float(np.float32(p.bulk()))
        a.interaction_coeff = float(np.float32(p.interaction()))
        a.motility_coeff = float(np.float32(p.motility()))
        a.V0 = p.volume0()
        a.volume_scale = p.volume()
        a.p_tumble = p.p_tumble()
        a.polarity_stream = p.polarity_stream()
        a.aging_steps = p.aging_steps()
        return a

    def set_launch_shape(self, shards: int, slabs: int) -> None:
        """Measurement shards and update slabs of the following steps (the
        shard partials were sized for the shard count at construction)."""
        if not 1 <= shards <= self.shards_alloc or not 1 <= slabs <= self.B:
            raise ValueError(f"shards must lie in [1, {self.shards_alloc}], slabs in "
                             f"[1, {self.B}]")
        self.shards, self.slabs = int(shards), int(slabs)
        self._args = [self._bake(s) for s in range(GRAPH_BODY)]
        self._graph = None

    def _sp(self):
        return C.c_void_p(self.stream.cuda_stream)

    # ---- fresh initial condition (src/pf3d/sim.cu:1099-1177) ----
    def host_cells(self) -> np.ndarray:
        p = self.p
        cells = np.zeros(self.N, abi.CELL3_DTYPE)
        stream = p.polarity_stream()
        for i in range(self.N):
            c = cells[i]
            c["global_id"] = i
            c["gamma"] = np.float32(p.gamma_normal)
            c["v_A"] = np.float32(p.v_A)
            c["R_tgt"] = np.float32(p.target_radius)
            pol = initial_polarity(i, stream)
            c["polarity_x"], c["polarity_y"], c["polarity_z"] = (np.float32(v) for v in pol)
        return cells

    def _init_fresh(self, wl: Workload3D) -> None:
        p = self.p
        cells = self.host_cells()
        self.cell.copy_(torch.from_numpy(cells.view(np.uint8).reshape(-1)))
        centres = np.stack([wl.cx, wl.cy, wl.cz], axis=1).astype(np.float64).reshape(-1)
        dc = torch.from_numpy(centres).to(self.dev)
        lib, sp = self.lib, self._sp()
        abi.check(lib.pfh3d_init_spheres(_ptr(self.phi[0]), _ptr(self.phi[1]), _ptr(self.cell),
                                         _ptr(dc), self.N, self.B, self.L,
                                         float(np.float32(p.lambda_)),
                                         float(np.float32(self.seed_radius)), sp),
                  "pfh3d_init_spheres")
        abi.check(lib.pfh3d_clear_S(_ptr(self.S[0]), self.L, self.P, sp), "pfh3d_clear_S")
        abi.check(lib.pfh3d_scatter_current(_ptr(self.phi[0]), _ptr(self.cell), _ptr(self.S[0]),
                                            _ptr(self.flags), self.N, self.B, self.L, self.P, sp),
                  "pfh3d_scatter_current")
        abi.check(lib.pfh3d_measure_cells(C.byref(self._args[0]), 0, 1, 1, sp),
                  "pfh3d_measure_cells")
        torch.cuda.synchronize(self.dev)
        self.steps_done = 0

    def load_state(self, phi: np.ndarray, S: np.ndarray, cells: np.ndarray,
                   counters=(0, 0)) -> None:
        """Install a state: phi/S into parity counters[1] % 2 (resume semantics)."""
        slot = int(counters[1]) % 2
        self.steps_done = int(counters[1])
        self.phi[slot].copy_(torch.from_numpy(np.ascontiguousarray(phi, np.float32).reshape(-1)))
        self.phi[1 - slot].zero_()
        self.S[slot].copy_(torch.from_numpy(np.ascontiguousarray(S).view(np.int32).reshape(-1)))
        self.S[1 - slot].zero_()
        self.cell.copy_(torch.from_numpy(np.ascontiguousarray(cells).view(np.uint8).reshape(-1)))
        self.counters.copy_(torch.tensor([int(counters[0]), int(counters[1])], dtype=torch.int64))
        self.flags.zero_()
        self.arrivals.zero_()
        torch.cuda.synchronize(self.dev)

    # ---- checkpoints (PF3D schema 1) ----
    def save_checkpoint(self, path, trajectory_interval: int = 0) -> None:
        """The current state (accepted step, cell records, current cubes) as
        a PF3D file; the shard count is recorded as the reduction contract."""
        from . import checkpoint3d
        self.sync()
        step = int(self.counters[0].item())
        rec = checkpoint3d.params_record(self.p, self.B, shards_policy=0,
                                         trajectory_interval=trajectory_interval)
        cells = self.cells()
        abi.check(self.lib.pfh3d_checkpoint_write(
            str(path).encode(), step, float(step) * self.p.dt, rec.ctypes.data,
            cells.ctypes.data, self.N, self.B, _ptr(self.phi[self.steps_done % 2]),
            self.shards, self._sp()), "pfh3d_checkpoint_write")

    def load_checkpoint(self, path) -> None:
        """Resume from a PF3D file written for this workload
        (src/pf3d/sim.cu:840-880): cubes and records installed at the file's
        step, the aggregate rebuilt from the cubes and the measurement
        refreshed (no tumble)."""
        from . import checkpoint3d
        info = checkpoint3d.probe(path)
        if info["N"] != self.N or info["B"] < self.B or int(info["params"]["Nx"]) != self.L:
            raise ValueError(f"{path}: checkpoint geometry (N={info['N']}, B={info['B']}) "
                             f"does not match this simulation (N={self.N}, B={self.B})")
        if info["B"] != self.B:
            # a population grown past this run's starting brick resumes at its edge
            if info["B"] > abi.MAX_EDGE_3D or info["B"] % 8:
                raise ValueError(f"{path}: brick edge {info['B']} is not supported")
            self.sync()
            self._set_brick(int(info["B"]))
        step = int(info["step"])
        slot = step % 2
        cells = np.zeros(self.N, abi.CELL3_DTYPE)
        pr = np.zeros(1, checkpoint3d.PARAMS3_DTYPE)
        st, t, sh = C.c_int64(), C.c_double(), C.c_uint64()
        n, b = C.c_int(), C.c_int()
        self.sync()
        abi.check(self.lib.pfh3d_checkpoint_read(
            str(path).encode(), pr.ctypes.data, C.byref(st), C.byref(t), C.byref(sh),
            C.byref(n), C.byref(b), cells.ctypes.data, _ptr(self.phi[slot]), self._sp()),
            "pfh3d_checkpoint_read")
        self.phi[1 - slot].zero_()
        self.S[1 - slot].zero_()
        self.cell.copy_(torch.from_numpy(cells.view(np.uint8).reshape(-1)))
        self.counters.copy_(torch.tensor([step, step], dtype=torch.int64))
        self.flags.zero_()
        self.arrivals.zero_()
        self.steps_done = step
        lib, sp = self.lib, self._sp()
        abi.check(lib.pfh3d_clear_S(_ptr(self.S[slot]), self.L, self.P, sp), "pfh3d_clear_S")
        abi.check(lib.pfh3d_scatter_current(_ptr(self.phi[slot]), _ptr(self.cell),
                                            _ptr(self.S[slot]), _ptr(self.flags), self.N, self.B,
                                            self.L, self.P, sp), "pfh3d_scatter_current")
        abi.check(lib.pfh3d_measure_cells(C.byref(self._args[slot]), 0, 1, 1, sp),
                  "pfh3d_measure_cells")
        self.sync()

    # ---- stepping (src/pf3d/sim.cu:1460-1651) ----
    def launch_one(self, slot: int) -> None:
        abi.check(self.lib.pfh3d_step(C.byref(self._args[slot]), self.slabs, self._sp()),
                  "pfh3d_step")

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            self.launch_one(self.steps_done % GRAPH_BODY)
            self.steps_done += 1

    def capture_graph(self) -> None:
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.dev)
        s.wait_stream(self.stream)
        with torch.cuda.stream(s):
            old = self.stream
            self.stream = s
            with torch.cuda.graph(g, stream=s):
                for slot in range(GRAPH_BODY):
                    self.launch_one(slot)
            self.stream = old
        self.stream.wait_stream(s)
        torch.cuda.synchronize(self.dev)
        self._graph = g

    def run(self, steps: int) -> None:
        end = self.steps_done + steps
        while self.steps_done < end:
            if (self._graph is not None and self.steps_done % GRAPH_BODY == 0
                    and self.steps_done + GRAPH_BODY <= end):
                self._graph.replay()
                self.steps_done += GRAPH_BODY
            else:
                self.step(1)

    def sync(self) -> None:
        torch.cuda.synchronize(self.dev)

    # ---- adaptive bricks (src/pf3d/sim.cu:1653-1733) ----
    def grow_bricks(self, B_new: int) -> None:
        """Move the whole population into cubes of edge B_new (every value at
        its world position, the states moved with their cubes, the support
        flags cleared; S is unchanged). The graph, if any, is dropped: the
        fields are new buffers."""
        if not (self.B < B_new <= abi.MAX_EDGE_3D and B_new % 8 == 0):
            raise ValueError(f"brick edge {B_new}: must exceed {self.B}, be a multiple of 8 "
                             f"and at most {abi.MAX_EDGE_3D}")
        self.sync()
        slot = self.steps_done % 2
        vox = B_new ** 3
        new = [torch.zeros(self.N * vox, dtype=torch.float32, device=self.dev) for _ in range(2)]
        abi.check(self.lib.pfh3d_grow_bricks(_ptr(self.phi[slot]), _ptr(new[slot]),
                                             _ptr(self.cell), _ptr(self.flags), self.N, self.B,
                                             B_new, self._sp()), "pfh3d_grow_bricks")
        self.sync()
        self._set_brick(B_new, new)

    def _set_brick(self, B: int, phi=None) -> None:
        """Fields of edge B (new zeroed buffers unless given) and the launch
        shape and arguments that follow from it."""
        self.B, self.voxels = int(B), int(B) ** 3
        self.phi = phi if phi is not None else [
            torch.zeros(self.N * self.voxels, dtype=torch.float32, device=self.dev)
            for _ in range(2)]
        dix = self.dev.index or 0
        self.shards = min(self.shards, max(1, self.B // 4))
        self.slabs = int(self.lib.pfh3d_update_slabs(self.N, self.B, dix))
        self._args = [self._bake(s) for s in range(GRAPH_BODY)]
        self._graph = None

    def recover_support(self) -> bool:
        """After steps rejected for SUPPORT_EXHAUSTED alone (the accepted state
        kept), grow the population to the next brick tier; False when another
        fatal flag is raised or no larger brick fits."""
        f = self.flag_counts()
        if f[abi.FLAG3_NAMES.index("support_exhausted")] == 0:
            return False
        if any(f[i] for i in range(abi.FLAG3_COUNT)
               if (abi.FATAL3_MASK >> i) & 1 and abi.FLAG3_NAMES[i] != "support_exhausted"):
            return False
        nb = int(self.lib.pfh3d_next_brick_edge(self.B, self.L))
        if nb <= self.B:
            return False
        self.grow_bricks(nb)
        return True

    def run_adaptive(self, steps: int, poll: int = 16) -> int:
        """run() with the reference's support recovery, polled every `poll`
        steps (src/pf3d/sim.cu:2142-2268): returns the number of growths;
        stops early on a fatal flag it cannot recover from."""
        end = self.steps_done + steps
        grown = 0
        graph = self._graph is not None
        while self.steps_done < end:
            self.run(min(poll, end - self.steps_done))
            self.sync()
            if self.fatal():
                if not self.recover_support():
                    break
                grown += 1
                if graph:
                    self.capture_graph()
        return grown

    # ---- readback ----
    def cells(self) -> np.ndarray:
        return self.cell.cpu().numpy().view(abi.CELL3_DTYPE).copy()

    def phi_now(self) -> np.ndarray:
        return self.phi[self.steps_done % 2].cpu().numpy().reshape(self.N, self.voxels)

    def S_now(self) -> np.ndarray:
        return self.S[self.steps_done % 2].cpu().numpy().view(np.uint32).copy()

    def flag_counts(self) -> np.ndarray:
        return self.flags.cpu().numpy().view(np.uint32).copy()

    def counter_values(self) -> tuple[int, int]:
        c = self.counters.cpu().numpy()
        return int(c[0]), int(c[1])

    def fatal(self) -> bool:
        f = self.flag_counts()
        return any(f[i] != 0 for i in range(abi.FLAG3_COUNT) if (abi.FATAL3_MASK >> i) & 1)
