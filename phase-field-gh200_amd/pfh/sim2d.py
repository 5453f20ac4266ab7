"""Host-side 2D simulation over the C ABI: buffers, initial condition, the
per-slot argument baking and graph replay of the reference's Sim
(src/sim.cu:159-370, 578-687, 1006-1055).

Device memory and streams come from PyTorch; every kernel is the gfx950
library's (lib/libpfh.so) -- there is no other implementation to fall back to.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import abi
from .params import SimParams, Workload2D, ic_theta, ic_v_A, s_pitch_for

GRAPH_BODY = 6   # two phi parities x three S rotation slots


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class Sim2D:
    """One 2D monolayer on one GPU (mirror of pf::Sim)."""

    def __init__(self, wl: Workload2D, device: str = "cuda:0"):
        if not torch.cuda.is_available():
            raise RuntimeError("Sim2D needs a GPU; the kernels have no CPU path")
        errs = wl.params.validate()
        if errs:
            raise ValueError("; ".join(errs))
        self.lib = abi.load()
        self.p: SimParams = wl.params
        self.dev = torch.device(device)
        torch.cuda.set_device(self.dev)
        self.stream = torch.cuda.current_stream(self.dev)
        self.N = int(self.p.num_cells)
        self.L = int(self.p.Nx)
        self.P = s_pitch_for(self.L)
        self.words = self.P * self.L
        N = self.N
        z = dict(device=self.dev)
        self.phi = [torch.zeros(N * abi.TILE_AREA, dtype=torch.float32, **z) for _ in range(2)]
        self.S = torch.zeros(3 * self.words, dtype=torch.int32, **z)
        self.cell = torch.zeros(N * 192, dtype=torch.uint8, **z)
        self.cls = torch.zeros(N, dtype=torch.uint8, **z)
        self.perm = torch.arange(N, dtype=torch.int32, **z)
        self.cursor = torch.zeros(2, dtype=torch.int64, **z)
        self.stepc = torch.zeros(2, dtype=torch.int64, **z)
        self.flags = torch.zeros(abi.FLAG_COUNT, dtype=torch.int32, **z)
        self.grid = self.lib.pfh2d_step_grid(self.dev.index or 0)
        if self.grid <= 0:
            abi.check(self.grid, "pfh2d_step_grid")
        abi.check(self.lib.pfh2d_configure_step(), "pfh2d_configure_step")
        self.steps_done = 0
        self.moments = None
        self._args = [self._bake(s) for s in range(GRAPH_BODY)]
        self._graph = None
        self._init_fresh(wl.cx, wl.cy)

    # ---- buffers ----
    def S_slot(self, k: int) -> torch.Tensor:
        return self.S[k * self.words:(k + 1) * self.words]

    def _s_ptr(self, k: int) -> int:
        return _ptr(self.S) + 4 * k * self.words

    def _bake(self, slot: int) -> abi.StepArgs:
        """args_for_slot (src/sim.cu:578-618)."""
        p = self.p
        pin, pout, rs = slot % 2, 1 - slot % 2, slot % 3
        a = abi.StepArgs()
        a.phi_in = _ptr(self.phi[pin])
        a.phi_out = _ptr(self.phi[pout])
        a.S_rd = self._s_ptr(rs)
        a.S_sc = self._s_ptr((rs + 1) % 3)
        a.S_cl = self._s_ptr((rs + 2) % 3)
        a.cell = _ptr(self.cell)
        a.cell_cls = _ptr(self.cls)
        a.perm = _ptr(self.perm)
        a.cursor_use = _ptr(self.cursor) + 8 * pin
        a.cursor_clear = _ptr(self.cursor) + 8 * pout
        a.step_rd = _ptr(self.stepc) + 8 * pin
        a.step_wr = _ptr(self.stepc) + 8 * pout
        a.flags = _ptr(self.flags)
        a.moments = None
        a.physical_dt = p.dt
        a.N, a.L, a.P, a.parity_out = self.N, self.L, self.P, pout
        a.dt = float(np.float32(p.dt))
        a.A0 = p.area0()
        a.vol_scale = p.volume()
        a.bulk_scale = float(np.float32(p.bulk()))
        a.rep_coeff = float(np.float32(p.interaction()))
        a.mot_coeff = float(np.float32(p.motility()))
        a.seed = p.seed
        a.polarity_seed = p.polarity_stream()
        a.p_tumble = p.p_tumble()
        a.full_moment_every = p.full_moment_every
        a.clear_ahead_words = self.words
        return a

    def _sp(self):
        return C.c_void_p(self.stream.cuda_stream)

    # ---- initial condition (src/sim.cu:308-370) ----
    def host_cells(self) -> np.ndarray:
        p = self.p
        cells = np.zeros(self.N, abi.CELL_DTYPE)
        n_cancer = int(round(p.cancer_fraction * self.N))
        for i in range(self.N):
            c = cells[i]
            c["global_id"] = i
            c["gamma"] = np.float32(p.gamma_cancer if i < n_cancer else p.gamma_normal)
            c["v_A"] = np.float32(ic_v_A(i, p.seed, p.v_A, p.v_A_sigma))
            c["M_pf"] = np.float32(p.phase_field_mobility)
            c["R_tgt"] = np.float32(p.target_radius)
            c["theta"] = ic_theta(i, p.polarity_stream())
        return cells

    def _init_fresh(self, cx: np.ndarray, cy: np.ndarray) -> None:
        cells = self.host_cells()
        self.cell.copy_(torch.from_numpy(cells.view(np.uint8).reshape(-1)))
        dcx = torch.from_numpy(np.ascontiguousarray(cx, np.float32)).to(self.dev)
        dcy = torch.from_numpy(np.ascontiguousarray(cy, np.float32)).to(self.dev)
        lib, sp = self.lib, self._sp()
        abi.check(lib.pfh2d_init_tiles(_ptr(self.phi[0]), _ptr(self.phi[1]), _ptr(self.cell),
                                       _ptr(self.cls), self.N, self.L, _ptr(dcx), _ptr(dcy),
                                       float(np.float32(self.p.init_radius())),
                                       float(np.float32(self.p.interface_k())), sp),
                  "pfh2d_init_tiles")
        abi.check(lib.pfh2d_init_moments(_ptr(self.cell), _ptr(self.cls), _ptr(self.phi[0]),
                                         self.N, sp), "pfh2d_init_moments")
        abi.check(lib.pfh2d_scatter_all(_ptr(self.phi[0]), _ptr(self.cell), _ptr(self.cls),
                                        self._s_ptr(0), self.N, self.L, self.P,
                                        _ptr(self.flags), sp), "pfh2d_scatter_all")
        torch.cuda.synchronize(self.dev)

    def load_state(self, phi: np.ndarray, S: np.ndarray, cells: np.ndarray,
                   cls: np.ndarray, step: int) -> None:
        """Install an arbitrary state at `step` (resume semantics,
        src/sim.cu:375-485): phi into parity step%2, S into rotation step%3."""
        self.steps_done = int(step)
        pin, rot = step % 2, step % 3
        self.phi[pin].copy_(torch.from_numpy(np.ascontiguousarray(phi, np.float32).reshape(-1)))
        self.phi[1 - pin].zero_()
        self.S.zero_()
        self.S_slot(rot).copy_(torch.from_numpy(np.ascontiguousarray(S).view(np.int32).reshape(-1)))
        cells = cells.copy()
        cells["cls_written"][:, 0] = cells["cls"]
        cells["cls_written"][:, 1] = cells["cls"]
        self.cell.copy_(torch.from_numpy(cells.view(np.uint8).reshape(-1)))
        self.cls.copy_(torch.from_numpy(np.ascontiguousarray(cls, np.uint8)))
        self.stepc.copy_(torch.tensor([step if pin == 0 else step + 1,
                                       step if pin == 1 else step + 1], dtype=torch.int64))
        self.cursor.zero_()
        self.flags.zero_()
        torch.cuda.synchronize(self.dev)

    # ---- velocity observation (src/sim.cu:648-669) ----
    def attach_moments(self) -> torch.Tensor:
        """Allocate the per-cell moment accumulators (zeroed) that spatial
        observation writes, and point every baked argument set at them."""
        self.moments = torch.zeros(self.N * abi.MOMENT_DTYPE.itemsize, dtype=torch.uint8,
                                   device=self.dev)
        for a in self._args:
            a.moments = _ptr(self.moments)
        return self.moments

    def moments_host(self) -> np.ndarray:
        return self.moments.cpu().numpy().view(abi.MOMENT_DTYPE)

    def observe_dense(self, out: torch.Tensor, physical_dt: float | None = None) -> None:
        """The independent dense observer on the next step's input
        (pfh2d_velocity_moments) into out (N accumulators)."""
        a = self._args[self.steps_done % GRAPH_BODY]
        abi.check(self.lib.pfh2d_velocity_moments(
            C.byref(a), _ptr(out), self.p.dt if physical_dt is None else physical_dt,
            self._sp()), "pfh2d_velocity_moments")

    # ---- stepping ----
    def launch_one(self, slot: int, observation: int = abi.OBSERVE_DISABLED) -> None:
        abi.check(self.lib.pfh2d_launch_step(C.byref(self._args[slot]), self.grid, self._sp(),
                                             None, 0, 0.0, observation), "pfh2d_launch_step")

    def step(self, n: int = 1, observation: int = abi.OBSERVE_DISABLED) -> None:
        for _ in range(n):
            self.launch_one(self.steps_done % GRAPH_BODY, observation)
            self.steps_done += 1

    def capture_graph(self) -> None:
        """Six steps per graph: both phi parities and all S rotation slots
        (src/sim.cu:671-687)."""
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
        """Advance `steps` steps: single launches up to a graph boundary,
        then graph replays, then single launches."""
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

    # ---- readback ----
    def cells(self) -> np.ndarray:
        return self.cell.cpu().numpy().view(abi.CELL_DTYPE).copy()

    def classes(self) -> np.ndarray:
        return self.cls.cpu().numpy().copy()

    def phi_now(self) -> np.ndarray:
        return self.phi[self.steps_done % 2].cpu().numpy().reshape(self.N, abi.TILE_AREA)

    def S_now(self) -> np.ndarray:
        return self.S_slot(self.steps_done % 3).cpu().numpy().view(np.uint32).copy()

    def flag_counts(self) -> np.ndarray:
        return self.flags.cpu().numpy().view(np.uint32).copy()

    def fatal(self) -> bool:
        f = self.flag_counts()
        return any(f[i] != 0 for i in range(abi.FLAG_COUNT) if i != abi.FLAG_SUPPORT_CLIP)
