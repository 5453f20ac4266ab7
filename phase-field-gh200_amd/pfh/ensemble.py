"""Batched 2D ensemble over the C ABI: many independent replicas (each its
own periodic side, population and Philox streams) advanced by one persistent
launch per step (pfh2d_launch_batch_step), replayed from six-step HIP graphs.

Layout: every replica's tiles, cell states and classes are concatenated in
one population (replica r's cells at [off_r, off_r + N_r)); each of the three
S rotation slots is one arena holding every replica's aggregate (replica r
at word S_offset_r, its own pitch P_r); the queue order runs over every cell.
Each replica is set up exactly as a single simulation (src/sim.cu:308-370)
on its slices, so replica r of the ensemble follows the same trajectory as a
lone simulation of that replica.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi
from .params import Workload2D, ic_theta, ic_v_A, s_pitch_for
from .sim2d import GRAPH_BODY

REPLICA_DTYPE = np.dtype([("S_offset", "<u8"), ("L", "<i4"), ("P", "<i4"),
                          ("seed", "<u8"), ("polarity_seed", "<u8")])
assert REPLICA_DTYPE.itemsize == 32
MIN_SIDE = 256


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class Ensemble2D:
    """Replicas of the 2D monolayer on one GPU, one launch per step."""

    def __init__(self, workloads: list[Workload2D], device: str = "cuda:0",
                 morton: bool = False):
        if not torch.cuda.is_available():
            raise RuntimeError("Ensemble2D needs a GPU; the kernels have no CPU path")
        if not workloads or len(workloads) > 65536:
            raise ValueError("an ensemble holds 1..65536 replicas")
        for wl in workloads:
            errs = wl.params.validate(min_side=MIN_SIDE)
            if errs:
                raise ValueError(f"{wl.name}: " + "; ".join(errs))
        # every replica shares the model coefficients (the step arguments'
        # scalars); only side, population and streams differ
        p0 = workloads[0].params
        for wl in workloads[1:]:
            p = wl.params
            same = (p.dt, p.lam, p.kappa, p.mu, p.xi, p.tau, p.target_radius,
                    p.full_moment_every)
            if same != (p0.dt, p0.lam, p0.kappa, p0.mu, p0.xi, p0.tau, p0.target_radius,
                        p0.full_moment_every):
                raise ValueError("replicas must share dt, lambda, kappa, mu, xi, tau, R and "
                                 "the full-moment cadence")
        self.lib = abi.load()
        self.wls = workloads
        self.p = p0
        self.R = len(workloads)
        self.dev = torch.device(device)
        torch.cuda.set_device(self.dev)
        self.stream = torch.cuda.current_stream(self.dev)
        self.Ns = [int(w.params.num_cells) for w in workloads]
        self.Ls = [int(w.params.Nx) for w in workloads]
        self.Ps = [s_pitch_for(L) for L in self.Ls]
        self.cell_off = np.concatenate([[0], np.cumsum(self.Ns)]).astype(np.int64)
        self.s_off = np.concatenate([[0], np.cumsum([P * L for P, L in
                                                     zip(self.Ps, self.Ls)])]).astype(np.int64)
        self.N = int(self.cell_off[-1])
        self.words = int(self.s_off[-1])            # one S slot arena
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
        reps = np.zeros(self.R, REPLICA_DTYPE)
        for r, wl in enumerate(workloads):
            reps[r] = (self.s_off[r], self.Ls[r], self.Ps[r], wl.params.seed,
                       wl.params.polarity_stream())
        self.reps_host = reps
        self.reps = torch.from_numpy(reps.view(np.uint8).copy()).to(self.dev)
        rep_of_cell = np.repeat(np.arange(self.R, dtype=np.uint16), self.Ns)
        self.cell_rep = torch.from_numpy(rep_of_cell.view(np.int16).copy()).to(self.dev)
        self.grid = self.lib.pfh2d_step_grid(self.dev.index or 0)
        if self.grid <= 0:
            abi.check(self.grid, "pfh2d_step_grid")
        abi.check(self.lib.pfh2d_configure_step(), "pfh2d_configure_step")
        self.max_side = max(self.Ls)
        self.max_pitch = max(self.Ps)
        self.steps_done = 0
        self._args = [self._bake(s) for s in range(GRAPH_BODY)]
        self._graph = None
        self._init_fresh()
        if morton:
            self.morton_order()

    # ---- buffers ----
    def _s_ptr(self, k: int) -> int:
        return _ptr(self.S) + 4 * k * self.words

    def _bake(self, slot: int) -> abi.StepArgs:
        """The ensemble's step arguments for one graph slot (src/sim.cu:578-618
        over the whole population): replica-specific L, P and streams come
        from the replica records."""
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
        a.N, a.L, a.P, a.parity_out = self.N, self.max_side, self.max_pitch, pout
        a.dt = float(np.float32(p.dt))
        a.A0 = p.area0()
        a.vol_scale = p.volume()
        a.bulk_scale = float(np.float32(p.bulk()))
        a.rep_coeff = float(np.float32(p.interaction()))
        a.mot_coeff = float(np.float32(p.motility()))
        a.seed = 0
        a.polarity_seed = 0
        a.p_tumble = p.p_tumble()
        a.full_moment_every = p.full_moment_every
        a.clear_ahead_words = self.words
        return a

    def _sp(self):
        return C.c_void_p(self.stream.cuda_stream)

    # ---- initial condition: every replica as a lone simulation ----
    def _init_fresh(self) -> None:
        lib, sp = self.lib, self._sp()
        cells = np.zeros(self.N, abi.CELL_DTYPE)
        for r, wl in enumerate(self.wls):
            p, o = wl.params, int(self.cell_off[r])
            n_soft = int(round(p.cancer_fraction * self.Ns[r]))
            for i in range(self.Ns[r]):
                c = cells[o + i]
                c["global_id"] = i
                c["gamma"] = np.float32(p.gamma_cancer if i < n_soft else p.gamma_normal)
                c["v_A"] = np.float32(ic_v_A(i, p.seed, p.v_A, p.v_A_sigma))
                c["M_pf"] = np.float32(p.phase_field_mobility)
                c["R_tgt"] = np.float32(p.target_radius)
                c["theta"] = ic_theta(i, p.polarity_stream())
        self.cell.copy_(torch.from_numpy(cells.view(np.uint8).reshape(-1)))
        cx = np.concatenate([np.ascontiguousarray(w.cx, np.float32) for w in self.wls])
        cy = np.concatenate([np.ascontiguousarray(w.cy, np.float32) for w in self.wls])
        dcx = torch.from_numpy(cx).to(self.dev)
        dcy = torch.from_numpy(cy).to(self.dev)
        ta = 4 * abi.TILE_AREA
        for r, wl in enumerate(self.wls):
            o, n, L, P = int(self.cell_off[r]), self.Ns[r], self.Ls[r], self.Ps[r]
            phi0, phi1 = _ptr(self.phi[0]) + ta * o, _ptr(self.phi[1]) + ta * o
            cell, cls = _ptr(self.cell) + 192 * o, _ptr(self.cls) + o
            abi.check(lib.pfh2d_init_tiles(phi0, phi1, cell, cls, n, L, _ptr(dcx) + 4 * o,
                                           _ptr(dcy) + 4 * o,
                                           float(np.float32(wl.params.init_radius())),
                                           float(np.float32(wl.params.interface_k())), sp),
                      "pfh2d_init_tiles")
            abi.check(lib.pfh2d_init_moments(cell, cls, phi0, n, sp), "pfh2d_init_moments")
            abi.check(lib.pfh2d_scatter_all(phi0, cell, cls, self._s_ptr(0) + 4 * int(self.s_off[r]),
                                            n, L, P, _ptr(self.flags), sp), "pfh2d_scatter_all")
        torch.cuda.synchronize(self.dev)

    def morton_order(self) -> None:
        """Queue order: each replica's cells in Morton order of their
        centroids (pfh2d_morton_sort per replica), replicas in turn."""
        lib, sp = self.lib, self._sp()
        for r in range(self.R):
            o, n = int(self.cell_off[r]), self.Ns[r]
            M = 1
            while M < n:
                M <<= 1
            abi.check(lib.pfh2d_morton_sort(_ptr(self.cell) + 192 * o, _ptr(self.perm) + 4 * o,
                                            n, M, self.Ls[r], sp), "pfh2d_morton_sort")
        torch.cuda.synchronize(self.dev)
        perm = self.perm.cpu().numpy()
        for r in range(self.R):
            o, n = int(self.cell_off[r]), self.Ns[r]
            perm[o:o + n] += o
        self.perm.copy_(torch.from_numpy(perm))

    def set_order(self, order: str) -> None:
        """Queue order: "identity" (replicas in turn), "morton" (Morton order
        within each replica) or "interleaved" (one cell of each replica in
        turn). Results do not depend on it."""
        if order == "identity":
            self.perm.copy_(torch.arange(self.N, dtype=torch.int32, device=self.dev))
        elif order == "morton":
            self.morton_order()
        elif order == "interleaved":
            idx = [int(self.cell_off[r]) + i for i in range(max(self.Ns))
                   for r in range(self.R) if i < self.Ns[r]]
            self.perm.copy_(torch.tensor(idx, dtype=torch.int32, device=self.dev))
        else:
            raise ValueError(f"unknown queue order {order!r}")
        self.order = order
        torch.cuda.synchronize(self.dev)

    # ---- stepping ----
    def launch_one(self, slot: int) -> None:
        abi.check(self.lib.pfh2d_launch_batch_step(
            C.byref(self._args[slot]), _ptr(self.reps), _ptr(self.cell_rep), self.R,
            self.max_side, self.max_pitch, self.grid, self._sp()), "pfh2d_launch_batch_step")

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            self.launch_one(self.steps_done % GRAPH_BODY)
            self.steps_done += 1

    def capture_graph(self) -> None:
        """Six batched steps per graph (both phi parities, all S slots)."""
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

    # ---- per-replica readback ----
    def replica_phi(self, r: int) -> np.ndarray:
        o, n = int(self.cell_off[r]), self.Ns[r]
        t = self.phi[self.steps_done % 2].view(self.N, abi.TILE_AREA)[o:o + n]
        return t.cpu().numpy()

    def replica_S(self, r: int) -> np.ndarray:
        base = (self.steps_done % 3) * self.words + int(self.s_off[r])
        w = self.Ps[r] * self.Ls[r]
        return self.S[base:base + w].cpu().numpy().view(np.uint32).copy()

    def replica_cells(self, r: int) -> np.ndarray:
        o, n = int(self.cell_off[r]), self.Ns[r]
        raw = self.cell[192 * o:192 * (o + n)].cpu().numpy()
        return raw.view(abi.CELL_DTYPE).copy()

    def replica_classes(self, r: int) -> np.ndarray:
        o, n = int(self.cell_off[r]), self.Ns[r]
        return self.cls[o:o + n].cpu().numpy().copy()

    def flag_counts(self) -> np.ndarray:
        return self.flags.cpu().numpy().view(np.uint32).copy()

    def fatal(self) -> bool:
        f = self.flag_counts()
        return any(f[i] != 0 for i in range(abi.FLAG_COUNT) if i != abi.FLAG_SUPPORT_CLIP)
