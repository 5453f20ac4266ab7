"""TEST INFRASTRUCTURE ONLY: Python binding of the CPU oracle
(oracle/_build/libpfo.so). Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; product code never does.

OracleSim2D mirrors the host-side Sim2D (same slot/parity/rotation
bookkeeping, src/sim.cu:578-618) on host arrays, so parity tests compare the
two state by state.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libpfo.so"

TILE_AREA = 288 * 288
ORDER_MI355X = 0
ORDER_REFERENCE = 1

CELL_DTYPE = np.dtype({
    "names": ["gx0", "gy0", "cls", "cls_written", "pad0", "global_id",
              "gamma", "v_A", "R_tgt", "theta", "V", "Cx", "Cy", "perim",
              "bb_lo_x", "bb_hi_x", "bb_lo_y", "bb_hi_y", "vx", "vy", "Ix",
              "Iy", "promote_ctr", "shift_ctr", "tumble_ctr", "phi_max",
              "reserved", "adv_q", "adv_count", "reserved_unused", "M_pf"],
    "formats": ["<i4", "<i4", "u1", ("u1", (2,)), "u1", "<i4", "<f4", "<f4",
                "<f4", "<f4", "<f8", "<f8", "<f8", "<f8", "<i4", "<i4", "<i4",
                "<i4", "<f4", "<f4", "<f8", "<f8", "<u4", "<u4", "<u4", "<f4",
                ("<u4", (2,)), ("<f8", (6,)), "<u8", "<u4", "<f4"],
    "offsets": [0, 4, 8, 9, 11, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 68,
                72, 76, 80, 84, 88, 96, 104, 108, 112, 116, 120, 128, 176,
                184, 188],
    "itemsize": 192,
})

# MomentAccum (common/velocity_moments.h:30-40): 240 bytes.
MOMENT_DTYPE = np.dtype([("q", "<f8", (12,)), ("count", "<u8"), ("held", "<f8", (6,)),
                         ("spatial_count", "<u8"), ("last_sample", "<u8"),
                         ("last_interval", "<u8"), ("last_totals", "<f8", (4,)),
                         ("correction", "<f8", (4,))])


class StepArgs(C.Structure):
    _fields_ = [
        ("phi_in", C.c_void_p), ("phi_out", C.c_void_p),
        ("S_rd", C.c_void_p), ("S_sc", C.c_void_p), ("S_cl", C.c_void_p),
        ("cell", C.c_void_p), ("cell_cls", C.c_void_p), ("perm", C.c_void_p),
        ("flags", C.c_void_p), ("vel_f64", C.c_void_p),
        ("step", C.c_uint64),
        ("N", C.c_int), ("L", C.c_int), ("P", C.c_int), ("parity_out", C.c_int),
        ("dt", C.c_float), ("A0", C.c_double), ("vol_scale", C.c_double),
        ("bulk_scale", C.c_float), ("rep_coeff", C.c_float), ("mot_coeff", C.c_float),
        ("dt_d", C.c_double), ("bulk_d", C.c_double), ("rep_d", C.c_double),
        ("mot_d", C.c_double),
        ("polarity_seed", C.c_uint64), ("p_tumble", C.c_double),
        ("full_moment_every", C.c_int), ("clear_ahead_words", C.c_uint64),
        ("observation", C.c_int), ("moments", C.c_void_p), ("physical_dt", C.c_double),
    ]


_lib = None
_P = C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    sig = {
        "pfo2d_step": (C.c_int, [C.POINTER(StepArgs), C.c_int, C.c_int]),
        "pfo2d_step_f64": (C.c_int, [C.POINTER(StepArgs), C.c_int]),
        "pfo2d_init_tiles": (None, [_P, _P, _P, _P, C.c_int, C.c_int, _P, _P, C.c_float, C.c_float]),
        "pfo2d_init_moments": (None, [_P, _P, _P, C.c_int, C.c_int]),
        "pfo2d_init_moments_f64": (None, [_P, _P, _P, C.c_int]),
        "pfo2d_scatter_all": (None, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P]),
        "pfo2d_scatter_all_f64": (None, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P]),
        "pfo2d_verify_cells": (None, [_P, _P, _P, C.c_int, _P, _P]),
        "pfo2d_verify_S": (C.c_uint32, [_P, C.c_size_t]),
        "pfo2d_pack_traj": (None, [_P, _P, _P, C.c_int, C.c_int]),
        "pfo2d_velocity_moments": (None, [_P, _P, C.c_double]),
        "pfo2d_morton_perm": (None, [_P, C.c_int, C.c_int, _P]),
        "pfo_philox": (None, [_P, _P, _P]),
        "pfo2d_tumble": (C.c_int, [C.c_uint64, C.c_int32, C.c_uint64, C.c_double, _P]),
        "pfo_u53": (C.c_double, [C.c_uint32, C.c_uint32]),
        "pfo_qof": (C.c_uint32, [C.c_float]),
        "pfo_ic_theta": (C.c_float, [C.c_int32, C.c_uint64]),
        "pfo_ic_v_A": (C.c_double, [C.c_int32, C.c_uint64, C.c_double, C.c_double]),
        "pfo_sincosf": (None, [C.c_float, _P, _P]),
        "pfo_tanh": (C.c_double, [C.c_double]),
        "pfo_rhs": (C.c_float, [C.c_float] * 12),
        "pfo_lap9": (C.c_float, [_P, C.c_int]),
        "pfo_class_containing": (C.c_int, [C.c_int] * 3),
        "pfo_domain_side_for": (C.c_int, [C.c_int, C.c_double, C.c_double]),
        "pfo_s_pitch_for": (C.c_int, [C.c_int]),
        "pfo2d_grid_jitter_centres": (None, [C.c_int, C.c_int, C.c_uint64, _P, _P]),
        "pfo2d_harness_centres": (None, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, _P, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def philox(ctr, key) -> tuple:
    lib = load()
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib.pfo_philox(_p(c), _p(k), _p(out))
    return tuple(int(v) for v in out)


def sincosf(theta: float) -> tuple[np.float32, np.float32]:
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    load().pfo_sincosf(float(theta), _p(s), _p(c))
    return s[0], c[0]


def s_pitch_for(side: int) -> int:
    return 32 * ((side + 31) // 32)


class OracleSim2D:
    """CPU oracle run of a 2D workload (binary32, or FP64 with f64=True)."""

    def __init__(self, wl, order: int = ORDER_MI355X, f64: bool = False,
                 threads: int = 0, init_phi: np.ndarray | None = None):
        self.lib = load()
        p = wl.params
        self.p = p
        self.N, self.L = int(p.num_cells), int(p.Nx)
        self.P = s_pitch_for(self.L)
        self.words = self.P * self.L
        self.order = order
        self.f64 = f64
        self.threads = threads
        rt = np.float64 if f64 else np.float32
        self.phi = [np.zeros((self.N, TILE_AREA), rt) for _ in range(2)]
        self.S = [np.zeros(self.words, np.uint32) for _ in range(3)]
        self.cells = np.zeros(self.N, CELL_DTYPE)
        self.cls = np.zeros(self.N, np.uint8)
        self.perm = np.arange(self.N, dtype=np.uint32)
        self.flags = np.zeros(8, np.uint32)
        self.vel = np.zeros(2 * self.N, np.float64)
        # velocity observation of the next steps (0 none, 1 advection,
        # 2 spatial into self.moments)
        self.observation = 0
        self.moments = np.zeros(self.N, MOMENT_DTYPE)
        self.steps_done = 0
        self._init(wl, init_phi)

    def _init(self, wl, init_phi) -> None:
        from_params = _host_cells(self.p, self.N)
        self.cells[:] = from_params
        lib = self.lib
        phi32 = [np.zeros((self.N, TILE_AREA), np.float32) for _ in range(2)]
        cx = np.ascontiguousarray(wl.cx, np.float32)
        cy = np.ascontiguousarray(wl.cy, np.float32)
        lib.pfo2d_init_tiles(_p(phi32[0]), _p(phi32[1]), _p(self.cells), _p(self.cls),
                             self.N, self.L, _p(cx), _p(cy),
                             float(np.float32(self.p.init_radius())),
                             float(np.float32(self.p.interface_k())))
        if init_phi is not None:
            phi32[0][:] = init_phi
        if self.f64:
            self.phi[0][:] = phi32[0]
            lib.pfo2d_init_moments_f64(_p(self.cells), _p(self.cls), _p(self.phi[0]), self.N)
            lib.pfo2d_scatter_all_f64(_p(self.phi[0]), _p(self.cells), _p(self.cls),
                                      _p(self.S[0]), self.N, self.L, self.P, _p(self.flags))
        else:
            self.phi[0][:] = phi32[0]
            lib.pfo2d_init_moments(_p(self.cells), _p(self.cls), _p(self.phi[0]), self.N,
                                   self.order)
            lib.pfo2d_scatter_all(_p(self.phi[0]), _p(self.cells), _p(self.cls),
                                  _p(self.S[0]), self.N, self.L, self.P, _p(self.flags))

    def load_state(self, phi, S, cells, cls, step) -> None:
        self.steps_done = int(step)
        pin = step % 2
        self.phi[pin][:] = phi
        self.phi[1 - pin][:] = 0
        for k in range(3):
            self.S[k][:] = 0
        self.S[step % 3][:] = S
        self.cells[:] = cells
        self.cells["cls_written"][:, 0] = self.cells["cls"]
        self.cells["cls_written"][:, 1] = self.cells["cls"]
        self.cls[:] = cls
        self.flags[:] = 0

    def _args(self) -> StepArgs:
        p = self.p
        s = self.steps_done
        pin, pout, rs = s % 2, 1 - s % 2, s % 3
        a = StepArgs()
        a.phi_in = _p(self.phi[pin])
        a.phi_out = _p(self.phi[pout])
        a.S_rd = _p(self.S[rs])
        a.S_sc = _p(self.S[(rs + 1) % 3])
        a.S_cl = _p(self.S[(rs + 2) % 3])
        a.cell = _p(self.cells)
        a.cell_cls = _p(self.cls)
        a.perm = _p(self.perm)
        a.flags = _p(self.flags)
        a.vel_f64 = _p(self.vel)
        a.step = s
        a.N, a.L, a.P, a.parity_out = self.N, self.L, self.P, pout
        a.dt = float(np.float32(p.dt))
        a.A0 = p.area0()
        a.vol_scale = p.volume()
        a.bulk_scale = float(np.float32(p.bulk()))
        a.rep_coeff = float(np.float32(p.interaction()))
        a.mot_coeff = float(np.float32(p.motility()))
        a.dt_d, a.bulk_d, a.rep_d, a.mot_d = p.dt, p.bulk(), p.interaction(), p.motility()
        a.polarity_seed = p.polarity_stream()
        a.p_tumble = p.p_tumble()
        a.full_moment_every = p.full_moment_every
        a.clear_ahead_words = self.words
        a.observation = self.observation
        a.moments = _p(self.moments)
        a.physical_dt = p.dt
        return a

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            a = self._args()
            if self.f64:
                rc = self.lib.pfo2d_step_f64(C.byref(a), self.threads)
            else:
                rc = self.lib.pfo2d_step(C.byref(a), self.order, self.threads)
            if rc != 0:
                raise RuntimeError(f"oracle step failed ({rc})")
            self.steps_done += 1

    def phi_now(self) -> np.ndarray:
        return self.phi[self.steps_done % 2]

    def S_now(self) -> np.ndarray:
        return self.S[self.steps_done % 3]


def _host_cells(p, N: int) -> np.ndarray:
    """Per-cell host fill of Sim::init (src/sim.cu:324-340)."""
    lib = load()
    cells = np.zeros(N, CELL_DTYPE)
    n_cancer = int(round(p.cancer_fraction * N))
    for i in range(N):
        cells[i]["global_id"] = i
        cells[i]["gamma"] = np.float32(p.gamma_cancer if i < n_cancer else p.gamma_normal)
        cells[i]["v_A"] = np.float32(lib.pfo_ic_v_A(i, p.seed, p.v_A, p.v_A_sigma))
        cells[i]["M_pf"] = np.float32(p.phase_field_mobility)
        cells[i]["R_tgt"] = np.float32(p.target_radius)
        cells[i]["theta"] = np.float32(lib.pfo_ic_theta(i, p.polarity_stream()))
    return cells


def world_field(phi: np.ndarray, cells: np.ndarray, L: int) -> np.ndarray:
    """Render each cell's active window into world coordinates: (N, L, L)."""
    CLASSES = [(144, 144, 64, 64), (176, 144, 32, 64), (144, 176, 64, 32),
               (160, 160, 32, 32), (224, 224, 32, 32), (286, 286, 1, 1)]
    N = len(cells)
    out = np.zeros((N, L, L), phi.dtype)
    for n in range(N):
        wx, wy, tx0, ty0 = CLASSES[int(cells[n]["cls"])]
        tile = phi[n].reshape(288, 288)[ty0:ty0 + wy, tx0:tx0 + wx]
        ys = (int(cells[n]["gy0"]) + np.arange(wy)) % L
        xs = (int(cells[n]["gx0"]) + np.arange(wx)) % L
        out[n][np.ix_(ys, xs)] = tile
    return out
