"""TEST INFRASTRUCTURE ONLY: Python binding of the 3D CPU oracle
(oracle/_build/libpfo.so). Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; product code never does.

OracleSim3D mirrors the host-side Sim3D step sequence (buffer parity every
begun step, src/pf3d/sim.cu:1460-1651) on host arrays. The workload inputs
are a params object with the SimParams3D fields and the binary32 centres;
the seed radius, the initial polarities and every kernel are the oracle's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pfo

ORDER_MI355X = 0
ORDER_REFERENCE = 1
FLAG3_COUNT = 10
FATAL3_MASK = 0x3BF

CELL3_DTYPE = np.dtype({
    "names": ["global_id", "origin_x", "origin_y", "origin_z",
              "polarity_x", "polarity_y", "polarity_z",
              "velocity_x", "velocity_y", "velocity_z", "gamma", "v_A", "R_tgt",
              "phi_max", "shift_ctr", "tumble_ctr", "flags", "pad0",
              "bb_lo_x", "bb_hi_x", "bb_lo_y", "bb_hi_y", "bb_lo_z", "bb_hi_z",
              "pending_shift_x", "pending_shift_y", "pending_shift_z", "pad1",
              "V", "Cx", "Cy", "Cz", "surface", "Ix", "Iy", "Iz",
              "storage_edge", "pad2", "wall_overlap", "outside_slit_volume", "reserved"],
    "formats": ["<i8", "<i8", "<i8", "<i8"] + ["<f4"] * 10 + ["<u4"] * 4 + ["<i4"] * 9
               + ["<u4"] + ["<f8"] * 8 + ["<u4", "<u4", "<f8", "<f8", ("<u4", (10,))],
    "offsets": [0, 8, 16, 24, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80, 84,
                88, 92, 96, 100, 104, 108, 112, 116, 120, 124, 128, 136, 144, 152, 160,
                168, 176, 184, 192, 196, 200, 208, 216],
    "itemsize": 256,
})


class StepArgs3(C.Structure):
    _fields_ = [
        ("phi_in", C.c_void_p), ("phi_out", C.c_void_p),
        ("S_in", C.c_void_p), ("S_out", C.c_void_p),
        ("cells", C.c_void_p), ("accepted", C.c_void_p),
        ("flags", C.c_void_p), ("counters", C.c_void_p),
        ("partials", C.c_void_p), ("arrivals", C.c_void_p),
        ("N", C.c_int), ("B", C.c_int), ("L", C.c_int), ("P", C.c_int),
        ("shards", C.c_int), ("max_shift", C.c_int), ("support_margin", C.c_int),
        ("full_moment_every", C.c_int),
        ("dt", C.c_float), ("bulk_scale", C.c_float), ("interaction_coeff", C.c_float),
        ("motility_coeff", C.c_float),
        ("V0", C.c_double), ("volume_scale", C.c_double), ("p_tumble", C.c_double),
        ("polarity_stream", C.c_uint64), ("aging_steps", C.c_uint64),
    ]


assert C.sizeof(StepArgs3) == 168

_P = C.c_void_p
_I = C.c_int
_sig_done = False


def load() -> C.CDLL:
    global _sig_done
    lib = pfo.load()
    if _sig_done:
        return lib
    SA = C.POINTER(StepArgs3)
    sig = {
        "pfo3d_init_spheres": (None, [_P, _P, _P, _P, _I, _I, C.c_float, C.c_float]),
        "pfo3d_scatter": (None, [_P, _P, _P, _P, _I, _I, _I, _I]),
        "pfo3d_scatter_f64": (None, [_P, _P, _P, _P, _I, _I, _I, _I]),
        "pfo3d_measure": (None, [SA, _I, _I, _I, _I, _I]),
        "pfo3d_measure_f64": (None, [SA, _P, _P, _P, _I, _I, _I, _I]),
        "pfo3d_step": (_I, [SA, _I, _I]),
        "pfo3d_step_f64": (_I, [SA, _P, _P, _P, _P, _I]),
        "pfo3d_calibrate_seed_radius": (C.c_double, [C.c_double, C.c_double, _I, C.c_double,
                                                     _I, C.POINTER(C.c_int)]),
        "pfo3d_initial_polarity": (None, [C.c_int64, C.c_uint64, _P]),
        "pfo3d_tumble_sample": (_I, [C.c_uint64, C.c_int64, C.c_uint64, C.c_double, _P, _P]),
        "pfo3d_lap27": (C.c_float, [_P, _I]),
        "pfo3d_brick_edge": (_I, [C.c_double, C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _sig_done = True
    return lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def s_pitch_for(side: int) -> int:
    return 32 * ((side + 31) // 32)


def initial_polarity(gid: int, stream: int) -> np.ndarray:
    out = np.zeros(3, np.float64)
    load().pfo3d_initial_polarity(gid, stream, _p(out))
    return out


def seed_radius(radius: float, lam: float, B: int) -> float:
    it = C.c_int(0)
    return float(load().pfo3d_calibrate_seed_radius(radius, lam, B, 1e-10, 80, C.byref(it)))


class OracleSim3D:
    """CPU oracle run of a periodic 3D workload.

    order: ORDER_MI355X (the kernels' operation order, bit-exact) or
    ORDER_REFERENCE (the reference's); f64=True evaluates the field, the
    measurement and the velocities in double."""

    def __init__(self, params, cx, cy, cz, shards: int, order: int = ORDER_MI355X,
                 f64: bool = False, threads: int = 0):
        self.lib = load()
        p = params
        self.p = p
        self.N, self.L = int(p.num_cells), int(p.Nx)
        self.P = s_pitch_for(self.L)
        auto = int(self.lib.pfo3d_brick_edge(p.target_radius, p.lambda_))
        explicit = int(getattr(p, "brick_edge_override", 0) or 0)
        self.B = explicit if explicit > 0 else auto
        self.words = self.L * self.L * self.P
        self.voxels = self.B ** 3
        self.order, self.f64, self.threads = order, f64, threads
        self.shards = shards
        rt = np.float64 if f64 else np.float32
        self.phi = [np.zeros((self.N, self.voxels), rt) for _ in range(2)]
        self.S = [np.zeros(self.words, np.uint32) for _ in range(2)]
        self.cells = np.zeros(self.N, CELL3_DTYPE)
        self.accepted = np.zeros(self.N, CELL3_DTYPE)
        self.flags = np.zeros(FLAG3_COUNT, np.uint32)
        self.counters = np.zeros(2, np.uint64)
        self.vel = np.zeros(3 * self.N, np.float64)
        self.coef = np.array([p.dt, p.bulk(), p.interaction(), p.motility()], np.float64)
        self.steps_done = 0
        self.seed_radius = seed_radius(p.target_radius, p.lambda_, self.B)
        self._init(cx, cy, cz)

    def _host_cells(self) -> None:
        p = self.p
        stream = p.polarity_seed if p.polarity_seed else p.seed
        for i in range(self.N):
            c = self.cells[i]
            c["global_id"] = i
            c["gamma"] = np.float32(p.gamma_normal)
            c["v_A"] = np.float32(p.v_A)
            c["R_tgt"] = np.float32(p.target_radius)
            pol = initial_polarity(i, stream)
            c["polarity_x"], c["polarity_y"], c["polarity_z"] = (np.float32(v) for v in pol)

    def _init(self, cx, cy, cz) -> None:
        self._host_cells()
        centres = np.stack([cx, cy, cz], axis=1).astype(np.float64).reshape(-1)
        phi32 = [np.zeros((self.N, self.voxels), np.float32) for _ in range(2)]
        lib = self.lib
        lib.pfo3d_init_spheres(_p(phi32[0]), _p(phi32[1]), _p(self.cells), _p(centres),
                               self.N, self.B, float(np.float32(self.p.lambda_)),
                               float(np.float32(self.seed_radius)))
        self.phi[0][:] = phi32[0]
        a = self._args(0)
        if self.f64:
            lib.pfo3d_scatter_f64(_p(self.phi[0]), _p(self.cells), _p(self.S[0]),
                                  _p(self.flags), self.N, self.B, self.L, self.P)
            lib.pfo3d_measure_f64(C.byref(a), _p(self.phi[0]), _p(self.vel), _p(self.coef),
                                  0, 1, 1, self.threads)
        else:
            lib.pfo3d_scatter(_p(self.phi[0]), _p(self.cells), _p(self.S[0]), _p(self.flags),
                              self.N, self.B, self.L, self.P)
            lib.pfo3d_measure(C.byref(a), self.order, 0, 1, 1, self.threads)

    def load_state(self, phi, S, cells, counters=(0, 0)) -> None:
        slot = int(counters[1]) % 2
        self.steps_done = int(counters[1])
        self.phi[slot][:] = np.asarray(phi).reshape(self.N, self.voxels)
        self.phi[1 - slot][:] = 0
        self.S[slot][:] = S
        self.S[1 - slot][:] = 0
        self.cells[:] = cells
        self.counters[:] = counters
        self.flags[:] = 0

    def _args(self, slot: int) -> StepArgs3:
        p = self.p
        a = StepArgs3()
        a.phi_in = _p(self.phi[slot])
        a.phi_out = _p(self.phi[1 - slot])
        a.S_in = _p(self.S[slot])
        a.S_out = _p(self.S[1 - slot])
        a.cells = _p(self.cells)
        a.accepted = _p(self.accepted)
        a.flags = _p(self.flags)
        a.counters = _p(self.counters)
        a.N, a.B, a.L, a.P = self.N, self.B, self.L, self.P
        a.shards = self.shards
        a.max_shift = 4
        a.support_margin = 8
        a.full_moment_every = p.full_moment_every
        a.dt = float(np.float32(p.dt))
        a.bulk_scale = float(np.float32(p.bulk()))
        a.interaction_coeff = float(np.float32(p.interaction()))
        a.motility_coeff = float(np.float32(p.motility()))
        a.V0 = p.volume0()
        a.volume_scale = p.volume()
        a.p_tumble = p.p_tumble()
        a.polarity_stream = p.polarity_seed if p.polarity_seed else p.seed
        a.aging_steps = p.aging_steps()
        return a

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            slot = self.steps_done % 2
            a = self._args(slot)
            if self.f64:
                rc = self.lib.pfo3d_step_f64(C.byref(a), _p(self.phi[slot]),
                                             _p(self.phi[1 - slot]), _p(self.vel),
                                             _p(self.coef), self.threads)
            else:
                rc = self.lib.pfo3d_step(C.byref(a), self.order, self.threads)
            if rc != 0:
                raise RuntimeError(f"oracle 3D step failed ({rc})")
            self.steps_done += 1

    def phi_now(self) -> np.ndarray:
        return self.phi[self.steps_done % 2]

    def S_now(self) -> np.ndarray:
        return self.S[self.steps_done % 2]

    def fatal(self) -> bool:
        return any(self.flags[i] != 0 for i in range(FLAG3_COUNT) if (FATAL3_MASK >> i) & 1)
