"""3D parameter convention of the reference (include/pf3d/params.cuh:236-330,
common/model_coefficients.cuh:16-53) for the periodic cube, and the named
3D workloads.

Derived coefficients are computed in IEEE double in the reference's
operation order; the binary32 copies handed to the kernels are the
reference's bits. Geometry helpers that the reference computes on the host
(brick edge, seed-radius bisection, centre placement, initial polarity) run
in the library's host code (lib/libpfh.so).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import abi

K_PI = 3.141592653589793238462643383279502884
LAP27_ROW_SUM_BOUND = 256.0 / 30.0      # include/pf3d/params.cuh:77
SUPPORT_MARGIN = 8                      # kBrickSafetyMargin, params.cuh:105


def s_pitch_for(side: int) -> int:
    """src/pf3d/sim.cu:276: 32-word aligned S rows."""
    return 32 * ((side + 31) // 32)


@dataclass
class SimParams3D:
    """Mirror of pf3d::SimParams3D for the periodic-XYZ geometry, with the
    reference defaults."""
    Nx: int = 0
    num_cells: int = 288
    rho: float = 0.90
    dt: float = 0.01
    t_end: float = 100.0
    aging_time: float = 0.0
    lambda_: float = 7.0
    target_radius: float = 49.0
    kappa: float = 10.0
    mu: float = 1.0
    xi: float = 1500.0
    tau: float = 1.0e4
    v_A: float = 1.0e-2
    gamma_normal: float = 1.0
    gamma_cancer: float = 0.35
    cancer_fraction: float = 0.0
    v_A_sigma: float = 0.0
    seed: int = 1234
    polarity_seed: int = 0
    full_moment_every: int = 100
    brick_edge_override: int = 0    # --brick-edge: explicit aligned edge, 0 = automatic

    def polarity_stream(self) -> int:
        return self.polarity_seed if self.polarity_seed else self.seed

    def volume0(self) -> float:
        r = self.target_radius
        return 4.0 * K_PI * r * r * r / 3.0

    def bulk(self) -> float:
        return 30.0 / (self.lambda_ * self.lambda_)

    def interaction(self) -> float:
        return 60.0 * self.kappa / (self.lambda_ * self.lambda_)

    def motility(self) -> float:
        return self.interaction() / self.xi

    def volume(self) -> float:
        return 2.0 * self.mu / self.volume0()

    def p_tumble(self) -> float:
        return -math.expm1(-self.dt / self.tau)

    def aging_steps(self) -> int:
        s = self.aging_time / self.dt + 0.5
        return int(s) if math.isfinite(s) and 0.0 <= s <= 9.0e18 else -1

    def automatic_brick_edge(self) -> int:
        return int(abi.load().pfh3d_brick_edge(self.target_radius, self.lambda_))

    def brick_edge(self) -> int:
        """Sim3D::allocate (src/pf3d/sim.cu:253-262): the explicit edge when
        given, else the automatic one."""
        return self.brick_edge_override if self.brick_edge_override > 0 \
            else self.automatic_brick_edge()

    def validate(self) -> list[str]:
        """The periodic-geometry subset of validate (include/pf3d/params.cuh:332-441)."""
        errs = []
        B = self.brick_edge()
        if self.Nx <= 0:
            errs.append("the periodic domain must be a positive cube")
        if self.num_cells <= 0:
            errs.append("num_cells must be positive")
        if not (self.dt > 0.0):
            errs.append("dt must be positive")
        if B <= 0 or self.Nx <= B:
            errs.append("the brick must fit strictly inside the domain")
        if self.brick_edge_override and (self.brick_edge_override < self.automatic_brick_edge()
                                         or self.brick_edge_override % 8):
            errs.append("an explicit brick edge must be an aligned edge >= the automatic one")
        if B > abi.MAX_EDGE_3D:
            errs.append("bricks above 64^3 (adaptive support) are outside this build")
        g = max(self.gamma_normal, self.gamma_cancer)
        if self.dt * g * LAP27_ROW_SUM_BOUND >= 1.0:
            errs.append("explicit-Euler Laplacian stability bound is violated")
        if self.dt * g * self.bulk() >= 1.0:
            errs.append("explicit-Euler bulk-reaction timestep bound is violated")
        if self.dt * 2.0 * self.mu >= 1.0:
            errs.append("explicit-Euler volume-restoring timestep bound is violated")
        if self.cancer_fraction != 0.0 or self.v_A_sigma != 0.0:
            errs.append("heterogeneous cell populations are outside this build")
        return errs


@dataclass
class Workload3D:
    name: str
    params: SimParams3D
    cx: np.ndarray   # float32 centres (binary32, as the placement stores them)
    cy: np.ndarray
    cz: np.ndarray


def sequential_centres(n: int, side: float, radius: float, seed: int):
    """palmieri3d_sequential_centres (include/pf3d/palmieri_initializer.hpp:72-123),
    evaluated by the library's host code."""
    lib = abi.load()
    x = np.zeros(n, np.float32)
    y = np.zeros(n, np.float32)
    z = np.zeros(n, np.float32)
    st = lib.pfh3d_sequential_centres(n, float(side), float(radius), int(seed),
                                      x.ctypes.data, y.ctypes.data, z.ctypes.data)
    abi.check(st if st < 0 else 0, "pfh3d_sequential_centres")
    return x, y, z


def seed_radius(p: SimParams3D) -> float:
    """calibrate_seed_radius (include/pf3d/reference.hpp:712-745) at the brick edge."""
    r = C.c_double(0.0)
    st = abi.load().pfh3d_calibrate_seed_radius(p.target_radius, p.lambda_, p.brick_edge(),
                                                C.byref(r))
    abi.check(st if st < 0 else 0, "pfh3d_calibrate_seed_radius")
    return r.value


def initial_polarity(gid: int, stream: int) -> np.ndarray:
    out = np.zeros(3, np.float64)
    abi.check(abi.load().pfh3d_initial_polarity(gid, stream, out.ctypes.data),
              "pfh3d_initial_polarity")
    return out


def workload(name: str, n: int, radius: float, lam: float, side: int, seed: int = 1234,
             **kw) -> Workload3D:
    p = SimParams3D(Nx=side, num_cells=n, target_radius=radius, lambda_=lam, seed=seed, **kw)
    p.rho = n * p.volume0() / float(side) ** 3
    x, y, z = sequential_centres(n, float(side), radius, seed)
    return Workload3D(name, p, x, y, z)


def workload_c3(brick_edge: int = 56) -> Workload3D:
    """256^3 periodic, 128 spheres R=12, lambda=3.5, rho 0.05555, sequential
    placement at minimum separation R, seed 1234.

    The automatic brick edge is 48, with no slack beyond the 8-voxel support
    margin: the relaxing diffuse tail raises SUPPORT_EXHAUSTED within tens of
    steps, which the reference answers by promoting cells to larger bricks
    (here Sim3D.run_adaptive grows the population to 64). The throughput
    workload uses the reference's explicit --brick-edge 56, which holds the
    relaxed support over the whole run."""
    lib = abi.load()
    side = int(lib.pfh3d_domain_side(128, 12.0, 0.05555))
    w = workload(f"pf3d_L256_N128_R12_B{brick_edge}", 128, 12.0, 3.5, side,
                 brick_edge_override=brick_edge)
    w.params.rho = 0.05555
    return w


def workload_small() -> Workload3D:
    """96^3 periodic, 8 spheres R=8, lambda=3.5 (brick 40): the oracle-sized
    3D parity case (non-square plane groups: 400 lanes in 7 waves)."""
    return workload("pf3d_L96_N8_R8", 8, 8.0, 3.5, 96, seed=11)
