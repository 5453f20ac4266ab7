"""2D parameter convention of the reference (include/params.cuh:499-565,
common/model_coefficients.cuh:16-53) and the host-side initial-condition
draws (include/kernels.cuh:146-172, src/sim.cu:82-156).

All derived coefficients are computed in IEEE double in the reference's
operation order so that the float copies handed to the kernels are the
reference's bits.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

K_PI = 3.14159265358979323846
PHASE_FIELD_M0 = 0.5
LAP_ROW_SUM_BOUND = (4.0 * 4 + 4.0 * 1 - (-20)) / 6.0


def philox4x32_10(c0: int, c1: int, c2: int, c3: int, k0: int, k1: int):
    """Philox4x32-10 (common/philox.cuh:32-58) on Python integers."""
    M = 0xFFFFFFFF
    for rnd in range(10):
        if rnd:
            k0 = (k0 + 0x9E3779B9) & M
            k1 = (k1 + 0xBB67AE85) & M
        p0 = 0xD2511F53 * c0
        p1 = 0xCD9E8D57 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & M, p1 & M, ((p0 >> 32) ^ c3 ^ k1) & M, p0 & M
    return (c0, c1, c2, c3)


def uniform53(a: int, b: int) -> float:
    """common/philox.cuh:61-67."""
    return float(((a << 32) | b) >> 11) * (1.0 / 9007199254740992.0)


def ic_theta(gid: int, stream: int) -> np.float32:
    """include/kernels.cuh:153-158: counter (gid, 2, 0, 0)."""
    r = philox4x32_10(gid & 0xFFFFFFFF, 2, 0, 0, stream & 0xFFFFFFFF, stream >> 32)
    return np.float32(2.0 * K_PI * uniform53(r[0], r[1]))


def ic_v_A(gid: int, seed: int, v_A: float, sigma: float) -> float:
    """include/kernels.cuh:161-172 (lognormal, median v_A)."""
    if not sigma > 0.0:
        return v_A
    r = philox4x32_10(gid & 0xFFFFFFFF, 3, 0, 0, seed & 0xFFFFFFFF, seed >> 32)
    u1 = max(1e-300, uniform53(r[0], r[1]))
    u2 = uniform53(r[2], r[3])
    z = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * K_PI * u2)
    return v_A * math.exp(sigma * z)


def domain_side_for(n: int, radius: float, rho: float) -> int:
    """include/params.cuh:568-574."""
    a = float(n) * (K_PI * radius * radius) / rho
    side = math.ceil(math.sqrt(a))
    return int(side) if math.isfinite(side) and 1 <= side <= 2**31 - 1 else -1


def s_pitch_for(side: int) -> int:
    """include/params.cuh:578-584: 32-word aligned S rows."""
    if side <= 0:
        return -1
    p = 32 * ((side + 31) // 32)
    return p if p <= 2**31 - 1 else -1


@dataclass
class SimParams:
    """Mirror of pf::SimParams with the reference defaults."""
    Nx: int = 0
    Ny: int = 0
    dx: float = 1.0
    dy: float = 1.0
    num_cells: int = 288
    rho: float = 0.90
    dt: float = 0.01
    t_end: float = 100.0
    lam: float = 7.0
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
    phase_field_mobility: float = PHASE_FIELD_M0
    seed: int = 1234
    polarity_seed: int = 0
    print_interval: int = 100
    full_moment_every: int = 100
    verify_every: int = 4096

    def polarity_stream(self) -> int:
        return self.polarity_seed if self.polarity_seed else self.seed

    def area0(self) -> float:
        return K_PI * self.target_radius * self.target_radius

    def bulk(self) -> float:
        return 30.0 / (self.lam * self.lam)

    def interaction(self) -> float:
        return 60.0 * self.kappa / (self.lam * self.lam)

    def motility(self) -> float:
        return self.interaction() / self.xi

    def volume(self) -> float:
        return 2.0 * self.mu / self.area0()

    def interface_k(self) -> float:
        return 2.7386127875258306 / self.lam

    def init_radius(self) -> float:
        return self.target_radius + 0.5 / self.interface_k()

    def p_tumble(self) -> float:
        return -math.expm1(-self.dt / self.tau)

    def total_steps(self) -> int:
        n = self.t_end / self.dt + 0.5
        return int(n) if 0.0 <= n <= 9.0e18 else -1

    def validate(self, min_side: int = 287) -> list[str]:
        """Subset of validate() (include/params.cuh:596-722) that the kernels
        rely on; returns the fatal messages (empty when valid). min_side is
        the smallest side accepted: 287 (every class window fits), or 256
        for a replica of a batched ensemble (no 286-wide fallback class)."""
        errs = []
        if not (self.dx == 1.0 and self.dy == 1.0):
            errs.append("dx and dy must both be exactly 1.0")
        if self.Nx != self.Ny:
            errs.append("Nx must equal Ny")
        if self.Nx <= 0 or s_pitch_for(self.Nx) < 0:
            errs.append("domain side is invalid or exceeds the supported integer range")
        if self.num_cells <= 0:
            errs.append("num_cells must be positive")
        if self.Nx > 0 and self.Nx < min_side:
            errs.append(f"domain side {self.Nx} must exceed the largest rect dimension "
                        f"{min_side - 1}.")
        gmax = max(self.gamma_normal, self.gamma_cancer)
        mrel = self.phase_field_mobility / PHASE_FIELD_M0
        if self.dt * gmax * mrel * LAP_ROW_SUM_BOUND >= 1.0:
            errs.append("dt*(M/M0)*gamma*laplacian_bound exceeds the safety limit 1")
        if not self.rho > 0.0 or not self.rho < 1.0:
            errs.append("rho must be finite and lie in (0,1)")
        return errs


def grid_jitter_centres(N: int, L: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Default placement, src/sim.cu:108-129."""
    nx = int(math.ceil(math.sqrt(float(N))))
    sp = float(L) / float(nx)
    cx = np.empty(N, np.float32)
    cy = np.empty(N, np.float32)
    for i in range(N):
        gxi, gyi = i % nx, i // nx
        r = philox4x32_10(i, 1, 0, 0, seed & 0xFFFFFFFF, seed >> 32)
        jx = (uniform53(r[0], r[1]) - 0.5) * 0.10 * sp
        jy = (uniform53(r[2], r[3]) - 0.5) * 0.10 * sp
        x = (gxi + 0.5 + (0.5 if gyi & 1 else 0.0)) * sp + jx
        y = (gyi + 0.5) * sp + jy
        x -= math.floor(x / L) * L
        y -= math.floor(y / L) * L
        cx[i] = np.float32(x)
        cy[i] = np.float32(y)
    return cx, cy


HARNESS_DOMAIN = 0x4A495454


def harness_centres(N: int, L: int, g: int, amp: float, seed: int):
    """Seeded jittered grid of the benchmark workloads: cell i at
    ((i%g)+0.5)L/g, ((i/g)+0.5)L/g plus uniform jitter in [-amp, amp] from
    Philox counter (i, 0x4A495454, 0, 0), key seed."""
    sp = float(L) / float(g)
    cx = np.empty(N, np.float32)
    cy = np.empty(N, np.float32)
    for i in range(N):
        r = philox4x32_10(i, HARNESS_DOMAIN, 0, 0, seed & 0xFFFFFFFF, seed >> 32)
        jx = (2.0 * uniform53(r[0], r[1]) - 1.0) * amp
        jy = (2.0 * uniform53(r[2], r[3]) - 1.0) * amp
        x = ((i % g) + 0.5) * sp + jx
        y = ((i // g) + 0.5) * sp + jy
        x -= math.floor(x / L) * L
        y -= math.floor(y / L) * L
        cx[i] = np.float32(x)
        cy[i] = np.float32(y)
    return cx, cy


@dataclass
class Workload2D:
    """A named 2D workload: parameters plus the centre table."""
    name: str
    params: SimParams
    cx: np.ndarray = field(repr=False)
    cy: np.ndarray = field(repr=False)


def monolayer(N: int, L: int, R: float, g: int, amp: float, seed: int,
              name: str | None = None, **kw) -> Workload2D:
    """Seeded jittered-grid monolayer with reference default coefficients."""
    rho = N * (K_PI * R * R) / float(L * L)
    p = SimParams(Nx=L, Ny=L, num_cells=N, rho=rho, target_radius=R,
                  seed=seed, polarity_seed=seed, **kw)
    cx, cy = harness_centres(N, L, g, amp, seed)
    return Workload2D(name or f"monolayer_L{L}_N{N}_R{R:g}", p, cx, cy)


def workload_c1() -> Workload2D:
    """Periodic 384^2, 16 cells R=20 on a 4x4 grid (+-4 jitter), seed 1."""
    return monolayer(16, 384, 20.0, 4, 4.0, 1, name="pf2d_L384_N16_R20")


def workload_c2() -> Workload2D:
    """Periodic 2048^2, 1024 cells R=20 on a 32x32 grid (+-4 jitter), seed 7."""
    return monolayer(1024, 2048, 20.0, 32, 4.0, 7, name="pf2d_L2048_N1024_R20")


ENSEMBLE_DOMAIN = 0x43344534
ENSEMBLE_SIDES = (256, 512, 1024)


def ensemble_max_cells(L: int, R: float = 20.0) -> int:
    """Largest replica population: 90% of the lattice covered by disks of
    radius R, floor(0.9 L^2 / (pi R^2))."""
    return int(math.floor(0.9 * L * L / (K_PI * R * R)))


def ensemble_replica(L: int, N: int, seed: int, R: float = 20.0) -> Workload2D:
    """One replica of the batched ensemble: N cells of radius R on an L x L
    periodic lattice, the reference's default grid + jitter placement and
    Philox streams of `seed` (placement, polarity and tumbles)."""
    rho = N * (K_PI * R * R) / float(L * L)
    p = SimParams(Nx=L, Ny=L, num_cells=N, rho=rho, target_radius=R, seed=seed,
                  polarity_seed=seed)
    cx, cy = grid_jitter_centres(N, L, seed)
    return Workload2D(f"replica_L{L}_N{N}_s{seed}", p, cx, cy)


def ensemble_c4(replicas: int = 64, seed0: int = 100, R: float = 20.0) -> list[Workload2D]:
    """The batched-ensemble workload: replica r (seed seed0 + r) has side L_r
    uniform in {256, 512, 1024} and N_r uniform in [16, N_max(L_r)], both
    drawn from Philox counter (r, 0x43344534, 0, 0) under key seed0 + r."""
    out = []
    for r in range(replicas):
        seed = seed0 + r
        v = philox4x32_10(r, ENSEMBLE_DOMAIN, 0, 0, seed & 0xFFFFFFFF, seed >> 32)
        L = ENSEMBLE_SIDES[min(2, int(3.0 * uniform53(v[0], v[1])))]
        nmax = ensemble_max_cells(L, R)
        N = 16 + min(nmax - 16, int((nmax - 15) * uniform53(v[2], v[3])))
        out.append(ensemble_replica(L, N, seed, R))
    return out
