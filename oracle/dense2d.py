"""TEST INFRASTRUCTURE ONLY: an independent dense-lattice FP64 evaluator of
the 2D per-cell phase-field step, used to pin the windowed oracle
(pfo2d.c / pfo2d_step.inc). Only tests/ import it; product code never does.

It shares nothing with the windowed oracle but the equations: every cell's
field lives on the whole periodic L x L lattice (no tiles, windows, zero
rings, recentring shifts or shape classes), the aggregate is the exact FP64
sum of phi_m^2 (no Q5.27 quantisation), gradients and the Laplacian wrap
periodically, and the polarity draw takes a Philox callable from the caller
(the tests pass the reference's own common/philox.cuh compiled under
oracle/_ref/ where it is built).

Equations followed (reference file:line):
  tumble draw: counter (step_lo, step_hi, gid, 0xA5A5A5A5), key = polarity
    stream; tumble when u53(v0, v1) < p_tumble, theta = (float)(2 pi
    u53(v2, v3))                                   src/kernels.cu:98-112
  interaction integral I = sum phi_n * grad(phi_n) * sum_{m!=n} phi_m^2,
    grad = central difference / 2                   src/kernels.cu:254-298
                                                    (sign gate src/main.cu:179-214)
  velocity v = v_A (cos theta, sin theta) + (60 kappa / (xi lambda^2)) I
                                                    src/kernels.cu:300-342
  9-point Laplacian (4 edges + diagonals - 20 centre) / 6
                                                    include/params.cuh:75-92
  rhs = mrel (gamma lap - (30/lambda^2) gamma phi (1-phi)(1-2phi)
        + (2 mu / A0)(A0 - V) phi - (60 kappa/lambda^2) phi S_other)
        - (v . grad phi),  mrel = M_pf / 0.5         include/kernels.cuh:121-138
  explicit Euler phi' = phi + dt rhs                 src/kernels.cu:354-443
  V = sum phi^2 (the area entering the next step's constraint)
                                                    src/kernels.cu:447-592
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

K_PI = 3.14159265358979323846
MASK32 = 0xFFFFFFFF


def philox_py(ctr, key):
    """Philox4x32-10 on Python integers (Salmon et al., SC'11): multipliers
    0xD2511F53 / 0xCD9E8D57, Weyl increments 0x9E3779B9 / 0xBB67AE85."""
    c0, c1, c2, c3 = (int(v) & MASK32 for v in ctr)
    k0, k1 = (int(v) & MASK32 for v in key)
    for rnd in range(10):
        if rnd:
            k0 = (k0 + 0x9E3779B9) & MASK32
            k1 = (k1 + 0xBB67AE85) & MASK32
        p0 = 0xD2511F53 * c0
        p1 = 0xCD9E8D57 * c2
        c0, c1, c2, c3 = (((p1 >> 32) ^ c1 ^ k0) & MASK32, p1 & MASK32,
                          ((p0 >> 32) ^ c3 ^ k1) & MASK32, p0 & MASK32)
    return (c0, c1, c2, c3)


def u53(a: int, b: int) -> float:
    return float(((int(a) << 32) | int(b)) >> 11) * 2.0 ** -53


@dataclass
class Coeffs:
    dt: float
    gamma: np.ndarray      # (N,) per-cell stiffness
    v_A: np.ndarray        # (N,) per-cell active speed
    M_pf: np.ndarray       # (N,) per-cell mobility
    lam: float
    kappa: float
    xi: float
    mu: float
    target_radius: float
    p_tumble: float
    polarity_seed: int


def coeffs_from(p, cells) -> Coeffs:
    """Coefficients of a SimParams-like object and a CellState array."""
    return Coeffs(dt=float(p.dt), gamma=cells["gamma"].astype(np.float64),
                  v_A=cells["v_A"].astype(np.float64),
                  M_pf=cells["M_pf"].astype(np.float64), lam=float(p.lam),
                  kappa=float(p.kappa), xi=float(p.xi), mu=float(p.mu),
                  target_radius=float(p.target_radius), p_tumble=float(p.p_tumble()),
                  polarity_seed=int(p.polarity_stream()))


def _sh(f, dy, dx):
    """f sampled at (y + dy, x + dx), periodic."""
    return np.roll(f, shift=(-dy, -dx), axis=(-2, -1))


def laplacian9(f: np.ndarray) -> np.ndarray:
    edges = _sh(f, 1, 0) + _sh(f, -1, 0) + _sh(f, 0, 1) + _sh(f, 0, -1)
    diag = _sh(f, 1, 1) + _sh(f, 1, -1) + _sh(f, -1, 1) + _sh(f, -1, -1)
    return (4.0 * edges + diag - 20.0 * f) / 6.0


def gradients(f: np.ndarray):
    return 0.5 * (_sh(f, 0, 1) - _sh(f, 0, -1)), 0.5 * (_sh(f, 1, 0) - _sh(f, -1, 0))


class DenseSim2D:
    """All cells as full periodic fields, FP64 throughout."""

    def __init__(self, phi_world: np.ndarray, theta: np.ndarray, gid: np.ndarray,
                 co: Coeffs, step: int, philox=philox_py):
        self.phi = np.array(phi_world, np.float64)        # (N, L, L), [y, x]
        self.theta = np.array(theta, np.float32)
        self.gid = np.array(gid, np.int64)
        self.co = co
        self.step_index = int(step)
        self.philox = philox
        self.I = np.zeros((len(self.gid), 2))
        self.v = np.zeros((len(self.gid), 2))
        self.tumbles = np.zeros(len(self.gid), np.int64)

    def _draw(self, n: int) -> None:
        s, seed = self.step_index, self.co.polarity_seed
        r = self.philox((s & MASK32, s >> 32, int(self.gid[n]) & MASK32, 0xA5A5A5A5),
                        (seed & MASK32, seed >> 32))
        if u53(r[0], r[1]) < self.co.p_tumble:
            self.theta[n] = np.float32(2.0 * K_PI * u53(r[2], r[3]))
            self.tumbles[n] += 1

    def step(self, k: int = 1) -> None:
        co = self.co
        lam2 = co.lam * co.lam
        bulk = 30.0 / lam2
        rep = 60.0 * co.kappa / lam2
        mot = rep / co.xi
        A0 = K_PI * co.target_radius * co.target_radius
        for _ in range(k):
            phi = self.phi
            sq = phi * phi
            total = sq.sum(axis=0)
            new = np.empty_like(phi)
            for n in range(phi.shape[0]):
                self._draw(n)
                f = phi[n]
                others = total - sq[n]
                gx, gy = gradients(f)
                Ix = float(np.sum(f * gx * others))
                Iy = float(np.sum(f * gy * others))
                th = float(self.theta[n])
                vx = co.v_A[n] * math.cos(th) + mot * Ix
                vy = co.v_A[n] * math.sin(th) + mot * Iy
                V = float(np.sum(sq[n]))
                g = co.gamma[n]
                passive = (g * laplacian9(f) - bulk * g * f * (1.0 - f) * (1.0 - 2.0 * f)
                           + (2.0 * co.mu / A0) * (A0 - V) * f - rep * f * others)
                rhs = (co.M_pf[n] / 0.5) * passive - (vx * gx + vy * gy)
                new[n] = f + co.dt * rhs
                self.I[n] = (Ix, Iy)
                self.v[n] = (vx, vy)
            self.phi = new
            self.step_index += 1

    def volumes(self) -> np.ndarray:
        return (self.phi * self.phi).sum(axis=(1, 2))

    def centroids_in_frame(self, gx0: np.ndarray, gy0: np.ndarray) -> np.ndarray:
        """Centroid of phi^2 in each cell's frame: offsets (x - gx0) mod L."""
        N, L = self.phi.shape[0], self.phi.shape[-1]
        out = np.zeros((N, 2))
        for n in range(N):
            w = self.phi[n] ** 2
            ax = (np.arange(L) - int(gx0[n])) % L
            ay = (np.arange(L) - int(gy0[n])) % L
            V = w.sum()
            out[n] = ((w * ax[None, :]).sum() / V, (w * ay[:, None]).sum() / V)
        return out
