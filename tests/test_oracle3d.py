"""Pins the 3D CPU oracle (oracle/pfo3d.c):
  - closed forms of the 27-point stencil (include/pf3d/params.cuh:73-90,
    tests/3d/model_one_step_cpu_test.cpp:54-66) and its weights, in both
    operation orders;
  - the reference's own host code compiled into oracle/_ref/ (oracle/ref.mk):
    the dense FP64 one-step reference (include/pf3d/reference.hpp:256-331),
    the Philox polarity draws (include/pf3d/rng.cuh), the seed-radius
    bisection, the brick sizing, q_of and the sequential centre placement;
  - invariants: an isolated sphere has zero interaction, integer translation
    commutes with the step, a rejected step restores the accepted state.
"""
import ctypes as C
import random
from pathlib import Path

import numpy as np
import pytest

from pfh import abi, params3d

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref"
SO3 = REF / "libref3d.so"
ONE_STEP_TEST = REF / "model_one_step_cpu_test"
needs_ref = pytest.mark.skipif(not SO3.exists(), reason="oracle/_ref not built (no reference tree)")


@pytest.fixture(scope="module")
def o3():
    from oracle import pfo3
    pfo3.load()
    return pfo3


@pytest.fixture(scope="module")
def ref3():
    lib = C.CDLL(str(SO3))
    D, I, P = C.c_double, C.c_int, C.c_void_p
    lib.ref3d_calibrate.argtypes = [D, D, I, D, I, P]
    lib.ref3d_calibrate.restype = I
    lib.ref3d_initial_polarity.argtypes = [C.c_int64, C.c_uint64, P]
    lib.ref3d_tumble.argtypes = [C.c_uint64, C.c_int64, C.c_uint64, D, P, P]
    lib.ref3d_tumble.restype = I
    lib.ref3d_brick_edge.argtypes = [D, D]
    lib.ref3d_brick_edge.restype = I
    lib.ref3d_domain_side.argtypes = [I, D, D]
    lib.ref3d_domain_side.restype = I
    lib.ref3d_q_of.argtypes = [C.c_float]
    lib.ref3d_q_of.restype = C.c_uint32
    lib.ref3d_sequential_centres.argtypes = [I, D, D, C.c_uint64, P, P, P]
    lib.ref3d_sequential_centres.restype = I
    lib.ref3d_one_step.argtypes = [I, I, P, P, P, P, D, D, D, D, D, D, P, P, P, P, P]
    lib.ref3d_one_step.restype = I
    return lib


def _p(a):
    return a.ctypes.data


# ---- stencil ----
def _nbhd(f):
    """3x3x3 neighbourhood n[(dz+1)*9 + (dy+1)*3 + (dx+1)] of f(dx, dy, dz)."""
    return np.array([f(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                     for dx in (-1, 0, 1)], np.float32)


@pytest.mark.parametrize("order", [0, 1])
def test_lap27_closed_forms(o3, order):
    lib = o3.load()
    const = _nbhd(lambda dx, dy, dz: 3.25)
    assert lib.pfo3d_lap27(_p(const), order) == 0.0
    quad = _nbhd(lambda dx, dy, dz: float((4 + dx) ** 2 + (4 + dy) ** 2 + (4 + dz) ** 2))
    assert lib.pfo3d_lap27(_p(quad), order) == pytest.approx(6.0, abs=2e-6)
    # one impulse at a time recovers the weights 14 / 3 / 1 and -128, over 30
    for k in range(27):
        e = np.zeros(27, np.float32)
        e[k] = 1.0
        dz, dy, dx = k // 9 - 1, (k // 3) % 3 - 1, k % 3 - 1
        d = (dx != 0) + (dy != 0) + (dz != 0)
        w = [-128, 14, 3, 1][d]
        assert lib.pfo3d_lap27(_p(e), order) == np.float32(np.float32(w) * np.float32(1.0 / 30.0))


def test_lap27_orders_agree_on_random_fields(o3):
    lib = o3.load()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        n = rng.random(27).astype(np.float32)
        a = lib.pfo3d_lap27(_p(n), 0)
        b = lib.pfo3d_lap27(_p(n), 1)
        assert abs(a - b) <= 4e-6 * (1 + np.abs(n).sum() * 5)


# ---- reference host code ----
@needs_ref
def test_reference_one_step_cpu_test_passes():
    import subprocess
    out = subprocess.run([str(ONE_STEP_TEST)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert "passed" in out.stdout


@needs_ref
def test_initial_polarity_and_tumbles_match_reference(o3, ref3):
    lib = o3.load()
    rng = random.Random(5)
    a = np.zeros(3)
    b = np.zeros(3)
    for _ in range(3000):
        gid = rng.getrandbits(40)
        stream = rng.getrandbits(64)
        ref3.ref3d_initial_polarity(gid, stream, _p(a))
        lib.pfo3d_initial_polarity(gid, stream, _p(b))
        np.testing.assert_allclose(b, a, rtol=0, atol=4e-16)
        assert np.float32(b).tolist() == np.float32(a).tolist() or \
            np.abs(np.float32(b) - np.float32(a)).max() <= np.float32(6e-8)
        assert abs(np.linalg.norm(b) - 1.0) < 1e-15
    cur = np.array([0.0, 0.0, 1.0])
    hits = 0
    for _ in range(4000):
        step, gid, stream = rng.getrandbits(40), rng.getrandbits(20), rng.getrandbits(64)
        p = rng.choice([1e-6, 0.3, 0.999])
        ta = ref3.ref3d_tumble(step, gid, stream, p, _p(cur), _p(a))
        tb = lib.pfo3d_tumble_sample(step, gid, stream, p, _p(cur), _p(b))
        assert ta == tb
        hits += ta
        np.testing.assert_allclose(b, a, rtol=0, atol=4e-16)
    assert 0 < hits < 4000


@needs_ref
def test_sizing_q_and_calibration_match_reference(o3, ref3):
    lib = o3.load()
    plib = abi.load()
    for R, lam in ((12.0, 3.5), (12.0, 7.0), (49.0, 7.0), (8.0, 3.5), (4.0, 2.0), (20.0, 5.25)):
        e = ref3.ref3d_brick_edge(R, lam)
        assert lib.pfo3d_brick_edge(R, lam) == e == plib.pfh3d_brick_edge(R, lam)
    assert ref3.ref3d_brick_edge(12.0, 3.5) == 48
    for n, R, rho in ((128, 12.0, 0.05555), (8, 8.0, 0.02), (500, 30.0, 0.4)):
        assert plib.pfh3d_domain_side(n, R, rho) == ref3.ref3d_domain_side(n, R, rho)
    assert plib.pfh3d_domain_side(128, 12.0, 0.05555) == 256
    rng = np.random.default_rng(9)
    vals = np.concatenate([rng.normal(0, 0.8, 20000), [0.0, 2.0 ** -14, 2.0, 2.5, -2.1]])
    for v in vals.astype(np.float32):
        assert ref3.ref3d_q_of(float(v)) == o3.pfo.load().pfo_qof(float(v))
    for R, lam, B in ((12.0, 3.5, 48), (8.0, 3.5, 40), (4.0, 2.0, 24)):
        out = np.zeros(3)
        it = ref3.ref3d_calibrate(R, lam, B, 1e-10, 80, _p(out))
        assert it > 0
        assert o3.seed_radius(R, lam, B) == out[0]
        r = C.c_double(0.0)
        assert plib.pfh3d_calibrate_seed_radius(R, lam, B, C.byref(r)) == it
        assert r.value == out[0]


@needs_ref
def test_sequential_centres_match_reference(ref3):
    for n, side, R, seed in ((128, 256.0, 12.0, 1234), (8, 96.0, 8.0, 11), (40, 70.0, 6.0, 99)):
        a = [np.zeros(n, np.float32) for _ in range(3)]
        drawn = ref3.ref3d_sequential_centres(n, side, R, seed, *[_p(v) for v in a])
        assert drawn > 0
        b = params3d.sequential_centres(n, side, R, seed)
        for u, v in zip(a, b):
            assert u.tobytes() == v.tobytes()


def _world(sim, phi, L):
    """Render every cell's brick into its own periodic world array (N, L, L, L)."""
    B = sim.B
    out = np.zeros((sim.N, L, L, L), np.float64)
    for n in range(sim.N):
        c = sim.cells[n]
        ix = (int(c["origin_x"]) + np.arange(B)) % L
        iy = (int(c["origin_y"]) + np.arange(B)) % L
        iz = (int(c["origin_z"]) + np.arange(B)) % L
        out[n][np.ix_(iz, iy, ix)] = phi[n].reshape(B, B, B)
    return out


@needs_ref
@pytest.mark.parametrize("f64", [True, False])
def test_oracle_step_matches_reference_dense_one_step(o3, ref3, f64):
    """One oracle step from a seeded state against the reference's dense
    periodic FP64 one-step on the same fields (tumbling off, binary32
    polarities and coefficients as the kernels receive them)."""
    w = params3d.workload("t", 3, 5.0, 2.0, 36, seed=21, tau=1e300)
    p = w.params
    # pull the three cells into contact so the interaction term is exercised
    w.cx[:] = [18.0, 24.5, 18.0]
    w.cy[:] = [18.0, 18.0, 24.0]
    w.cz[:] = [18.0, 18.5, 19.0]
    sim = o3.OracleSim3D(p, w.cx, w.cy, w.cz, shards=1, f64=f64, threads=4)
    L, N = sim.L, sim.N
    before = _world(sim, sim.phi_now(), L)
    cells0 = sim.cells.copy()
    pol = np.stack([cells0["polarity_x"], cells0["polarity_y"], cells0["polarity_z"]],
                   axis=1).astype(np.float64).reshape(-1)
    gamma = cells0["gamma"].astype(np.float64)
    speed = cells0["v_A"].astype(np.float64)
    nxt = np.zeros(N * L ** 3)
    vol = np.zeros(N)
    integ = np.zeros(3 * N)
    vel = np.zeros(3 * N)
    agg = np.zeros(L ** 3, np.uint32)
    rc = ref3.ref3d_one_step(L, N, _p(np.ascontiguousarray(before.reshape(-1))), _p(gamma),
                             _p(speed), _p(pol), p.dt, p.lambda_, p.target_radius, p.kappa,
                             p.mu, p.xi, _p(nxt), _p(vol), _p(integ), _p(vel), _p(agg))
    assert rc == 0
    # the oracle's measurement of the same state
    np.testing.assert_allclose(cells0["V"], vol, rtol=1e-13)
    I = np.stack([cells0["Ix"], cells0["Iy"], cells0["Iz"]], axis=1).reshape(-1)
    assert np.abs(I).max() > 1e-3
    tol = 1e-12 if f64 else 2e-6
    np.testing.assert_allclose(I, integ, rtol=tol * 10, atol=tol)
    sim.step(1)
    after = _world(sim, sim.phi_now(), L).reshape(N, -1)
    err = np.abs(after - nxt.reshape(N, -1)).max()
    assert err <= (5e-9 if f64 else 1e-6), err
    assert np.abs(nxt).max() > 0.9


def test_isolated_sphere_has_zero_interaction(o3):
    w = params3d.workload("iso", 1, 8.0, 3.5, 64, seed=3)
    sim = o3.OracleSim3D(w.params, w.cx, w.cy, w.cz, shards=2, threads=4)
    c = sim.cells[0]
    assert c["Ix"] == 0.0 and c["Iy"] == 0.0 and c["Iz"] == 0.0
    sim.step(3)
    assert sim.cells[0]["Ix"] == 0.0 and not sim.fatal()


def test_integer_translation_commutes_with_steps(o3):
    base = params3d.workload("tr", 4, 6.0, 3.0, 64, seed=5)
    # quarter-voxel centres: the translated binary32 centres stay exact
    for v in (base.cx, base.cy, base.cz):
        v[:] = np.round(v * 4) / 4
    a = o3.OracleSim3D(base.params, base.cx, base.cy, base.cz, shards=3, threads=4)
    d = np.float32(8.0)   # even: origins round half to even
    b = o3.OracleSim3D(base.params, (base.cx + d) % 64, (base.cy + 2 * d) % 64,
                       (base.cz + 3 * d) % 64, shards=3, threads=4)
    a.step(4)
    b.step(4)
    assert a.phi_now().tobytes() == b.phi_now().tobytes()
    for f in ("V", "Ix", "Iy", "Iz", "velocity_x", "pending_shift_x", "bb_lo_z"):
        assert a.cells[f].tobytes() == b.cells[f].tobytes(), f
    L = 64
    S_a = a.S_now().reshape(L, L, a.P)[:, :, :L]
    S_b = b.S_now().reshape(L, L, b.P)[:, :, :L]
    assert np.array_equal(np.roll(S_a, (24, 16, 8), axis=(0, 1, 2)), S_b)


def test_rejected_step_restores_accepted_state(o3):
    w = params3d.workload("nan", 3, 6.0, 3.0, 64, seed=8)
    sim = o3.OracleSim3D(w.params, w.cx, w.cy, w.cz, shards=2, threads=4)
    sim.step(2)
    phi = sim.phi_now().copy()
    S = sim.S_now().copy()
    phi[1, 1000] = np.nan
    sim.load_state(phi, S, sim.cells.copy(), counters=(2, 2))
    cells0 = sim.cells.copy()
    sim.step(1)
    assert sim.flags[3] >= 1 and sim.fatal()
    assert sim.counters.tolist() == [2, 3]
    np.testing.assert_array_equal(sim.phi_now(), phi)
    assert np.array_equal(sim.S_now(), S)
    assert sim.cells.tobytes() == cells0.tobytes()
