"""Pins the oracle's scalar primitives against the reference's golden vectors
and against closed forms, so that a wrong constant, index or operation order
in the oracle fails here before it can vouch for a kernel.

Golden vectors: tests/common_primitives_cpu_test.cpp:54-61 (Philox4x32-10,
the Random123 known-answer set) and the phase_field_rhs fixtures of
tests/mobility_model_cpu_test.cpp:61-75 of the reference.
"""
import ctypes as C
import math
import random
from fractions import Fraction

import numpy as np
import pytest

from pfh import params

# (counter, key, expected) -- reference tests/common_primitives_cpu_test.cpp:54-61
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((1, 2, 3, 4), (5, 6), (0xC0C839BC, 0x889C87C5, 0x61986739, 0x2D4623D0)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2,
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
]


@pytest.mark.parametrize("ctr,key,expected", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, expected):
    assert oracle.philox(ctr, key) == expected
    assert params.philox4x32_10(*ctr, *key) == expected


def test_philox_matches_independent_python(oracle):
    rng = random.Random(5)
    for _ in range(2000):
        ctr = tuple(rng.getrandbits(32) for _ in range(4))
        key = tuple(rng.getrandbits(32) for _ in range(2))
        assert oracle.philox(ctr, key) == params.philox4x32_10(*ctr, *key)


def test_uniform53_closed_form(oracle):
    lib = oracle.load()
    assert lib.pfo_u53(0, 0) == 0.0
    assert lib.pfo_u53(0, 0x800) == 2.0 ** -53
    assert lib.pfo_u53(0x80000000, 0) == 0.5
    assert lib.pfo_u53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert lib.pfo_u53(0, 0x7FF) == 0.0          # the low 11 bits are dropped


def _q_exact(phi: np.float32) -> int:
    """RN_even(min(fl32(phi*phi), 4) * 2^27) in exact rationals."""
    sq = np.float32(phi) * np.float32(phi)
    if math.isnan(sq):
        return 4 << 27          # fminf(NaN, 4) = 4
    v = min(Fraction(float(sq)), Fraction(4)) * (1 << 27)
    return round(v)             # Fraction rounds half to even


def test_q_of_edges_and_random(oracle):
    lib = oracle.load()
    f32 = np.float32
    edges = {f32(0): 0, f32(1): 1 << 27, f32(2): 1 << 29, f32(3): 1 << 29,
             f32(-0.5): 1 << 25, f32(2.0 ** -14): 0,          # 0.5 quantum: ties to even
             f32(3 * 2.0 ** -15): 1, f32(np.nan): 1 << 29}
    for phi, q in edges.items():
        assert lib.pfo_qof(float(phi)) == q, phi
    rng = np.random.default_rng(7)
    vals = np.concatenate([rng.uniform(-2.5, 2.5, 4000), rng.uniform(-1e-3, 1e-3, 2000),
                           rng.uniform(-1e-4, 1e-4, 2000)]).astype(np.float32)
    for phi in vals:
        assert lib.pfo_qof(float(phi)) == _q_exact(phi), phi


def test_q_of_tie_rounds_to_even(oracle):
    lib = oracle.load()
    # phi^2 * 2^27 = k + 1/2 exactly for phi = sqrt((k+1/2) 2^-27) when the
    # square is exact: phi = 2^-14 gives 1/2 -> 0; phi = 3*2^-14 -> 9/2 -> 4.
    assert lib.pfo_qof(float(np.float32(3 * 2.0 ** -14))) == 4
    assert lib.pfo_qof(float(np.float32(5 * 2.0 ** -14))) == 12   # 25/2 -> 12


def test_sincos_matches_libm(oracle):
    rng = np.random.default_rng(11)
    thetas = np.concatenate([rng.uniform(0, 2 * math.pi, 20000), np.float32([0, math.pi / 2,
                             math.pi, 1.5 * math.pi, 2 * math.pi, 1e-7])]).astype(np.float32)
    mism = 0
    for th in thetas:
        s, c = oracle.sincosf(float(th))
        rs, rc = np.float32(math.sin(float(th))), np.float32(math.cos(float(th)))
        for got, ref, exact in ((s, rs, math.sin(float(th))), (c, rc, math.cos(float(th)))):
            if got != ref:
                mism += 1
                # only a value within a hair of a rounding midpoint may differ
                assert abs(float(got) - exact) <= float(np.spacing(np.float32(abs(exact)))), th
    assert mism <= 2


def test_tanh_matches_libm(oracle):
    lib = oracle.load()
    for y in np.concatenate([np.linspace(-25, 25, 5001), np.logspace(-12, 1.3, 400)]):
        got, ref = lib.pfo_tanh(float(y)), math.tanh(float(y))
        assert abs(got - ref) <= 8e-16 * max(abs(ref), 1e-300), y   # ~3.6 ulp


# ---- phase_field_rhs: fixtures of tests/mobility_model_cpu_test.cpp:61-75 ----
PIXELS = [(0.125, 0.5, 0.25, 0.03, -0.02), (-0.75, 0.9999, 0.0, -0.4, 0.4),
          (2.5e-3, 1.0e-5, 0.9, 1.0e-4, -1.0e-4), (-3.1, 0.31, 1.7, 0.11, 0.07),
          (0.0, 0.0, 0.0, 0.0, 0.0), (1.0, 1.0, 1.0, 1.0, 1.0),
          (-1.9e-4, 0.62, 3.9, -0.9, -0.8)]
COEFFS = [(1.0, 0.612244879, 1.3e-4, 12.2448978, 0.01, -0.02),
          (0.35, 0.214285716, -2.0e-4, 12.2448978, -0.004, 0.006),
          (2.0, 1.5, 0.5, 0.0, 0.0, 0.3)]


def legacy_rhs(lap, phi, So, gx, gy, gam, dwC, volC, repC, vx, vy):
    """The reference's pre-mobility expression in numpy binary32 (one IEEE
    operation per step, left to right, no contraction)."""
    f = np.float32
    lap, phi, So, gx, gy = map(f, (lap, phi, So, gx, gy))
    gam, dwC, volC, repC, vx, vy = map(f, (gam, dwC, volC, repC, vx, vy))
    one, two = f(1), f(2)
    dw = (phi * (one - phi)) * (one - two * phi)
    return (((gam * lap - dwC * dw) + volC * phi) - (repC * phi) * So) - (vx * gx + vy * gy)


def _bits(x) -> int:
    return int(np.float32(x).view(np.uint32))


def test_rhs_default_path_equals_legacy_bitwise(oracle):
    lib = oracle.load()
    for px in PIXELS:
        for cf in COEFFS:
            got = lib.pfo_rhs(*[float(np.float32(v)) for v in px + cf], 1.0)
            assert _bits(got) == _bits(legacy_rhs(*px, *cf)), (px, cf)


def test_rhs_mobility_scales_passive_only(oracle):
    lib = oracle.load()
    f = np.float32
    for px in PIXELS:
        for cf in COEFFS:
            for mrel in (0.5, 1.0 / 0.35, 2.857143, 4.0):
                mrel = f(mrel)
                a = [float(f(v)) for v in px + cf[:4]]
                p1 = f(lib.pfo_rhs(*a, 0.0, 0.0, 1.0))
                pm = f(lib.pfo_rhs(*a, 0.0, 0.0, float(mrel)))
                assert _bits(pm) == _bits(mrel * p1)
                adv = f(cf[4]) * f(px[3]) + f(cf[5]) * f(px[4])
                full = f(lib.pfo_rhs(*a, float(f(cf[4])), float(f(cf[5])), float(mrel)))
                assert _bits(full) == _bits(pm - adv)


def test_laplacian_closed_forms(oracle):
    lib = oracle.load()
    ys, xs = np.mgrid[0:7, 0:7].astype(np.float32)
    for field, want in ((xs * xs, 2.0), (ys * ys, 2.0), (xs * 0 + 3, 0.0),
                        (xs * ys, 0.0), (xs * xs + ys * ys, 4.0), (xs + 2 * ys, 0.0)):
        fld = np.ascontiguousarray(field, np.float32)
        ptr = fld.ctypes.data + 4 * (3 * 7 + 3)
        assert lib.pfo_lap9(C.c_void_p(ptr), 7) == pytest.approx(want, abs=0), want


def test_class_containing_matches_brute_force(oracle):
    lib = oracle.load()
    from pfh.abi import CLASSES
    for ex in range(0, 300, 3):
        for ey in range(0, 300, 7):
            for slack in (8, 20):
                fits = [c for c, (wx, wy, _, _) in enumerate(CLASSES)
                        if ex + slack <= wx and ey + slack <= wy]
                want = min(fits, key=lambda c: (CLASSES[c][0] * CLASSES[c][1], c)) if fits else -1
                assert lib.pfo_class_containing(ex, ey, slack) == want


def test_domain_side_and_pitch(oracle):
    lib = oracle.load()
    # rho windows (0.13635, 0.13707] -> 384 for 16 cells and (0.30680, 0.30710] -> 2048 for 1024 cells, R=20.
    assert lib.pfo_domain_side_for(16, 20.0, 0.13671) == 384
    assert lib.pfo_domain_side_for(1024, 20.0, 0.30695) == 2048
    assert params.domain_side_for(16, 20.0, 0.13671) == 384
    assert lib.pfo_s_pitch_for(1) == 32 and lib.pfo_s_pitch_for(2048) == 2048
    assert lib.pfo_s_pitch_for(385) == 416


def test_initial_draws_match_product_host(oracle):
    lib = oracle.load()
    for gid in range(0, 5000, 37):
        for stream in (1, 7, 1234, 2**40 + 3):
            assert np.float32(lib.pfo_ic_theta(gid, stream)) == params.ic_theta(gid, stream)
        assert lib.pfo_ic_v_A(gid, 99, 0.01, 0.3) == params.ic_v_A(gid, 99, 0.01, 0.3)
    for N, L, seed in ((16, 384, 1), (1024, 2048, 7), (37, 512, 5)):
        cx = np.zeros(N, np.float32)
        cy = np.zeros(N, np.float32)
        lib.pfo2d_grid_jitter_centres(N, L, seed, cx.ctypes.data, cy.ctypes.data)
        px, py = params.grid_jitter_centres(N, L, seed)
        assert np.array_equal(cx, px) and np.array_equal(cy, py)
        g = int(math.ceil(math.sqrt(N)))
        lib.pfo2d_harness_centres(N, L, g, 4.0, seed, cx.ctypes.data, cy.ctypes.data)
        px, py = params.harness_centres(N, L, g, 4.0, seed)
        assert np.array_equal(cx, px) and np.array_equal(cy, py)


def test_coefficients_bitwise_identities():
    """common_primitives_cpu_test.cpp:24-39: the helpers are the plain
    formulas, evaluated in the reference's order."""
    p = params.SimParams(lam=7.0, kappa=10.0, xi=1500.0, mu=1.0, target_radius=49.0)
    assert p.bulk() == 30.0 / (7.0 * 7.0)
    assert p.interaction() == 60.0 * 10.0 / (7.0 * 7.0)
    assert p.motility() == (60.0 * 10.0 / (7.0 * 7.0)) / 1500.0
    assert p.volume() == 2.0 * 1.0 / (3.14159265358979323846 * 49.0 * 49.0)
    assert p.interface_k() == 2.7386127875258306 / 7.0
    assert p.init_radius() == 49.0 + 0.5 / (2.7386127875258306 / 7.0)
    assert abs(p.interaction() / p.motility() - p.xi) <= 1e-9 * p.xi
    assert p.p_tumble() == -math.expm1(-0.01 / 1.0e4)
