"""Pins the surrogate network's CPU oracle (oracle/sur_ref.py): its
backward pass against central finite differences in float64, the
bfloat16 rounding rule against known bit patterns, the backward-data
weight flip against the direct adjoint definition, and the convolution's
tap offsets, channel order and zero padding against single impulses."""
import numpy as np
import pytest

from oracle import sur_ref as R


def test_bf16_rounding_known_patterns():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0,
                  3.0e38, 1e-40], np.float32)
    bits = R.bf16_bits(x)
    # ties go to even: 1 + 2^-8 -> 1.0 (0x3F80), 1 + 3*2^-8 -> 1 + 2^-6 (0x3F82)
    assert bits.tolist()[:6] == [0x3F80, 0x3F80, 0x3F81, 0x3F82, 0xC020, 0x0000]
    assert R.bf16_value(bits)[2] == np.float32(1.0 + 2 ** -7)
    assert np.isnan(R.bf16_value(R.bf16_bits(np.array([np.nan], np.float32))))[0]


def _tiny(rng, H=7, Wd=6, n=2):
    p = R.Params(rng.normal(0, 0.4, (32, 9, 8)), rng.normal(0, 0.1, 32),
                 rng.normal(0, 0.1, (32, 9, 32)), rng.normal(0, 0.1, 32),
                 rng.normal(0, 0.1, (32, 9, 32)), rng.normal(0, 0.1, 32),
                 rng.normal(0, 0.3, (9, 32)), 0.05, bf16=False)
    X = rng.normal(0, 1.0, (n, H, Wd, 8))
    T = rng.normal(0, 0.5, (n, H, Wd))
    return p, X, T


def _loss(p, X, T):
    return float(np.sum((R.forward(X, p)["dphi"] - T) ** 2) / T.size)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    p, X, T = _tiny(rng)
    g = R.backward(R.forward(X, p), T, p)
    h = 1e-6
    for name, idx in (("W1", (3, 4, 1)), ("W2", (7, 0, 5)), ("W3", (11, 8, 30)),
                      ("W4", (0, 2, 9)), ("b1", (5,)), ("b3", (17,))):
        arr = getattr(p, name)
        old = arr[idx]
        arr[idx] = old + h
        lp = _loss(p, X, T)
        arr[idx] = old - h
        lm = _loss(p, X, T)
        arr[idx] = old
        num = (lp - lm) / (2 * h)
        ana = g[name][idx[1:]] if name == "W4" else g[name][idx]
        assert abs(num - ana) <= 1e-6 * (1 + abs(num)), (name, num, ana)
    p.b4 += h
    lp = _loss(p, X, T)
    p.b4 -= 2 * h
    lm = _loss(p, X, T)
    p.b4 += h
    assert abs((lp - lm) / (2 * h) - g["b4"]) <= 1e-6


def test_dgrad_is_the_adjoint_of_conv():
    rng = np.random.default_rng(6)
    X = rng.normal(size=(1, 5, 6, 32))
    Y = rng.normal(size=(1, 5, 6, 32))
    W = rng.normal(size=(32, 9, 32))
    # <conv(X, W), Y> == <X, dgrad(Y, W)>
    assert np.isclose(np.sum(R.conv(X, W) * Y), np.sum(X * R.conv_dgrad(Y, W)), rtol=1e-12)
    # the flip the kernels use: Wt[ci][t][co] = W[co][8 - t][ci]
    Wt = np.zeros_like(W)
    for ci in range(32):
        for t in range(9):
            Wt[ci, t, :] = W[:, 8 - t, ci]
    assert np.allclose(R.conv_dgrad(Y, W), R.conv(Y, Wt))


def test_conv_impulse_fixes_tap_offsets_and_padding():
    """A single input impulse through a single non-zero weight lands at the
    pixel the definition out[p] = sum_t X[p + (ky-1, kx-1)] W[t], t = 3 ky + kx,
    names, in the named output channel, and nothing else; an impulse on the
    crop edge loses exactly the taps that would read outside (zero padding)."""
    from oracle import sur_ref as R
    H = 7
    for (y0, x0) in ((3, 3), (0, 0), (6, 2)):
        for t in range(9):
            ky, kx = divmod(t, 3)
            X = np.zeros((1, H, H, 4))
            X[0, y0, x0, 2] = 1.0
            W = np.zeros((5, 9, 4))
            W[3, t, 2] = 0.75
            out = R.conv(X, W)
            py, px = y0 - (ky - 1), x0 - (kx - 1)
            expect = np.zeros_like(out)
            if 0 <= py < H and 0 <= px < H:
                expect[0, py, px, 3] = 0.75
            assert np.array_equal(out, expect), (y0, x0, t)


def test_check_step_stability_rule():
    """A rejected cell may take the check step's field only where one
    explicit step of K dt is a stable integrator: on the monolayer's
    coefficients the 8-step block qualifies (rate x K dt = 1.46) and the
    16-step block does not (2.9), so the 16-step runner falls back to whole
    populations. Halving dt doubles the admissible K."""
    import dataclasses
    from pfh import params, surrogate as su
    p = params.workload_c2().params
    assert su.check_step_stable(p, 8)
    assert not su.check_step_stable(p, 16)
    assert max(k for k in range(1, 40) if su.check_step_stable(p, k)) == 10
    half = dataclasses.replace(p, dt=p.dt / 2)
    assert max(k for k in range(1, 40) if su.check_step_stable(half, k)) == 21
