"""GPU velocity observers (k_step_advection / k_step_moments variants of the
step and the dense observer, include/fused_moments.cuh:23-127,
src/velocity_moments.cu:43-128) against the oracle: the fused observers
bit-exact (advection accumulators and moment accumulators) on every class
path with forced tumbles, the observed steps bit-identical to unobserved
ones in every field, aggregate and cell-state byte but the accumulators
(tests/velocity_observer_test.cu:324-352), and the dense observer within
fp64 summation-order tolerance of the oracle's and of the fused rates."""
import numpy as np
import pytest

from pfh import abi, params
from test_gpu_parity_2d import _forced_state, assert_state_equal, say

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _adv_equal(gc, oc, where):
    assert np.array_equal(_bits(gc["adv_q"]), _bits(oc["adv_q"])), f"{where}: adv_q differs"
    assert np.array_equal(gc["adv_count"], oc["adv_count"]), f"{where}: adv_count differs"


def _cluster(tau=0.04):
    cx = np.array([143.0, 177.25, 144.5, 178.0], np.float32)
    cy = np.array([143.5, 144.0, 177.0, 178.75], np.float32)
    p = params.SimParams(Nx=320, Ny=320, num_cells=4, rho=0.05, target_radius=20.0,
                         tau=tau, seed=11, polarity_seed=12)
    return params.Workload2D("cluster", p, cx, cy)


@pytest.mark.parametrize("modes", [(2, 1, 2, 0, 2), (1, 1, 2)])
def test_observers_bit_exact_on_every_class_path(oracle, modes):
    from pfh import sim2d
    wl, phi, S, cells, cls = _forced_state(oracle)
    say(f"observers {modes} on the forced all-classes state L=701 N=12")
    sim = sim2d.Sim2D(wl)
    sim.attach_moments()
    orc = oracle.OracleSim2D(wl, threads=8)
    sim.load_state(phi, S, cells, cls, 99)
    orc.load_state(phi, S, cells, cls, 99)
    for k, mode in enumerate(modes):
        sim.step(1, observation=mode)
        orc.observation = mode
        orc.step(1)
        sim.sync()
        where = f"step {100 + k} observation {mode}"
        assert_state_equal(sim, orc, where)
        _adv_equal(sim.cells(), orc.cells, where)
        assert np.array_equal(_bits(sim.moments_host()), _bits(orc.moments)), \
            f"{where}: moment accumulators differ"
    m = sim.moments_host()
    assert (m["spatial_count"] == sum(1 for x in modes if x == 2)).all()
    assert (sim.cells()["adv_count"] == sum(1 for x in modes if x)).all()
    assert np.abs(m["held"][:, 0:4]).max() > 0.0


def test_observation_is_neutral_on_the_gpu(oracle):
    from pfh import sim2d
    wl, phi, S, cells, cls = _forced_state(oracle)
    out = []
    for modes in ((0, 0, 0, 0), (2, 1, 2, 2)):
        sim = sim2d.Sim2D(wl)
        sim.attach_moments()
        sim.load_state(phi, S, cells, cls, 5)
        for mode in modes:
            sim.step(1, observation=mode)
        sim.sync()
        c = sim.cells()
        c["adv_q"] = 0.0
        c["adv_count"] = 0
        out.append((sim.phi_now().view(np.uint32).copy(), sim.S_now().copy(), c.tobytes(),
                    sim.flag_counts().copy(), sim.classes().copy()))
    for a, b in zip(*out):
        assert np.array_equal(np.frombuffer(a, np.uint8) if isinstance(a, bytes) else a,
                              np.frombuffer(b, np.uint8) if isinstance(b, bytes) else b)


def test_dense_observer_matches_oracle_and_fused_rates(oracle):
    import torch
    from pfh import sim2d
    wl = _cluster()
    say("dense observer on the interacting 4-cell cluster L=320 after 20 steps")
    sim = sim2d.Sim2D(wl)
    sim.attach_moments()
    orc = oracle.OracleSim2D(wl, threads=4)
    sim.step(20)
    orc.step(20)
    dense = torch.zeros(sim.N * abi.MOMENT_DTYPE.itemsize, dtype=torch.uint8, device=sim.dev)
    sim.observe_dense(dense)
    od = np.zeros(sim.N, oracle.MOMENT_DTYPE)
    a = orc._args()
    orc.lib.pfo2d_velocity_moments(oracle.C.byref(a), od.ctypes.data, wl.params.dt)
    sim.step(1, observation=abi.OBSERVE_SPATIAL)
    sim.sync()
    gd = dense.cpu().numpy().view(abi.MOMENT_DTYPE)
    assert (gd["count"] == 1).all()
    # columns 8, 9 (the area source's moment) are zero up to rounding: they
    # are compared against the size of the other moments
    cols = [0, 1, 2, 3, 4, 5, 6, 7, 10, 11]
    scale = np.abs(od["q"][:, cols]).max(axis=0) + 1e-300
    assert (np.abs(gd["q"][:, cols] - od["q"][:, cols]) / scale).max() < 1e-12
    assert np.abs(gd["q"][:, 8:10]).max() < 1e-12 * np.abs(od["q"][:, 4:8]).max()
    np.testing.assert_array_equal(gd["q"][:, 0:2], od["q"][:, 0:2])   # same draws, same sincos
    dt = wl.params.dt
    m, c = sim.moments_host(), sim.cells()
    np.testing.assert_allclose(gd["q"][:, 0:4], c["adv_q"][:, 0:4], rtol=1e-13, atol=1e-18)
    for k in range(4):
        np.testing.assert_allclose(gd["q"][:, 4 + k] / dt, m["held"][:, k], rtol=1e-9,
                                   atol=1e-13)
    assert np.abs(c["Ix"]).max() > 1e-3
