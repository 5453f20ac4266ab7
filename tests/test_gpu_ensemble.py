"""The batched ensemble (pfh2d_launch_batch_step) against per-replica oracle
runs: each replica of one persistent launch per step, with its own side
(256 to 512), population and Philox streams, follows the oracle's lone
simulation of that replica bit for bit (field, aggregate, cell states,
classes); graph replay and Morton queue order leave the state unchanged; a
replica equals the lone single-population step of the same workload."""
import sys

import numpy as np
import pytest

from pfh import params

pytestmark = pytest.mark.gpu

MIXED = [(256, 16, 100), (512, 40, 101), (384, 12, 102), (256, 46, 103)]


def say(msg):
    print(f"[gpu-ensemble] {msg}", flush=True)
    sys.stdout.flush()


def _mixed(tau=None):
    wls = [params.ensemble_replica(L, N, s) for L, N, s in MIXED]
    if tau is not None:
        for w in wls:
            w.params.tau = tau
    return wls


def _check_replica(ens, r, orc, what):
    phi = ens.replica_phi(r).view(np.uint32)
    assert np.array_equal(phi, orc.phi_now().view(np.uint32)), f"{what}: phi of replica {r}"
    assert np.array_equal(ens.replica_S(r), orc.S_now()), f"{what}: S of replica {r}"
    got, want = ens.replica_cells(r), orc.cells
    for k in ("gx0", "gy0", "cls", "V", "Cx", "Cy", "perim", "bb_lo_x", "bb_hi_x", "bb_lo_y",
              "bb_hi_y", "vx", "vy", "Ix", "Iy", "theta", "tumble_ctr", "shift_ctr",
              "promote_ctr", "phi_max"):
        assert np.array_equal(got[k], want[k]), f"{what}: {k} of replica {r}"
    assert np.array_equal(ens.replica_classes(r), orc.cls), f"{what}: classes of replica {r}"


def test_mixed_batch_matches_per_replica_oracle():
    """Four replicas (sides 256, 512, 384, 256; 16 to 46 cells, the last at
    90% coverage) in one launch per step, tumbling often: bit-exact with four
    oracle runs after 3 and after 7 steps (the latter across a graph body)."""
    from oracle import pfo
    from pfh import ensemble
    wls = _mixed(tau=0.05)
    ens = ensemble.Ensemble2D(wls)
    orcs = [pfo.OracleSim2D(w) for w in wls]
    say(f"mixed batch: {ens.N} cells in {ens.R} replicas")
    ens.step(3)
    ens.sync()
    for r, o in enumerate(orcs):
        o.step(3)
        _check_replica(ens, r, o, "3 steps")
    ens.capture_graph()
    ens.run(4)      # steps 3..6: single launches up to 6, none replayed
    ens.sync()
    for r, o in enumerate(orcs):
        o.step(4)
        _check_replica(ens, r, o, "7 steps")
    assert not ens.fatal(), ens.flag_counts()


def test_graph_replay_and_morton_order_leave_state_unchanged():
    from pfh import ensemble
    wls = _mixed()
    a = ensemble.Ensemble2D(wls)
    b = ensemble.Ensemble2D(wls)
    c = ensemble.Ensemble2D(wls, morton=True)
    a.step(12)
    b.capture_graph()
    b.run(12)
    c.capture_graph()
    c.run(12)
    for e in (a, b, c):
        e.sync()
    for r in range(a.R):
        for other in (b, c):
            assert np.array_equal(a.replica_phi(r).view(np.uint32),
                                  other.replica_phi(r).view(np.uint32))
            assert np.array_equal(a.replica_S(r), other.replica_S(r))
            assert a.replica_cells(r).tobytes() == other.replica_cells(r).tobytes()


def test_replica_equals_single_population_step():
    """A replica with a side above 286 follows the single-population launch
    (pfh2d_launch_step) of the same workload bit for bit."""
    from pfh import ensemble, sim2d
    wls = _mixed()
    ens = ensemble.Ensemble2D(wls)
    lone = sim2d.Sim2D(wls[1])
    ens.step(5)
    lone.step(5)
    ens.sync()
    lone.sync()
    assert np.array_equal(ens.replica_phi(1).view(np.uint32), lone.phi_now().view(np.uint32))
    assert np.array_equal(ens.replica_S(1), lone.S_now())
    assert ens.replica_cells(1).tobytes() == lone.cells().tobytes()


def test_autotuned_ensemble_and_3d_shapes_keep_results():
    """The autotuner times every candidate and picks the fastest; the tuned
    ensemble order leaves every replica's state unchanged, and a tuned 3D
    launch shape matches the oracle run of the chosen shard count."""
    from oracle import pfo3
    from pfh import autotune, ensemble, params3d
    tuner = autotune.Autotuner()
    wls = _mixed()
    ens, best = autotune.tuned_ensemble(wls, tuner)
    assert best["order"] in autotune.Autotuner.ORDERS
    assert len(tuner.log) == len(autotune.Autotuner.ORDERS)
    assert best["ms"] == min(e["ms"] for e in tuner.log)
    ref = ensemble.Ensemble2D(wls)
    ens.step(4)
    ref.step(4)
    ens.sync()
    ref.sync()
    for r in range(ref.R):
        assert ens.replica_cells(r).tobytes() == ref.replica_cells(r).tobytes()
        assert np.array_equal(ens.replica_S(r), ref.replica_S(r))
    w3 = params3d.workload_small()
    sim, b3 = autotune.tuned_sim3d(w3, tuner)
    say(f"3D tuned shape {b3}")
    assert b3["candidates"] == len(tuner.log) > 1
    assert b3["ms"] == min(e["ms"] for e in tuner.log)
    o3 = pfo3.OracleSim3D(w3.params, w3.cx, w3.cy, w3.cz, shards=b3["shards"], threads=4)
    sim.step(2)
    o3.step(2)
    sim.sync()
    assert np.array_equal(sim.phi_now().view(np.uint32), o3.phi_now().view(np.uint32))
    assert sim.cells().tobytes() == o3.cells.tobytes()


def test_c4_long_run_S_exact():
    """The 64-replica ensemble (sides 256/512/1024, seeds 100..163) over its
    full 5000 steps (5004: whole six-step graphs): no fatal flag, every
    replica's field finite and bounded, and every replica's aggregate S equal
    bit for bit to the oracle's scatter of that replica's stored fields (no
    drift from the incremental scatter into the shared slot arenas or from
    the arena-wide clear-ahead)."""
    from oracle import pfo
    from pfh import ensemble
    lib = pfo.load()
    wls = params.ensemble_c4()
    ens = ensemble.Ensemble2D(wls)
    ens.capture_graph()
    say(f"C4: {ens.N} cells in {ens.R} replicas, 5004 steps from six-step graphs")
    ens.run(5004)
    ens.sync()
    assert ens.steps_done == 5004
    assert not ens.fatal(), ens.flag_counts()
    for r in range(ens.R):
        phi = np.ascontiguousarray(ens.replica_phi(r))
        assert np.isfinite(phi).all(), f"replica {r}"
        assert phi.max() <= 1.5 and phi.min() >= -0.5 and phi.max() > 0.9, f"replica {r}"
        cells = ens.replica_cells(r)     # bound: the oracle reads through their addresses
        cls = ens.replica_classes(r)
        L, P = ens.Ls[r], ens.Ps[r]
        S = np.zeros(L * P, np.uint32)
        flags = np.zeros(8, np.uint32)
        lib.pfo2d_scatter_all(phi.ctypes.data, cells.ctypes.data, cls.ctypes.data,
                              S.ctypes.data, ens.Ns[r], L, P, flags.ctypes.data)
        assert np.array_equal(S, ens.replica_S(r)), f"S of replica {r} (side {L})"
    say("C4: every replica's S exact")
