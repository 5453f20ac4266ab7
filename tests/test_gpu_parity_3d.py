"""GPU parity of the 3D step (gfx950 kernels through the C ABI) against the
CPU oracle on the same seeded inputs.

Bar: bit-exact phi, S (Q5.27), CellState3D, flag counts and step counters
against the oracle in the kernels' operation order; <= 1e-5 max-abs phi
after 100 steps against the oracle in the reference's operation order and
against the FP64 oracle. Cases: fresh initialisation, the 96^3 / 8-cell
workload (brick 40: 400-lane plane groups in 7 waves) across shard and slab
counts, the 256^3 / 128-cell throughput workload (brick 56), graph replay,
restart, a recentring shift with the clip check, and a rejected step.
"""
import os
import sys

import numpy as np
import pytest

from pfh import abi, params3d

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(f"[gpu-parity-3d] {msg}", flush=True)
    sys.stdout.flush()


def assert_cells_equal(a, b, where=""):
    for f in a.dtype.names:
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            ok = np.array_equal(x.view(np.uint32 if x.itemsize == 4 else np.uint64),
                                y.view(np.uint32 if y.itemsize == 4 else np.uint64))
        else:
            ok = np.array_equal(x, y)
        if not ok:
            idx = np.argwhere(np.atleast_1d(x != y).reshape(len(x), -1).any(axis=1)).ravel()[:5]
            raise AssertionError(f"{where}: CellState3D.{f} differs at cells {idx}: "
                                 f"gpu={x[idx]} oracle={y[idx]}")


def assert_state_equal(sim, orc, where=""):
    gphi, ophi = sim.phi_now(), orc.phi_now()
    d = np.argwhere(gphi.view(np.uint32) != ophi.view(np.uint32))
    assert d.size == 0, (f"{where}: phi differs at {d.shape[0]} voxels, first {d[:3]}, "
                         f"max|d|={np.abs(gphi - ophi).max():.3e}")
    gS, oS = sim.S_now(), orc.S_now()
    assert np.array_equal(gS, oS), f"{where}: S differs at {np.count_nonzero(gS != oS)} words"
    assert_cells_equal(sim.cells(), orc.cells, where)
    assert np.array_equal(sim.flag_counts(), orc.flags), \
        f"{where}: flags gpu={sim.flag_counts()} oracle={orc.flags}"
    assert list(sim.counter_values()) == [int(v) for v in orc.counters], where


@pytest.fixture(scope="module")
def sim3():
    from pfh import sim3d
    return sim3d


@pytest.fixture(scope="module")
def o3():
    from oracle import pfo3
    pfo3.load()
    return pfo3


def _pair(sim3, o3, wl, shards, slabs=0, order=0, f64=False):
    sim = sim3.Sim3D(wl, shards=shards, slabs=slabs)
    orc = o3.OracleSim3D(wl.params, wl.cx, wl.cy, wl.cz, shards=shards, order=order, f64=f64,
                         threads=THREADS)
    return sim, orc


def test_init_bit_exact(sim3, o3):
    for wl, shards in ((params3d.workload_small(), 2), (params3d.workload_c3(), 3)):
        say(f"init {wl.name} shards={shards}")
        sim, orc = _pair(sim3, o3, wl, shards)
        assert sim.seed_radius == orc.seed_radius
        assert_state_equal(sim, orc, f"init {wl.name}")


@pytest.mark.parametrize("shards,slabs", [(2, 0), (1, 1), (3, 7), (5, 10)])
def test_small_steps_bit_exact(sim3, o3, shards, slabs):
    wl = params3d.workload_small()
    say(f"{wl.name} 30 steps shards={shards} slabs={slabs}")
    sim, orc = _pair(sim3, o3, wl, shards, slabs)
    for k in range(3):
        sim.step(10)
        orc.step(10)
        sim.sync()
        assert_state_equal(sim, orc, f"step {sim.steps_done}")
    assert not sim.fatal()


def test_small_100_steps_tolerances(sim3, o3):
    wl = params3d.workload_small()
    say(f"{wl.name} 100 steps vs oracle (kernel order, reference order, FP64)")
    sim, orc = _pair(sim3, o3, wl, 2)
    sim.step(100)
    orc.step(100)
    sim.sync()
    assert_state_equal(sim, orc, "step 100")
    ref = o3.OracleSim3D(wl.params, wl.cx, wl.cy, wl.cz, shards=2, order=o3.ORDER_REFERENCE,
                         threads=THREADS)
    f64 = o3.OracleSim3D(wl.params, wl.cx, wl.cy, wl.cz, shards=2, f64=True, threads=THREADS)
    ref.step(100)
    f64.step(100)
    g = sim.phi_now().astype(np.float64)
    e_ref = float(np.abs(g - ref.phi_now()).max())
    e_64 = float(np.abs(g - f64.phi_now()).max())
    say(f"max|phi - reference order| = {e_ref:.3e}, max|phi - FP64| = {e_64:.3e}")
    assert e_ref <= 1e-5 and e_64 <= 1e-5


def test_c3_steps_bit_exact(sim3, o3):
    wl = params3d.workload_c3()
    say(f"{wl.name} 12 steps shards=2")
    sim, orc = _pair(sim3, o3, wl, 2)
    for k in range(2):
        sim.step(6)
        orc.step(6)
        sim.sync()
        assert_state_equal(sim, orc, f"C3 step {sim.steps_done}")
    c = sim.cells()
    assert c["shift_ctr"].sum() > 0, "the recentring path was not exercised"
    assert not sim.fatal()


def test_c3_long_run_S_exact(sim3, o3):
    """The 256^3 / 128-cell workload (brick 56) over its full 2000 steps from
    graphs: no flag at all (the brick holds the relaxed support), the
    aggregate S equals the oracle's scatter of the stored fields at the
    stored brick origins bit for bit (no drift over many recentring shifts),
    and the field stays finite and bounded."""
    wl = params3d.workload_c3()
    sim = sim3.Sim3D(wl)
    sim.capture_graph()
    say(f"{wl.name}: 2000 steps from graphs")
    sim.run(2000)
    sim.sync()
    assert sim.steps_done == 2000
    assert not sim.flag_counts().any(), sim.flag_counts()
    phi = np.ascontiguousarray(sim.phi_now())
    assert np.isfinite(phi).all()
    assert phi.max() <= 1.5 and phi.min() >= -0.5 and phi.max() > 0.9
    cells = sim.cells()          # bound: the oracle reads through their addresses
    S = np.zeros(sim.words, np.uint32)
    flags = np.zeros(16, np.uint32)
    o3.load().pfo3d_scatter(phi.ctypes.data, cells.ctypes.data, S.ctypes.data,
                            flags.ctypes.data, sim.N, sim.B, sim.L, sim.P)
    assert not flags.any()
    assert np.array_equal(S, sim.S_now())
    assert cells["shift_ctr"].sum() > 0


def test_graph_replay_matches_eager(sim3):
    wl = params3d.workload_small()
    say(f"{wl.name} graph replay vs eager")
    a = sim3.Sim3D(wl, shards=2)
    b = sim3.Sim3D(wl, shards=2)
    b.capture_graph()
    for k in range(4):
        a.step(2)
        b.run(2)
        b.sync()
        assert np.array_equal(a.phi_now().view(np.uint32), b.phi_now().view(np.uint32)), k
        assert np.array_equal(a.S_now(), b.S_now()), k
        assert a.cells().tobytes() == b.cells().tobytes(), k
        assert a.counter_values() == b.counter_values(), k


def test_restart_equals_continuous(sim3):
    wl = params3d.workload_small()
    say(f"{wl.name} restart at step 7")
    a = sim3.Sim3D(wl, shards=2)
    a.step(7)
    a.sync()
    snap = (a.phi_now().copy(), a.S_now().copy(), a.cells(), a.counter_values())
    a.step(9)
    b = sim3.Sim3D(wl, shards=2)
    b.load_state(*snap)
    b.step(9)
    b.sync()
    assert np.array_equal(a.phi_now().view(np.uint32), b.phi_now().view(np.uint32))
    assert np.array_equal(a.S_now(), b.S_now())
    assert a.cells().tobytes() == b.cells().tobytes()


def test_forced_recentring_and_clip(sim3, o3):
    """A cell moved 3 voxels inside its brick: the measurement plans a shift,
    the update recentres it and runs the clip check on the vacated faces."""
    wl = params3d.workload_small()
    say(f"{wl.name} forced recentring")
    sim, orc = _pair(sim3, o3, wl, 2)
    B = sim.B
    phi = sim.phi_now().copy().reshape(sim.N, B, B, B)
    phi[0] = np.roll(phi[0], (3, -2, 1), axis=(0, 1, 2))
    phi[3] = np.roll(phi[3], (0, 0, -3), axis=(0, 1, 2))
    cells = sim.cells()
    S = sim.S_now()   # the aggregate of the unmoved fields: S_other stays exact-positive
    # rebuild S consistently from the moved fields on the oracle side
    orc.load_state(phi.reshape(sim.N, -1), np.zeros_like(S), cells, (0, 0))
    orc.lib.pfo3d_scatter(orc.phi[0].ctypes.data, orc.cells.ctypes.data, orc.S[0].ctypes.data,
                          orc.flags.ctypes.data, orc.N, orc.B, orc.L, orc.P)
    S = orc.S[0].copy()
    a = orc._args(0)
    import ctypes as C
    orc.lib.pfo3d_measure(C.byref(a), 0, 0, 1, 1, THREADS)
    sim.load_state(phi.reshape(sim.N, -1), S, orc.cells.copy(), (0, 0))
    assert orc.cells["pending_shift_x"][3] != 0 or orc.cells["pending_shift_z"][0] != 0
    for k in range(3):
        sim.step(1)
        orc.step(1)
        sim.sync()
        assert_state_equal(sim, orc, f"recentring step {k}")
    assert sim.cells()["shift_ctr"][[0, 3]].min() >= 1


def test_rejected_step_matches_oracle(sim3, o3):
    wl = params3d.workload_small()
    say(f"{wl.name} rejected step (non-finite field)")
    sim, orc = _pair(sim3, o3, wl, 2)
    sim.step(4)
    orc.step(4)
    sim.sync()
    phi = sim.phi_now().copy()
    phi[2, 12345] = np.nan
    S = sim.S_now()
    cells = sim.cells()
    sim.load_state(phi, S, cells, (4, 4))
    orc.load_state(phi, S, cells, (4, 4))
    sim.step(2)
    orc.step(2)
    sim.sync()
    assert sim.fatal() and orc.fatal()
    assert list(sim.counter_values()) == [4, 6]
    g = sim.phi_now()
    assert np.array_equal(np.isnan(g), np.isnan(phi))
    assert np.array_equal(np.nan_to_num(g).view(np.uint32), np.nan_to_num(phi).view(np.uint32))
    assert np.array_equal(sim.S_now(), S)
    assert sim.cells().tobytes() == cells.tobytes()
    assert np.array_equal(sim.flag_counts() != 0, orc.flags != 0)
