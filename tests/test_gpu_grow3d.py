"""Brick growth of the 3D population on the GPU: the reference's support
recovery (Sim3D::recover_support_exhaustion, src/pf3d/sim.cu:1653-1733, and
resize_cells, src/pf3d/adaptive_storage.cu:143-266), applied to the whole
population at once.

Pins: every value keeps its world position and S is unchanged (each cube
rendered into the periodic domain at its origin, before and after); the
moved states follow translate_state (adaptive_storage.cu:85-106), restated
here in numpy; six steps after the growth equal the oracle's from the same
state bit for bit; the 256^3 / 128-cell workload started at the reference's
automatic brick 48 grows once (to 64, this build's largest brick) when its
support is exhausted and then runs its 2000 steps flag-free with S exact."""
import dataclasses
import os
import sys

import numpy as np
import pytest

from pfh import abi, params3d

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(f"[gpu-grow3d] {msg}", flush=True)
    sys.stdout.flush()


def world3(phi, cells, B, L):
    """Every cell's cube at its world position in the periodic domain."""
    out = np.zeros((len(cells), L, L, L), np.float32)
    for n in range(len(cells)):
        ax = [(int(cells[n][f"origin_{a}"]) + np.arange(B)) % L for a in "zyx"]
        out[n][np.ix_(*ax)] = phi[n].reshape(B, B, B)
    return out


@pytest.fixture(scope="module")
def sim3():
    from pfh import sim3d
    return sim3d


@pytest.fixture(scope="module")
def o3():
    from oracle import pfo3
    pfo3.load()
    return pfo3


def test_growth_keeps_world_field_and_moves_states(sim3, o3):
    wl = params3d.workload_small()
    sim = sim3.Sim3D(wl, shards=2)
    B0, L = sim.B, sim.L
    say(f"{wl.name}: 5 steps at B={B0}, then growth to 56 and 6 steps against the oracle")
    sim.step(5)
    sim.sync()
    phi0, cells0, S0 = sim.phi_now().copy(), sim.cells(), sim.S_now()
    sim.flags[abi.FLAG3_NAMES.index("support_edge")] = 3    # cleared by the growth
    sim.grow_bricks(56)
    phi1, cells1, S1 = sim.phi_now(), sim.cells(), sim.S_now()
    assert sim.B == 56 and phi1.shape == (sim.N, 56 ** 3)
    w0, w1 = world3(phi0, cells0, B0, L), world3(phi1, cells1, 56, L)
    assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32))
    assert np.count_nonzero(phi1) == np.count_nonzero(phi0)
    assert np.array_equal(S0, S1)
    off = (56 - B0) // 2
    for a in "xyz":
        assert np.array_equal(cells1[f"origin_{a}"], cells0[f"origin_{a}"] - off)
        want = cells0[f"C{a}"] + np.float64(off) * cells0["V"]
        assert np.array_equal(cells1[f"C{a}"].view(np.uint64), want.view(np.uint64))
        assert (cells1[f"pending_shift_{a}"] == 0).all()
        nonempty = cells0["bb_hi_x"] >= cells0["bb_lo_x"]
        for e in ("lo", "hi"):
            f = f"bb_{e}_{a}"
            assert np.array_equal(cells1[f], np.where(nonempty, cells0[f] + off, cells0[f]))
    support = (1 << abi.FLAG3_NAMES.index("support_exhausted")) | \
              (1 << abi.FLAG3_NAMES.index("support_edge"))
    assert np.array_equal(cells1["flags"], cells0["flags"] & ~np.uint32(support))
    assert not sim.flag_counts().any()
    for f in ("V", "Ix", "Iy", "Iz", "surface", "phi_max", "velocity_x", "polarity_x"):
        assert np.array_equal(cells1[f], cells0[f]), f

    p56 = dataclasses.replace(wl.params, brick_edge_override=56)
    orc = o3.OracleSim3D(p56, wl.cx, wl.cy, wl.cz, shards=2, threads=THREADS)
    orc.load_state(phi1, S1, cells1, sim.counter_values())
    sim.step(6)
    orc.step(6)
    sim.sync()
    d = np.argwhere(sim.phi_now().view(np.uint32) != orc.phi_now().view(np.uint32))
    assert d.size == 0, f"phi differs at {d.shape[0]} voxels after the growth"
    assert np.array_equal(sim.S_now(), orc.S_now())
    gc, oc = sim.cells(), orc.cells
    for f in gc.dtype.names:
        x, y = gc[f], oc[f]
        same = (np.array_equal(x.view(np.uint32 if x.itemsize == 4 else np.uint64),
                               y.view(np.uint32 if y.itemsize == 4 else np.uint64))
                if x.dtype.kind == "f" else np.array_equal(x, y))
        assert same, f"CellState3D.{f} differs after the growth"
    assert np.array_equal(sim.flag_counts(), orc.flags)


def test_c3_at_automatic_brick_grows_and_runs_2000_steps(sim3, o3):
    wl = params3d.workload_c3(brick_edge=48)
    sim = sim3.Sim3D(wl)
    sim.capture_graph()
    say(f"{wl.name}: 2000 steps from B=48 with support recovery every 16 steps")
    grown = sim.run_adaptive(2000)
    sim.sync()
    c = sim.counter_values()
    say(f"grown {grown}x to B={sim.B}; accepted {c[0]} of {c[1]} begun steps")
    assert grown == 1 and sim.B == 64
    assert sim.steps_done == 2000 and c[1] == 2000 and 1900 < c[0] < 2000
    assert not sim.flag_counts().any(), sim.flag_counts()
    phi = np.ascontiguousarray(sim.phi_now())
    assert np.isfinite(phi).all() and phi.max() <= 1.5 and phi.max() > 0.9
    cells = sim.cells()
    S = np.zeros(sim.words, np.uint32)
    flags = np.zeros(16, np.uint32)
    o3.load().pfo3d_scatter(phi.ctypes.data, cells.ctypes.data, S.ctypes.data,
                            flags.ctypes.data, sim.N, sim.B, sim.L, sim.P)
    assert not flags.any()
    assert np.array_equal(S, sim.S_now())
