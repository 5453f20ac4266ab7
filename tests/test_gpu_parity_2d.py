"""GPU parity of the 2D step (gfx950 kernels through the C ABI) against the
CPU oracle on the same seeded inputs.

Bar: bit-exact phi, S (Q5.27), CellState and flags against the oracle in the
kernel's reduction order; <= 1e-5 max-abs phi after 100 steps against the
oracle in the reference kernel's reduction order; the FP64 oracle's error is
reported beside it. Sizes keep every oracle run to seconds.
"""
import math
import sys

import numpy as np
import pytest

from pfh import abi, params

pytestmark = pytest.mark.gpu

COMPARE_FIELDS = ["gx0", "gy0", "cls", "cls_written", "global_id", "gamma", "v_A",
                  "R_tgt", "theta", "V", "Cx", "Cy", "perim", "bb_lo_x", "bb_hi_x",
                  "bb_lo_y", "bb_hi_y", "vx", "vy", "Ix", "Iy", "promote_ctr",
                  "shift_ctr", "tumble_ctr", "phi_max", "reserved", "M_pf"]


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


def assert_cells_equal(a, b, where=""):
    for f in COMPARE_FIELDS:
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            ok = np.array_equal(x.view(np.uint32 if x.itemsize == 4 else np.uint64),
                                y.view(np.uint32 if y.itemsize == 4 else np.uint64))
        else:
            ok = np.array_equal(x, y)
        if not ok:
            idx = np.argwhere(np.atleast_1d(x != y).reshape(len(x), -1).any(axis=1)).ravel()[:5]
            raise AssertionError(f"{where}: CellState.{f} differs at cells {idx}: "
                                 f"gpu={x[idx]} oracle={y[idx]}")


def assert_state_equal(sim, orc, where=""):
    gphi, ophi = sim.phi_now(), orc.phi_now()
    d = np.argwhere(gphi.view(np.uint32) != ophi.view(np.uint32))
    assert d.size == 0, (f"{where}: phi differs at {d.shape[0]} pixels, first {d[:3]}, "
                         f"max|d|={np.abs(gphi - ophi).max():.3e}")
    assert np.array_equal(sim.S_now(), orc.S_now()), f"{where}: S differs"
    assert_cells_equal(sim.cells(), orc.cells, where)
    assert np.array_equal(sim.classes(), orc.cls), f"{where}: cell_cls differs"
    assert np.array_equal(sim.flag_counts(), orc.flags), \
        f"{where}: flags gpu={sim.flag_counts()} oracle={orc.flags}"


@pytest.fixture(scope="module")
def sim_mod():
    from pfh import sim2d
    return sim2d


def test_init_bit_exact_c1(sim_mod, oracle):
    wl = params.workload_c1()
    say("init C1 (L=384 N=16 R=20)")
    sim = sim_mod.Sim2D(wl)
    orc = oracle.OracleSim2D(wl, threads=8)
    assert_state_equal(sim, orc, "init")


def test_c1_100_steps_bit_exact_and_tolerances(sim_mod, oracle):
    wl = params.workload_c1()
    say("C1 100 steps vs oracle (kernel order, reference order, FP64)")
    sim = sim_mod.Sim2D(wl)
    orc = oracle.OracleSim2D(wl, threads=8)
    for k in range(4):
        sim.step(25)
        orc.step(25)
        sim.sync()
        assert_state_equal(sim, orc, f"C1 step {sim.steps_done}")
    ref = oracle.OracleSim2D(wl, order=oracle.ORDER_REFERENCE, threads=8)
    ref.step(100)
    f64 = oracle.OracleSim2D(wl, f64=True, threads=8)
    f64.step(100)
    gphi = sim.phi_now()
    err_ref = float(np.abs(gphi - ref.phi_now()).max())
    err_f64 = float(np.abs(gphi.astype(np.float64) - f64.phi_now()).max())
    say(f"C1 100 steps: max|phi - reference-order| = {err_ref:.3e}, "
        f"max|phi - FP64| = {err_f64:.3e}")
    assert err_ref <= 1e-5
    assert err_f64 <= 1e-5
    assert not sim.fatal()


def test_c1_1000_steps_bit_exact(sim_mod, oracle):
    """The 384^2 / 16-cell workload over its full 1000 steps: bit-exact
    against the oracle in the kernel's order every 250 steps, and within
    1e-5 max-abs of the FP64 oracle at the end."""
    wl = params.workload_c1()
    say("C1 1000 steps vs oracle (kernel order every 250 steps, FP64 at the end)")
    sim = sim_mod.Sim2D(wl)
    sim.capture_graph()
    orc = oracle.OracleSim2D(wl, threads=8)
    for k in range(4):
        sim.run(250)
        orc.step(250)
        sim.sync()
        assert_state_equal(sim, orc, f"C1 step {sim.steps_done}")
    f64 = oracle.OracleSim2D(wl, f64=True, threads=8)
    f64.step(1000)
    err = float(np.abs(sim.phi_now().astype(np.float64) - f64.phi_now()).max())
    say(f"C1 1000 steps: max|phi - FP64| = {err:.3e}")
    assert err <= 1e-5
    assert not sim.fatal()


def _forced_state(oracle, L=701, seed=3):
    """All six shape classes, forced tumbles, heterogeneous stiffness and
    mobility, and scripted class changes and fatal conditions."""
    p = params.SimParams(Nx=L, Ny=L, num_cells=12, rho=0.05, target_radius=20.0,
                         tau=1e-9, seed=seed, polarity_seed=seed + 1,
                         gamma_cancer=0.35, cancer_fraction=0.25, v_A_sigma=0.3)
    g = 4
    cx, cy = params.harness_centres(12, L, g, 6.0, seed)
    wl = params.Workload2D("forced", p, cx, cy)
    lib = oracle.load()
    N = 12
    cls = np.array([0, 1, 2, 3, 4, 5, 0, 3, 5, 0, 1, 0], np.uint8)
    cells = oracle._host_cells(p, N)
    cells["M_pf"][3] = np.float32(0.5 / 0.35)
    cells["M_pf"][7] = np.float32(0.8)
    phi = [np.zeros((N, oracle.TILE_AREA), np.float32) for _ in range(2)]
    lib.pfo2d_init_tiles(phi[0].ctypes.data, phi[1].ctypes.data, cells.ctypes.data,
                         cls.ctypes.data, N, L, cx.ctypes.data, cy.ctypes.data,
                         float(np.float32(p.init_radius())), float(np.float32(p.interface_k())))
    lib.pfo2d_init_moments(cells.ctypes.data, cls.ctypes.data, phi[0].ctypes.data, N, 0)
    # scripted class logic
    cells["bb_lo_x"][6], cells["bb_hi_x"][6] = 0, 139     # promote 0 -> wide
    cells["promote_ctr"][7] = 199                        # demote 3 -> 0 on dwell
    cells["promote_ctr"][8] = 199                        # demote 5 -> smaller
    cells["bb_lo_x"][9], cells["bb_hi_x"][9] = 0, 279     # class_exhausted
    cells["bb_lo_y"][10], cells["bb_hi_y"][10] = 0, 150   # 1 -> bigger (tall+wide)
    S = np.zeros(params.s_pitch_for(L) * L, np.uint32)
    flags = np.zeros(8, np.uint32)
    lib.pfo2d_scatter_all(phi[0].ctypes.data, cells.ctypes.data, cls.ctypes.data,
                          S.ctypes.data, N, L, params.s_pitch_for(L), flags.ctypes.data)
    return wl, phi[0], S, cells, cls


@pytest.mark.parametrize("step0", [0, 99, 12345677])
def test_all_classes_forced_paths_bit_exact(sim_mod, oracle, step0):
    wl, phi, S, cells, cls = _forced_state(oracle)
    say(f"forced all-classes state L=701 N=12, start step {step0}")
    sim = sim_mod.Sim2D(wl)
    orc = oracle.OracleSim2D(wl, threads=8)
    sim.load_state(phi, S, cells, cls, step0)
    orc.load_state(phi, S, cells, cls, step0)
    for k in range(3):
        sim.step(1)
        orc.step(1)
        sim.sync()
        assert_state_equal(sim, orc, f"forced step {step0 + k + 1}")
    c = sim.cells()
    assert c["cls"][6] != 0, "scripted promotion did not happen"
    assert c["tumble_ctr"].min() >= 1, "tumbles were not forced"
    assert sim.flag_counts()[abi.FLAG_NAMES.index("class_exhausted")] > 0


@pytest.mark.parametrize("grid,step0", [(1, 0), (2, 99), (3, 5)])
def test_few_workgroups_mixed_classes_bit_exact(sim_mod, oracle, grid, step0):
    """The same forced state on 1-3 workgroups, so that each one takes several
    cells of mixed classes in queue order: class-0 windows copied in ahead of
    the next class-0 cell (and the slot of the double-buffered piece), the
    switch to and from the out-of-line classes, and at step 99 the perimeter
    pass of a full-moment step."""
    wl, phi, S, cells, cls = _forced_state(oracle)
    say(f"forced all-classes state L=701 N=12 on {grid} workgroup(s), start step {step0}")
    sim = sim_mod.Sim2D(wl)
    sim.grid = grid
    orc = oracle.OracleSim2D(wl, threads=8)
    sim.load_state(phi, S, cells, cls, step0)
    orc.load_state(phi, S, cells, cls, step0)
    for k in range(3):
        sim.step(1)
        orc.step(1)
        sim.sync()
        assert_state_equal(sim, orc, f"grid {grid}, step {step0 + k + 1}")


def test_c1_two_workgroups_bit_exact(sim_mod, oracle):
    """C1's 16 class-0 cells on two workgroups: eight cells in a row per
    workgroup, each window copied in during the previous cell."""
    wl = params.workload_c1()
    say("C1 on 2 workgroups, 20 steps vs oracle")
    sim = sim_mod.Sim2D(wl)
    sim.grid = 2
    orc = oracle.OracleSim2D(wl, threads=8)
    for k in range(2):
        sim.step(10)
        orc.step(10)
        sim.sync()
        assert_state_equal(sim, orc, f"C1 two workgroups, step {sim.steps_done}")


def test_interacting_cluster_bit_exact(sim_mod, oracle):
    """Four overlapping R=20 cells (the oracle's translation, transpose and
    dense-evaluator pins run on this state) with tumbles on most steps:
    10 steps bit-exact, with nonzero interaction integrals."""
    cx = np.array([143.0, 177.25, 144.5, 178.0], np.float32)
    cy = np.array([143.5, 144.0, 177.0, 178.75], np.float32)
    p = params.SimParams(Nx=320, Ny=320, num_cells=4, rho=0.05, target_radius=20.0,
                         tau=0.04, seed=11, polarity_seed=12)
    wl = params.Workload2D("cluster", p, cx, cy)
    say("interacting 4-cell cluster L=320, 10 steps vs oracle")
    sim = sim_mod.Sim2D(wl)
    orc = oracle.OracleSim2D(wl, threads=4)
    for k in range(2):
        sim.step(5)
        orc.step(5)
        sim.sync()
        assert_state_equal(sim, orc, f"cluster step {sim.steps_done}")
    c = sim.cells()
    assert np.abs(c["Ix"]).max() > 1e-3 and c["tumble_ctr"].sum() > 0


def test_fatal_flags_match_oracle(sim_mod, oracle):
    wl, phi, S, cells, cls = _forced_state(oracle)
    cells = cells.copy()
    cls = cls.copy()
    cells["V"][2] = 0.0                     # V_nonpositive
    cls[11] = 7                             # class_unsupported (cell skipped)
    S = S.copy()
    wy = 144
    # under-report S on part of cell 0's window: S_other_negative
    gx0, gy0 = int(cells["gx0"][0]), int(cells["gy0"][0])
    P = params.s_pitch_for(701)
    for j in range(60, 80):
        S[((gy0 + j) % 701) * P + (gx0 + 70) % 701] = 0
    say("fatal-flag state")
    sim = sim_mod.Sim2D(wl)
    orc = oracle.OracleSim2D(wl, threads=8)
    sim.load_state(phi, S, cells, cls, 5)
    orc.load_state(phi, S, cells, cls, 5)
    sim.step(1)
    orc.step(1)
    sim.sync()
    f = sim.flag_counts()
    assert np.array_equal(f, orc.flags), (f, orc.flags)
    for name in ("V_nonpositive", "class_unsupported", "S_other_negative"):
        assert f[abi.FLAG_NAMES.index(name)] > 0, name


def test_c2_parity_and_S_exactness(sim_mod, oracle):
    wl = params.workload_c2()
    say("C2 init + 3 steps vs oracle (L=2048 N=1024 R=20)")
    sim = sim_mod.Sim2D(wl)
    orc = oracle.OracleSim2D(wl, threads=16)
    assert_state_equal(sim, orc, "C2 init")
    sim.step(3)
    orc.step(3)
    sim.sync()
    assert_state_equal(sim, orc, "C2 step 3")
    say("C2 60 more steps: S must equal sum of q(phi) of the stored fields")
    sim.step(60)
    sim.sync()
    lib = oracle.load()
    phi = np.ascontiguousarray(sim.phi_now())
    cells = sim.cells()
    cls = sim.classes()
    S = np.zeros(sim.words, np.uint32)
    flags = np.zeros(8, np.uint32)
    lib.pfo2d_scatter_all(phi.ctypes.data, cells.ctypes.data, cls.ctypes.data, S.ctypes.data,
                          sim.N, sim.L, sim.P, flags.ctypes.data)
    assert np.array_equal(S, sim.S_now())
    assert not sim.fatal(), sim.flag_counts()


def test_c2_long_run_S_exact(sim_mod, oracle):
    """The 2048^2 / 1024-cell monolayer over its full 10,000 steps (10,002:
    whole six-step graphs): the aggregate S still equals the oracle's scatter
    of the stored fields bit for bit (no drift from the incremental scatter
    and clear-ahead rotation), no fatal flag is raised and the field stays
    finite and bounded."""
    wl = params.workload_c2()
    sim = sim_mod.Sim2D(wl)
    sim.capture_graph()
    say("C2: 10002 steps from six-step graphs")
    sim.run(10002)
    sim.sync()
    assert sim.steps_done == 10002
    assert not sim.fatal(), sim.flag_counts()
    phi = np.ascontiguousarray(sim.phi_now())
    assert np.isfinite(phi).all()
    assert phi.max() <= 1.5 and phi.min() >= -0.5 and phi.max() > 0.9
    lib = oracle.load()
    cells = sim.cells()          # bound: the oracle reads through their addresses
    cls = sim.classes()
    S = np.zeros(sim.words, np.uint32)
    flags = np.zeros(8, np.uint32)
    lib.pfo2d_scatter_all(phi.ctypes.data, cells.ctypes.data, cls.ctypes.data, S.ctypes.data,
                          sim.N, sim.L, sim.P, flags.ctypes.data)
    assert np.array_equal(S, sim.S_now())


def test_graph_replay_equals_launches(sim_mod):
    wl = params.workload_c1()
    say("graph replay vs single launches, 24 steps")
    a = sim_mod.Sim2D(wl)
    b = sim_mod.Sim2D(wl)
    a.capture_graph()
    for _ in range(3):              # replays interleaved with eager work
        a.run(6)
        a.sync()
        b.step(6)
        b.sync()
        assert np.array_equal(a.phi_now().view(np.uint32), b.phi_now().view(np.uint32))
    a.run(6)
    b.step(6)
    a.sync()
    b.sync()
    assert np.array_equal(a.S_now(), b.S_now())
    assert_cells_equal(a.cells(), b.cells(), "graph")


def test_restart_equivalence(sim_mod):
    """Split runs give the same state (tests/run_gpu_smoke.py:68-109)."""
    wl = params.workload_c1()
    say("restart equivalence 7 + 5 vs 12")
    a = sim_mod.Sim2D(wl)
    a.step(12)
    a.sync()
    b = sim_mod.Sim2D(wl)
    b.step(7)
    b.sync()
    c = sim_mod.Sim2D(wl)
    c.load_state(b.phi_now().copy(), b.S_now(), b.cells(), b.classes(), 7)
    c.step(5)
    c.sync()
    assert np.array_equal(a.phi_now().view(np.uint32), c.phi_now().view(np.uint32))
    assert np.array_equal(a.S_now(), c.S_now())
    assert_cells_equal(a.cells(), c.cells(), "restart")
