"""Pins of the 2D oracle's composed step (not only its primitives), each
chosen so that a plausible mistake shared by the oracle and the kernels
fails it:

- the reference's repulsion-sign gate (src/main.cu:179-214) run through the
  oracle step: a swapped gx/gy, a sign error in the interaction integral or
  in its combination into the velocity fails it;
- integer translation of a whole interacting state commutes with 10 steps,
  bit for bit: the S read and scatter coordinates, the recentring store and
  the window origin update must agree with one another and with the
  lattice's (row = y, column = x) layout;
- transposing a state swaps (Ix, Iy) and (vx, vy) and transposes the
  fields: an x/y mix-up in the integral, the stencil or the moments fails it;
- an independent dense-lattice FP64 evaluator (oracle/dense2d.py: whole
  periodic fields, no windows, shifts or classes, exact FP64 aggregate)
  agrees with the windowed FP64 oracle within 1e-6 for 1 and 10 steps, on
  C1 and on an interacting cluster, across a forced recentring shift, a
  class promotion and forced tumbles drawn by the reference's own Philox;
- the tumble decision and angle over 10^4 (step, gid) pairs equal the
  counter layout of src/kernels.cu:102-112 evaluated with the reference's
  compiled common/philox.cuh.

CPU only; every case finishes in seconds.
"""
import ctypes as C
import random
from pathlib import Path

import numpy as np
import pytest

from oracle import dense2d, pfo
from pfh import params

REF_SO = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libref_common.so"
_ref = None


def ref_philox():
    """The reference's own Philox (common/philox.cuh built under oracle/_ref),
    or None where the reference tree was absent at build time."""
    global _ref
    if _ref is None and REF_SO.exists():
        lib = C.CDLL(str(REF_SO))
        lib.ref_philox.argtypes = [C.c_void_p] * 3
        out = np.zeros(4, np.uint32)

        def philox(ctr, key):
            c = np.array([int(v) & 0xFFFFFFFF for v in ctr], np.uint32)
            k = np.array([int(v) & 0xFFFFFFFF for v in key], np.uint32)
            lib.ref_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
            return tuple(int(v) for v in out)
        _ref = philox
    return _ref


def _p(a):
    return a.ctypes.data


def cluster_workload(tau=0.04, v_A=1e-2, L=320):
    """Four R=20 cells in a 2x2 cluster at spacing ~34 (overlapping interfaces)."""
    cx = np.array([143.0, 177.25, 144.5, 178.0], np.float32)
    cy = np.array([143.5, 144.0, 177.0, 178.75], np.float32)
    p = params.SimParams(Nx=L, Ny=L, num_cells=4, rho=0.05, target_radius=20.0,
                         tau=tau, v_A=v_A, seed=11, polarity_seed=12)
    return params.Workload2D("cluster", p, cx, cy)


def pair_workload(sep, vertical=False, L=192):
    """The sign gate's two R=20 disks at separation `sep` (src/main.cu:184-199)."""
    a = np.array([0.5 * L - 0.5 * sep, 0.5 * L + 0.5 * sep], np.float32)
    b = np.array([0.5 * L, 0.5 * L], np.float32)
    cx, cy = (b, a) if vertical else (a, b)
    p = params.SimParams(Nx=L, Ny=L, num_cells=2, rho=0.05, target_radius=20.0,
                         v_A=0.0, seed=5, polarity_seed=5)
    return params.Workload2D(f"pair{sep}", p, cx, cy)


def snapshot(o):
    return (o.phi_now().copy(), o.S_now().copy(), o.cells.copy(), o.cls.copy(),
            o.steps_done)


def rebuild(o, phi, cells, cls, step):
    """Moments and aggregate of an edited field (k_init_moments /
    k_scatter_all semantics), installed into oracle `o` at `step`."""
    lib = o.lib
    phi = np.ascontiguousarray(phi, np.float64 if o.f64 else np.float32)
    cells = cells.copy()
    cls = np.ascontiguousarray(cls, np.uint8)
    S = np.zeros(o.words, np.uint32)
    flags = np.zeros(8, np.uint32)
    if o.f64:
        lib.pfo2d_init_moments_f64(_p(cells), _p(cls), _p(phi), o.N)
        lib.pfo2d_scatter_all_f64(_p(phi), _p(cells), _p(cls), _p(S), o.N, o.L, o.P, _p(flags))
    else:
        lib.pfo2d_init_moments(_p(cells), _p(cls), _p(phi), o.N, o.order)
        lib.pfo2d_scatter_all(_p(phi), _p(cells), _p(cls), _p(S), o.N, o.L, o.P, _p(flags))
    o.load_state(phi, S, cells, cls, step)
    return o


# ---------------------------------------------------------------- (i) sign gate

@pytest.mark.parametrize("sep", [26, 30, 34, 38, 40])
def test_repulsion_sign_gate_through_the_step(oracle, sep):
    """src/main.cu:179-214: for n left of m the interaction integral and the
    velocity of n point away from m. Run through one oracle step (v_A = 0, so
    v = motility * I), for the horizontal pair and its vertical mirror."""
    got = {}
    for vertical in (False, True):
        wl = pair_workload(sep, vertical)
        o = oracle.OracleSim2D(wl, threads=2)
        world = oracle.world_field(o.phi_now(), o.cells, o.L).astype(np.float64)
        o.step(1)
        c = o.cells
        main, other = ("Iy", "Ix") if vertical else ("Ix", "Iy")
        vmain = "vy" if vertical else "vx"
        assert c[main][0] < 0.0 < c[main][1], (sep, vertical, c[main])
        assert c[vmain][0] < 0.0 < c[vmain][1]
        # the mirror symmetry of the pair: equal and opposite, no cross term
        assert abs(c[main][0] + c[main][1]) <= 1e-6 * abs(c[main][0])
        assert np.all(np.abs(c[other]) <= 1e-6 * abs(c[main][0]))
        mot = np.float32(wl.params.motility())
        np.testing.assert_allclose(c[vmain], (float(mot) * c[main]).astype(np.float32),
                                   rtol=1e-6)
        # the gate's own integral on the same fields: sum phi_n * d(phi_n) * phi_m^2
        ax = 0 if vertical else 1
        for n, m in ((0, 1), (1, 0)):
            f = world[n]
            g = 0.5 * (np.roll(f, -1, axis=ax) - np.roll(f, 1, axis=ax))
            gate = float(np.sum(f * g * world[m] ** 2))
            assert c[main][n] == pytest.approx(gate, rel=2e-5), (n, c[main][n], gate)
        got[vertical] = c
    # the vertical pair is the transpose of the horizontal one
    np.testing.assert_allclose(got[True]["Iy"], got[False]["Ix"], rtol=1e-12)
    np.testing.assert_allclose(got[True]["Ix"], got[False]["Iy"], atol=1e-6 * abs(got[False]["Ix"][0]))


def test_sign_gate_fails_a_planted_sign_error(oracle):
    """The gate's criterion rejects a plausible mistake: using +gradient of the
    other cell (or a flipped difference) reverses the sign."""
    wl = pair_workload(30)
    o = oracle.OracleSim2D(wl, threads=2)
    world = oracle.world_field(o.phi_now(), o.cells, o.L).astype(np.float64)
    f, m = world[0], world[1]
    g_wrong = 0.5 * (np.roll(f, 1, axis=1) - np.roll(f, -1, axis=1))   # W - E
    assert float(np.sum(f * g_wrong * m * m)) > 0.0                    # gate would fail


# ------------------------------------------------------------- (ii) translation

def _translate(snap, L, P, dx, dy):
    phi, S, cells, cls, step = snap
    cells = cells.copy()
    cells["gx0"] = (cells["gx0"] + dx) % L
    cells["gy0"] = (cells["gy0"] + dy) % L
    S2 = S.reshape(L, P).copy()
    S2[:, :L] = np.roll(S2[:, :L], shift=(dy, dx), axis=(0, 1))
    return phi, S2.reshape(-1), cells, cls, step


@pytest.mark.parametrize("dx,dy", [(37, -53), (-101, 8)])
def test_integer_translation_commutes_with_ten_steps(oracle, dx, dy):
    wl = cluster_workload()
    base = oracle.OracleSim2D(wl, threads=4)
    base.step(3)
    snap = snapshot(base)
    L, P = base.L, base.P
    a = oracle.OracleSim2D(wl, threads=4)
    b = oracle.OracleSim2D(wl, threads=4)
    a.load_state(*snap)
    b.load_state(*_translate(snap, L, P, dx, dy))
    a.step(10)
    b.step(10)
    assert int(a.cells["shift_ctr"].sum()) > 0          # recentring happened
    assert int(a.cells["tumble_ctr"].sum()) > 0         # tumbles happened
    assert np.abs(a.cells["Ix"]).max() > 1e-3           # the cells interact
    assert np.array_equal(a.phi_now().view(np.uint32), b.phi_now().view(np.uint32))
    _, S_t, cells_t, _, _ = _translate(snapshot(a), L, P, dx, dy)
    assert np.array_equal(b.S_now(), S_t)
    assert b.cells.tobytes() == cells_t.tobytes()
    assert np.array_equal(a.flags, b.flags)


# --------------------------------------------------------------- (iii) transpose

def _transpose(snap, L, P):
    phi, S, cells, cls, step = snap
    assert np.all(cls == 0) and np.all(cells["cls"] == 0)   # square class-0 windows
    phiT = phi.reshape(-1, 288, 288).transpose(0, 2, 1).reshape(phi.shape).copy()
    c = cells.copy()
    for a, b in (("gx0", "gy0"), ("Cx", "Cy"), ("bb_lo_x", "bb_lo_y"),
                 ("bb_hi_x", "bb_hi_y"), ("Ix", "Iy"), ("vx", "vy")):
        c[a], c[b] = cells[b], cells[a]
    S2 = S.reshape(L, P).copy()
    S2[:, :L] = S2[:, :L].T
    return phiT, S2.reshape(-1), c, cls, step


@pytest.mark.parametrize("f64", [True, False])
def test_transpose_swaps_force_and_velocity(oracle, f64):
    wl = cluster_workload(v_A=0.0)
    base = oracle.OracleSim2D(wl, f64=f64, threads=4)
    snap = snapshot(base)
    L, P = base.L, base.P
    a = oracle.OracleSim2D(wl, f64=f64, threads=4)
    b = oracle.OracleSim2D(wl, f64=f64, threads=4)
    a.load_state(*snap)
    b.load_state(*_transpose(snap, L, P))
    tol_f = 1e-12 if f64 else 2e-6
    tol_i = 1e-10 if f64 else 1e-5
    for k in (1, 4):
        a.step(k)
        b.step(k)
        ca, cb = a.cells, b.cells
        scale = np.abs(ca["Ix"]).max()
        assert scale > 1e-3
        np.testing.assert_allclose(cb["Ix"], ca["Iy"], rtol=0, atol=tol_i * scale)
        np.testing.assert_allclose(cb["Iy"], ca["Ix"], rtol=0, atol=tol_i * scale)
        vs = np.abs(ca["vx"]).max()
        np.testing.assert_allclose(cb["vx"], ca["vy"], rtol=0, atol=tol_i * vs)
        np.testing.assert_allclose(cb["vy"], ca["vx"], rtol=0, atol=tol_i * vs)
        wa = oracle.world_field(a.phi_now(), ca, L)
        wb = oracle.world_field(b.phi_now(), cb, L)
        assert np.abs(wb.transpose(0, 2, 1) - wa).max() <= tol_f
        # binary32: the fused advection term vy gy + (vx gx) rounds the two
        # axes differently, so the transposed run differs in last bits
        np.testing.assert_allclose(cb["Cx"], ca["Cy"], rtol=1e-9 if f64 else 1e-8)
        assert np.array_equal(cb["gx0"], ca["gy0"]) and np.array_equal(cb["gy0"], ca["gx0"])


# ------------------------------------------------- (iv) dense FP64 evaluator

def _dense_from(o, philox):
    world = pfo.world_field(o.phi_now(), o.cells, o.L).astype(np.float64)
    co = dense2d.coeffs_from(o.p, o.cells)
    return dense2d.DenseSim2D(world, o.cells["theta"], o.cells["global_id"], co,
                              o.steps_done, philox=philox)


def _compare_dense(o, d, tol_phi, where):
    c = o.cells
    L = o.L
    world = pfo.world_field(o.phi_now(), c, L).astype(np.float64)
    err = float(np.abs(world - d.phi).max())
    assert err <= tol_phi, f"{where}: max|windowed - dense| = {err:.3e}"
    assert np.array_equal(c["theta"].view(np.uint32), d.theta.view(np.uint32)), where
    np.testing.assert_allclose(c["V"], d.volumes(), rtol=1e-9, err_msg=where)
    cen = d.centroids_in_frame(c["gx0"], c["gy0"])
    np.testing.assert_allclose(c["Cx"] / c["V"], cen[:, 0], rtol=0, atol=1e-6, err_msg=where)
    np.testing.assert_allclose(c["Cy"] / c["V"], cen[:, 1], rtol=0, atol=1e-6, err_msg=where)
    scale = max(1.0, float(np.abs(d.I).max()))
    np.testing.assert_allclose(c["Ix"], d.I[:, 0], rtol=0, atol=1e-6 * scale, err_msg=where)
    np.testing.assert_allclose(c["Iy"], d.I[:, 1], rtol=0, atol=1e-6 * scale, err_msg=where)
    np.testing.assert_allclose(c["vx"], d.v[:, 0].astype(np.float32), rtol=0, atol=1e-8,
                               err_msg=where)
    np.testing.assert_allclose(c["vy"], d.v[:, 1].astype(np.float32), rtol=0, atol=1e-8,
                               err_msg=where)
    return err


def _forced_shift_and_promotion(o):
    """Cell 0's window origin moved by (+3, -2) with its field kept in place
    (so the next step recentres by (-3, +2)); cell 1's support box widened
    to 140 columns (so the next step promotes it to the 176 x 144 class)."""
    phi, _, cells, cls, step = snapshot(o)
    t = phi[0].reshape(288, 288)
    win = t[64:208, 64:208].copy()
    moved = np.zeros_like(win)
    moved[2:, :-3] = win[:-2, 3:]          # window pixel (a, b) <- old (a+3, b-2)
    t[64:208, 64:208] = moved
    cells["gx0"][0] = (cells["gx0"][0] + 3) % o.L
    cells["gy0"][0] = (cells["gy0"][0] - 2) % o.L
    rebuild(o, phi, cells, cls, step)
    o.cells["bb_lo_x"][1], o.cells["bb_hi_x"][1] = 0, 139


@pytest.mark.parametrize("name", ["c1", "cluster"])
def test_dense_fp64_evaluator_agrees_with_windowed_oracle(oracle, name):
    philox = ref_philox() or dense2d.philox_py
    if name == "c1":
        wl = params.workload_c1()
        wl.params.tau = 0.05                    # tumbles on most steps of 10
    else:
        wl = cluster_workload()
    o = oracle.OracleSim2D(wl, f64=True, threads=8)
    _forced_shift_and_promotion(o)
    d = _dense_from(o, philox)
    g0 = (int(o.cells["gx0"][0]), int(o.cells["gy0"][0]))
    o.step(1)
    d.step(1)
    assert o.cells["shift_ctr"][0] == 1
    assert (int(o.cells["gx0"][0]) - g0[0]) % o.L == o.L - 3
    assert (int(o.cells["gy0"][0]) - g0[1]) % o.L == 2
    assert o.cells["cls"][1] == 1 and o.cls[1] == 1
    e1 = _compare_dense(o, d, 1e-6, f"{name} step 1")
    o.step(9)
    d.step(9)
    e10 = _compare_dense(o, d, 1e-6, f"{name} step 10")
    assert int(o.cells["tumble_ctr"].sum()) == int(d.tumbles.sum()) > 0
    print(f"{name}: dense vs windowed FP64 max-abs after 1 step {e1:.2e}, 10 steps {e10:.2e}")
    # the binary32 oracle stays within 1e-5 of the dense FP64 evaluator
    o32 = oracle.OracleSim2D(wl, threads=8)
    _forced_shift_and_promotion(o32)
    d32 = _dense_from(o32, philox)
    o32.step(10)
    d32.step(10)
    w32 = pfo.world_field(o32.phi_now(), o32.cells, o32.L).astype(np.float64)
    assert float(np.abs(w32 - d32.phi).max()) <= 1e-5


def test_dense_evaluator_recentring_sign(oracle):
    """A planted store with the opposite shift sign (window pixel a reads
    source a - s) leaves the field displaced by 2 s: the dense comparison
    rejects it."""
    wl = cluster_workload()
    o = oracle.OracleSim2D(wl, f64=True, threads=4)
    _forced_shift_and_promotion(o)
    d = _dense_from(o, dense2d.philox_py)
    o.step(1)
    d.step(1)
    phi = o.phi_now().copy()
    t = phi[0].reshape(288, 288)
    win = t[64:208, 64:208].copy()
    wrong = np.zeros_like(win)
    wrong[:-4, 6:] = win[4:, :-6]          # displaced by twice the (-3, +2) shift
    t[64:208, 64:208] = wrong
    world = pfo.world_field(phi, o.cells, o.L)
    assert float(np.abs(world - d.phi).max()) > 1e-2


# ------------------------------------------------------------- (v) tumble draws

@pytest.mark.skipif(not REF_SO.exists(), reason="oracle/_ref not built (no reference tree)")
def test_tumble_draws_match_reference_philox_counter_layout(oracle):
    philox = ref_philox()
    lib = oracle.load()
    rng = random.Random(2024)
    th = np.zeros(1, np.float32)
    n_tumbles = 0
    for i in range(10_000):
        step = rng.getrandbits(40) if i % 2 else rng.randrange(100_000)
        gid = rng.randrange(1 << 20)
        seed = rng.getrandbits(64) if i % 3 else 7
        p = (0.0, 1e-6, 0.37, 1.0)[i % 4]
        th0 = np.float32(rng.random() * 6.0)
        th[0] = th0
        got = lib.pfo2d_tumble(step, gid, seed, p, th.ctypes.data)
        r = philox((step & 0xFFFFFFFF, step >> 32, gid, 0xA5A5A5A5),
                   (seed & 0xFFFFFFFF, seed >> 32))
        want = dense2d.u53(r[0], r[1]) < p
        assert bool(got) == want, (step, gid, seed, p)
        want_th = np.float32(2.0 * dense2d.K_PI * dense2d.u53(r[2], r[3])) if want else th0
        assert th[0].view(np.uint32) == np.float32(want_th).view(np.uint32)
        n_tumbles += want
    assert 2_000 < n_tumbles < 8_000
