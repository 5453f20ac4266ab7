"""Velocity observers in the oracle (k_step_advection / k_step_moments and
the dense observer, include/fused_moments.cuh:23-127,
src/velocity_moments.cu:43-128), pinned by what the mathematics fixes
rather than by a second copy of their formulas:

- observation is neutral: fields, aggregate and every cell-state byte other
  than the advection accumulator are identical with it off, on, or spatial;
- the advection record agrees with the step's own stored outputs: the
  active part with the stored (tumbled) polarity, the interaction part with
  the stored interaction integral, and active + interaction + rounding with
  the stored binary32 velocity;
- the spatial rates are centroid derivatives: each held rate equals the
  central finite difference of the centroid of phi^2 under phi + eps x its
  source (the reference's quotient check, tests/velocity_observer_test.cu:
  130-160), and their sum with the advection moment predicts the centroid
  motion of the actual oracle step (first order in dt);
- the dense observer (an independent formulation: whole windows, one
  accumulator per quantity) agrees with the fused rates of the same step.

CPU only.
"""
import dataclasses

import numpy as np

from oracle import pfo
from pfh import params

ADV = ("adv_q", "adv_count")


def cluster(tau=0.05):
    """Four R=20 cells in a 2x2 cluster with overlapping interfaces and
    tumbles forced often (p = 1 - exp(-dt / tau))."""
    cx = np.array([143.0, 177.25, 144.5, 178.0], np.float32)
    cy = np.array([143.5, 144.0, 177.0, 178.75], np.float32)
    p = params.SimParams(Nx=320, Ny=320, num_cells=4, rho=0.05, target_radius=20.0,
                         tau=tau, v_A=1e-2, seed=11, polarity_seed=12)
    return params.Workload2D("cluster", p, cx, cy)


def _strip(cells):
    c = cells.copy()
    c["adv_q"] = 0.0
    c["adv_count"] = 0
    return c.tobytes()


def _sim(wl, warm=20):
    o = pfo.OracleSim2D(wl)
    o.step(warm)
    return o


def _clone(o, wl):
    c = pfo.OracleSim2D(wl)
    c.steps_done = o.steps_done
    for k in range(2):
        c.phi[k][:] = o.phi[k]
    for k in range(3):
        c.S[k][:] = o.S[k]
    c.cells[:] = o.cells
    c.cls[:] = o.cls
    c.flags[:] = o.flags
    return c


def test_observation_is_neutral():
    wl = cluster()
    base = _sim(wl)
    runs = []
    for mode in (0, 1, 2):
        o = _clone(base, wl)
        o.observation = mode
        o.step(6)
        runs.append(o)
    for o in runs[1:]:
        assert np.array_equal(o.phi_now().view(np.uint32), runs[0].phi_now().view(np.uint32))
        assert np.array_equal(o.S_now(), runs[0].S_now())
        assert _strip(o.cells) == _strip(runs[0].cells)
        assert np.array_equal(o.flags, runs[0].flags)
    assert (runs[1].cells["adv_count"] == 6).all() and (runs[2].cells["adv_count"] == 6).all()
    assert (runs[2].moments["spatial_count"] == 6).all()
    # the advection record is the same whichever observer writes it
    assert np.array_equal(runs[1].cells["adv_q"], runs[2].cells["adv_q"])


def test_advection_record_matches_the_stored_step():
    wl = cluster(tau=0.02)
    o = _sim(wl)
    o.observation = 1
    before = o.cells.copy()
    o.step(1)
    c = o.cells
    p = wl.params
    dt = p.dt
    q = c["adv_q"]
    assert (c["adv_count"] == 1).all()
    tumbled = c["tumble_ctr"] != before["tumble_ctr"]
    assert tumbled.any() and (~tumbled).any()       # both kinds of cell-step
    th = c["theta"].astype(np.float64)
    vA = c["v_A"].astype(np.float64)
    np.testing.assert_allclose(q[:, 0] / dt, vA * np.cos(th), rtol=0, atol=2e-9)
    np.testing.assert_allclose(q[:, 1] / dt, vA * np.sin(th), rtol=0, atol=2e-9)
    mot = float(np.float32(p.motility()))
    np.testing.assert_allclose(q[:, 2] / dt, mot * c["Ix"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(q[:, 3] / dt, mot * c["Iy"], rtol=1e-14, atol=0)
    assert np.abs(c["Ix"]).max() > 0.0                # the cluster interacts
    vx = (q[:, 0] + q[:, 2] + q[:, 4]) / dt
    vy = (q[:, 1] + q[:, 3] + q[:, 5]) / dt
    np.testing.assert_allclose(vx, c["vx"].astype(np.float64), rtol=0, atol=1e-15)
    np.testing.assert_allclose(vy, c["vy"].astype(np.float64), rtol=0, atol=1e-15)


# shape classes (wx, wy, tx0, ty0), include/params.cuh:138-147
CLASSES = ((144, 144, 64, 64), (176, 144, 32, 64), (144, 176, 64, 32), (160, 160, 32, 32),
           (224, 224, 32, 32), (286, 286, 1, 1))


def _window(o, n):
    wx, wy, tx0, ty0 = CLASSES[int(o.cls[n])]
    t = o.phi_now()[n].reshape(288, 288)
    return t[ty0:ty0 + wy, tx0:tx0 + wx].astype(np.float64)


def _padded(f):
    g = np.zeros((f.shape[0] + 2, f.shape[1] + 2))
    g[1:-1, 1:-1] = f
    return g


def _sources(o, n, p):
    """Per-pixel interfacial and overlap sources of cell n (fp64 from the
    binary32 field, S_other from the aggregate), as the rates weight them."""
    c = o.cells[n]
    f = _window(o, n)
    g = _padded(f)
    wy, wx = f.shape
    C_ = g[1:-1, 1:-1]
    N_, S_ = g[2:, 1:-1], g[:-2, 1:-1]
    E_, W_ = g[1:-1, 2:], g[1:-1, :-2]
    NE, NW, SE, SW = g[2:, 2:], g[2:, :-2], g[:-2, 2:], g[:-2, :-2]
    lap = (4.0 * (N_ + S_ + E_ + W_) + NE + NW + SE + SW - 20.0 * C_) / 6.0
    ys, xs = np.mgrid[0:wy, 0:wx]
    gy = (int(c["gy0"]) + ys) % o.L
    gx = (int(c["gx0"]) + xs) % o.L
    S = o.S_now()[gy * o.P + gx].astype(np.int64)
    q = np.rint(np.minimum(f * f, 4.0) * 2.0 ** 27).astype(np.int64)
    other = np.maximum(S - q, 0) * 2.0 ** -27
    dw = float(np.float32(np.float32(p.bulk()) * np.float32(c["gamma"])))
    mob = float(np.float32(c["M_pf"]) / np.float32(0.5))
    rep = float(np.float32(p.interaction()))
    s_int = mob * (float(c["gamma"]) * lap - dw * C_ * (1 - C_) * (1 - 2 * C_))
    s_ov = -mob * rep * C_ * other
    return f, xs, ys, s_int, s_ov


def _centroid(f, xs, ys):
    w = f * f
    return (w * xs).sum() / w.sum(), (w * ys).sum() / w.sum()


def test_spatial_rates_are_centroid_derivatives():
    wl = cluster()
    p = wl.params
    o = _sim(wl, warm=30)
    pre = _clone(o, wl)
    o.observation = 2
    o.step(1)
    m = o.moments
    eps = 1e-6
    worst = 0.0
    for n in range(o.N):
        f, xs, ys, s_int, s_ov = _sources(pre, n, p)
        for k, s in ((0, s_int), (2, s_ov)):
            cp = _centroid(f + eps * s, xs, ys)
            cm = _centroid(f - eps * s, xs, ys)
            for xy in range(2):
                d = (cp[xy] - cm[xy]) / (2 * eps)
                rate = m["held"][n, k + xy]
                scale = max(abs(d), 1e-6)
                worst = max(worst, abs(rate - d) / scale)
    assert worst < 1e-5, worst
    assert np.abs(m["held"][:, 2:4]).max() > 0.0       # the overlap rates are exercised


def _world_centroids(cells, L):
    cx = cells["gx0"] + cells["Cx"] / cells["V"]
    cy = cells["gy0"] + cells["Cy"] / cells["V"]
    return np.stack([cx, cy], 1)


def test_spatial_rates_predict_the_step_centroid_motion():
    """The centroid of phi^2 (world coordinates) before and after one actual
    oracle step: its displacement over dt approaches interfacial + overlap
    + advection rate (the area source has a zero moment) at first order in
    dt, i.e. the mismatch halves when the step is halved. No tumbles: the
    tumble probability depends on dt, and a polarity that differs between
    the two steps would move the centroid differently."""
    wl = cluster(tau=1e12)
    wl = dataclasses.replace(wl, params=dataclasses.replace(wl.params, v_A=5e-2))
    base = _sim(wl, warm=30)
    c0 = _world_centroids(base.cells, base.L)
    errs, preds = [], None
    for frac in (1.0, 0.5):
        wlh = dataclasses.replace(wl, params=dataclasses.replace(wl.params,
                                                                 dt=wl.params.dt * frac))
        o = _clone(base, wlh)
        o.observation = 2
        o.step(1)
        m, after = o.moments, o.cells
        d = _world_centroids(after, o.L) - c0
        d -= o.L * np.round(d / o.L)
        if preds is None:
            v = np.stack([after["vx"], after["vy"]], 1).astype(np.float64)
            preds = m["held"][:, 0:2] + m["held"][:, 2:4] + v * m["held"][:, 4:6]
        errs.append(d / wlh.params.dt - preds)
    assert np.abs(errs[0]).max() <= 0.02 * np.abs(preds).max()
    ratio = np.abs(errs[1]) / np.maximum(np.abs(errs[0]), 1e-30)
    big = np.abs(errs[0]) > 1e-3 * np.abs(preds).max()
    assert big.any()
    assert ((ratio[big] > 0.4) & (ratio[big] < 0.6)).all(), ratio[big]


def test_dense_observer_agrees_with_the_fused_rates():
    wl = cluster()
    o = _sim(wl, warm=30)
    dense = np.zeros(o.N, pfo.MOMENT_DTYPE)
    a = o._args()
    o.lib.pfo2d_velocity_moments(pfo.C.byref(a), dense.ctypes.data, wl.params.dt)
    o.observation = 2
    o.step(1)
    dt = wl.params.dt
    m, c = o.moments, o.cells
    assert (dense["count"] == 1).all()
    np.testing.assert_allclose(dense["q"][:, 0:4], c["adv_q"][:, 0:4], rtol=1e-13, atol=1e-18)
    for k in range(4):
        np.testing.assert_allclose(dense["q"][:, 4 + k] / dt, m["held"][:, k], rtol=1e-9,
                                   atol=1e-13)
    # the area source has an exactly zero centroid moment (rounding only)
    assert np.abs(dense["q"][:, 8:10]).max() < 1e-12 * np.abs(dense["q"][:, 4:8]).max() + 1e-18
    # advection moment - (active + interaction) = (beta - 1)(A + I) + beta rounding
    v = c["adv_q"]
    for xy in range(2):
        beta = m["held"][:, 4 + xy]
        fused = (beta - 1.0) * (v[:, xy] + v[:, 2 + xy]) + beta * v[:, 4 + xy]
        np.testing.assert_allclose(dense["q"][:, 10 + xy], fused, rtol=1e-9, atol=1e-15)
