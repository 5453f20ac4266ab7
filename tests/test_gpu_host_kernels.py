"""GPU parity of the 2D and 3D set-up, diagnostic, trajectory and ordering
entry points (initial tiles, moments, aggregate scatter, fills, verify,
trajectory packing, Morton queue order) against the CPU oracle, through the
C ABI, on the same inputs. Integer, index and record outputs are bit-exact;
fp64 volumes are bit-exact too (same reduction order as the oracle's).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from pfh import abi, params

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-host-kernels] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    return torch


def to_dev(torch, a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def from_dev(t, dtype, shape=None):
    a = t.cpu().numpy().view(dtype)
    return a.reshape(shape) if shape is not None else a.copy()


def ok(status, what):
    abi.check(status, what)


def _mixed_state(oracle, L=701, N=12, seed=3):
    cls = np.array([0, 1, 2, 3, 4, 5, 0, 3, 5, 0, 1, 0], np.uint8)[:N]
    cx, cy = params.harness_centres(N, L, 4, 6.0, seed)
    # cells 0 and 11 wrap across the periodic boundary
    cx[0], cy[0] = np.float32(L - 3.25), np.float32(2.5)
    if N > 11:
        cx[11], cy[11] = np.float32(1.75), np.float32(L - 10.0)
    p = params.SimParams(Nx=L, Ny=L, num_cells=N, rho=0.05, target_radius=20.0, seed=seed)
    cells = oracle._host_cells(p, N)
    return p, cls, cells, cx, cy


def test_init_tiles_moments_scatter_all_classes(dev, oracle):
    torch = dev
    lib, olib = abi.load(), oracle.load()
    p, cls, cells, cx, cy = _mixed_state(oracle)
    N, L, P = len(cls), p.Nx, params.s_pitch_for(p.Nx)
    R, k = float(np.float32(p.init_radius())), float(np.float32(p.interface_k()))
    say(f"init tiles/moments/scatter, all six classes, L={L} N={N}")
    # oracle
    oa = np.full((N, abi.TILE_AREA), 7.0, np.float32)
    ob = np.full((N, abi.TILE_AREA), 7.0, np.float32)
    oc = cells.copy()
    olib.pfo2d_init_tiles(oa.ctypes.data, ob.ctypes.data, oc.ctypes.data, cls.ctypes.data,
                          N, L, cx.ctypes.data, cy.ctypes.data, R, k)
    olib.pfo2d_init_moments(oc.ctypes.data, cls.ctypes.data, oa.ctypes.data, N, 0)
    oS = np.zeros(P * L, np.uint32)
    of = np.zeros(8, np.uint32)
    olib.pfo2d_scatter_all(oa.ctypes.data, oc.ctypes.data, cls.ctypes.data, oS.ctypes.data,
                           N, L, P, of.ctypes.data)
    # GPU (tiles start as garbage: init must write every word)
    ga = torch.full((N * abi.TILE_AREA,), 7.0, dtype=torch.float32, device="cuda")
    gb = torch.full((N * abi.TILE_AREA,), 7.0, dtype=torch.float32, device="cuda")
    gc = to_dev(torch, cells)
    gcls = to_dev(torch, cls)
    gcx, gcy = to_dev(torch, cx), to_dev(torch, cy)
    gS = torch.zeros(P * L, dtype=torch.int32, device="cuda")
    gf = torch.zeros(8, dtype=torch.int32, device="cuda")
    ok(lib.pfh2d_init_tiles(ga.data_ptr(), gb.data_ptr(), gc.data_ptr(), gcls.data_ptr(), N, L,
                            gcx.data_ptr(), gcy.data_ptr(), R, k, None), "init_tiles")
    ok(lib.pfh2d_init_moments(gc.data_ptr(), gcls.data_ptr(), ga.data_ptr(), N, None),
       "init_moments")
    ok(lib.pfh2d_scatter_all(ga.data_ptr(), gc.data_ptr(), gcls.data_ptr(), gS.data_ptr(),
                             N, L, P, gf.data_ptr(), None), "scatter_all")
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(ga, np.uint32), oa.view(np.uint32).reshape(-1))
    assert not from_dev(gb, np.uint32).any()
    assert from_dev(gc, abi.CELL_DTYPE).tobytes() == oc.tobytes()
    assert np.array_equal(from_dev(gS, np.uint32), oS)
    assert np.array_equal(from_dev(gf, np.uint32), of)
    assert oS.max() > 0


def test_scatter_flags(dev, oracle):
    """Q_CLAMP (phi^2 > 4) and S_OVERFLOW (an S word near 2^32) are raised
    as in the oracle, and the wrapped sums equal the oracle's."""
    torch = dev
    lib, olib = abi.load(), oracle.load()
    p, cls, cells, cx, cy = _mixed_state(oracle, N=4)
    N, L, P = len(cls), p.Nx, params.s_pitch_for(p.Nx)
    oa = np.zeros((N, abi.TILE_AREA), np.float32)
    ob = np.zeros_like(oa)
    olib.pfo2d_init_tiles(oa.ctypes.data, ob.ctypes.data, cells.ctypes.data, cls.ctypes.data,
                          N, L, cx.ctypes.data, cy.ctypes.data,
                          float(np.float32(p.init_radius())), float(np.float32(p.interface_k())))
    oa[1].reshape(288, 288)[100, 100] = 2.5               # q clamp
    S0 = np.zeros(P * L, np.uint32)
    gx, gy = int(cells["gx0"][2]), int(cells["gy0"][2])
    S0[((gy + 70) % L) * P + (gx + 72) % L] = 0xFFFFFFF0  # overflow under cell 2's core
    oS, of = S0.copy(), np.zeros(8, np.uint32)
    olib.pfo2d_scatter_all(oa.ctypes.data, cells.ctypes.data, cls.ctypes.data, oS.ctypes.data,
                           N, L, P, of.ctypes.data)
    ga, gc, gcls = to_dev(torch, oa), to_dev(torch, cells), to_dev(torch, cls)
    gS, gf = to_dev(torch, S0), torch.zeros(8, dtype=torch.int32, device="cuda")
    ok(lib.pfh2d_scatter_all(ga.data_ptr(), gc.data_ptr(), gcls.data_ptr(), gS.data_ptr(),
                             N, L, P, gf.data_ptr(), None), "scatter_all")
    torch.cuda.synchronize()
    f = from_dev(gf, np.uint32)
    assert np.array_equal(f, of)
    assert f[abi.FLAG_NAMES.index("S_overflow")] == 1 and f[abi.FLAG_NAMES.index("q_clamp")] == 1
    assert np.array_equal(from_dev(gS, np.uint32), oS)


@pytest.mark.parametrize("n,offset", [(0, 0), (1, 1), (3, 2), (5, 3), (17, 1), (1_000_003, 0),
                                      (1_000_003, 3)])
def test_zero_u32_any_alignment(dev, n, offset):
    torch = dev
    lib = abi.load()
    buf = torch.full((n + 8,), -1, dtype=torch.int32, device="cuda")
    ok(lib.pfh2d_zero_u32(buf.data_ptr() + 4 * offset, n, None), "zero_u32")
    torch.cuda.synchronize()
    h = from_dev(buf, np.uint32)
    assert not h[offset:offset + n].any()
    assert np.all(h[:offset] == 0xFFFFFFFF) and np.all(h[offset + n:] == 0xFFFFFFFF)


def test_verify_cells_and_S_after_steps(dev, oracle):
    """Sim::verify's diagnostics (src/sim.cu:817-849) after 5 steps of C1: V
    of every window bit-equal to the oracle's, the outside-window maximum
    (here planted) and the largest S word (planted at the head and tail of
    the buffer, at three alignments) equal."""
    from pfh import sim2d
    torch = dev
    lib, olib = abi.load(), oracle.load()
    wl = params.workload_c1()
    sim = sim2d.Sim2D(wl)
    sim.step(5)
    sim.sync()
    phi = sim.phi_now().copy()
    phi[3].reshape(288, 288)[10, 5] = -0.125       # outside the class-0 window
    phi[7].reshape(288, 288)[250, 280] = 3.0e-3
    cells, cls = sim.cells(), sim.classes()
    N = sim.N
    say("verify cells / S on C1 after 5 steps")
    oV, oo = np.zeros(N), np.zeros(N, np.float32)
    olib.pfo2d_verify_cells(phi.ctypes.data, cells.ctypes.data, cls.ctypes.data, N,
                            oV.ctypes.data, oo.ctypes.data)
    gphi = to_dev(torch, phi)
    gc, gcls = to_dev(torch, cells), to_dev(torch, cls)
    gV = torch.zeros(N, dtype=torch.float64, device="cuda")
    go = torch.zeros(N, dtype=torch.float32, device="cuda")
    ok(lib.pfh2d_verify_cells(gphi.data_ptr(), gc.data_ptr(), gcls.data_ptr(), N, gV.data_ptr(),
                              go.data_ptr(), None), "verify_cells")
    S = sim.S_now()
    for off in (0, 1, 3):
        Sx = S.copy()
        Sx[off] = 0x7ABCDEF1 + off                     # head word
        Sx[-1] = 0x7ABCDEF0
        gS = to_dev(torch, np.concatenate([np.zeros(off, np.uint32), Sx]))
        gm = torch.zeros(1, dtype=torch.int32, device="cuda")
        ok(lib.pfh2d_verify_S(gS.data_ptr() + 4 * off, Sx.size, gm.data_ptr(), None), "verify_S")
        torch.cuda.synchronize()
        assert int(from_dev(gm, np.uint32)[0]) == int(olib.pfo2d_verify_S(Sx.ctypes.data, Sx.size))
    gm = torch.zeros(1, dtype=torch.int32, device="cuda")
    ok(lib.pfh2d_verify_S(to_dev(torch, S).data_ptr(), S.size, gm.data_ptr(), None), "verify_S")
    torch.cuda.synchronize()
    assert int(from_dev(gm, np.uint32)[0]) == int(S.max())
    gv = from_dev(gV, np.float64)
    assert np.array_equal(gv.view(np.uint64), oV.view(np.uint64))
    assert np.array_equal(from_dev(go, np.float32), oo)
    assert oo[3] == np.float32(0.125) and oo[7] == np.float32(3.0e-3)
    rel = np.abs(gv - cells["V"]) / np.maximum(1e-12, np.abs(cells["V"]))
    assert rel.max() < 1e-12


def _random_cells(rng, N, L, ties=False):
    c = np.zeros(N, abi.CELL_DTYPE)
    c["global_id"] = np.arange(N)
    c["gx0"] = rng.integers(0, L, N)
    c["gy0"] = rng.integers(0, L, N)
    c["V"] = rng.uniform(100.0, 2000.0, N)
    c["V"][::17] = 0.0                                  # V <= 0 uses 1
    c["Cx"] = c["V"] * rng.uniform(0.0, 143.0, N)
    c["Cy"] = c["V"] * rng.uniform(0.0, 143.0, N)
    if ties:
        c["gx0"][1::2] = c["gx0"][0::2][: N // 2]
        c["gy0"][1::2] = c["gy0"][0::2][: N // 2]
        c["Cx"][1::2] = c["Cx"][0::2][: N // 2]
        c["Cy"][1::2] = c["Cy"][0::2][: N // 2]
        c["V"][1::2] = c["V"][0::2][: N // 2]
    for f in ("vx", "vy", "theta", "gamma", "v_A", "phi_max", "M_pf"):
        c[f] = rng.standard_normal(N).astype(np.float32)
    c["perim"] = rng.uniform(0, 300, N)
    return c


@pytest.mark.parametrize("N,L,ties", [(1, 384, False), (16, 384, False), (1000, 2048, True),
                                      (1024, 2048, False), (5000, 1 << 14, True),
                                      (16384, 70000, False)])
def test_morton_sort_matches_oracle(dev, oracle, N, L, ties):
    torch = dev
    lib, olib = abi.load(), oracle.load()
    rng = np.random.default_rng(N + L)
    cells = _random_cells(rng, N, L, ties)
    say(f"morton sort N={N} L={L} ties={ties}")
    want = np.zeros(N, np.uint32)
    olib.pfo2d_morton_perm(cells.ctypes.data, N, L, want.ctypes.data)
    M = 1
    while M < N:
        M <<= 1
    gc = to_dev(torch, cells)
    gp = torch.full((N,), -1, dtype=torch.int32, device="cuda")
    ok(lib.pfh2d_morton_sort(gc.data_ptr(), gp.data_ptr(), N, M, L, None), "morton_sort")
    torch.cuda.synchronize()
    got = from_dev(gp, np.uint32)
    assert np.array_equal(got, want)
    assert np.array_equal(np.sort(got), np.arange(N, dtype=np.uint32))


def test_morton_sort_rejects_bad_sizes(dev):
    lib = abi.load()
    invalid = -1
    assert lib.pfh2d_morton_sort(1, 1, 5, 6, 384, None) == invalid      # M not a power of 2
    assert lib.pfh2d_morton_sort(1, 1, 5, 4, 384, None) == invalid      # M < N
    assert lib.pfh2d_morton_sort(1, 1, 20000, 32768, 384, None) == invalid   # N > 16384


@pytest.mark.parametrize("N", [1, 63, 64, 65, 130])
def test_pack_traj_matches_oracle(dev, oracle, N):
    torch = dev
    lib, olib = abi.load(), oracle.load()
    rng = np.random.default_rng(N)
    L = 701
    cells = _random_cells(rng, N, L)
    cls = rng.integers(0, 6, N).astype(np.uint8)
    want = np.zeros(N, abi.TRAJ_DTYPE)
    olib.pfo2d_pack_traj(cells.ctypes.data, cls.ctypes.data, want.ctypes.data, N, L)
    gc, gcls = to_dev(torch, cells), to_dev(torch, cls)
    gout = torch.full((N * 12,), -1, dtype=torch.int32, device="cuda")
    ok(lib.pfh2d_pack_traj(gc.data_ptr(), gcls.data_ptr(), gout.data_ptr(), N, L, None),
       "pack_traj")
    torch.cuda.synchronize()
    assert from_dev(gout, np.uint8).tobytes() == want.tobytes()


@pytest.mark.parametrize("N", [1, 64, 77])
def test_pack_trajectory_3d(dev, N):
    """k_pack_trajectory (src/pf3d/kernels.cu:2323-2356, periodic: no height
    offset): centroid = origin + C / V (0 when V <= 0), fields copied."""
    torch = dev
    lib = abi.load()
    rng = np.random.default_rng(100 + N)
    c = np.zeros(N, abi.CELL3_DTYPE)
    c["global_id"] = rng.integers(0, 1 << 40, N)
    for f in ("origin_x", "origin_y", "origin_z"):
        c[f] = rng.integers(-5, 300, N)
    c["V"] = rng.uniform(10, 1e4, N)
    c["V"][::5] = 0.0
    for f in ("Cx", "Cy", "Cz", "surface", "wall_overlap", "outside_slit_volume"):
        c[f] = rng.uniform(0, 1e5, N)
    for f in ("velocity_x", "velocity_y", "velocity_z", "polarity_x", "polarity_y",
              "polarity_z", "gamma", "v_A", "phi_max"):
        c[f] = rng.standard_normal(N).astype(np.float32)
    want = np.zeros(N, abi.TRAJ3_DTYPE)
    inv = np.where(c["V"] > 0.0, 1.0 / np.where(c["V"] > 0, c["V"], 1.0), 0.0)
    want["global_id"] = c["global_id"]
    want["x"] = c["origin_x"].astype(np.float64) + c["Cx"] * inv
    want["y"] = c["origin_y"].astype(np.float64) + c["Cy"] * inv
    want["z"] = c["origin_z"].astype(np.float64) + c["Cz"] * inv
    for f in ("surface", "wall_overlap", "outside_slit_volume", "velocity_x", "velocity_y",
              "velocity_z", "polarity_x", "polarity_y", "polarity_z", "gamma", "v_A", "phi_max"):
        want[f] = c[f]
    want["volume"] = c["V"]
    gc = to_dev(torch, c)
    gout = torch.full((N * 26,), -1, dtype=torch.int32, device="cuda")
    ok(lib.pfh3d_pack_trajectory(gc.data_ptr(), gout.data_ptr(), N, None), "pack_trajectory")
    torch.cuda.synchronize()
    assert from_dev(gout, np.uint8).tobytes() == want.tobytes()
