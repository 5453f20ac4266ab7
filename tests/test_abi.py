"""The C-ABI library loads, exports every entry point include/pfh.h declares,
keeps the reference's record layouts, and rejects bad arguments before any
GPU work (no device is needed for any of this)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from pfh import abi

HEADER = Path(__file__).resolve().parent.parent / "include" / "pfh.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfh\w*)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = declared_functions()
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(abi.EXPORTED) == names


def test_abi_version():
    assert abi.load().pfh_abi_version() == 10000


def test_cell_state_layout_matches_reference():
    # include/kernels.cuh:53-57 pins these; CellState is 192 B.
    d = abi.CELL_DTYPE
    assert d.itemsize == 192
    assert d.fields["advection" if "advection" in d.fields else "adv_q"][1] == 128
    assert d.fields["reserved"][1] == 120
    assert d.fields["M_pf"][1] == 188
    assert abi.TRAJ_DTYPE.itemsize == 48
    assert C.sizeof(abi.StepArgs) == 216


def test_invalid_arguments_rejected_without_gpu():
    lib = abi.load()
    assert lib.pfh2d_launch_step(None, 1, None, None, 0, 0.0, 0) == -1
    a = abi.StepArgs()
    assert lib.pfh2d_launch_step(C.byref(a), 1, None, None, 0, 0.0, 0) == -1
    assert lib.pfh2d_launch_step(C.byref(a), 1, None, None, 0, 0.0, 3) == -1   # no such mode
    assert b"observation" in lib.pfh_last_error()
    # a spatial observation without accumulators
    assert lib.pfh2d_launch_step(C.byref(a), 1, None, None, 0, 0.0, 2) == -1
    assert b"moments" in lib.pfh_last_error()
    assert lib.pfh2d_velocity_moments(C.byref(a), None, 0.01, None) == -1
    assert lib.pfh2d_init_tiles(None, None, None, None, 4, 64, None, None, 1.0, 1.0, None) == -1
    assert lib.pfh2d_init_moments(None, None, None, -1, None) == -1
    assert lib.pfh2d_scatter_all(None, None, None, None, 2, 100, 50, None, None) == -1
    assert lib.pfh2d_morton_sort(None, None, 5, 6, 64, None) == -1      # M not a power of 2
    assert lib.pfh2d_verify_S(None, 4, None, None) == -1
    assert lib.pfh2d_pack_traj(None, None, None, 3, 0, None) == -1
    # zero-sized requests are no-ops that touch nothing
    assert lib.pfh2d_init_tiles(None, None, None, None, 0, 64, None, None, 1.0, 1.0, None) == 0
    assert lib.pfh2d_zero_u32(None, 0, None) == 0


def test_launch_rejects_small_domain():
    lib = abi.load()
    a = abi.StepArgs()
    dummy = np.zeros(16, np.uint64)
    for f in ("phi_in", "phi_out", "S_rd", "S_sc", "S_cl", "cell", "cell_cls", "perm",
              "cursor_use", "cursor_clear", "step_rd", "step_wr", "flags"):
        setattr(a, f, dummy.ctypes.data)
    a.N, a.L, a.P = 4, 256, 256
    # L <= 286: a window would alias onto itself (include/params.cuh:666-680)
    assert lib.pfh2d_launch_step(C.byref(a), 1, None, None, 0, 0.0, 0) == -1
    assert b"286" in lib.pfh_last_error()


def _fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_host_setup_draws_match_the_restated_placement():
    """The library's host-side draws (Philox, u53, ic_theta, ic_v_A, the grid
    + jitter placement) equal the package's own restatement, which the
    reference's Philox known-answer vectors pin; FNV-1a follows its
    published definition (the 'a' vector) over x0, y0, x1, y1 little-endian."""
    lib = abi.load()
    from pfh import params
    out = (C.c_uint32 * 4)()
    for ctr, key in [((0, 0, 0, 0), (0, 0)), ((1, 2, 3, 4), (5, 6)),
                     ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2)]:
        assert lib.pfh_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out) == 0
        assert tuple(out) == tuple(params.philox4x32_10(*ctr, *key))
    assert lib.pfh_uniform53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert lib.pfh_uniform53(0, 0x800) == 2.0 ** -53
    for gid in (0, 1, 7, 1023):
        assert np.float32(lib.pfh2d_ic_theta(gid, 7)) == params.ic_theta(gid, 7)
        assert lib.pfh2d_ic_v_A(gid, 1234, 0.01, 0.3) == params.ic_v_A(gid, 1234, 0.01, 0.3)
    assert lib.pfh2d_ic_v_A(3, 1234, 0.01, 0.0) == 0.01
    for N, L, seed in [(16, 384, 1), (4, 301, 20260823), (1000, 2048, 1234)]:
        cx = np.zeros(N, np.float32)
        cy = np.zeros(N, np.float32)
        assert lib.pfh2d_grid_jitter_centres(N, L, seed, cx.ctypes.data, cy.ctypes.data) == 0
        ex, ey = params.grid_jitter_centres(N, L, seed)
        assert cx.tobytes() == ex.tobytes() and cy.tobytes() == ey.tobytes()
        xy = np.stack([cx, cy], axis=1).astype("<f4").tobytes()
        assert lib.pfh_centre_table_fnv1a64(cx.ctypes.data, cy.ctypes.data, N) == _fnv1a64(xy)
    assert _fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert lib.pfh2d_domain_side(16, 20.0, 0.13671) == 384
    assert lib.pfh2d_domain_side(4, 12.0, 0.02) == 301
    assert lib.pfh2d_s_pitch(301) == 320 and lib.pfh2d_s_pitch(2048) == 2048
    assert lib.pfh2d_domain_side(0, 20.0, 0.5) < 0


def _reference_next_brick_edge(current, maximum):
    """next_brick_edge, include/pf3d/brick_classes.hpp:23-32 (alignment 8,
    growth 32, tier quantum 16), restated."""
    if current <= 0 or current % 8 or maximum <= current:
        return 0
    limit = maximum - maximum % 8
    if limit <= current:
        return 0
    rounded = (current + 32 + 15) // 16 * 16
    return min(rounded, limit)


def test_next_brick_edge_follows_the_reference_tiers_within_64():
    lib = abi.load()
    for L in (49, 64, 65, 96, 97, 200, 256, 2048):
        maximum = (L - 1) // 8 * 8            # src/pf3d/sim.cu:291-293
        for cur in range(8, 72, 8):
            want = _reference_next_brick_edge(cur, min(maximum, abi.MAX_EDGE_3D))
            assert lib.pfh3d_next_brick_edge(cur, L) == want, (cur, L)
    assert lib.pfh3d_next_brick_edge(48, 256) == 64      # the reference's 80, capped
    assert lib.pfh3d_next_brick_edge(16, 256) == 48
    assert lib.pfh3d_next_brick_edge(64, 256) == 0
    assert lib.pfh3d_next_brick_edge(44, 256) == 0       # not an aligned edge
    assert lib.pfh3d_grow_bricks(None, None, None, None, 1, 40, 56, None) < 0
