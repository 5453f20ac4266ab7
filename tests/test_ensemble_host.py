"""Host side of the batched ensemble without a GPU: the workload draw, the
replica records, and the oracle's treatment of a 256-pixel replica (windows
as wide as the side are unavailable)."""
import numpy as np

from pfh import abi, params


def test_c4_workload_draw():
    wls = params.ensemble_c4()
    assert len(wls) == 64
    assert [w.params.seed for w in wls] == list(range(100, 164))
    for w in wls:
        L, N = w.params.Nx, w.params.num_cells
        assert L in (256, 512, 1024)
        assert 16 <= N <= params.ensemble_max_cells(L)
        assert not w.params.validate(min_side=256)
        cx, cy = params.grid_jitter_centres(N, L, w.params.seed)
        assert np.array_equal(w.cx, cx) and np.array_equal(w.cy, cy)
    assert params.ensemble_max_cells(256) == 46
    assert params.ensemble_max_cells(512) == 187
    assert params.ensemble_max_cells(1024) == 750
    # the draw is a pure function of the replica index and seed
    assert [(w.params.Nx, w.params.num_cells) for w in params.ensemble_c4(4)] == \
        [(w.params.Nx, w.params.num_cells) for w in wls[:4]]
    assert params.SimParams(Nx=256, Ny=256, rho=0.1).validate()          # lone: rejected


def test_replica_record_layout():
    from pfh import ensemble
    assert ensemble.REPLICA_DTYPE.itemsize == 32
    assert ensemble.REPLICA_DTYPE.fields["L"][1] == 8
    assert ensemble.REPLICA_DTYPE.fields["polarity_seed"][1] == 24


def test_oracle_256_replica_has_no_fallback_class():
    """A cell whose support needs the 286-wide class on a 256 side raises
    CLASS_EXHAUSTED and keeps its class; on a 384 side it is promoted."""
    from oracle import pfo
    for L, want_cls, want_flag in [(256, 4, 1), (384, 5, 0)]:
        wl = params.ensemble_replica(L, 1, 7)
        orc = pfo.OracleSim2D(wl)
        cells = orc.cells.copy()
        cls = np.full(1, 4, np.uint8)
        cells["cls"] = 4
        cells["cls_written"] = 4
        # support box of 220 x 220 in the 224-wide class: 220 + 8 > 224
        cells["bb_lo_x"], cells["bb_hi_x"] = 2, 221
        cells["bb_lo_y"], cells["bb_hi_y"] = 2, 221
        phi = np.zeros((1, abi.TILE_AREA), np.float32)
        wx, wy, tx0, ty0 = abi.CLASSES[4]
        tile = phi[0].reshape(abi.TILE_PITCH, abi.TILE_PITCH)
        tile[ty0 + 100:ty0 + 120, tx0 + 100:tx0 + 120] = 0.5
        S = np.zeros(params.s_pitch_for(L) * L, np.uint32)
        orc.load_state(phi, S, cells, cls, 0)
        orc.step(1)
        assert int(orc.cells["cls"][0]) == want_cls, L
        assert int(orc.flags[3] != 0) == want_flag, (L, orc.flags)
