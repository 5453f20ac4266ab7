"""The 3D launch API in the reference's own records (InitArgs3D,
ScatterArgs3D, MeasureArgs3D, UpdateArgs3D) and the 3D verification
kernels: a step composed from the record-based launches equals pfh3d_step
bit for bit, record-based initialisation equals the plain entry points, and
verify_cells / verify_S agree with a host evaluation of the same field."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from pfh import abi, params3d

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-boundary3d] {msg}", flush=True)
    sys.stdout.flush()


def _layout(sim):
    return abi.Layout3(sim.L, sim.L, sim.L, sim.P, 7)


def test_record_launches_equal_the_step():
    from pfh import sim3d
    w = params3d.workload_small()
    a = sim3d.Sim3D(w, device="cuda:0", shards=2)
    b = sim3d.Sim3D(w, device="cuda:0", shards=2)
    lib = abi.load()
    sp = b._sp()
    for k in range(3):
        say(f"step {k}")
        a.step(1)
        slot = b.steps_done % 2
        args = b._args[slot]
        m = abi.MeasureArgs3()
        m.phi, m.S, m.cells = args.phi_in, args.S_in, args.cells
        m.global_flags, m.step = args.flags, args.counters
        m.layout = _layout(b)
        m.N, m.B, m.partials, m.shards = b.N, b.B, args.partials, b.shards
        m.polarity_stream, m.aging_steps = args.polarity_stream, args.aging_steps
        m.p_tumble, m.motility_coeff = args.p_tumble, args.motility_coeff
        m.max_shift, m.support_margin = args.max_shift, args.support_margin
        m.apply_tumble = 1
        m.compute_surface = 1 if (b.steps_done % w.params.full_moment_every) == 0 else 0
        # the step snapshots the accepted state before measuring
        b.accepted.copy_(b.cell)
        abi.check(lib.pfh3d_launch_measure(C.byref(m), args.arrivals, 1, args.dt, sp), "measure")
        abi.check(lib.pfh3d_clear_S(args.S_out, b.L, b.P, sp), "clear")
        u = abi.UpdateArgs3()
        u.phi_in, u.phi_out, u.S_in, u.S_out = args.phi_in, args.phi_out, args.S_in, args.S_out
        u.cells, u.global_flags = args.cells, args.flags
        u.layout = _layout(b)
        u.N, u.B, u.dt, u.V0, u.volume_scale = b.N, b.B, args.dt, args.V0, args.volume_scale
        u.bulk_scale, u.interaction_coeff = args.bulk_scale, args.interaction_coeff
        abi.check(lib.pfh3d_launch_update(C.byref(u), b.slabs, sp), "update")
        abi.check(lib.pfh3d_finish_step(C.byref(args), sp), "finish")
        b.steps_done += 1
        a.sync()
        b.sync()
        assert np.array_equal(a.phi_now().view(np.uint32), b.phi_now().view(np.uint32))
        assert np.array_equal(a.S_now(), b.S_now())
        assert a.cells().tobytes() == b.cells().tobytes()
        assert a.counter_values() == b.counter_values()
    # records outside this build's scope are refused, not run
    u.tile_pass = 1
    assert lib.pfh3d_launch_update(C.byref(u), b.slabs, sp) == -3
    m.layout = abi.Layout3(b.L, b.L, 40, b.P, 3)
    assert lib.pfh3d_launch_measure(C.byref(m), args.arrivals, 1, args.dt, sp) == -3


def test_record_initialisation_equals_entry_points():
    from pfh import sim3d
    w = params3d.workload_small()
    a = sim3d.Sim3D(w, device="cuda:0", shards=1)
    lib = abi.load()
    phi = torch.full_like(a.phi[0], 3.0)
    phi_odd = torch.full_like(a.phi[0], 3.0)
    cells = torch.from_numpy(a.host_cells().view(np.uint8).reshape(-1)).to(a.dev)
    centres = torch.from_numpy(np.stack([w.cx, w.cy, w.cz], 1).astype(np.float64)
                               .reshape(-1)).to(a.dev)
    rec = abi.InitArgs3(phi.data_ptr(), phi_odd.data_ptr(), cells.data_ptr(),
                        centres.data_ptr(), a.N, a.B, _layout(a),
                        float(np.float32(w.params.lambda_)), float(np.float32(a.seed_radius)),
                        abi.Storage3(0))
    abi.check(lib.pfh3d_launch_initialize_spheres(C.byref(rec), a._sp()), "init")
    S = torch.zeros_like(a.S[0])
    flags = torch.zeros_like(a.flags)
    sc = abi.ScatterArgs3(phi.data_ptr(), cells.data_ptr(), S.data_ptr(), flags.data_ptr(),
                          _layout(a), a.N, a.B, abi.Selection3(None, 0), abi.Storage3(0))
    abi.check(lib.pfh3d_launch_scatter_current(C.byref(sc), a._sp()), "scatter")
    a.sync()
    assert torch.equal(phi, a.phi[0]) and int(phi_odd.count_nonzero()) == 0
    assert torch.equal(S, a.S[0])


def test_verify_cells_and_S_against_host():
    from pfh import sim3d
    w = params3d.workload_small()
    sim = sim3d.Sim3D(w, device="cuda:0", shards=2)
    sim.step(3)
    sim.sync()
    lib = abi.load()
    out = torch.zeros(sim.N * 32, dtype=torch.uint8, device=sim.dev)
    lay = _layout(sim)
    abi.check(lib.pfh3d_verify_cells(sim.phi[sim.steps_done % 2].data_ptr(), sim.cell.data_ptr(),
                                     out.data_ptr(), sim.N, sim.B, C.byref(lay), sim._sp()),
              "verify_cells")
    smax = torch.zeros(1, dtype=torch.int32, device=sim.dev)
    abi.check(lib.pfh3d_verify_S(sim.S[sim.steps_done % 2].data_ptr(), C.byref(lay),
                                 smax.data_ptr(), sim._sp()), "verify_S")
    sim.sync()
    r = out.cpu().numpy().view(abi.VERIFY3_DTYPE)
    phi = sim.phi_now().reshape(sim.N, sim.B, sim.B, sim.B)
    cells = sim.cells()
    B = sim.B
    for n in range(sim.N):
        f = phi[n].astype(np.float64)
        assert abs(r["measured_V"][n] - (f * f).sum()) <= 1e-12 * (f * f).sum()
        assert r["state_V"][n] == cells["V"][n]
        assert r["max_abs_phi"][n] == np.abs(phi[n]).max()
        face = np.zeros((B, B, B), bool)
        face[0], face[-1], face[:, 0], face[:, -1], face[:, :, 0], face[:, :, -1] = (True,) * 6
        assert r["support_edge_count"][n] == int((face & (np.abs(phi[n]) > 1e-5)).sum())
        assert r["nonfinite_count"][n] == 0 and r["verified"][n] == 1
    assert int(smax.cpu().numpy().view(np.uint32)[0]) == int(sim.S_now().max())
