"""PF3D checkpoints of the 3D periodic path on the GPU: the file the library
writes after two steps equals, byte for byte, the file of the oracle's state
after the same steps; three steps in one run equal 1+1+1 steps resumed from
checkpoint files; a corrupted payload fails the CRC."""
import sys

import numpy as np
import pytest

from pfh import abi, checkpoint3d, params3d

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-checkpoint3d] {msg}", flush=True)
    sys.stdout.flush()


def _wl(full_moment_every=100):
    w = params3d.workload_small()
    w.params.full_moment_every = full_moment_every
    return w


def test_written_file_equals_oracle_state_file(tmp_path):
    from oracle import pfo3
    from pfh import sim3d
    w = _wl()
    sim = sim3d.Sim3D(w, device="cuda:0", shards=2)
    o = pfo3.OracleSim3D(w.params, w.cx, w.cy, w.cz, shards=2, threads=4)
    sim.step(2)
    o.step(2)
    got = tmp_path / "gpu.pf3d"
    sim.save_checkpoint(got)
    want = tmp_path / "oracle.pf3d"
    B = w.params.brick_edge()
    checkpoint3d.write_host(want, checkpoint3d.params_record(w.params, B), 2,
                            2.0 * w.params.dt, o.cells, o.phi_now(), B, 2)
    say(f"file sizes {got.stat().st_size} / {want.stat().st_size}")
    assert got.read_bytes() == want.read_bytes()


def test_split_restart_is_byte_identical(tmp_path):
    from pfh import sim3d
    w = _wl(full_moment_every=1)
    direct = sim3d.Sim3D(w, device="cuda:0", shards=2)
    direct.step(3)
    d = tmp_path / "direct.pf3d"
    direct.save_checkpoint(d)
    path = tmp_path / "split.pf3d"
    leg = sim3d.Sim3D(w, device="cuda:0", shards=2)
    leg.step(1)
    leg.save_checkpoint(path)
    for k in (2, 3):
        leg = sim3d.Sim3D(w, device="cuda:0", shards=2)
        leg.load_checkpoint(path)
        assert leg.steps_done == k - 1
        leg.step(1)
        leg.save_checkpoint(path)
    assert checkpoint3d.read_header(path)["step"] == 3
    assert path.read_bytes() == d.read_bytes(), "3D state differs after split continuation"
    assert not leg.fatal()


def test_corrupted_cube_fails_the_crc(tmp_path):
    from pfh import sim3d
    w = _wl()
    sim = sim3d.Sim3D(w, device="cuda:0", shards=2)
    sim.step(1)
    p = tmp_path / "ok.pf3d"
    sim.save_checkpoint(p)
    data = bytearray(p.read_bytes())
    data[-1] ^= 0x01                       # one bit of the last cube
    bad = tmp_path / "bad.pf3d"
    bad.write_bytes(bytes(data))
    other = sim3d.Sim3D(w, device="cuda:0", shards=2)
    with pytest.raises(abi.PfhError, match="CRC-64/ECMA mismatch"):
        other.load_checkpoint(bad)
