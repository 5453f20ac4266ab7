"""The pf2d command-line driver on the GPU: its checkpoints equal the oracle's
state byte for byte, a 3-step run equals a 1+1+1 continuation through
checkpoint files, a bumped schema is rejected with the reference's message,
and graph replay and Morton ordering leave the state unchanged.

Mirrors the reference's own driver contract (tests/run_gpu_smoke.py:36-132
with examples/smoke.args) on this build's executable.
"""
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from pfh import checkpoint, params

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parent.parent
EXE = REPO / "phase-field-gh200_amd" / "lib" / "pf2d"

# examples/smoke.args: the reference's small smoke configuration.
SMOKE = ["--N", "4", "--radius", "12", "--rho", "0.02", "--dt", "0.01", "--t-end", "0.02",
         "--lambda", "7", "--kappa", "10", "--mu", "1", "--xi", "1500", "--tau", "10000",
         "--v-A", "0.01", "--gamma", "1.0", "--seed", "20260823", "--polarity-seed",
         "20260823", "--self-test", "--strict", "--verify-every", "1",
         "--trajectory-interval", "1", "--out", "trajectory.txt", "--checkpoint-dir",
         "checkpoints", "--checkpoint-interval", "1"]


def say(msg):
    print(f"[gpu-pf2d-driver] {msg}", flush=True)
    sys.stdout.flush()


def pf2d(args, cwd, expect_ok=True):
    say("pf2d " + " ".join(str(a) for a in args))
    r = subprocess.run([str(EXE), *map(str, args)], cwd=cwd, capture_output=True, text=True,
                       timeout=150)
    if expect_ok and r.returncode != 0:
        raise AssertionError(f"pf2d exited {r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
    return r


@pytest.fixture(scope="module", autouse=True)
def built():
    if not EXE.exists():
        pytest.fail(f"{EXE} is missing: run the build first")


def test_checkpoint_equals_oracle_state_bytes(tmp_path):
    """16 cells R=20 at rho 0.13671 (L=384), default grid + jitter placement
    of seed 1: after 3 steps the driver's checkpoint equals, byte for byte,
    the CELL v1 file of the oracle's state after the same 3 steps."""
    from oracle import pfo
    pf2d(["--N", 16, "--radius", 20, "--rho", 0.13671, "--seed", 1, "--polarity-seed", 1,
          "--t-end", 0.03, "--print-interval", 0, "--checkpoint-dir", "ck"], tmp_path)
    got = (tmp_path / "ck" / "checkpoint.bin").read_bytes()
    L = 384
    p = params.SimParams(Nx=L, Ny=L, num_cells=16, rho=0.13671, target_radius=20.0, seed=1,
                         polarity_seed=1, t_end=0.03, print_interval=0)
    cx, cy = params.grid_jitter_centres(16, L, 1)
    orc = pfo.OracleSim2D(params.Workload2D("grid16", p, cx, cy))
    orc.step(3)
    from pfh import abi
    p.initialization_hash = int(abi.load().pfh_centre_table_fnv1a64(cx.ctypes.data,
                                                                     cy.ctypes.data, 16))
    want = tmp_path / "oracle.bin"
    checkpoint.write(want, p, 3, orc.cells, orc.cls, orc.phi_now(), traj_samples=100,
                     traj_interval=1)
    w = want.read_bytes()
    assert len(got) == len(w)
    if got != w:
        a, b = checkpoint.read(tmp_path / "ck" / "checkpoint.bin"), checkpoint.read(want)
        for k in RECORD_KEYS:
            assert np.array_equal(a["records"][k], b["records"][k]), k
        assert np.array_equal(a["tiles"].view(np.uint32), b["tiles"].view(np.uint32))
        assert a["params"].tobytes() == b["params"].tobytes()
        raise AssertionError("checkpoint bytes differ outside records, tiles and parameters")


RECORD_KEYS = [n for n in checkpoint.RECORD.names if not n.startswith("reserved")]


def test_smoke_example_resume_and_schema_rejection(tmp_path):
    """examples/smoke.args, a resume leg from its checkpoint, and a checkpoint
    whose schema word is bumped (tests/run_gpu_smoke.py:36-132)."""
    run = tmp_path / "run"
    run.mkdir()
    r = pf2d(SMOKE, run)
    assert "self-test passed" in r.stdout
    assert "[verify] step 2" in r.stdout
    traj = (run / "trajectory.txt").read_text().splitlines()
    assert traj[0] == "# Trajectory data" and traj[2].startswith("# trajectory_schema=1")
    assert len(traj) == 3 + 2 * 4      # two frames of four cells
    ck = run / "checkpoints" / "checkpoint.bin"
    assert ck.stat().st_size > 0
    pf2d(["--checkpoint", ck, "--t-end", "0.03", "--out", run / "resumed" / "trajectory.txt",
          "--checkpoint-dir", run / "resumed" / "checkpoints", "--checkpoint-interval", "1"],
         run)
    assert (run / "resumed" / "checkpoints" / "checkpoint.bin").stat().st_size > 0
    bad = tmp_path / "unsupported.bin"
    data = bytearray(ck.read_bytes())
    struct.pack_into("<I", data, 4, struct.unpack_from("<I", data, 4)[0] + 1)
    bad.write_bytes(bytes(data))
    r = subprocess.run([str(EXE), "--checkpoint", str(bad), "--t-end", "0.03"], text=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=60)
    assert r.returncode != 0
    assert "unsupported 2D checkpoint schema" in r.stdout


def test_split_restart_is_byte_identical(tmp_path):
    """Three steps in one run equal 1+1+1 steps resumed from checkpoint files
    (tumbling every step, full moments every step, strict checks); the
    middle leg writes no trajectory, so the last leg appends frame 3 after
    frame 1 and its rows equal the direct run's frame 3."""
    base = [a for a in SMOKE]
    common = ["--tau", "0.001", "--full-moment", "1", "--strict"]
    direct, split = tmp_path / "direct", tmp_path / "split"
    direct.mkdir()
    split.mkdir()
    pf2d(base + common + ["--t-end", "0.03", "--out", direct / "trajectory.txt",
                          "--checkpoint-dir", direct / "checkpoints"], direct)
    pf2d(base + common + ["--t-end", "0.01", "--out", split / "trajectory.txt",
                          "--checkpoint-dir", split / "checkpoints"], split)
    sc = split / "checkpoints" / "checkpoint.bin"
    pf2d(["--checkpoint", sc, "--t-end", "0.02", "--strict", "--checkpoint-dir",
          split / "checkpoints"], split)
    pf2d(["--checkpoint", sc, "--t-end", "0.03", "--strict", "--out", split / "trajectory.txt",
          "--checkpoint-dir", split / "checkpoints"], split)
    d = (direct / "checkpoints" / "checkpoint.bin").read_bytes()
    assert d == sc.read_bytes(), "three-step state differs after split checkpoint continuation"
    hdr = 3
    dt = (direct / "trajectory.txt").read_text().splitlines()
    st = (split / "trajectory.txt").read_text().splitlines()
    assert dt[:hdr] == st[:hdr]
    assert len(dt) == hdr + 3 * 4 and len(st) == hdr + 2 * 4
    assert st[hdr:hdr + 4] == dt[hdr:hdr + 4]            # frame 1
    assert st[hdr + 4:] == dt[hdr + 8:]                   # frame 3
    info = checkpoint.read(sc)
    assert int(info["prefix"]["step"]) == 3
    assert int(info["params"]["trajectory_interval"]) == 1


def test_graph_replay_and_morton_order_leave_state_unchanged(tmp_path):
    """13 steps replayed from six-step graphs, launched one by one, and with
    the Morton-ordered queue all end in the same checkpoint bytes (the
    per-cell update and the integer scatter do not depend on queue order)."""
    args = ["--N", 64, "--radius", 20, "--rho", 0.3, "--seed", 11, "--tau", 0.5,
            "--t-end", 0.13, "--print-interval", 0, "--full-moment", 5]
    out = {}
    for name, extra in [("graph", []), ("eager", ["--no-graph"]), ("morton", ["--morton"])]:
        d = tmp_path / name
        d.mkdir()
        r = pf2d(args + extra + ["--checkpoint-dir", d / "ck"], d)
        assert "integrity checks: all clear" in r.stdout
        out[name] = (d / "ck" / "checkpoint.bin").read_bytes()
    assert out["graph"] == out["eager"]
    assert out["morton"] == out["eager"]


def _read_pfvmom4(path):
    """PFVMOM4 (common/velocity_format.h): 56-byte header, then frames of an
    int64 step and cells x columns float64."""
    raw = Path(path).read_bytes()
    magic, cells, cols, dt, start, side, stride, quad, dense = struct.unpack(
        "<8sIIdqdIIQ", raw[:56])
    assert magic == b"PFVMOM4\0" and cols == 26 and quad == 2
    frame = 8 + 8 * cells * cols
    assert (len(raw) - 56) % frame == 0
    frames = []
    for k in range((len(raw) - 56) // frame):
        off = 56 + k * frame
        step = struct.unpack("<q", raw[off:off + 8])[0]
        rows = np.frombuffer(raw[off + 8:off + frame], np.float64).reshape(cells, cols)
        frames.append((step, rows))
    return dict(cells=cells, dt=dt, start=start, side=side, stride=stride, dense=dense,
                frames=frames)


def test_velocity_moments_fused_equal_dense_reference(tmp_path):
    """The PFVMOM4 output of a run observed by the fused step variants
    (spatial moments every step) and of the same run observed by the
    independent dense observer (--velocity-moments-reference) integrate the
    same velocity components: frames at the same steps, same cell columns,
    cumulative integrals within 1e-9 relative. A run with a spatial stride
    of 5 after a 10-step dense start records every step, and its physics is
    unchanged by the observation (its final checkpoint equals an unobserved
    run's)."""
    base = ["--N", 16, "--radius", 20, "--rho", 0.24, "--seed", 3, "--polarity-seed", 4,
            "--tau", 0.5, "--t-end", 0.3, "--print-interval", 0, "--trajectory-interval", 10]
    pf2d(base + ["--out", "a.txt", "--velocity-moments", "a.vm", "--velocity-spatial-stride", 1,
                 "--velocity-dense-start", 100], tmp_path)
    pf2d(base + ["--out", "b.txt", "--velocity-moments", "b.vm",
                 "--velocity-moments-reference"], tmp_path)
    pf2d(base + ["--out", "c.txt", "--velocity-moments", "c.vm", "--velocity-spatial-stride", 5,
                 "--velocity-dense-start", 10, "--checkpoint-dir", "ckc"], tmp_path)
    pf2d(base + ["--out", "d.txt", "--checkpoint-dir", "ckd"], tmp_path)
    a, b, c = (_read_pfvmom4(tmp_path / f) for f in ("a.vm", "b.vm", "c.vm"))
    assert a["cells"] == 16 and a["stride"] == 1 and b["stride"] == 1 and c["stride"] == 5
    steps = [s for s, _ in a["frames"]]
    assert steps == [0, 10, 20, 30] == [s for s, _ in b["frames"]] == [s for s, _ in c["frames"]]
    for (s, ra), (_, rb), (_, rc) in zip(a["frames"], b["frames"], c["frames"]):
        assert np.isfinite(ra).all() and np.isfinite(rb).all()
        np.testing.assert_array_equal(ra[:, :9], rb[:, :9])        # id, centroid, ..., count
        assert (ra[:, 8] == s).all()
        q_a, q_b = ra[:, 9:21], rb[:, 9:21]
        # q columns 8, 9: the area source's moment, exactly zero for the
        # fused observer and rounding noise for the dense one
        cols = [0, 1, 2, 3, 4, 5, 6, 7, 10, 11]
        scale = np.abs(q_b[:, cols]).max(axis=0) + 1e-300
        assert (np.abs(q_a[:, cols] - q_b[:, cols]) / scale).max() < 1e-9, s
        # (cancelling sums: their noise is relative to the sum of magnitudes)
        assert np.abs(q_a[:, 8:10] - q_b[:, 8:10]).max() <= 1e-8 * np.abs(q_b[:, 4:8]).max()
        np.testing.assert_array_equal(rc[:, :9], ra[:, :9])
        assert (rc[:, 21] <= ra[:, 21]).all()                       # fewer spatial samples
    assert (a["frames"][-1][1][:, 21] == 30).all()
    # observation changes nothing in the physics
    assert (tmp_path / "ckc" / "checkpoint.bin").read_bytes() == \
        (tmp_path / "ckd" / "checkpoint.bin").read_bytes()
    assert (tmp_path / "c.txt").read_bytes() == (tmp_path / "d.txt").read_bytes()
