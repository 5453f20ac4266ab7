"""The pf3d command-line driver (C++ over the C ABI) on the GPU: its run of
the 96^3 / 8-sphere workload equals, cell record and cube for cube, the
Python host's run of the same workload; 5 steps equal 3 + 2 steps resumed
from its own PF3D checkpoint; its trajectory keeps the reference's periodic
header and column contract (include/pf3d/trajectory_header.hpp,
src/pf3d/sim.cu:1969-2034) and appends across a resume."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from pfh import checkpoint3d, params3d

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parent.parent
EXE = REPO / "phase-field-gh200_amd" / "lib" / "pf3d"


def say(msg):
    print(f"[gpu-pf3d-driver] {msg}", flush=True)
    sys.stdout.flush()


def pf3d(args, cwd, expect_ok=True):
    say("pf3d " + " ".join(str(a) for a in args))
    r = subprocess.run([str(EXE), *map(str, args)], cwd=cwd, capture_output=True, text=True,
                       timeout=150)
    if expect_ok and r.returncode != 0:
        raise AssertionError(f"pf3d exited {r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
    return r


def state_bytes(path):
    """(step, cell records, cubes) of a PF3D file, without its parameters."""
    raw = Path(path).read_bytes()
    h = np.frombuffer(raw[:96], checkpoint3d.HEADER3_DTYPE)[0]
    N, B = int(h["num_cells"]), int(h["brick_edge"])
    body = raw[96 + 288:]
    rec = 256 + 4 * B ** 3
    assert len(body) == N * rec
    recs = b"".join(body[i * rec:i * rec + 256] for i in range(N))
    cubes = b"".join(body[i * rec + 256:(i + 1) * rec] for i in range(N))
    return int(h["step"]), recs, cubes


def small_args():
    """The driver's flags for the Python host's 96^3 / 8-sphere workload. The
    driver derives the side from rho (the smallest cube holding the target
    fraction, include/pf3d/params.cuh:224-233), so rho is the workload's realized
    fraction raised by one part in 10^9: exactly that fraction rounds up to
    97."""
    w = params3d.workload_small()
    p = w.params
    return w, ["--N", p.num_cells, "--radius", p.target_radius, "--lambda", p.lambda_,
               "--rho", repr(p.rho * (1 + 1e-9)), "--seed", p.seed, "--print-interval", 0]


def test_driver_equals_python_host(tmp_path):
    from pfh import sim3d
    w, args = small_args()
    r = pf3d(args + ["--t-end", 0.05, "--checkpoint-dir", "ck", "--measure-shards", 2],
             tmp_path)
    assert "periodic cube L=96, N=8" in r.stdout
    sim = sim3d.Sim3D(w, device="cuda:0", shards=2)
    sim.step(5)
    sim.save_checkpoint(tmp_path / "py.pf3d")
    d = state_bytes(tmp_path / "ck" / "checkpoint.pf3d")
    q = state_bytes(tmp_path / "py.pf3d")
    assert d[0] == q[0] == 5
    assert d[1] == q[1], "cell records differ"
    assert d[2] == q[2], "cubes differ"


def test_resume_equals_straight_run_and_trajectory_appends(tmp_path):
    _, args = small_args()
    pf3d(args + ["--t-end", 0.05, "--checkpoint-dir", "a", "--out", "a.txt",
                 "--trajectory-interval", 1], tmp_path)
    pf3d(args + ["--t-end", 0.03, "--checkpoint-dir", "b", "--out", "b.txt",
                 "--trajectory-interval", 1], tmp_path)
    r = pf3d(["--checkpoint", "b/checkpoint.pf3d", "--t-end", 0.05, "--checkpoint-dir", "b",
              "--out", "b.txt", "--trajectory-interval", 1], tmp_path)
    assert "[ckpt] loaded" in r.stdout
    a, b = state_bytes(tmp_path / "a" / "checkpoint.pf3d"), \
        state_bytes(tmp_path / "b" / "checkpoint.pf3d")
    assert a == b
    ta = (tmp_path / "a.txt").read_text().splitlines()
    tb = (tmp_path / "b.txt").read_text().splitlines()
    assert ta[0].startswith("# phase_field_gh200 trajectory schema=1 dim=3 "
                            "geometry=periodic-xyz N=8 L=96 B=40 ")
    assert ta[1] == ("# columns: time cell_id x y z vx vy vz px py pz gamma v_A "
                     "surface sphericity volume phi_max")
    assert len(ta) == 2 + 6 * 8 and ta == tb
    assert all(len(l.split()) == 17 for l in ta[2:])
    # a run whose metadata differs does not append to the file
    r = pf3d(args + ["--t-end", 0.02, "--kappa", 11, "--out", "a.txt"], tmp_path,
             expect_ok=False)
    assert r.returncode != 0 and "refusing to append" in r.stderr
