"""The pf2d driver's host side without a GPU: command-line validation, the
self-test gates, and the CELL v1 reader's acceptance and rejection rules
(src/checkpoint.cu:348-615) on files written from an oracle state."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pfh import checkpoint, params

REPO = Path(__file__).resolve().parent.parent
EXE = REPO / "phase-field-gh200_amd" / "lib" / "pf2d"


def pf2d(args, cwd):
    return subprocess.run([str(EXE), *map(str, args)], cwd=cwd, text=True,
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)


@pytest.fixture(scope="module")
def c1_file(tmp_path_factory):
    """A valid step-0 checkpoint of 16 cells (L=384) written from the oracle's
    initial state."""
    from oracle import pfo
    L = 384
    p = params.SimParams(Nx=L, Ny=L, num_cells=16, rho=0.13671, target_radius=20.0, seed=1,
                         polarity_seed=1, t_end=0.03)
    cx, cy = params.grid_jitter_centres(16, L, 1)
    orc = pfo.OracleSim2D(params.Workload2D("grid16", p, cx, cy))
    path = tmp_path_factory.mktemp("ck") / "c1.bin"
    checkpoint.write(path, p, 0, orc.cells, orc.cls, orc.phi_now())
    return path


def test_help_and_bad_options(tmp_path):
    assert pf2d(["--help"], tmp_path).returncode == 0
    r = pf2d(["--bogus"], tmp_path)
    assert r.returncode == 2 and "unknown option" in r.stdout
    r = pf2d(["--boundary-out", "b.bin"], tmp_path)
    assert r.returncode == 2 and "not supported" in r.stdout
    # the reference's rule: velocity recording needs an output path and --out
    # (src/main.cu:512-517)
    r = pf2d(["--velocity-moments", "v.bin"], tmp_path)
    assert r.returncode == 2 and "requires --velocity-moments and --out" in r.stdout
    r = pf2d(["--velocity-spatial-stride", "5", "--out", "t.txt"], tmp_path)
    assert r.returncode == 2 and "requires --velocity-moments and --out" in r.stdout
    r = pf2d(["--velocity-spatial-stride", "1001"], tmp_path)
    assert r.returncode == 2
    r = pf2d(["--N", "0"], tmp_path)
    assert r.returncode == 2
    r = pf2d(["--N", "4", "--rho", "0.5"], tmp_path)   # L = 18: smaller than a window
    assert r.returncode == 3 and "must exceed the largest rect dimension 286" in r.stdout


def test_self_test_gates_pass(tmp_path):
    r = pf2d(["--N", "4", "--radius", "12", "--rho", "0.02", "--self-test", "--t-end", "0"],
             tmp_path)
    assert "--- self-test passed ---" in r.stdout, r.stdout
    assert r.stdout.count("[PASS] repulsive sign") == 4


def test_reader_accepts_oracle_written_checkpoint(c1_file, tmp_path):
    r = pf2d(["--checkpoint", c1_file, "--t-end", "0.01", "--no-final-checkpoint"], tmp_path)
    assert "[ckpt] loaded" in r.stdout and "step 0" in r.stdout, r.stdout
    assert "classes round/wide/tall/square/large/fallback 16/0/0/0/0/0" in r.stdout
    assert r.returncode in (0, 5)      # 5: no GPU in this container


def _mutated(src, tmp_path, name, fn):
    data = bytearray(Path(src).read_bytes())
    fn(data)
    out = tmp_path / name
    out.write_bytes(bytes(data))
    return out


@pytest.mark.parametrize("name,mutate,message", [
    ("schema", lambda d: struct.pack_into("<I", d, 4, 2), "unsupported 2D checkpoint schema 2"),
    ("magic", lambda d: struct.pack_into("<I", d, 0, 0x12345678), "is not a cell checkpoint"),
    ("pitch", lambda d: struct.pack_into("<i", d, 44 + 192, 256), "tile pitch 256"),
    ("time", lambda d: struct.pack_into("<d", d, 12, 0.5), "checkpoint time does not match"),
    ("ranks", lambda d: struct.pack_into("<i", d, 44 + 192 + 4, 2),
     "is not a complete single-GPU checkpoint"),
    ("cell_id", lambda d: struct.pack_into("<i", d, 44 + 192 + 16, 5), "carries global id 5"),
    ("phi_max", lambda d: struct.pack_into("<f", d, 44 + 192 + 16 + 88, 0.5),
     "disagrees with its stored class or moments"),
    ("truncated", lambda d: d.__delitem__(slice(len(d) - 16 * 4 - 8, None)),
     "missing required sidecars: MOBI"),
])
def test_reader_rejects(c1_file, tmp_path, name, mutate, message):
    bad = _mutated(c1_file, tmp_path, name + ".bin", mutate)
    r = pf2d(["--checkpoint", bad, "--t-end", "0.01"], tmp_path)
    if name == "truncated":
        # dropping the optional MOBI block leaves a valid file (M defaults to 0.5)
        assert "[ckpt] loaded" in r.stdout, r.stdout
        return
    assert r.returncode == 8, r.stdout
    assert message in r.stdout, r.stdout


def test_writer_layout_round_trips(c1_file):
    info = checkpoint.read(c1_file)
    assert info["pitch"] == 288 and int(info["prefix"]["num_cells_local"]) == 16
    assert int(info["params"]["Nx"]) == 384
    assert sorted(info["sidecars"]) == sorted(n for _, n in checkpoint.SIDECARS)
    rec = info["records"]
    # tile origin + class offset = active-window origin; support inside the window
    assert np.all((rec["support_lo_x"] >= 0) & (rec["support_hi_x"] < 144))
