"""The pf3d command line without a GPU: help, unknown and unsupported
options, and the parameter validation that runs before any device call."""
import subprocess
from pathlib import Path

REPO = Path(__file__).resolve().parent.parent
EXE = REPO / "phase-field-gh200_amd" / "lib" / "pf3d"


def pf3d(args, cwd):
    return subprocess.run([str(EXE), *map(str, args)], cwd=cwd, capture_output=True,
                          text=True, timeout=60)


def test_help_and_rejected_options(tmp_path):
    r = pf3d(["--help"], tmp_path)
    assert r.returncode == 0 and "--brick-edge" in r.stdout and "--checkpoint" in r.stdout
    r = pf3d(["--bogus"], tmp_path)
    assert r.returncode == 2 and "unknown option" in r.stdout
    for flag in (["--geometry", "slab"], ["--memory-mode", "compact"], ["--peer-device", "1"],
                 ["--boundary-out", "b.bin"]):
        r = pf3d(flag, tmp_path)
        assert r.returncode == 2 and "not supported" in r.stdout, flag
    r = pf3d(["--N", "0"], tmp_path)
    assert r.returncode == 2


def test_validation_before_the_device(tmp_path):
    # a misaligned explicit brick, a heterogeneous population and an unstable
    # dt are refused by the parameter checks (exit 5: initialisation failed)
    for extra, message in ((["--brick-edge", "44"], "aligned edge"),
                           (["--cancer-fraction", "0.2"], "heterogeneous"),
                           (["--dt", "0.2"], "stability bound")):
        r = pf3d(["--N", "8", "--radius", "8", "--lambda", "3.5", "--rho", "0.02"] + extra,
                 tmp_path)
        assert r.returncode == 5 and message in r.stdout, (extra, r.stdout)
