"""Build recipe of the gfx950 library (lib/libpfh.so) and the two drivers.

Compiles every HIP source under csrc/ for gfx950 into one shared library
exporting the C ABI of include/pfh.h, then the C++ host drivers under host/
(the pf2d and pf3d command-line simulators) into lib/pf2d and lib/pf3d,
linked against that library. Runs on a machine without a GPU (hipcc
cross-compiles).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
CSRC = HERE / "csrc"
LIB = HERE / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: binary32 arithmetic in source order, as the reference's
# host semantics define it. Packed FP32 instructions are disabled on purpose.
# The atomic optimizer is off: it turns a single lane's queue-ticket atomic
# into a wave-wide broadcast that waits for the atomic's return at once (and
# with it for every load in flight); no kernel here has many lanes adding to
# one address on a hot path.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-I", str(REPO / "include"), "-I", str(CSRC),
]

# Per-source additions. The 2D step kernel runs at the 128-VGPR limit of a
# 1024-thread workgroup; sinking definitions towards their uses removes most
# of its scratch spills (each spill reload waits for every load in flight).
FILE_FLAGS = {
    "pf2d_kernels.hip": ["-mllvm", "-sink-insts-to-avoid-spills"],
}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:4])} ...")


def _stale(out: Path, deps: list[Path]) -> bool:
    if not out.exists():
        return True
    t = out.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build(force: bool = True) -> Path:
    """Compile csrc/ into lib/libpfh.so and host/ into lib/pf2d and lib/pf3d.
    Returns the .so path."""
    LIB.mkdir(exist_ok=True)
    hip_srcs = sorted(CSRC.glob("*.hip"))
    cpp_srcs = sorted(CSRC.glob("*.cpp"))
    headers = sorted(CSRC.glob("*.h")) + sorted(CSRC.glob("*.hpp")) + [REPO / "include" / "pfh.h"]
    objs = []
    for src in hip_srcs + cpp_srcs:
        obj = LIB / (src.stem + (".hip.o" if src.suffix == ".hip" else ".o"))
        if force or _stale(obj, [src] + headers):
            lang = [] if src.suffix == ".hip" else ["-x", "hip"]
            _run([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src.name, []), *lang, "-c", str(src),
                  "-o", str(obj)])
        objs.append(obj)
    so = LIB / "libpfh.so"
    tmp = LIB / "libpfh.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, so)
    # The host drivers: host-only C++ that reaches the GPU through the C ABI
    # (host/pf2d_*.cpp -> lib/pf2d, host/pf3d_*.cpp -> lib/pf3d).
    host_hdrs = sorted((HERE / "host").glob("*.hpp")) + [REPO / "include" / "pfh.h"]
    for name in ("pf2d", "pf3d"):
        host_srcs = sorted((HERE / "host").glob(f"{name}_*.cpp"))
        exe = LIB / name
        if host_srcs and (force or _stale(exe, host_srcs + host_hdrs + [so])):
            tmp_exe = LIB / f"{name}.tmp"
            _run([HIPCC, "-O2", "-std=c++17", "-I", str(REPO / "include"),
                  *map(str, host_srcs), "-o", str(tmp_exe), "-L", str(LIB), "-lpfh",
                  "-Wl,-rpath,$ORIGIN"])
            os.replace(tmp_exe, exe)
    return so


if __name__ == "__main__":
    print(build(force="--incremental" not in sys.argv))
