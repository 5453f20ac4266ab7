"""Checks the oracle's restated primitives against the reference's own
common/ sources compiled into oracle/_ref/ (oracle/ref.mk). Skipped where
the reference tree was not available at build time."""
import ctypes as C
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pfh import params

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref"
SO = REF / "libref_common.so"
KAT = REF / "common_primitives_cpu_test"

needs_ref = pytest.mark.skipif(not SO.exists(), reason="oracle/_ref not built (no reference tree)")


@needs_ref
def test_reference_known_answer_test_passes():
    out = subprocess.run([str(KAT)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert "passed" in out.stdout


@needs_ref
def test_oracle_philox_and_u53_match_reference_build(oracle):
    ref = C.CDLL(str(SO))
    ref.ref_u53.restype = C.c_double
    ref.ref_u53.argtypes = [C.c_uint32, C.c_uint32]
    lib = oracle.load()
    rng = random.Random(17)
    out = np.zeros(4, np.uint32)
    for _ in range(5000):
        ctr = np.array([rng.getrandbits(32) for _ in range(4)], np.uint32)
        key = np.array([rng.getrandbits(32) for _ in range(2)], np.uint32)
        ref.ref_philox(ctr.ctypes.data_as(C.c_void_p), key.ctypes.data_as(C.c_void_p),
                       out.ctypes.data_as(C.c_void_p))
        assert tuple(int(v) for v in out) == oracle.philox(tuple(ctr), tuple(key))
        a, b = int(ctr[0]), int(ctr[1])
        assert ref.ref_u53(a, b) == lib.pfo_u53(a, b)


@needs_ref
def test_coefficients_match_reference_build():
    ref = C.CDLL(str(SO))
    out = np.zeros(6, np.float64)
    for lam, kappa, xi, mu, R in ((7.0, 10.0, 1500.0, 1.0, 20.0), (3.5, 10.0, 1500.0, 1.0, 12.0),
                                  (5.25, 3.7, 977.0, 0.125, 49.0)):
        p = params.SimParams(lam=lam, kappa=kappa, xi=xi, mu=mu, target_radius=R)
        ref.ref_coeffs(C.c_double(lam), C.c_double(kappa), C.c_double(xi), C.c_double(mu),
                       C.c_double(p.area0()), C.c_double(R), out.ctypes.data_as(C.c_void_p))
        mine = [p.bulk(), p.interaction(), p.motility(), p.volume(), p.interface_k(),
                p.init_radius()]
        assert [float(v) for v in out] == mine
