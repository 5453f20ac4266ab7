"""PF3D checkpoint files without a GPU: the CRC-64/ECMA-182 check value,
the record layouts, and the library reader's header and parameter checks
(src/pf3d/checkpoint.cu:116-286) on a file written from an oracle state."""
import struct

import numpy as np
import pytest

from pfh import abi, checkpoint3d, params3d


def test_crc64_ecma_check_value():
    lib = abi.load()
    # the published check value of CRC-64/ECMA-182 over "123456789"
    assert lib.pfh_crc64_ecma(0, b"123456789", 9) == 0x6C40DF5F0B497347
    # continuation equals one pass
    a = lib.pfh_crc64_ecma(0, b"1234", 4)
    assert lib.pfh_crc64_ecma(a, b"56789", 5) == 0x6C40DF5F0B497347
    assert lib.pfh_crc64_ecma(0, b"", 0) == 0


@pytest.fixture(scope="module")
def small_file(tmp_path_factory):
    from oracle import pfo3
    w = params3d.workload_small()
    o = pfo3.OracleSim3D(w.params, w.cx, w.cy, w.cz, shards=2, threads=4)
    B = w.params.brick_edge()
    path = tmp_path_factory.mktemp("pf3d") / "small.pf3d"
    rec = checkpoint3d.params_record(w.params, B)
    checkpoint3d.write_host(path, rec, 0, 0.0, o.cells, o.phi_now(), B, 2)
    return path, w


def test_probe_accepts_oracle_written_file(small_file):
    path, w = small_file
    info = checkpoint3d.probe(path)
    assert info["N"] == w.params.num_cells and info["B"] == w.params.brick_edge()
    assert info["step"] == 0 and info["shards"] == 2
    assert int(info["params"]["Nx"]) == w.params.Nx
    assert int(info["params"]["polarity_seed"]) == w.params.polarity_stream()


@pytest.mark.parametrize("offset,fmt,value,message", [
    (0, "<I", 0x12345678, "not a PF3D checkpoint"),
    (4, "<H", 2, "unsupported PF3D checkpoint format"),
    (20, "<I", 1, "missing required checksum flag"),
    (88, "<Q", 65, "shard count is outside 0..64"),
    (96 + 52, "<I", 1, "only the periodic-XYZ cube"),
    (96 + 208, "<I", 2, "integrator, stencil, or RNG contract"),
    (96 + 240, "<Q", 65, "invalid promoted-measurement reduction contract"),
])
def test_probe_rejects(small_file, tmp_path, offset, fmt, value, message):
    path, _ = small_file
    data = bytearray(path.read_bytes())
    struct.pack_into(fmt, data, offset, value)
    bad = tmp_path / "bad.pf3d"
    bad.write_bytes(bytes(data))
    with pytest.raises(abi.PfhError, match=message):
        checkpoint3d.probe(bad)


def test_probe_rejects_truncated_payload(small_file, tmp_path):
    path, _ = small_file
    bad = tmp_path / "short.pf3d"
    bad.write_bytes(path.read_bytes()[:-4])
    with pytest.raises(abi.PfhError, match="file size"):
        checkpoint3d.probe(bad)


def test_header_and_record_offsets():
    h = checkpoint3d.HEADER3_DTYPE
    assert h.fields["step"][1] == 24 and h.fields["file_crc64"][1] == 80
    r = checkpoint3d.RECORD3_DTYPE
    assert r.fields["volume"][1] == 72 and r.fields["pending_shift_x"][1] == 172
    assert r.fields["reserved64"][1] == 192
    p = checkpoint3d.PARAMS3_DTYPE
    assert p.fields["trajectory_interval"][1] == 220 and p.fields["aging_time"][1] == 232
