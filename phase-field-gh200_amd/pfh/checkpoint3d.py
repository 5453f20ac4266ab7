"""PF3D schema-1 checkpoints of the 3D periodic path.

The library writes and reads them (pfh3d_checkpoint_write / _read, with the
CRC-64/ECMA-182 over the whole file); this module builds the parameter
record from SimParams3D (src/pf3d/checkpoint.cu:194-239: resolved polarity
stream, explicit Euler / 27-point / Philox ids, fixed shard policy) and holds
a numpy writer of the same bytes for host states (a CPU oracle's), so a GPU
run's file can be compared with the oracle's byte for byte.

Resume (src/pf3d/sim.cu:840-880): the cubes and cell records are installed,
the accepted-step counter restored, the aggregate rebuilt from the cubes and
the measurement refreshed (no tumble), exactly as after a fresh start.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import abi

PARAMS3_DTYPE = np.dtype([
    ("Nx", "<i8"), ("Ny", "<i8"), ("Nz", "<i8"), ("dx", "<f8"), ("dy", "<f8"), ("dz", "<f8"),
    ("brick_edge", "<u4"), ("boundary_flags", "<u4"), ("dt", "<f8"), ("t_end", "<f8"),
    ("lambda", "<f8"), ("gamma_normal", "<f8"), ("gamma_soft", "<f8"),
    ("soft_fraction", "<f8"), ("kappa", "<f8"), ("target_radius", "<f8"), ("mu", "<f8"),
    ("v_A", "<f8"), ("v_A_sigma", "<f8"), ("xi", "<f8"), ("tau", "<f8"),
    ("volume_fraction", "<f8"), ("seed", "<u8"), ("polarity_seed", "<u8"),
    ("print_interval", "<i8"), ("full_moment_every", "<i8"), ("verify_every", "<i8"),
    ("integrator_id", "<u4"), ("stencil_id", "<u4"), ("rng_id", "<u4"),
    ("trajectory_interval", "<u4"), ("initialization_hash", "<u8"), ("aging_time", "<f8"),
    ("promoted_measure_policy", "<u8"), ("promoted_measure_auto_wave_ctas", "<u8"),
    ("channel_height", "<i8"), ("channel_padding", "<i8"), ("wall_kappa", "<f8"),
    ("wall_width", "<f8")])
HEADER3_DTYPE = np.dtype([
    ("magic", "<u4"), ("version", "<u2"), ("header_bytes", "<u2"), ("endian_marker", "<u4"),
    ("dimensions", "<u4"), ("scalar_format", "<u4"), ("flags", "<u4"), ("step", "<i8"),
    ("time", "<f8"), ("num_cells", "<u8"), ("params_bytes", "<u4"),
    ("cell_record_bytes", "<u4"), ("brick_edge", "<u4"), ("phase_order", "<u4"),
    ("phase_values_per_cell", "<u8"), ("phase_bytes_per_cell", "<u8"), ("file_crc64", "<u8"),
    ("base_measure_shards", "<u8")])
RECORD3_DTYPE = np.dtype([
    ("global_id", "<i8"), ("origin_x", "<i8"), ("origin_y", "<i8"), ("origin_z", "<i8"),
    ("polarity_x", "<f4"), ("polarity_y", "<f4"), ("polarity_z", "<f4"),
    ("velocity_x", "<f4"), ("velocity_y", "<f4"), ("velocity_z", "<f4"), ("gamma", "<f4"),
    ("v_A", "<f4"), ("target_radius", "<f4"), ("phi_max", "<f4"), ("volume", "<f8"),
    ("moment_x", "<f8"), ("moment_y", "<f8"), ("moment_z", "<f8"), ("surface", "<f8"),
    ("interaction_x", "<f8"), ("interaction_y", "<f8"), ("interaction_z", "<f8"),
    ("shift_counter", "<u4"), ("tumble_counter", "<u4"), ("state_flags", "<u4"),
    ("bbox_lo_x", "<i4"), ("bbox_hi_x", "<i4"), ("bbox_lo_y", "<i4"), ("bbox_hi_y", "<i4"),
    ("bbox_lo_z", "<i4"), ("bbox_hi_z", "<i4"), ("pending_shift_x", "<i4"),
    ("pending_shift_y", "<i4"), ("pending_shift_z", "<i4"), ("storage_edge", "<u4"),
    ("reserved32", "<u4"), ("reserved64", "<u8", (8,))])
assert PARAMS3_DTYPE.itemsize == 288 and HEADER3_DTYPE.itemsize == 96
assert RECORD3_DTYPE.itemsize == 256
assert PARAMS3_DTYPE.fields["seed"][1] == 168 and PARAMS3_DTYPE.fields["wall_kappa"][1] == 272
assert RECORD3_DTYPE.fields["storage_edge"][1] == 184

_RECORD_FROM_CELL = [
    ("global_id", "global_id"), ("origin_x", "origin_x"), ("origin_y", "origin_y"),
    ("origin_z", "origin_z"), ("polarity_x", "polarity_x"), ("polarity_y", "polarity_y"),
    ("polarity_z", "polarity_z"), ("velocity_x", "velocity_x"), ("velocity_y", "velocity_y"),
    ("velocity_z", "velocity_z"), ("gamma", "gamma"), ("v_A", "v_A"),
    ("target_radius", "R_tgt"), ("phi_max", "phi_max"), ("volume", "V"), ("moment_x", "Cx"),
    ("moment_y", "Cy"), ("moment_z", "Cz"), ("surface", "surface"), ("interaction_x", "Ix"),
    ("interaction_y", "Iy"), ("interaction_z", "Iz"), ("shift_counter", "shift_ctr"),
    ("tumble_counter", "tumble_ctr"), ("state_flags", "flags"), ("bbox_lo_x", "bb_lo_x"),
    ("bbox_hi_x", "bb_hi_x"), ("bbox_lo_y", "bb_lo_y"), ("bbox_hi_y", "bb_hi_y"),
    ("bbox_lo_z", "bb_lo_z"), ("bbox_hi_z", "bb_hi_z"), ("pending_shift_x", "pending_shift_x"),
    ("pending_shift_y", "pending_shift_y"), ("pending_shift_z", "pending_shift_z")]


def params_record(p, B: int, shards_policy: int = 0, trajectory_interval: int = 0,
                  print_interval: int = 100, verify_every: int = 4096,
                  initialization_hash: int = 0) -> np.ndarray:
    """ParamsRecord3D of a periodic cube (src/pf3d/checkpoint.cu:194-239)."""
    r = np.zeros(1, PARAMS3_DTYPE)
    L = int(p.Nx)
    for k, v in [("Nx", L), ("Ny", L), ("Nz", L), ("dx", 1.0), ("dy", 1.0), ("dz", 1.0),
                 ("brick_edge", B), ("boundary_flags", 7), ("dt", p.dt), ("t_end", p.t_end),
                 ("lambda", p.lambda_), ("gamma_normal", p.gamma_normal),
                 ("gamma_soft", p.gamma_cancer), ("soft_fraction", p.cancer_fraction),
                 ("kappa", p.kappa), ("target_radius", p.target_radius), ("mu", p.mu),
                 ("v_A", p.v_A), ("v_A_sigma", p.v_A_sigma), ("xi", p.xi), ("tau", p.tau),
                 ("volume_fraction", p.rho), ("seed", p.seed),
                 ("polarity_seed", p.polarity_stream()), ("print_interval", print_interval),
                 ("full_moment_every", p.full_moment_every), ("verify_every", verify_every),
                 ("integrator_id", 1), ("stencil_id", 2), ("rng_id", 1),
                 ("trajectory_interval", trajectory_interval),
                 ("initialization_hash", initialization_hash), ("aging_time", p.aging_time),
                 ("promoted_measure_policy", shards_policy)]:
        r[k] = v
    return r


def write_host(path, params: np.ndarray, step: int, time: float, cells: np.ndarray,
               phi: np.ndarray, B: int, base_measure_shards: int) -> None:
    """The bytes pfh3d_checkpoint_write produces, from a host state."""
    lib = abi.load()
    N = len(cells)
    h = np.zeros(1, HEADER3_DTYPE)
    for k, v in [("magic", 0x44334650), ("version", 1), ("header_bytes", 96),
                 ("endian_marker", 0x01020304), ("dimensions", 3), ("scalar_format", 1),
                 ("flags", 3), ("step", step), ("time", time), ("num_cells", N),
                 ("params_bytes", 288), ("cell_record_bytes", 256), ("brick_edge", B),
                 ("phase_order", 1), ("phase_values_per_cell", B ** 3),
                 ("phase_bytes_per_cell", 4 * B ** 3),
                 ("base_measure_shards", base_measure_shards)]:
        h[k] = v
    recs = np.zeros(N, RECORD3_DTYPE)
    for kr, kc in _RECORD_FROM_CELL:
        recs[kr] = cells[kc]
    recs["storage_edge"] = B
    cubes = np.ascontiguousarray(phi, "<f4").reshape(N, B ** 3)
    body = bytearray(params.tobytes())
    for i in range(N):
        body += recs[i:i + 1].tobytes()
        body += cubes[i].tobytes()
    buf = bytes(body)
    crc = lib.pfh_crc64_ecma(0, h.tobytes(), 96)
    crc = lib.pfh_crc64_ecma(crc, buf, len(buf))
    h["file_crc64"] = crc
    Path(path).write_bytes(h.tobytes() + buf)


def read_header(path) -> np.ndarray:
    return np.frombuffer(Path(path).read_bytes()[:96], HEADER3_DTYPE)[0]


def probe(path) -> dict:
    """Header facts of a PF3D file, validated by the library (no GPU)."""
    lib = abi.load()
    pr = np.zeros(1, PARAMS3_DTYPE)
    step, t, sh = C.c_int64(), C.c_double(), C.c_uint64()
    n, b = C.c_int(), C.c_int()
    abi.check(lib.pfh3d_checkpoint_read(str(path).encode(), pr.ctypes.data, C.byref(step),
                                        C.byref(t), C.byref(sh), C.byref(n), C.byref(b),
                                        None, None, None), "pfh3d_checkpoint_read (probe)")
    return {"params": pr[0], "step": step.value, "time": t.value, "shards": sh.value,
            "N": n.value, "B": b.value}
