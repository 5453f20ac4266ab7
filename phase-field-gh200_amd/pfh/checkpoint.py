"""CELL v1 checkpoint files of the 2D simulator, in numpy.

The byte layout of common/checkpoint_format.h:8-147 (the same one the pf2d
driver writes and reads): a 44 B fixed prefix, the 192 B parameter record,
the int32 tile pitch, a 12 B rank trailer, N x (96 B cell record + 288^2
binary32 tile), then the POLR, GAMA, VA_A, RADI and MOBI sidecars (8 B
header + N binary32 each). Writer semantics follow src/checkpoint.cu:823-1011:
the record holds the tile origin (active-window origin minus the class
offset, wrapped), the centroid as binary32 of the wrapped double
gx0 + Cx * (1 / V), and the resolved polarity stream.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

from . import abi
from .params import SimParams

MAGIC = 0x43454C4C
FORMAT = 1
SIDECARS = [(0x504F4C52, "theta"), (0x47414D41, "gamma"), (0x56415F41, "v_A"),
            (0x52414449, "R_tgt"), (0x4D4F4249, "M_pf")]

PREFIX = np.dtype([("magic", "<u4"), ("version", "<u4"), ("step", "<i4"),
                   ("cur_time", "<f8"), ("num_cells_local", "<i4"), ("save_interval", "<i4"),
                   ("reserved", "<i4"), ("trajectory_samples", "<i4"), ("bools", "u1", (4,)),
                   ("sp_sz", "<u4")])
PARAMS = np.dtype([("Nx", "<i4"), ("Ny", "<i4"), ("dx", "<f8"), ("dy", "<f8"), ("dt", "<f8"),
                   ("t_end", "<f8"), ("rho", "<f8"), ("lambda", "<f8"),
                   ("gamma_normal", "<f8"), ("gamma_soft", "<f8"), ("soft_fraction", "<f8"),
                   ("kappa", "<f8"), ("target_radius", "<f8"), ("mu", "<f8"), ("v_A", "<f8"),
                   ("v_A_sigma", "<f8"), ("xi", "<f8"), ("tau", "<f8"), ("seed", "<u8"),
                   ("polarity_seed", "<u8"), ("initialization_hash", "<u8"),
                   ("print_interval", "<i4"), ("full_moment_every", "<i4"),
                   ("verify_every", "<i4"), ("save_interval", "<i4"),
                   ("trajectory_samples", "<i4"), ("reserved", "<i4"),
                   ("trajectory_interval", "<i8")])
TRAILER = np.dtype([("num_ranks", "<i4"), ("rank_id", "<i4"), ("num_cells_global", "<i4")])
RECORD = np.dtype([("cell_id", "<i4"), ("origin_x", "<i4"), ("origin_y", "<i4"),
                   ("shape_class", "<i4"), ("promote_ctr", "<u4"), ("reserved0", "<u4"),
                   ("cx", "<f4"), ("cy", "<f4"), ("vx", "<f4"), ("vy", "<f4"),
                   ("volume_moment", "<f8"), ("moment_x", "<f8"), ("moment_y", "<f8"),
                   ("perimeter", "<f8"), ("support_lo_x", "<i4"), ("support_hi_x", "<i4"),
                   ("support_lo_y", "<i4"), ("support_hi_y", "<i4"), ("phi_max", "<f4"),
                   ("reserved1", "<u4")])
assert PREFIX.itemsize == 44 and PARAMS.itemsize == 192
assert TRAILER.itemsize == 12 and RECORD.itemsize == 96
assert PARAMS.fields["tau"][1] == 128 and PARAMS.fields["trajectory_interval"][1] == 184
assert RECORD.fields["volume_moment"][1] == 40 and RECORD.fields["phi_max"][1] == 88


def _wrapd(v: float, L: float) -> float:
    m = math.fmod(v, L)
    return m + L if m < 0.0 else m


def params_record(p: SimParams, traj_samples: int, traj_interval: int,
                  save_interval: int) -> np.ndarray:
    r = np.zeros(1, PARAMS)
    for k, v in [("Nx", p.Nx), ("Ny", p.Ny), ("dx", p.dx), ("dy", p.dy), ("dt", p.dt),
                 ("t_end", p.t_end), ("rho", p.rho), ("lambda", p.lam),
                 ("gamma_normal", p.gamma_normal), ("gamma_soft", p.gamma_cancer),
                 ("soft_fraction", p.cancer_fraction), ("kappa", p.kappa),
                 ("target_radius", p.target_radius), ("mu", p.mu), ("v_A", p.v_A),
                 ("v_A_sigma", p.v_A_sigma), ("xi", p.xi), ("tau", p.tau), ("seed", p.seed),
                 ("polarity_seed", p.polarity_stream()),
                 ("initialization_hash", getattr(p, "initialization_hash", 0)),
                 ("print_interval", p.print_interval),
                 ("full_moment_every", p.full_moment_every), ("verify_every", p.verify_every),
                 ("save_interval", save_interval), ("trajectory_samples", traj_samples),
                 ("trajectory_interval", traj_interval)]:
        r[k] = v
    return r


def write(path, p: SimParams, step: int, cells: np.ndarray, cls: np.ndarray,
          phi: np.ndarray, traj_samples: int = 100, traj_interval: int = 1,
          save_interval: int = 0) -> None:
    """Write a CELL v1 file from a host state (CellState records, classes and
    tiles of phi^step), byte for byte as the pf2d driver writes it."""
    N, L = len(cells), int(p.Nx)
    pre = np.zeros(1, PREFIX)
    pre["magic"], pre["version"], pre["step"] = MAGIC, FORMAT, step
    pre["cur_time"] = float(step) * p.dt
    pre["num_cells_local"], pre["save_interval"] = N, save_interval
    pre["trajectory_samples"], pre["sp_sz"] = traj_samples, PARAMS.itemsize
    tr = np.array([(1, 0, N)], TRAILER)
    recs = np.zeros(N, RECORD)
    for i in range(N):
        c = cells[i]
        wx, wy, tx0, ty0 = abi.CLASSES[int(c["cls"])]
        r = recs[i]
        r["cell_id"] = c["global_id"]
        r["origin_x"] = (int(c["gx0"]) - tx0) % L
        r["origin_y"] = (int(c["gy0"]) - ty0) % L
        r["shape_class"] = c["cls"]
        r["promote_ctr"] = c["promote_ctr"]
        inv = 1.0 / float(c["V"])
        r["cx"] = np.float32(_wrapd(float(c["gx0"]) + float(c["Cx"]) * inv, float(L)))
        r["cy"] = np.float32(_wrapd(float(c["gy0"]) + float(c["Cy"]) * inv, float(L)))
        for k_rec, k_cell in [("vx", "vx"), ("vy", "vy"), ("volume_moment", "V"),
                              ("moment_x", "Cx"), ("moment_y", "Cy"), ("perimeter", "perim"),
                              ("support_lo_x", "bb_lo_x"), ("support_hi_x", "bb_hi_x"),
                              ("support_lo_y", "bb_lo_y"), ("support_hi_y", "bb_hi_y"),
                              ("phi_max", "phi_max")]:
            r[k_rec] = c[k_cell]
    tiles = np.ascontiguousarray(phi, "<f4").reshape(N, abi.TILE_AREA)
    with open(path, "wb") as f:
        f.write(pre.tobytes())
        f.write(params_record(p, traj_samples, traj_interval, save_interval).tobytes())
        f.write(np.int32(abi.TILE_PITCH).tobytes())
        f.write(tr.tobytes())
        for i in range(N):
            f.write(recs[i:i + 1].tobytes())
            f.write(tiles[i].tobytes())
        for magic, field in SIDECARS:
            f.write(np.array([magic, N], "<u4").tobytes())
            f.write(np.ascontiguousarray(cells[field], "<f4").tobytes())


def read(path) -> dict:
    """Parse a CELL v1 file (no validation beyond the layout): prefix,
    params, records, tiles [N, 288*288] and sidecars by field name."""
    data = Path(path).read_bytes()
    o = 0

    def take(dt, n=1):
        nonlocal o
        a = np.frombuffer(data, dt, n, o)
        o += dt.itemsize * n
        return a

    pre = take(PREFIX)[0]
    sp = take(PARAMS)[0]
    pitch = int(take(np.dtype("<i4"))[0])
    tr = take(TRAILER)[0]
    N = int(pre["num_cells_local"])
    recs = np.zeros(N, RECORD)
    tiles = np.zeros((N, pitch * pitch), np.float32)
    for i in range(N):
        recs[i] = take(RECORD)[0]
        tiles[i] = take(np.dtype("<f4"), pitch * pitch)
    side = {}
    while o < len(data):
        magic, count = take(np.dtype("<u4"), 2)
        name = dict(SIDECARS).get(int(magic), hex(int(magic)))
        side[name] = take(np.dtype("<f4"), int(count)).copy()
    return {"prefix": pre, "params": sp, "pitch": pitch, "trailer": tr, "records": recs,
            "tiles": tiles, "sidecars": side}
