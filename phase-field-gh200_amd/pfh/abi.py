"""ctypes view of the C ABI in include/pfh.h.

Loads lib/libpfh.so (built by build.py) and declares every entry point and
POD record. Raises ImportError when the library is missing: there is no
fallback implementation of any kernel.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent.parent / "lib" / "libpfh.so"

TILE_PITCH = 288
TILE_AREA = TILE_PITCH * TILE_PITCH
NUM_CLASSES = 6
CLASSES = [(144, 144, 64, 64), (176, 144, 32, 64), (144, 176, 64, 32),
           (160, 160, 32, 32), (224, 224, 32, 32), (286, 286, 1, 1)]

FLAG_NAMES = ["S_overflow", "q_clamp", "support_clip", "class_exhausted",
              "S_other_negative", "phi_nonfinite", "V_nonpositive",
              "class_unsupported"]
FLAG_COUNT = 8
FLAG_SUPPORT_CLIP = 2

PFH_OK = 0

# CellState (include/kernels.cuh:27-57): 192 bytes.
CELL_DTYPE = np.dtype({
    "names": ["gx0", "gy0", "cls", "cls_written", "pad0", "global_id",
              "gamma", "v_A", "R_tgt", "theta", "V", "Cx", "Cy", "perim",
              "bb_lo_x", "bb_hi_x", "bb_lo_y", "bb_hi_y", "vx", "vy", "Ix",
              "Iy", "promote_ctr", "shift_ctr", "tumble_ctr", "phi_max",
              "reserved", "adv_q", "adv_count", "reserved_unused", "M_pf"],
    "formats": ["<i4", "<i4", "u1", ("u1", (2,)), "u1", "<i4", "<f4", "<f4",
                "<f4", "<f4", "<f8", "<f8", "<f8", "<f8", "<i4", "<i4", "<i4",
                "<i4", "<f4", "<f4", "<f8", "<f8", "<u4", "<u4", "<u4", "<f4",
                ("<u4", (2,)), ("<f8", (6,)), "<u8", "<u4", "<f4"],
    "offsets": [0, 4, 8, 9, 11, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 68,
                72, 76, 80, 84, 88, 96, 104, 108, 112, 116, 120, 128, 176,
                184, 188],
    "itemsize": 192,
})

# MomentAccum (common/velocity_moments.h:30-40): 240 bytes.
MOMENT_DTYPE = np.dtype([("q", "<f8", (12,)), ("count", "<u8"), ("held", "<f8", (6,)),
                         ("spatial_count", "<u8"), ("last_sample", "<u8"),
                         ("last_interval", "<u8"), ("last_totals", "<f8", (4,)),
                         ("correction", "<f8", (4,))])
assert MOMENT_DTYPE.itemsize == 240

# Velocity observation modes of pfh2d_launch_step
OBSERVE_DISABLED, OBSERVE_ADVECTION, OBSERVE_SPATIAL = 0, 1, 2

# TrajPackedCell (include/kernels.cuh:60-70): 48 bytes.
TRAJ_DTYPE = np.dtype([("global_id", "<i4"), ("cls", "<i4"), ("cx", "<f4"),
                       ("cy", "<f4"), ("vx", "<f4"), ("vy", "<f4"),
                       ("theta", "<f4"), ("volume", "<f4"), ("perim", "<f4"),
                       ("gamma", "<f4"), ("v_A", "<f4"), ("phi_max", "<f4")])


class StepArgs(C.Structure):
    """pfh2d_step_args (mirror of StepArgs, include/kernels.cuh:74-116)."""
    _fields_ = [
        ("phi_in", C.c_void_p), ("phi_out", C.c_void_p),
        ("S_rd", C.c_void_p), ("S_sc", C.c_void_p), ("S_cl", C.c_void_p),
        ("cell", C.c_void_p), ("cell_cls", C.c_void_p), ("perm", C.c_void_p),
        ("cursor_use", C.c_void_p), ("cursor_clear", C.c_void_p),
        ("step_rd", C.c_void_p), ("step_wr", C.c_void_p),
        ("flags", C.c_void_p), ("moments", C.c_void_p),
        ("physical_dt", C.c_double),
        ("N", C.c_int), ("L", C.c_int), ("P", C.c_int), ("parity_out", C.c_int),
        ("dt", C.c_float), ("A0", C.c_double), ("vol_scale", C.c_double),
        ("bulk_scale", C.c_float), ("rep_coeff", C.c_float),
        ("mot_coeff", C.c_float),
        ("seed", C.c_ulonglong), ("polarity_seed", C.c_ulonglong),
        ("p_tumble", C.c_double), ("full_moment_every", C.c_int),
        ("clear_ahead_words", C.c_ulonglong),
    ]


class KernelStats(C.Structure):
    _fields_ = [("regs", C.c_int), ("agprs", C.c_int), ("sgprs", C.c_int),
                ("local_bytes", C.c_size_t), ("static_smem", C.c_int),
                ("dynamic_smem", C.c_int), ("ctas_per_sm", C.c_int),
                ("warps_per_sm", C.c_int), ("occupancy", C.c_double)]


assert C.sizeof(StepArgs) == 216, C.sizeof(StepArgs)

# ---- 3D records (include/pf3d/kernels.cuh) ----
FLAG3_COUNT = 10
FLAG3_NAMES = ["S_overflow", "q_clamp", "S_negative", "nonfinite", "V_nonpositive",
               "support_exhausted", "support_edge", "destination_clip",
               "invalid_geometry", "slab_top_contact"]
FATAL3_MASK = 0x3BF
MAX_SHIFT_3D = 4
MAX_EDGE_3D = 64

# CellState3D (include/pf3d/kernels.cuh:98-137): 256 bytes.
CELL3_DTYPE = np.dtype({
    "names": ["global_id", "origin_x", "origin_y", "origin_z",
              "polarity_x", "polarity_y", "polarity_z",
              "velocity_x", "velocity_y", "velocity_z", "gamma", "v_A", "R_tgt",
              "phi_max", "shift_ctr", "tumble_ctr", "flags", "pad0",
              "bb_lo_x", "bb_hi_x", "bb_lo_y", "bb_hi_y", "bb_lo_z", "bb_hi_z",
              "pending_shift_x", "pending_shift_y", "pending_shift_z", "pad1",
              "V", "Cx", "Cy", "Cz", "surface", "Ix", "Iy", "Iz",
              "storage_edge", "pad2", "wall_overlap", "outside_slit_volume", "reserved"],
    "formats": ["<i8", "<i8", "<i8", "<i8"] + ["<f4"] * 10 + ["<u4"] * 4 + ["<i4"] * 9
               + ["<u4"] + ["<f8"] * 8 + ["<u4", "<u4", "<f8", "<f8", ("<u4", (10,))],
    "offsets": [0, 8, 16, 24, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80, 84,
                88, 92, 96, 100, 104, 108, 112, 116, 120, 124, 128, 136, 144, 152, 160,
                168, 176, 184, 192, 196, 200, 208, 216],
    "itemsize": 256,
})

# TrajPackedCell3D (include/pf3d/kernels.cuh:318-328): 104 bytes.
TRAJ3_DTYPE = np.dtype([("global_id", "<i8"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                        ("volume", "<f8"), ("surface", "<f8"), ("wall_overlap", "<f8"),
                        ("outside_slit_volume", "<f8"),
                        ("velocity_x", "<f4"), ("velocity_y", "<f4"), ("velocity_z", "<f4"),
                        ("polarity_x", "<f4"), ("polarity_y", "<f4"), ("polarity_z", "<f4"),
                        ("gamma", "<f4"), ("v_A", "<f4"), ("phi_max", "<f4"),
                        ("_pad", "<u4")])
PARTIAL3_BYTES = 96


class StepArgs3(C.Structure):
    """pfh3d_step_args (MeasureArgs3D + UpdateArgs3D, include/pf3d/kernels.cuh:232-316)."""
    _fields_ = [
        ("phi_in", C.c_void_p), ("phi_out", C.c_void_p),
        ("S_in", C.c_void_p), ("S_out", C.c_void_p),
        ("cells", C.c_void_p), ("accepted", C.c_void_p),
        ("flags", C.c_void_p), ("counters", C.c_void_p),
        ("partials", C.c_void_p), ("arrivals", C.c_void_p),
        ("N", C.c_int), ("B", C.c_int), ("L", C.c_int), ("P", C.c_int),
        ("shards", C.c_int), ("max_shift", C.c_int), ("support_margin", C.c_int),
        ("full_moment_every", C.c_int),
        ("dt", C.c_float), ("bulk_scale", C.c_float), ("interaction_coeff", C.c_float),
        ("motility_coeff", C.c_float),
        ("V0", C.c_double), ("volume_scale", C.c_double), ("p_tumble", C.c_double),
        ("polarity_stream", C.c_uint64), ("aging_steps", C.c_uint64),
    ]


assert C.sizeof(StepArgs3) == 168, C.sizeof(StepArgs3)


# The reference's 3D launch records (include/pf3d/kernels.cuh:153-338).
class Layout3(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("pitch_x", C.c_int),
                ("boundary_flags", C.c_uint32)]


class WallField3(C.Structure):
    _fields_ = [("values", C.c_void_p), ("planes", C.c_int)]


class Storage3(C.Structure):
    _fields_ = [("z_cap", C.c_int)]


class Selection3(C.Structure):
    _fields_ = [("ids", C.c_void_p), ("count", C.c_int)]


class ExchangeRegion3(C.Structure):
    _fields_ = [("y_lo", C.c_int * 4), ("y_hi", C.c_int * 4), ("y_count", C.c_int),
                ("z_lo", C.c_int), ("z_hi", C.c_int)]


class InitArgs3(C.Structure):
    _fields_ = [("phi_even", C.c_void_p), ("phi_odd", C.c_void_p), ("cells", C.c_void_p),
                ("centres", C.c_void_p), ("N", C.c_int), ("B", C.c_int), ("layout", Layout3),
                ("lambda_", C.c_float), ("seed_radius", C.c_float), ("storage", Storage3)]


class ScatterArgs3(C.Structure):
    _fields_ = [("phi", C.c_void_p), ("cells", C.c_void_p), ("S", C.c_void_p),
                ("global_flags", C.c_void_p), ("layout", Layout3), ("N", C.c_int),
                ("B", C.c_int), ("selection", Selection3), ("storage", Storage3)]


class MeasureArgs3(C.Structure):
    _fields_ = [("phi", C.c_void_p), ("S", C.c_void_p), ("wall", WallField3),
                ("cells", C.c_void_p), ("global_flags", C.c_void_p), ("step", C.c_void_p),
                ("layout", Layout3), ("N", C.c_int), ("B", C.c_int), ("partials", C.c_void_p),
                ("shards", C.c_int), ("polarity_stream", C.c_uint64),
                ("aging_steps", C.c_uint64), ("p_tumble", C.c_double),
                ("motility_coeff", C.c_float), ("max_shift", C.c_int),
                ("support_margin", C.c_int), ("apply_tumble", C.c_uint8),
                ("compute_surface", C.c_uint8), ("selection", Selection3),
                ("storage", Storage3)]


class UpdateArgs3(C.Structure):
    _fields_ = [("phi_in", C.c_void_p), ("phi_out", C.c_void_p), ("S_in", C.c_void_p),
                ("S_out", C.c_void_p), ("wall", WallField3), ("cells", C.c_void_p),
                ("global_flags", C.c_void_p), ("work_cursor", C.c_void_p),
                ("layout", Layout3), ("N", C.c_int), ("B", C.c_int), ("dt", C.c_float),
                ("V0", C.c_double), ("volume_scale", C.c_double),
                ("bulk_scale", C.c_float), ("interaction_coeff", C.c_float),
                ("compute_surface", C.c_uint8), ("selection", Selection3),
                ("tile_pass", C.c_int), ("exchange_region", ExchangeRegion3),
                ("storage", Storage3)]


VERIFY3_DTYPE = np.dtype([("measured_V", "<f8"), ("state_V", "<f8"), ("max_abs_phi", "<f4"),
                          ("nonfinite_count", "<u4"), ("support_edge_count", "<u4"),
                          ("verified", "<u4")])
assert (C.sizeof(InitArgs3), C.sizeof(ScatterArgs3), C.sizeof(MeasureArgs3),
        C.sizeof(UpdateArgs3), VERIFY3_DTYPE.itemsize) == (72, 88, 168, 208, 32)
assert TRAJ3_DTYPE.itemsize == 104

_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_SZ = C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "pfh_abi_version": (_I, []),
    "pfh_last_error": (C.c_char_p, []),
    "pfh_stream_copy": (_I, [_P, _P, _SZ, _P]),
    "pfh2d_configure_step": (_I, []),
    "pfh2d_configure_morton": (_I, [_I]),
    "pfh2d_step_grid": (_I, [_I]),
    "pfh2d_launch_step": (_I, [C.POINTER(StepArgs), _I, _P, _P, _SZ, _F, _I]),
    "pfh2d_velocity_moments": (_I, [C.POINTER(StepArgs), _P, C.c_double, _P]),
    "pfh2d_launch_step_if": (_I, [C.POINTER(StepArgs), _P, _I, _P]),
    "pfh2d_launch_check_step_if": (_I, [C.POINTER(StepArgs), _P, _I, _P]),
    "pfh2d_launch_batch_step": (_I, [C.POINTER(StepArgs), _P, _P, _I, _I, _I, _I, _P]),
    "pfh2d_query_step_stats": (_I, [_I, C.POINTER(KernelStats)]),
    "pfh2d_init_tiles": (_I, [_P, _P, _P, _P, _I, _I, _P, _P, _F, _F, _P]),
    "pfh2d_init_moments": (_I, [_P, _P, _P, _I, _P]),
    "pfh2d_scatter_all": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P]),
    "pfh2d_zero_u32": (_I, [_P, _SZ, _P]),
    "pfh2d_verify_cells": (_I, [_P, _P, _P, _I, _P, _P, _P]),
    "pfh2d_verify_S": (_I, [_P, _SZ, _P, _P]),
    "pfh2d_pack_traj": (_I, [_P, _P, _P, _I, _I, _P]),
    "pfh2d_morton_sort": (_I, [_P, _P, _I, _I, _I, _P]),
    "pfh_philox4x32_10": (_I, [_P, _P, _P]),
    "pfh_uniform53": (C.c_double, [C.c_uint32, C.c_uint32]),
    "pfh2d_domain_side": (_I, [_I, C.c_double, C.c_double]),
    "pfh2d_s_pitch": (_I, [_I]),
    "pfh2d_ic_theta": (_F, [C.c_int32, C.c_uint64]),
    "pfh2d_ic_v_A": (C.c_double, [C.c_int32, C.c_uint64, C.c_double, C.c_double]),
    "pfh2d_grid_jitter_centres": (_I, [_I, _I, C.c_uint64, _P, _P]),
    "pfh_centre_table_fnv1a64": (C.c_uint64, [_P, _P, _I]),
    "pfh3d_init_spheres": (_I, [_P, _P, _P, _P, _I, _I, _I, _F, _F, _P]),
    "pfh3d_clear_S": (_I, [_P, _I, _I, _P]),
    "pfh3d_scatter_current": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _P]),
    "pfh3d_measure_cells": (_I, [C.POINTER(StepArgs3), _I, _I, _I, _P]),
    "pfh3d_update": (_I, [C.POINTER(StepArgs3), _I, _P]),
    "pfh3d_finish_step": (_I, [C.POINTER(StepArgs3), _P]),
    "pfh3d_step": (_I, [C.POINTER(StepArgs3), _I, _P]),
    "pfh3d_step_noalloc": (_I, [C.POINTER(StepArgs3), _I, _P]),
    "pfh3d_pack_trajectory": (_I, [_P, _P, _I, _P]),
    "pfh3d_launch_initialize_spheres": (_I, [C.POINTER(InitArgs3), _P]),
    "pfh3d_launch_scatter_current": (_I, [C.POINTER(ScatterArgs3), _P]),
    "pfh3d_launch_measure": (_I, [C.POINTER(MeasureArgs3), _P, _I, _F, _P]),
    "pfh3d_launch_update": (_I, [C.POINTER(UpdateArgs3), _I, _P]),
    "pfh3d_verify_cells": (_I, [_P, _P, _P, _I, _I, C.POINTER(Layout3), _P]),
    "pfh3d_verify_S": (_I, [_P, C.POINTER(Layout3), _P, _P]),
    "pfh_crc64_ecma": (C.c_uint64, [C.c_uint64, _P, _SZ]),
    "pfh3d_checkpoint_write": (_I, [C.c_char_p, C.c_int64, C.c_double, _P, _P, _I, _I, _P,
                                    C.c_uint64, _P]),
    "pfh3d_checkpoint_read": (_I, [C.c_char_p, _P, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), _P, _P, _P]),
    "pfh3d_update_slabs": (_I, [_I, _I, _I]),
    "pfh3d_grow_bricks": (_I, [_P, _P, _P, _P, _I, _I, _I, _P]),
    "pfh3d_next_brick_edge": (_I, [_I, _I]),
    "pfh3d_brick_edge": (_I, [C.c_double, C.c_double]),
    "pfh3d_domain_side": (_I, [_I, C.c_double, C.c_double]),
    "pfh3d_calibrate_seed_radius": (_I, [C.c_double, C.c_double, _I, C.POINTER(C.c_double)]),
    "pfh3d_sequential_centres": (_I, [_I, C.c_double, C.c_double, C.c_uint64, _P, _P, _P]),
    "pfh3d_initial_polarity": (_I, [C.c_int64, C.c_uint64, _P]),
    "pfh_sur_crop": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "pfh_sur_recrop": (_I, [_P, _P, _P, _P, _I, _I, _P, _P]),
    "pfh_sur_conv": (_I, [_I, _P, _P, _P, _P, _P, _I, _I, _P]),
    "pfh_sur_out": (_I, [_P, _P, _F, _P, _P, _P, _I, _P]),
    "pfh_sur_check": (_I, [_P, _P, _P, _P, _I, _I, _P, _F, _P, _P, _P, _P, _P]),
    "pfh_sur_commit": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "pfh_sur_zero2": (_I, [_P, _P, C.c_ulonglong, _P, _P]),
    "pfh_sur_check_cells": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _F, _F, _F, _P, _P,
                                 _P, _P, _P, _P]),
    "pfh_sur_commit_cells": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _P, _P,
                                  _P]),
    "pfh_sur_advance": (_I, [_P, _P, _I, _I, _P, _P, _P, _P, _P]),
    "pfh_sur_prep": (_I, [_P, _P, _P, _P, _P, _P, _I, _P, C.c_ulonglong, _P, _P, _P, _P]),
    "pfh_sur_round_bf16": (_I, [_P, _P, _I, _P]),
    "pfh_sur_infer_cells": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _F,
                                _P, _P, _P, _P, _P]),
    "pfh_sur_infer": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P, _I, _P, _P]),
    "pfh_sur_loss": (_I, [_P, _P, _P, _P, _I, _F, _P]),
    "pfh_sur_out_bwd": (_I, [_P, _P, _P, _P, _P, _I, _P]),
    "pfh_sur_wgrad": (_I, [_I, _P, _P, _P, _I, _P]),
    "pfh_sur_reduce": (_I, [_P, _I, _I, _P, _P]),
    "pfh_sur_flip": (_I, [_P, _P, _P]),
    "pfh_sur_adam": (_I, [_P, _P, _P, _P, _P, _I, _F, _F, _F, _F, _F, _F, _P]),
    "pfh_sur_rel_l2": (_I, [_P, _P, _P, _I, _P]),
    "pfh_sur_writeback": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _P]),
    "pfh_sur_tumble": (_I, [_P, _I, _P, _I, C.c_ulonglong, C.c_double, _P, _P]),
    "pfh_sur_finish": (_I, [_P, _I, _P, _I, C.c_ulonglong, C.c_double, _P, _P, _I, _P, _P,
                            _P, _P, _P]),
}

# Entry points the ABI header declares; tests check the library exports all.
EXPORTED = list(SIGNATURES)

_lib = None


def load() -> C.CDLL:
    """Load lib/libpfh.so once. Raises ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with "
            "`python phase-field-gh200_amd/build.py` (there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PfhError(RuntimeError):
    pass


def check(status: int, what: str) -> None:
    if status != PFH_OK:
        msg = load().pfh_last_error().decode(errors="replace")
        raise PfhError(f"{what} failed with status {status}: {msg}")
