"""ctypes binding of the C ABI in include/his_amd.h (libhis_amd.so, built for gfx950).

The library is the only compute path: if it is missing, or no ROCm GPU is visible,
every solver entry raises instead of computing anything elsewhere.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch

LIBPATH = Path(__file__).resolve().parent / "_lib" / "libhis_amd.so"

HIS_F64, HIS_F32 = 0, 1
HIS_E_WORKSPACE = -3
SWE2D_NOALLOC_MAX_BYTES = 64 * 1024 * 1024     # HIS_SWE2D_NOALLOC_MAX_BYTES of include/his_amd.h
BC = {"closed": 0, "transmissive": 1, "stage": 2, "discharge": 3}
MAX_GROUPS = 4
VP = C.c_void_p


class Grid2D(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("B", C.c_int32), ("ny", C.c_int32), ("nx", C.c_int32),
                ("order", C.c_int32), ("flux", C.c_int32), ("bc", C.c_int32 * 4),
                ("dx", C.c_double), ("dy", C.c_double), ("g", C.c_double), ("h_dry", C.c_double),
                ("kp_eps", C.c_double), ("horton_t0", C.c_double),
                ("z", VP), ("manning", VP), ("manning_bs", C.c_int64), ("storage", VP), ("storage_bs", C.c_int64),
                ("f0", VP), ("fc", VP), ("k", VP), ("horton_bs", C.c_int64),
                ("rain_weight", VP), ("rain_weight_bs", C.c_int64)]


class State2D(C.Structure):
    _fields_ = [("h", VP), ("hu", VP), ("hv", VP), ("s_filled", VP), ("infil_cum", VP)]


class Net1D(C.Structure):
    _fields_ = [("B", C.c_int32), ("nc", C.c_int32), ("nf", C.c_int32), ("ngh", C.c_int32),
                ("g", C.c_double), ("h_dry", C.c_double), ("cfl_1d", C.c_double), ("dt_max", C.c_double),
                ("z", VP), ("b", VP), ("yb", VP), ("ts", VP), ("bs", C.c_int64),
                ("m", VP), ("nman", VP), ("dx", VP),
                ("fb", VP), ("fyb", VP), ("fts", VP), ("fbs", C.c_int64), ("fm", VP),
                ("fl", VP), ("fr", VP), ("kind", VP), ("g_idx", VP), ("g_src", VP), ("g_sign", VP),
                ("lf_ptr", VP), ("lf_face", VP), ("rf_ptr", VP), ("rf_face", VP), ("zbank", VP),
                ("nparts", C.c_int32), ("part_width", C.c_int32), ("p_cptr", VP), ("p_cell", VP), ("p_fptr", VP), ("p_face", VP),
                ("p_gptr", VP), ("p_ghost", VP)]


class Links(C.Structure):
    _fields_ = [("nl", C.c_int32), ("nu", C.c_int32), ("cell1d", VP), ("inv", VP), ("uniq", VP),
                ("zbank", VP), ("length", VP), ("z2", VP), ("u_ptr", VP), ("u_link", VP), ("c_ptr", VP),
                ("c_link", VP), ("area2", C.c_double), ("weir_cw", C.c_double), ("limit", C.c_double),
                ("relax", C.c_double)]


class SeriesC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("K", C.c_int32), ("t0", C.c_double), ("step", C.c_double),
                ("vals", VP), ("base", VP), ("shift", VP), ("on", VP), ("amp", VP), ("w", VP), ("phase", VP)]


class Group(C.Structure):
    _fields_ = [("grid", Grid2D), ("state", State2D), ("net", C.POINTER(Net1D)), ("links", C.POINTER(Links)),
                ("A", VP), ("Q", VP), ("rain", SeriesC), ("edge", SeriesC * 4), ("inflow", SeriesC),
                ("stage", SeriesC), ("dt_group_size", C.c_int32)]


class Control(C.Structure):
    _fields_ = [("t", VP), ("t_limit", VP), ("t_end", VP), ("dt", VP), ("n_sub", VP), ("steps", VP),
                ("clamped", VP), ("acc", VP), ("dt_min_seen", VP), ("dt_max_seen", VP), ("n_sub_max", VP),
                ("error", VP), ("n_sub_limit", C.c_int32)]


class EngineDesc(C.Structure):
    _fields_ = [("ngroups", C.c_int32), ("group", Group * MAX_GROUPS), ("ctl", Control), ("cfl", C.c_double),
                ("dt_min", C.c_double), ("dt_max", C.c_double), ("ws", VP), ("ws_bytes", C.c_size_t),
                ("use_graph", C.c_int32)]


SYMBOLS = {
    "his_swe2d_workspace_bytes": (C.c_size_t, [C.POINTER(Grid2D)]),
    "his_swe2d_step": (C.c_int, [C.POINTER(Grid2D), State2D, VP, VP, VP, VP, VP, C.c_size_t, VP]),
    "his_swe2d_step_noalloc": (C.c_int, [C.POINTER(Grid2D), State2D, VP, VP, VP, VP, VP]),
    "his_swe2d_max_rate": (C.c_int, [C.POINTER(Grid2D), VP, VP, VP, VP, VP, C.c_size_t, VP]),
    "his_swe2d_velocities": (C.c_int, [C.POINTER(Grid2D), VP, VP, VP, VP, VP, VP]),
    "his_swe2d_volumes": (C.c_int, [C.POINTER(Grid2D), VP, VP, VP, VP, C.c_size_t, VP]),
    "his_plausibility": (C.c_int, [C.POINTER(Grid2D), VP, C.POINTER(Net1D), VP, VP, VP]),
    "his_swe1d_workspace_bytes": (C.c_size_t, [C.POINTER(Net1D)]),
    "his_swe1d_step": (C.c_int, [C.POINTER(Net1D), VP, VP, VP, VP, VP, VP, C.c_size_t, VP]),
    "his_swe1d_max_rate": (C.c_int, [C.POINTER(Net1D), VP, VP, VP, VP, VP]),
    "his_exchange_workspace_bytes": (C.c_size_t, [C.POINTER(Links), C.POINTER(Net1D)]),
    "his_exchange": (C.c_int, [C.POINTER(Links), C.POINTER(Net1D), C.c_int32, C.c_int64, VP, VP, VP, VP, VP, VP,
                               VP, VP, C.c_size_t, VP]),
    "his_engine_workspace_bytes": (C.c_size_t, [C.POINTER(EngineDesc)]),
    "his_engine_create": (C.c_int, [C.POINTER(EngineDesc), C.POINTER(VP), VP]),
    "his_engine_step": (C.c_int, [VP, C.c_int32, VP]),
    "his_engine_refresh": (C.c_int, [VP, VP]),
    "his_engine_refresh_members": (C.c_int, [VP, VP, VP]),
    "his_engine_fold_max": (C.c_int, [VP, C.c_int32, VP, VP]),
    "his_engine_track_step_max": (C.c_int, [VP, C.c_int32, VP]),
    "his_engine_set_branch_order": (C.c_int, [VP, C.c_int32]),
    "his_engine_destroy": (C.c_int, [VP]),
    "his_set_tile_shape": (C.c_int, [C.c_int32, C.c_int32]),
    "his_tile_shape": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]),
    "his_last_error": (C.c_char_p, []),
    "his_version": (C.c_char_p, []),
}

_lib = None


def load():
    """Load libhis_amd.so and declare every exported entry point (no GPU needed)."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(f"his_amd: {LIBPATH} is not built; run python -m his_amd.csrc.build")
        lib = C.CDLL(str(LIBPATH))
        for name, (res, args) in SYMBOLS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().his_last_error().decode(errors="replace")
        raise RuntimeError(f"his_amd {what} failed ({rc}): {msg}")


def require_gpu(device) -> torch.device:
    """The engine runs on a ROCm GPU only; anything else is an error, never a fallback."""
    dev = torch.device(device) if device is not None else (torch.device("cuda") if torch.cuda.is_available()
                                                          else torch.device("cpu"))
    if dev.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError("his_amd computes on a ROCm GPU through libhis_amd.so; no GPU device is available "
                           f"(asked for {dev})")
    load()
    # always carry an explicit index: torch.device("cuda") != torch.device("cuda:0"), and state tensors are
    # compared by device before they are handed to the library as raw pointers
    return torch.device("cuda", dev.index if dev.index is not None else torch.cuda.current_device())


def stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def dtype_code(dt: torch.dtype) -> int:
    if dt == torch.float64:
        return HIS_F64
    if dt == torch.float32:
        return HIS_F32
    raise ValueError(f"unsupported dtype {dt}")


def set_tile_shape(dtype: torch.dtype, shape: int) -> None:
    check(load().his_set_tile_shape(dtype_code(dtype), int(shape)), "set_tile_shape")


def num_tile_shapes() -> int:
    """Number of compiled 2-D kernel shapes (the same for both precisions)."""
    tx, ty = C.c_int32(), C.c_int32()
    n = 0
    while load().his_tile_shape(HIS_F32, n, C.byref(tx), C.byref(ty), None) == 0:
        n += 1
    return n


def tile_shape(dtype: torch.dtype, shape: int):
    """(tx, ty, variant) of a compiled 2-D kernel shape; variant 0 tile, 1 LDS march, 2 register march."""
    tx, ty, var = C.c_int32(), C.c_int32(), C.c_int32()
    check(load().his_tile_shape(dtype_code(dtype), int(shape), C.byref(tx), C.byref(ty), C.byref(var)),
          "tile_shape")
    return tx.value, ty.value, var.value
