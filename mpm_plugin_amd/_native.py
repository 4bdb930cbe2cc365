"""ctypes binding of libmpm_hip.so (declarations in include/mpm_hip.h).

The library is loaded from this package's folder. A missing or stale build raises at import of a caller
that needs it; there is no fallback path.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_hip.so")

ABI_VERSION = 1
OK, ERR_VALUE, ERR_NOT_IMPLEMENTED, ERR_MEMORY, ERR_RUNTIME = 0, -1, -2, -3, -4
K_P2G, K_GRID, K_G2P, K_BIN, K_COLLIDERS, K_OTHER = range(6)
OPT_P2G_MODE, OPT_SORT_INTERVAL, OPT_P2G_BLOCKS, OPT_PRECISION, OPT_P2G_WAVES, OPT_P2G_FLUSH = range(6)
P2G_FLUSH_STAGED, P2G_FLUSH_ATOMIC = 0, 1
P2G_BINNED, P2G_ATOMIC, P2G_FIXED = 0, 1, 2
PRECISION_FP32, PRECISION_FALLBACK, PRECISION_FP64 = 0, 1, 2
(BATCH_OPT_P2G_BLOCKS, BATCH_OPT_GRAPH, BATCH_OPT_SORT_INTERVAL, BATCH_OPT_SUBSTEPS_LAUNCHED, BATCH_OPT_G2P_BLOCKS,
 BATCH_OPT_GRID_BLOCKS, BATCH_OPT_G2P_TILE, BATCH_OPT_P2G_WAVES) = range(8)
KERNEL_NAMES = ("p2g", "grid", "g2p", "bin", "colliders", "other")
# kernel classes of a batched substep (MPM_BK_*), in their order
BATCH_KERNEL_NAMES = ("plan", "sort", "drift", "p2g", "grid_scan", "grid", "g2p", "fixup", "advance")


class DomainDesc(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("origin", C.c_float * 3), ("dx", C.c_float), ("inv_dx", C.c_float),
                ("p_vol", C.c_float), ("gravity", C.c_float * 3), ("wall_lo_node", C.c_int32 * 3),
                ("wall_hi_node", C.c_int32 * 3), ("lo_on", C.c_int32 * 3), ("hi_on", C.c_int32 * 3),
                ("wall_lo", C.c_float * 3), ("wall_hi", C.c_float * 3), ("clamp_lo", C.c_float * 3),
                ("clamp_hi", C.c_float * 3), ("wall_friction", C.c_float), ("cell_vol", C.c_float),
                ("remove_lo", C.c_float * 3), ("remove_hi", C.c_float * 3), ("dx_f64", C.c_double)]


class MaterialDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("mu", C.c_float), ("lam", C.c_float), ("yield_stress", C.c_float),
                ("alpha", C.c_float), ("hardening", C.c_float), ("max_hardening", C.c_float),
                ("crit_c", C.c_float), ("crit_s", C.c_float), ("p_mass", C.c_float), ("wave_speed", C.c_double)]


class ColliderPose(C.Structure):
    _fields_ = [("t0", C.c_float * 3), ("t1", C.c_float * 3), ("q0", C.c_float * 4), ("q1", C.c_float * 4),
                ("S0", C.c_float * 9), ("S1", C.c_float * 9), ("margin", C.c_float)]


class AdvanceStats(C.Structure):
    _fields_ = [("substeps", C.c_int64), ("dt", C.c_double), ("removed", C.c_int64)]


class SurfaceDesc(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("origin", C.c_float * 3), ("cell", C.c_float), ("h", C.c_float),
                ("inv_rest", C.c_float), ("iso", C.c_float)]


class NativeError(RuntimeError):
    pass


_lib = None


def _declare(L: C.CDLL) -> None:
    ctx = C.c_void_p
    fp, ip, u8, i64p = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint8), C.POINTER(C.c_int64)
    i32, i64, f64 = C.c_int32, C.c_int64, C.c_double
    table = {
        "mpm_abi_version": ([], i32),
        "mpm_last_error": ([], C.c_char_p),
        "mpm_create": ([i32, C.POINTER(DomainDesc), C.POINTER(MaterialDesc), i32, C.POINTER(ctx)], i32),
        "mpm_destroy": ([ctx], i32),
        "mpm_add_particles": ([ctx, i64, fp, fp, ip, ip], i32),
        "mpm_load_state": ([ctx, i64, fp, fp, fp, fp, fp, ip, ip], i32),
        "mpm_particle_count": ([ctx], i64),
        "mpm_get_state": ([ctx, fp, fp, fp, fp, fp, ip, ip], i32),
        "mpm_set_colliders": ([ctx, i32, ip, ip, fp, ip, ip, fp], i32),
        "mpm_set_collider_poses": ([ctx, i32, C.POINTER(ColliderPose)], i32),
        "mpm_set_collider_transforms": ([ctx, i32, fp, fp], i32),
        "mpm_decompose_transform": ([fp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)], i32),
        "mpm_advance": ([ctx, f64, C.POINTER(AdvanceStats)], i32),
        "mpm_substep": ([ctx, f64, f64, f64], i32),
        "mpm_is_occupied": ([ctx, i64, fp, u8], i32),
        "mpm_is_occupied_noalloc": ([ctx, i64, fp, u8], i32),
        "mpm_lattice_inside": ([ctx, C.POINTER(C.c_double), i64, ip, i64, f64, i64, C.POINTER(C.c_double), i64p],
                               i32),
        "mpm_debug_p2g": ([ctx, f64], i32),
        "mpm_debug_grid_update": ([ctx, f64, f64, f64], i32),
        "mpm_debug_g2p": ([ctx, f64, f64, f64], i32),
        "mpm_get_grid": ([ctx, i32, fp, fp], i32),
        "mpm_set_grid": ([ctx, fp, fp], i32),
        "mpm_get_binning": ([ctx, ip, ip, ip, i64p], i32),
        "mpm_debug_sort": ([ctx], i32),
        "mpm_set_option": ([ctx, i32, i64], i32),
        "mpm_get_option": ([ctx, i32], i64),
        "mpm_set_timing": ([ctx, i32], i32),
        "mpm_get_kernel_time": ([ctx, i32, C.POINTER(C.c_double), i64p], i32),
        "mpm_reset_timing": ([ctx], i32),
        "mpm_synchronize": ([ctx], i32),
        "mpm_batch_create": ([i32, C.POINTER(ctx)], i32),
        "mpm_batch_destroy": ([ctx], i32),
        "mpm_batch_set_scenes": ([ctx, C.POINTER(ctx), i32], i32),
        "mpm_batch_scene_count": ([ctx], i32),
        "mpm_batch_advance": ([ctx, C.POINTER(f64), C.POINTER(AdvanceStats)], i32),
        "mpm_batch_begin": ([ctx, C.POINTER(f64)], i32),
        "mpm_batch_run": ([ctx, i64, i64p, ip, C.POINTER(AdvanceStats)], i32),
        "mpm_batch_progress": ([ctx, ip, ip], i32),
        "mpm_batch_set_option": ([ctx, i32, i64], i32),
        "mpm_batch_get_option": ([ctx, i32], i64),
        "mpm_batch_set_timing": ([ctx, i32], i32),
        "mpm_batch_get_kernel_time": ([ctx, i32, C.POINTER(f64), i64p], i32),
        "mpm_batch_reset_timing": ([ctx], i32),
        "mpm_surface_create": ([i32, C.POINTER(ctx)], i32),
        "mpm_surface_destroy": ([ctx], i32),
        "mpm_surface_build": ([ctx, C.POINTER(SurfaceDesc), i64, fp, fp, i64p, i64p], i32),
        "mpm_surface_upload": ([ctx, i64, fp, fp, fp, fp], i32),
        "mpm_surface_build_uploaded": ([ctx, C.POINTER(SurfaceDesc), i64p, i64p], i32),
        "mpm_surface_fetch": ([ctx, fp, ip, fp], i32),
        "mpm_surface_get_field": ([ctx, fp, fp], i32),
        "mpm_surface_times": ([ctx, fp], i32),
        "mpm_surface_tables": ([ip, C.POINTER(C.c_int8)], i32),
        "mpm_device_memory": ([i32, i64p, i64p], i32),
    }
    for name, (args, res) in table.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def exported_names() -> list[str]:
    """Entry points this binding expects the library to export."""
    L = C.CDLL(LIB_PATH) if _lib is None else _lib
    return [n for n in dir(L) if n.startswith("mpm_")]


def lib() -> C.CDLL:
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"{LIB_PATH} is missing: build it with `python -m mpm_plugin_amd.build`")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        if L.mpm_abi_version() != ABI_VERSION:
            raise NativeError("libmpm_hip.so ABI version mismatch: rebuild it")
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Map a C-ABI return code to the exception the reference raises for it."""
    if rc == OK:
        return
    msg = lib().mpm_last_error().decode("utf-8", "replace")
    if rc == ERR_VALUE:
        raise ValueError(msg)
    if rc == ERR_NOT_IMPLEMENTED:
        raise NotImplementedError(msg)
    if rc == ERR_MEMORY:
        raise MemoryError(msg)
    raise NativeError(msg)


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def u8ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))
