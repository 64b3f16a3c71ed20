"""ctypes binding of include/bendy2d.h (libbendy2d.so, built in-tree).

This module never falls back to anything: if the HIP library is missing it
raises, and every solver call runs the library's HIP code.
"""
import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_float, c_int32, c_int64,  # noqa: F401
                    c_uint8, c_uint32, c_uint64, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbendy2d.so")

BD_OK = 0
BD_ERR_ARG = -1
BD_ERR_LINK = -2
BD_ERR_INDEX = -3
BD_ERR_DEVICE = -4
BD_ERR_NOMEM = -5
BD_ERR_CAPACITY = -6
BD_ERR_COLLISION = -7
BD_ERR_NODEVICE = -8
BD_ERR_BUDGET = -9
BD_ERR_STATE = -10

BD_ENGINE_AUTO = 0
BD_ENGINE_GLOBAL = 1
BD_ENGINE_SCENE = 2
BD_ENGINE_TILED = 3


class Vec2(Structure):
    _fields_ = [("x", c_float), ("y", c_float)]


class Particle(Structure):
    _fields_ = [("pos", Vec2), ("prev_pos", Vec2), ("acc", Vec2)]


class Link(Structure):
    _fields_ = [("particle_a", c_uint64), ("particle_b", c_uint64), ("target_distance", c_float)]


class Circle(Structure):
    _fields_ = [("point", Particle), ("radius", c_float)]


class Bounds(Structure):
    _fields_ = [("pos", Vec2), ("size", Vec2)]


class PolygonIn(Structure):
    _fields_ = [("particles", POINTER(Particle)), ("n_particles", c_uint64),
                ("particle_links", POINTER(Link)), ("n_links", c_uint64),
                ("is_static", c_int32), ("center", Vec2), ("scale", c_float)]


class PolygonInfo(Structure):
    _fields_ = [("n_particles", c_uint64), ("n_links", c_uint64), ("is_static", c_int32),
                ("center", Vec2), ("scale", c_float)]


class GridDesc(Structure):
    _fields_ = [("nx", c_uint32), ("ny", c_uint32), ("spacing", c_float), ("x0", c_float), ("y0", c_float),
                ("jitter", c_float), ("seed", c_uint64), ("order", c_int32), ("pin_top", c_int32),
                ("as_circles", c_int32), ("radius", c_float)]


class BatchConfig(Structure):
    _fields_ = [("wg_small", c_uint32), ("wg_large", c_uint32), ("split_points", c_uint32),
                ("lds_small", c_uint32), ("lds_large", c_uint32), ("steps_per_launch", c_uint32),
                ("use_graph", c_int32), ("split_tiny", c_uint32), ("wg_tiny", c_uint32), ("lds_tiny", c_uint32)]


class ServerConfig(Structure):
    _fields_ = [("memory_budget", c_uint64), ("steps_per_launch", c_uint32), ("max_active", c_uint32),
                ("launch", BatchConfig), ("launch_budget_us", c_uint32), ("heavy_update_us", c_uint32)]


class RequestDone(Structure):
    _fields_ = [("request_id", c_int64), ("scene_id", c_int64), ("steps", c_uint32), ("status", c_int32),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("particle_updates", c_uint64)]


class ServerStats(Structure):
    _fields_ = [("resident_scenes", c_uint64), ("known_scenes", c_uint64), ("bytes_used", c_uint64),
                ("bytes_budget", c_uint64), ("launches", c_uint64), ("particle_updates", c_uint64),
                ("requests_done", c_uint64), ("requests_pending", c_uint64), ("evictions", c_uint64),
                ("readmissions", c_uint64), ("waiting_admission", c_uint64), ("busy_s", ctypes.c_double),
                ("heavy_launches", c_uint64)]


BD_ORDER_COLOUR = 0
BD_ORDER_PARTICLE = 1


class SceneImageInfo(Structure):
    _fields_ = [("image_bytes", c_uint64), ("workspace_bytes", c_uint64), ("n_bodies", c_uint32),
                ("n_circles", c_uint32), ("workgroup", c_uint32), ("lds_bytes", c_uint32)]


class Stats(Structure):
    _fields_ = [("n_points", c_uint64), ("n_links", c_uint64), ("link_layers", c_uint64),
                ("updates", c_uint64), ("substeps", c_uint64), ("kernel_launches", c_uint64),
                ("cc_rounds", c_uint64), ("cc_redos", c_uint64), ("cc_pairs", c_uint64),
                ("cc_margin", c_float), ("engine", c_int32), ("tile_passes", c_uint32)]


# (name, restype, argtypes) for every symbol the header declares that the
# bindings use; the ABI test checks the header and the library agree.
_S = c_void_p
_SIGS = [
    ("bd_error_string", c_char_p, [c_int32]),
    ("bd_last_error", c_char_p, []),
    ("bd_version", c_char_p, []),
    ("bd_device_count", c_int32, []),
    ("bd_solver_new", _S, []),
    ("bd_solver_free", None, [_S]),
    ("bd_solver_clone", _S, [_S]),
    ("bd_solver_save", c_int64, [_S, c_void_p, c_uint64]),
    ("bd_solver_load", _S, [c_void_p, c_uint64]),
    ("bd_set_gravity", None, [_S, Vec2]),
    ("bd_get_gravity", Vec2, [_S]),
    ("bd_set_bounds", None, [_S, Bounds]),
    ("bd_get_bounds", Bounds, [_S]),
    ("bd_set_bounds_active", None, [_S, c_int32]),
    ("bd_get_bounds_active", c_int32, [_S]),
    ("bd_add_particle", c_int64, [_S, Vec2]),
    ("bd_add_circle", c_int64, [_S, POINTER(Circle)]),
    ("bd_add_polygon", c_int64, [_S, POINTER(PolygonIn)]),
    ("bd_add_particle_link", c_int64, [_S, POINTER(Link)]),
    ("bd_add_circle_link", c_int64, [_S, POINTER(Link)]),
    ("bd_get_particle_len", c_uint64, [_S]),
    ("bd_get_circles_len", c_uint64, [_S]),
    ("bd_get_polygons_len", c_uint64, [_S]),
    ("bd_get_particle_links", c_uint64, [_S, c_void_p, c_uint64]),
    ("bd_get_circle_links", c_uint64, [_S, c_void_p, c_uint64]),
    ("bd_get_particles", c_uint64, [_S, c_void_p, c_uint64]),
    ("bd_get_circles", c_uint64, [_S, c_void_p, c_uint64]),
    ("bd_get_particle", c_int32, [_S, c_uint64, POINTER(Particle)]),
    ("bd_get_circle", c_int32, [_S, c_uint64, POINTER(Circle)]),
    ("bd_get_polygon", c_int32, [_S, c_uint64, POINTER(PolygonInfo)]),
    ("bd_get_polygon_particles", c_uint64, [_S, c_uint64, c_void_p, c_uint64]),
    ("bd_get_polygon_links", c_uint64, [_S, c_uint64, c_void_p, c_uint64]),
    ("bd_update", c_int32, [_S, c_float]),
    ("bd_polygon_circle", c_int32, [c_float, Vec2, c_uint64, c_int32, c_void_p, c_void_p, POINTER(Vec2)]),
    ("bd_polygon_new", c_int32, [c_void_p, c_uint64, c_int32, c_void_p, c_void_p, POINTER(Vec2)]),
    ("bd_ext_set_sub_steps", c_int32, [_S, c_uint32]),
    ("bd_ext_get_sub_steps", c_uint32, [_S]),
    ("bd_ext_set_link_iterations", c_int32, [_S, c_uint32]),
    ("bd_ext_get_link_iterations", c_uint32, [_S]),
    ("bd_ext_set_pinned", c_int32, [_S, c_uint64, c_int32]),
    ("bd_ext_add_particle_with_prev", c_int64, [_S, Vec2, Vec2]),
    ("bd_ext_add_static_circle", c_int64, [_S, Vec2, c_float]),
    ("bd_ext_add_static_polygon", c_int64, [_S, c_void_p, c_uint64]),
    ("bd_ext_get_static_collider_len", c_uint64, [_S]),
    ("bd_ext_add_particles", c_int64, [_S, c_void_p, c_void_p, c_uint64]),
    ("bd_ext_add_particle_links", c_int64, [_S, c_void_p, c_void_p, c_void_p, c_uint64]),
    ("bd_ext_add_circles", c_int64, [_S, c_void_p, c_uint64]),
    ("bd_ext_add_circle_links", c_int64, [_S, c_void_p, c_void_p, c_void_p, c_uint64]),
    ("bd_ext_set_pinned_range", c_int32, [_S, c_uint64, c_void_p, c_uint64]),
    ("bd_ext_set_engine", c_int32, [_S, c_int32]),
    ("bd_ext_get_engine_used", c_int32, [_S]),
    ("bd_ext_set_tiling", c_int32, [_S, c_uint32, c_uint32]),
    ("bd_ext_tiled_selfcheck", c_int64, [_S, c_uint32, c_uint32, POINTER(c_uint32)]),
    ("bd_ext_sync", c_int32, [_S]),
    ("bd_ext_set_readback", c_int32, [_S, c_int32]),
    ("bd_ext_get_stats", c_int32, [_S, POINTER(Stats)]),
    ("bd_ext_get_static_collider", c_int32, [_S, c_uint64, POINTER(Vec2), POINTER(c_float), c_void_p, c_uint64,
                                              POINTER(c_uint64)]),
    ("bd_ext_get_pinned_range", c_int32, [_S, c_uint64, c_void_p, c_uint64]),
    ("bd_scene_add_grid", c_int64, [_S, POINTER(GridDesc)]),
    ("bd_scene_add_gas", c_int64, [_S, c_uint64, c_float, c_float, c_float, c_float, c_float, c_uint64]),
    ("bd_scene_gas_box", c_int32, [_S, c_uint64, c_float, c_float, c_uint32, c_uint64]),
    ("bd_scene_softbody", c_int64, [_S, c_uint64, c_uint32]),
    ("bd_batch_new", _S, []),
    ("bd_batch_free", None, [_S]),
    ("bd_batch_add_scene", c_int64, [_S, _S]),
    ("bd_batch_len", c_uint64, [_S]),
    ("bd_batch_total_points", c_uint64, [_S]),
    ("bd_batch_update", c_int32, [_S, c_float, c_uint32]),
    ("bd_batch_sync", c_int32, [_S]),
    ("bd_batch_read_scene", c_int32, [_S, c_int64, _S]),
    ("bd_batch_set_config", c_int32, [_S, POINTER(BatchConfig)]),
    ("bd_batch_get_config", c_int32, [_S, POINTER(BatchConfig)]),
    ("bd_batch_autotune", c_int32, [_S, c_float, c_uint32]),
    ("bd_batch_last_timing", c_int32, [_S, POINTER(c_float), POINTER(c_uint32)]),
    ("bd_batch_cc_stats", c_int32, [_S, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
    ("bd_batch_cc_fallbacks", c_int32, [_S, POINTER(c_uint64), POINTER(c_uint64)]),
    ("bd_batch_scene_times", c_int32, [_S, c_int64] + [POINTER(ctypes.c_double)] * 4),
    ("bd_ext_copy_calibration", c_int32, [c_uint64, POINTER(c_float)]),
    ("bd_batch_scene_lists", c_int32, [_S, c_int64, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                                       POINTER(c_float), POINTER(c_float)] + [POINTER(ctypes.c_double)] * 3
                                      + [POINTER(c_uint32)] * 3),
    ("bd_batch_scene_status", c_int32, [_S, c_int64, POINTER(c_float), POINTER(c_uint32), POINTER(c_uint32),
                                        POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)]),
    ("bd_server_new", _S, [POINTER(ServerConfig)]),
    ("bd_server_free", None, [_S]),
    ("bd_server_now", ctypes.c_double, [_S]),
    ("bd_server_submit_create", c_int64, [_S, _S, c_uint32, c_float, POINTER(c_int64)]),
    ("bd_server_submit_continue", c_int64, [_S, c_int64, c_uint32, c_float]),
    ("bd_server_poll", c_int32, [_S, POINTER(RequestDone), c_int32]),
    ("bd_server_wait_idle", c_int32, [_S, ctypes.c_double]),
    ("bd_server_read_scene", c_int32, [_S, c_int64, _S]),
    ("bd_server_get_stats", c_int32, [_S, POINTER(ServerStats)]),
    ("bd_plan_layers", c_int64, [c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_void_p]),
    ("bd_scene_image_describe", c_int32, [_S, POINTER(SceneImageInfo)]),
    ("bd_scene_image_write", c_int32, [_S, c_void_p, c_uint64]),
    ("bd_scene_image_update", c_int32, [POINTER(SceneImageInfo), c_void_p, c_void_p, c_uint64, c_float, c_uint32,
                                        c_void_p]),
    ("bd_scene_image_read", c_int32, [_S, c_void_p, c_uint64]),
    ("bd_scene_image_update_noalloc", c_int32, [POINTER(SceneImageInfo), c_void_p, c_float, c_uint32, c_void_p]),
]

_lib = None


def lib():
    """Load libbendy2d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(python bendy2d_amd/build.py); there is no fallback path")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return [n for n, _, _ in _SIGS]


class BendyError(RuntimeError):
    def __init__(self, code, what=""):
        L = lib()
        msg = L.bd_error_string(code).decode()
        detail = L.bd_last_error().decode()
        super().__init__(f"{what}: {msg} ({code}) {detail}".strip())
        self.code = code


def check(code, what=""):
    if code < 0:
        raise BendyError(code, what)
    return code
