"""Python mirror of the crate's Solver API over the C ABI.

Method names follow the Rust API of Murrent/bendy2d (src/solver.rs:34-116);
bulk getters return numpy arrays:
  particles: (n, 6) float32  pos.xy, prev_pos.xy, acc.xy      (src/particle.rs:4-9)
  circles:   (n, 7) float32  point pos.xy, prev.xy, acc.xy, radius (src/circle.rs:4-8)
  links:     (a uint64[n], b uint64[n], target float32[n])      (src/link.rs:4-10)
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (BD_ENGINE_AUTO, BD_ENGINE_GLOBAL, BD_ENGINE_SCENE, BD_ENGINE_TILED, BD_ERR_STATE, BatchConfig, Bounds,
                   Circle, Link, Particle, PolygonIn, PolygonInfo, RequestDone, ServerConfig, ServerStats, Stats,
                   Vec2, check, lib)

LINK_DTYPE = np.dtype([("a", "<u8"), ("b", "<u8"), ("t", "<f4"), ("pad", "<u4")])
assert LINK_DTYPE.itemsize == ctypes.sizeof(Link)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _links_out(raw):
    return raw["a"].copy(), raw["b"].copy(), raw["t"].copy()


def polygon_circle(radius, pos, point_count, is_static=False):
    """Polygon::circle (src/polygon.rs:17-82): (particles (n,6), links, center)."""
    n = int(point_count)
    pts = np.zeros((max(n, 1), 6), np.float32)
    links = np.zeros(max(2 * n, 1), LINK_DTYPE)
    c = Vec2()
    check(lib().bd_polygon_circle(float(radius), Vec2(*pos), n, int(bool(is_static)), _ptr(pts), _ptr(links),
                                  ctypes.byref(c)), "bd_polygon_circle")
    return pts[:n], _links_out(links[:2 * n]), (c.x, c.y)


def polygon_new(points, is_static=False):
    """Polygon::new (src/polygon.rs:84-123): (particles (n,6), links, center)."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
    n = len(p)
    pts = np.zeros((max(n, 1), 6), np.float32)
    links = np.zeros(max(n, 1), LINK_DTYPE)
    c = Vec2()
    check(lib().bd_polygon_new(_ptr(p), n, int(bool(is_static)), _ptr(pts), _ptr(links), ctypes.byref(c)),
          "bd_polygon_new")
    return pts[:n], _links_out(links[:n]), (c.x, c.y)


class Solver:
    """Solver (src/solver.rs:19-31) backed by the MI355X library."""

    def __init__(self, _handle=None):
        L = lib()
        self._h = _handle if _handle is not None else L.bd_solver_new()
        if not self._h:
            raise MemoryError("bd_solver_new failed")

    def close(self):
        if getattr(self, "_h", None):
            lib().bd_solver_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clone(self):
        h = lib().bd_solver_clone(self._h)
        if not h:
            raise RuntimeError("bd_solver_clone failed")
        return Solver(h)

    def save(self):
        """The solver's snapshot (bd_solver_save) as bytes."""
        n = check(lib().bd_solver_save(self._h, None, 0), "bd_solver_save")
        buf = (ctypes.c_uint8 * n)()
        check(lib().bd_solver_save(self._h, buf, n), "bd_solver_save")
        return bytes(buf)

    @staticmethod
    def load(data):
        """A new solver from a snapshot made by save()."""
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        h = lib().bd_solver_load(buf, len(data))
        if not h:
            raise _lib.BendyError(_lib.BD_ERR_ARG, "bd_solver_load")
        return Solver(h)

    # --- public fields (src/solver.rs:21-23) ---
    @property
    def gravity(self):
        g = lib().bd_get_gravity(self._h)
        return (g.x, g.y)

    @gravity.setter
    def gravity(self, g):
        lib().bd_set_gravity(self._h, Vec2(float(g[0]), float(g[1])))

    @property
    def bounds(self):
        b = lib().bd_get_bounds(self._h)
        return ((b.pos.x, b.pos.y), (b.size.x, b.size.y))

    @bounds.setter
    def bounds(self, b):
        (px, py), (sx, sy) = b
        lib().bd_set_bounds(self._h, Bounds(Vec2(px, py), Vec2(sx, sy)))

    @property
    def bounds_active(self):
        return bool(lib().bd_get_bounds_active(self._h))

    @bounds_active.setter
    def bounds_active(self, v):
        lib().bd_set_bounds_active(self._h, int(bool(v)))

    # --- adders (src/solver.rs:52-67) ---
    def add_particle(self, pos):
        return check(lib().bd_add_particle(self._h, Vec2(float(pos[0]), float(pos[1]))), "add_particle")

    def add_circle(self, pos, radius, prev_pos=None, acc=(0.0, 0.0)):
        prev_pos = pos if prev_pos is None else prev_pos
        c = Circle(Particle(Vec2(*pos), Vec2(*prev_pos), Vec2(*acc)), float(radius))
        return check(lib().bd_add_circle(self._h, ctypes.byref(c)), "add_circle")

    def add_particle_link(self, a, b, target):
        l = Link(int(a), int(b), float(target))
        return check(lib().bd_add_particle_link(self._h, ctypes.byref(l)), "add_particle_link")

    def add_circle_link(self, a, b, target):
        l = Link(int(a), int(b), float(target))
        return check(lib().bd_add_circle_link(self._h, ctypes.byref(l)), "add_circle_link")

    def add_polygon(self, particles, links, is_static=False, center=(0.0, 0.0), scale=1.0):
        pts = np.ascontiguousarray(particles, np.float32).reshape(-1, 6)
        a, b, t = links
        raw = np.zeros(len(a), LINK_DTYPE)
        raw["a"], raw["b"], raw["t"] = a, b, t
        p = PolygonIn(ctypes.cast(_ptr(pts), ctypes.POINTER(Particle)), len(pts),
                      ctypes.cast(_ptr(raw), ctypes.POINTER(Link)), len(raw),
                      int(bool(is_static)), Vec2(*center), float(scale))
        return check(lib().bd_add_polygon(self._h, ctypes.byref(p)), "add_polygon")

    def add_polygon_circle(self, radius, pos, point_count, is_static=False):
        pts, links, c = polygon_circle(radius, pos, point_count, is_static)
        return self.add_polygon(pts, links, is_static, c, 1.0)

    def add_polygon_new(self, points, is_static=False):
        pts, links, c = polygon_new(points, is_static)
        return self.add_polygon(pts, links, is_static, c, 1.0)

    # --- bulk / extension adders ---
    def add_particles(self, pos, prev=None):
        p = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
        q = None if prev is None else np.ascontiguousarray(prev, np.float32).reshape(-1, 2)
        return check(lib().bd_ext_add_particles(self._h, _ptr(p), _ptr(q), len(p)), "add_particles")

    def add_particle_with_prev(self, pos, prev):
        return check(lib().bd_ext_add_particle_with_prev(self._h, Vec2(*pos), Vec2(*prev)), "add_particle_with_prev")

    def add_particle_links(self, a, b, t):
        a = np.ascontiguousarray(a, np.uint32); b = np.ascontiguousarray(b, np.uint32)
        t = np.ascontiguousarray(t, np.float32)
        return check(lib().bd_ext_add_particle_links(self._h, _ptr(a), _ptr(b), _ptr(t), len(a)), "add_particle_links")

    def add_circles(self, circles):
        c = np.ascontiguousarray(circles, np.float32).reshape(-1, 7)
        return check(lib().bd_ext_add_circles(self._h, _ptr(c), len(c)), "add_circles")

    def add_circle_links(self, a, b, t):
        a = np.ascontiguousarray(a, np.uint32); b = np.ascontiguousarray(b, np.uint32)
        t = np.ascontiguousarray(t, np.float32)
        return check(lib().bd_ext_add_circle_links(self._h, _ptr(a), _ptr(b), _ptr(t), len(a)), "add_circle_links")

    def set_sub_steps(self, n):
        check(lib().bd_ext_set_sub_steps(self._h, int(n)), "set_sub_steps")

    def set_link_iterations(self, k):
        check(lib().bd_ext_set_link_iterations(self._h, int(k)), "set_link_iterations")

    def set_pinned(self, i, pinned=True):
        check(lib().bd_ext_set_pinned(self._h, int(i), int(bool(pinned))), "set_pinned")

    def set_pinned_range(self, first, flags):
        f = np.ascontiguousarray(flags, np.uint8)
        check(lib().bd_ext_set_pinned_range(self._h, int(first), _ptr(f), len(f)), "set_pinned_range")

    def add_static_circle(self, center, radius):
        return check(lib().bd_ext_add_static_circle(self._h, Vec2(*center), float(radius)), "add_static_circle")

    def add_static_polygon(self, vertices):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 2)
        return check(lib().bd_ext_add_static_polygon(self._h, _ptr(v), len(v)), "add_static_polygon")

    def set_engine(self, engine):
        check(lib().bd_ext_set_engine(self._h, int(engine)), "set_engine")

    def set_tiling(self, tile_points=2048, workgroup=1024):
        check(lib().bd_ext_set_tiling(self._h, int(tile_points), int(workgroup)), "set_tiling")

    def tiled_selfcheck(self, tile_points=4096, region_cap=16384):
        """Host check of the TILED schedule: (mismatching values, passes per sweep)."""
        passes = ctypes.c_uint32()
        bad = check(lib().bd_ext_tiled_selfcheck(self._h, int(tile_points), int(region_cap), ctypes.byref(passes)),
                    "tiled_selfcheck")
        return bad, passes.value

    @property
    def engine_used(self):
        return lib().bd_ext_get_engine_used(self._h)

    # --- getters (src/solver.rs:69-104) ---
    def get_particle_len(self):
        return lib().bd_get_particle_len(self._h)

    def get_circles_len(self):
        return lib().bd_get_circles_len(self._h)

    def get_polygons_len(self):
        return lib().bd_get_polygons_len(self._h)

    def get_particles(self):
        n = self.get_particle_len()
        out = np.zeros((max(n, 1), 6), np.float32)
        if lib().bd_get_particles(self._h, _ptr(out), n) != n:
            raise RuntimeError("bd_get_particles: " + lib().bd_last_error().decode())
        return out[:n]

    def get_circles(self):
        n = self.get_circles_len()
        out = np.zeros((max(n, 1), 7), np.float32)
        if lib().bd_get_circles(self._h, _ptr(out), n) != n:
            raise RuntimeError("bd_get_circles: " + lib().bd_last_error().decode())
        return out[:n]

    def get_particle(self, i):
        p = Particle()
        if not lib().bd_get_particle(self._h, int(i), ctypes.byref(p)):
            return None
        return np.array([p.pos.x, p.pos.y, p.prev_pos.x, p.prev_pos.y, p.acc.x, p.acc.y], np.float32)

    def get_circle(self, i):
        c = Circle()
        if not lib().bd_get_circle(self._h, int(i), ctypes.byref(c)):
            return None
        q = c.point
        return np.array([q.pos.x, q.pos.y, q.prev_pos.x, q.prev_pos.y, q.acc.x, q.acc.y, c.radius], np.float32)

    def get_particle_links(self):
        n = lib().bd_get_particle_links(self._h, None, 0)
        raw = np.zeros(max(n, 1), LINK_DTYPE)
        lib().bd_get_particle_links(self._h, _ptr(raw), n)
        return _links_out(raw[:n])

    def get_circle_links(self):
        n = lib().bd_get_circle_links(self._h, None, 0)
        raw = np.zeros(max(n, 1), LINK_DTYPE)
        lib().bd_get_circle_links(self._h, _ptr(raw), n)
        return _links_out(raw[:n])

    def get_polygon(self, i):
        info = PolygonInfo()
        if not lib().bd_get_polygon(self._h, int(i), ctypes.byref(info)):
            return None
        n, nl = info.n_particles, info.n_links
        pts = np.zeros((max(n, 1), 6), np.float32)
        lib().bd_get_polygon_particles(self._h, int(i), _ptr(pts), n)
        raw = np.zeros(max(nl, 1), LINK_DTYPE)
        lib().bd_get_polygon_links(self._h, int(i), _ptr(raw), nl)
        return {"particles": pts[:n], "links": _links_out(raw[:nl]), "is_static": bool(info.is_static),
                "center": (info.center.x, info.center.y), "scale": info.scale}

    # --- the hot path ---
    def update(self, dt):
        """Solver::update (src/solver.rs:106-116) on the GPU."""
        check(lib().bd_update(self._h, float(dt)), "update")

    def sync(self):
        check(lib().bd_ext_sync(self._h), "sync")

    def set_readback(self, on=True):
        """Copy every update's state to pinned host memory as it finishes
        (bd_ext_set_readback), so getters after an update only wait for it."""
        check(lib().bd_ext_set_readback(self._h, int(bool(on))), "set_readback")

    def stats(self):
        s = Stats()
        check(lib().bd_ext_get_stats(self._h, ctypes.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}


def plan_layers(a, b, n_points):
    """Order-preserving dependency layers of a link sequence: (layer[], order[], depth)."""
    a = np.ascontiguousarray(a, np.uint32); b = np.ascontiguousarray(b, np.uint32)
    layer = np.zeros(max(len(a), 1), np.uint32)
    order = np.zeros(max(len(a), 1), np.uint64)
    d = check(lib().bd_plan_layers(_ptr(a), _ptr(b), len(a), int(n_points), _ptr(layer), _ptr(order)), "plan_layers")
    return layer[:len(a)], order[:len(a)], d


class Batch:
    """Many independent scenes advanced together (bd_batch_*)."""

    def __init__(self):
        self._h = lib().bd_batch_new()
        if not self._h:
            raise MemoryError("bd_batch_new failed")

    def close(self):
        if getattr(self, "_h", None):
            lib().bd_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return lib().bd_batch_len(self._h)

    def total_points(self):
        return lib().bd_batch_total_points(self._h)

    def add_scene(self, solver):
        return check(lib().bd_batch_add_scene(self._h, solver._h), "batch_add_scene")

    def update(self, dt, steps=1):
        check(lib().bd_batch_update(self._h, float(dt), int(steps)), "batch_update")

    def sync(self):
        check(lib().bd_batch_sync(self._h), "batch_sync")

    def read_scene(self, i, solver):
        check(lib().bd_batch_read_scene(self._h, int(i), solver._h), "batch_read_scene")

    @property
    def config(self):
        c = BatchConfig()
        check(lib().bd_batch_get_config(self._h, ctypes.byref(c)), "batch_get_config")
        return {k: getattr(c, k) for k, _ in BatchConfig._fields_}

    @config.setter
    def config(self, d):
        c = BatchConfig()
        check(lib().bd_batch_get_config(self._h, ctypes.byref(c)), "batch_get_config")
        for k, v in d.items():
            setattr(c, k, int(v))
        check(lib().bd_batch_set_config(self._h, ctypes.byref(c)), "batch_set_config")

    def autotune(self, dt, steps=2):
        return check(lib().bd_batch_autotune(self._h, float(dt), int(steps)), "batch_autotune")

    def last_timing(self):
        ms = ctypes.c_float()
        n = ctypes.c_uint32()
        check(lib().bd_batch_last_timing(self._h, ctypes.byref(ms), ctypes.byref(n)), "batch_last_timing")
        return ms.value, n.value

    def scene_status(self, i):
        f = ctypes.c_float()
        u = [ctypes.c_uint32() for _ in range(5)]
        check(lib().bd_batch_scene_status(self._h, int(i), ctypes.byref(f), *[ctypes.byref(x) for x in u]),
              "batch_scene_status")
        return dict(zip(("skin", "err", "rounds", "redos", "rebuilds", "pairs"), [f.value] + [x.value for x in u]))

    def scene_times(self, i):
        v = [ctypes.c_double() for _ in range(4)]
        check(lib().bd_batch_scene_times(self._h, int(i), *[ctypes.byref(x) for x in v]), "batch_scene_times")
        return dict(zip(("links", "build", "sweep", "bodies"), [x.value for x in v]))

    def scene_lists(self, i):
        u = [ctypes.c_uint32() for _ in range(3)]
        f = [ctypes.c_float() for _ in range(2)]
        d = [ctypes.c_double() for _ in range(3)]
        k = [ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()]
        check(lib().bd_batch_scene_lists(self._h, int(i), *[ctypes.byref(x) for x in u + f + d + k]),
              "batch_scene_lists")
        names = ("n_out", "n_far", "max_list", "s_cell", "smax", "t_count", "t_lists", "t_sort", "kept", "unchecked",
                 "fallbacks")
        return dict(zip(names, [x.value for x in u + f + d + k]))

    def cc_stats(self):
        v = [ctypes.c_uint64() for _ in range(6)]
        check(lib().bd_batch_cc_stats(self._h, *[ctypes.byref(x) for x in v[:4]]), "batch_cc_stats")
        check(lib().bd_batch_cc_fallbacks(self._h, *[ctypes.byref(x) for x in v[4:]]), "batch_cc_fallbacks")
        return dict(zip(("rounds", "redos", "rebuilds", "err_scenes", "fallbacks", "fallback_scenes"),
                        [x.value for x in v]))


class Server:
    """Continuous-batching server of scene-advance requests (bd_server_*)."""

    def __init__(self, memory_budget=64 << 30, steps_per_launch=8, max_active=0, launch=None, launch_budget_us=0,
                 heavy_update_us=0):
        cfg = ServerConfig()
        cfg.memory_budget = int(memory_budget)
        cfg.steps_per_launch = int(steps_per_launch)
        cfg.max_active = int(max_active)
        cfg.launch_budget_us = int(launch_budget_us)
        cfg.heavy_update_us = int(heavy_update_us)
        if launch:
            for k, v in launch.items():
                setattr(cfg.launch, k, int(v))
        self._h = lib().bd_server_new(ctypes.byref(cfg))
        if not self._h:
            raise RuntimeError("bd_server_new failed: " + lib().bd_last_error().decode())
        self._buf = (RequestDone * 4096)()

    def close(self):
        if getattr(self, "_h", None):
            lib().bd_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self):
        return lib().bd_server_now(self._h)

    def create(self, solver, steps, dt):
        sid = ctypes.c_int64()
        rid = check(lib().bd_server_submit_create(self._h, solver._h, int(steps), float(dt), ctypes.byref(sid)),
                    "server_submit_create")
        return rid, sid.value

    def advance(self, scene_id, steps, dt):
        return check(lib().bd_server_submit_continue(self._h, int(scene_id), int(steps), float(dt)),
                     "server_submit_continue")

    def poll(self):
        out = []
        while True:
            n = check(lib().bd_server_poll(self._h, self._buf, len(self._buf)), "server_poll")
            out += [{k: getattr(self._buf[i], k) for k, _ in RequestDone._fields_} for i in range(n)]
            if n < len(self._buf):
                return out

    def wait_idle(self, timeout=60.0):
        """True once no request is pending; False if `timeout` seconds pass first."""
        rc = lib().bd_server_wait_idle(self._h, float(timeout))
        if rc == BD_ERR_STATE:
            return False
        check(rc, "server_wait_idle")
        return True

    def read_scene(self, scene_id, solver):
        check(lib().bd_server_read_scene(self._h, int(scene_id), solver._h), "server_read_scene")

    def stats(self):
        s = ServerStats()
        check(lib().bd_server_get_stats(self._h, ctypes.byref(s)), "server_get_stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}


class SceneImage:
    """One scene in caller-owned device memory (bd_scene_image_*): `info`
    describes its sizes and launch shape, `host` holds the image bytes to copy
    to a 256-byte-aligned device buffer.  update() queues the kernels on a
    caller stream and does not synchronise or allocate."""

    def __init__(self, solver):
        self.info = _lib.SceneImageInfo()
        check(lib().bd_scene_image_describe(solver._h, ctypes.byref(self.info)), "bd_scene_image_describe")
        self.host = np.zeros(self.info.image_bytes, np.uint8)
        check(lib().bd_scene_image_write(solver._h, _ptr(self.host), self.host.nbytes), "bd_scene_image_write")

    @property
    def workspace_bytes(self):
        return self.info.workspace_bytes

    def update(self, image_ptr, workspace_ptr, workspace_bytes, dt, steps, stream=None):
        check(lib().bd_scene_image_update(ctypes.byref(self.info), image_ptr, workspace_ptr, int(workspace_bytes),
                                          float(dt), int(steps), stream), "bd_scene_image_update")

    @staticmethod
    def read(solver, host_bytes):
        """Positions from a host copy of an image into `solver` (holding the same scene)."""
        check(lib().bd_scene_image_read(solver._h, _ptr(host_bytes), host_bytes.nbytes), "bd_scene_image_read")
