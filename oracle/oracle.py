"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product library never
imports it.

OracleSolver(32) is the bit-exact f32 restatement of Murrent/bendy2d 0.0.1;
OracleSolver(64) is its FP64 twin (error reporting only).
"""
import ctypes
import os
import subprocess
from ctypes import c_double, c_float, c_int, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_library(path=None):
    """Load the oracle from `path` for the calls that follow (None: the built
    liboracle.so).  The mutation test points the known-answer tests at copies
    of the oracle with a planted slip this way."""
    global _lib, _lib_path
    _lib = None
    _lib_path = path


_lib_path = None


def lib():
    global _lib
    if _lib is None:
        path = _lib_path or LIB_PATH
        if path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        for pfx in ("or32_", "or64_"):
            sig = {
                "new": (c_void_p, []), "free": (None, [c_void_p]), "clone": (c_void_p, [c_void_p]),
                "set_gravity": (None, [c_void_p, c_float, c_float]),
                "set_bounds": (None, [c_void_p, c_float, c_float, c_float, c_float]),
                "set_bounds_active": (None, [c_void_p, c_int]),
                "set_sub_steps": (c_int, [c_void_p, c_int]),
                "set_link_iterations": (c_int, [c_void_p, c_int]),
                "set_cc_mode": (None, [c_void_p, c_int, c_float]),
                "cc_redos": (c_int64, [c_void_p]),
                "add_particle": (c_int64, [c_void_p, c_float, c_float]),
                "add_particle_ex": (c_int64, [c_void_p, c_float, c_float, c_float, c_float]),
                "pin": (c_int, [c_void_p, c_int64, c_int]),
                "add_particle_link": (c_int64, [c_void_p, c_int64, c_int64, c_float]),
                "add_circle": (c_int64, [c_void_p] + [c_float] * 7),
                "add_circle_link": (c_int64, [c_void_p, c_int64, c_int64, c_float]),
                "add_polygon": (c_int64, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                          c_void_p, c_void_p, c_int, c_float, c_float, c_float]),
                "add_polygon_circle": (c_int64, [c_void_p, c_float, c_float, c_float, c_int64, c_int]),
                "add_polygon_new": (c_int64, [c_void_p, c_int64, c_void_p, c_int]),
                "add_static_circle": (c_int64, [c_void_p, c_float, c_float, c_float]),
                "add_static_polygon": (c_int64, [c_void_p, c_int64, c_void_p]),
                "validate": (c_int, [c_void_p]),
                "update": (c_int, [c_void_p, c_float]),
                "apply_particle_links_in_order": (c_int, [c_void_p, c_void_p, c_int64]),
                "get_particle_len": (c_int64, [c_void_p]),
                "get_circles_len": (c_int64, [c_void_p]),
                "get_polygons_len": (c_int64, [c_void_p]),
                "get_polygon_point_len": (c_int64, [c_void_p, c_int64]),
                "get_particles": (None, [c_void_p, c_void_p]),
                "get_circles": (None, [c_void_p, c_void_p]),
                "get_polygon": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
                "get_polygon_links": (c_int64, [c_void_p, c_int64, c_void_p, c_void_p]),
            }
            c_real = c_float if pfx == "or32_" else c_double
            sig.update({
                "law_particle_link": (None, [c_void_p, c_real, c_int, c_int]),
                "law_circle_link": (None, [c_void_p, c_real, c_real, c_real]),
                "law_circle_pair": (c_int, [c_void_p, c_real, c_real]),
                "law_bounds": (None, [c_void_p, c_void_p, c_int, c_real]),
                "law_particle_update": (None, [c_void_p, c_real]),
                "law_line_intersection": (c_int, [c_void_p, c_void_p]),
                "law_resolve_line": (c_int, [c_void_p, c_void_p]),
                "all_pairs_rows": (c_int64, [c_void_p, c_int64, c_int64]),
            })
            for name, (res, args) in sig.items():
                f = getattr(L, pfx + name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


class OracleSolver:
    """The reference Solver restated on the CPU (f32: bits == reference; f64: twin)."""

    def __init__(self, bits=32, _h=None):
        self.bits = bits
        self.pfx = "or32_" if bits == 32 else "or64_"
        self._L = lib()
        self._h = _h if _h is not None else self._f("new")()

    def _f(self, name):
        return getattr(self._L, self.pfx + name)

    def close(self):
        if self._h:
            self._f("free")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clone(self):
        return OracleSolver(self.bits, self._f("clone")(self._h))

    def set_gravity(self, g):
        self._f("set_gravity")(self._h, float(g[0]), float(g[1]))

    def set_bounds(self, pos, size):
        self._f("set_bounds")(self._h, float(pos[0]), float(pos[1]), float(size[0]), float(size[1]))

    def set_sub_steps(self, n):
        assert self._f("set_sub_steps")(self._h, int(n)) == 0

    def set_link_iterations(self, k):
        assert self._f("set_link_iterations")(self._h, int(k)) == 0

    def set_candidate_mode(self, on=True, margin=0.25):
        """Circle pairs by the candidate-list sweep (equal results, for large scenes)."""
        self._f("set_cc_mode")(self._h, 1 if on else 0, float(margin))

    def cc_redos(self):
        return self._f("cc_redos")(self._h)

    def all_pairs_rows(self, first, last):
        """Run rows [first, last) of the reference's all-pairs circle loop on
        this solver's circles (moving them); returns the pair tests made."""
        return self._f("all_pairs_rows")(self._h, int(first), int(last))

    def add_particle(self, pos, prev=None):
        if prev is None:
            return self._f("add_particle")(self._h, float(pos[0]), float(pos[1]))
        return self._f("add_particle_ex")(self._h, float(pos[0]), float(pos[1]), float(prev[0]), float(prev[1]))

    def add_particles(self, pos, prev=None):
        pos = np.asarray(pos, np.float32).reshape(-1, 2)
        prev = pos if prev is None else np.asarray(prev, np.float32).reshape(-1, 2)
        f = self._f("add_particle_ex")
        for (x, y), (px, py) in zip(pos.tolist(), prev.tolist()):
            f(self._h, x, y, px, py)

    def pin(self, i, flag=True):
        assert self._f("pin")(self._h, int(i), int(bool(flag))) == 0

    def add_particle_link(self, a, b, t):
        return self._f("add_particle_link")(self._h, int(a), int(b), float(t))

    def add_particle_links(self, a, b, t):
        f = self._f("add_particle_link")
        for x, y, z in zip(np.asarray(a).tolist(), np.asarray(b).tolist(), np.asarray(t, np.float32).tolist()):
            f(self._h, x, y, z)

    def add_circle(self, pos, radius, prev_pos=None, acc=(0.0, 0.0)):
        prev_pos = pos if prev_pos is None else prev_pos
        return self._f("add_circle")(self._h, float(pos[0]), float(pos[1]), float(prev_pos[0]), float(prev_pos[1]),
                                     float(acc[0]), float(acc[1]), float(radius))

    def add_circles(self, rows):
        f = self._f("add_circle")
        for r in np.asarray(rows, np.float32).reshape(-1, 7).tolist():
            f(self._h, *r)

    def add_circle_link(self, a, b, t):
        return self._f("add_circle_link")(self._h, int(a), int(b), float(t))

    def add_circle_links(self, a, b, t):
        f = self._f("add_circle_link")
        for x, y, z in zip(np.asarray(a).tolist(), np.asarray(b).tolist(), np.asarray(t, np.float32).tolist()):
            f(self._h, x, y, z)

    def add_polygon(self, particles, links, is_static, center, scale=1.0):
        pts = np.ascontiguousarray(particles, np.float32).reshape(-1, 6)
        pos = np.ascontiguousarray(pts[:, 0:2]); prev = np.ascontiguousarray(pts[:, 2:4])
        acc = np.ascontiguousarray(pts[:, 4:6])
        a, b, t = links
        a = np.ascontiguousarray(a, np.int64); b = np.ascontiguousarray(b, np.int64)
        t = np.ascontiguousarray(t, np.float32)
        return self._f("add_polygon")(self._h, len(pts), _p(pos), _p(prev), _p(acc), len(a), _p(a), _p(b), _p(t),
                                      int(bool(is_static)), float(center[0]), float(center[1]), float(scale))

    def add_polygon_circle(self, radius, pos, n, is_static=False):
        return self._f("add_polygon_circle")(self._h, float(radius), float(pos[0]), float(pos[1]), int(n),
                                             int(bool(is_static)))

    def add_polygon_new(self, points, is_static=False):
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
        return self._f("add_polygon_new")(self._h, len(p), _p(p), int(bool(is_static)))

    def add_static_circle(self, c, r):
        return self._f("add_static_circle")(self._h, float(c[0]), float(c[1]), float(r))

    def add_static_polygon(self, verts):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1, 2)
        return self._f("add_static_polygon")(self._h, len(v), _p(v))

    def update(self, dt):
        return self._f("update")(self._h, float(dt))

    def apply_particle_links_in_order(self, order):
        o = np.ascontiguousarray(order, np.int64)
        return self._f("apply_particle_links_in_order")(self._h, _p(o), len(o))

    def get_particles(self):
        n = self._f("get_particle_len")(self._h)
        out = np.zeros((max(n, 1), 6), np.float64)
        self._f("get_particles")(self._h, _p(out))
        return out[:n]

    def get_circles(self):
        n = self._f("get_circles_len")(self._h)
        out = np.zeros((max(n, 1), 7), np.float64)
        self._f("get_circles")(self._h, _p(out))
        return out[:n]

    def get_polygons_len(self):
        return self._f("get_polygons_len")(self._h)

    def get_polygon(self, i):
        n = self._f("get_polygon_point_len")(self._h, int(i))
        out = np.zeros((max(n, 1), 6), np.float64)
        c = np.zeros(2, np.float64)
        self._f("get_polygon")(self._h, int(i), _p(out), _p(c))
        ab = np.zeros((max(4 * n, 2), 2), np.int64)
        t = np.zeros(max(4 * n, 1), np.float64)
        nl = self._f("get_polygon_links")(self._h, int(i), _p(ab), _p(t))
        return {"particles": out[:n], "center": (c[0], c[1]), "links": (ab[:nl, 0], ab[:nl, 1], t[:nl])}


# ---- single laws (the exact functions the solver step uses), for known-answer tests ----

def _real(bits):
    return np.float32 if bits == 32 else np.float64


def _pfx(bits):
    return "or32_" if bits == 32 else "or64_"


def law_particle_link(a, b, t, pinned_a=False, pinned_b=False, bits=32):
    """ParticleLink::solve on one link; returns the new (a, b)."""
    ab = np.array([a[0], a[1], b[0], b[1]], _real(bits))
    getattr(lib(), _pfx(bits) + "law_particle_link")(_p(ab), t, int(pinned_a), int(pinned_b))
    return ab[:2].copy(), ab[2:].copy()


def law_circle_link(a, b, ra, rb, t, bits=32):
    """CircleLink::solve on one link of circles of radii ra, rb; returns the new (a, b)."""
    ab = np.array([a[0], a[1], b[0], b[1]], _real(bits))
    getattr(lib(), _pfx(bits) + "law_circle_link")(_p(ab), ra, rb, t)
    return ab[:2].copy(), ab[2:].copy()


def law_circle_pair(a, b, ra, rb, bits=32):
    """Circle::solve_circle on one pair; returns (hit, new a, new b)."""
    ab = np.array([a[0], a[1], b[0], b[1]], _real(bits))
    hit = getattr(lib(), _pfx(bits) + "law_circle_pair")(_p(ab), ra, rb)
    return bool(hit), ab[:2].copy(), ab[2:].copy()


def law_bounds(pos, prev, bounds_pos, bounds_size, radius=None, bits=32):
    """Particle::solve_bounds (radius None) or Circle::solve_bounds; returns (pos, prev)."""
    pp = np.array([pos[0], pos[1], prev[0], prev[1]], _real(bits))
    bb = np.array([bounds_pos[0], bounds_pos[1], bounds_size[0], bounds_size[1]], _real(bits))
    getattr(lib(), _pfx(bits) + "law_bounds")(_p(pp), _p(bb), int(radius is not None),
                                               0.0 if radius is None else radius)
    return pp[:2].copy(), pp[2:].copy()


def law_particle_update(pos, prev, acc, dt, bits=32):
    """Particle::update (Verlet); returns (pos, prev, acc)."""
    p = np.array([pos[0], pos[1], prev[0], prev[1], acc[0], acc[1]], _real(bits))
    getattr(lib(), _pfx(bits) + "law_particle_update")(_p(p), dt)
    return p[:2].copy(), p[2:4].copy(), p[4:].copy()


def law_line_intersection(p1, p2, p3, p4, bits=32):
    """common::line_intersection of segments p1-p2 and p3-p4; the point or None."""
    p = np.array([p1[0], p1[1], p2[0], p2[1], p3[0], p3[1], p4[0], p4[1]], _real(bits))
    out = np.zeros(2, _real(bits))
    hit = getattr(lib(), _pfx(bits) + "law_line_intersection")(_p(p), _p(out))
    return out if hit else None


def law_resolve_line(center, a, b, o, other_center, bits=32):
    """Polygon::resolve_line_intersection of edge a-b (polygon centre `center`)
    against point o (its polygon's centre other_center); (new_a, new_b, new_o)
    or None."""
    v = np.array([*center, *a, *b, *o, *other_center], _real(bits))
    out = np.zeros(6, _real(bits))
    hit = getattr(lib(), _pfx(bits) + "law_resolve_line")(_p(v), _p(out))
    return (out[0:2], out[2:4], out[4:6]) if hit else None
