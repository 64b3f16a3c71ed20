"""Test helpers: mirror a product Solver's initial state into the oracle and
compare results."""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from bendy2d_amd import Solver  # noqa: E402
from bendy2d_amd._lib import Vec2, lib  # noqa: E402
from oracle import OracleSolver  # noqa: E402


def gpu_available():
    try:
        return lib().bd_device_count() > 0
    except Exception:
        return False


def to_oracle(s: Solver, bits=32, candidate_mode=False):
    """Build an oracle solver holding exactly the state of product solver s."""
    o = OracleSolver(bits)
    o.set_gravity(s.gravity)
    pos, size = s.bounds
    o.set_bounds(pos, size)
    o.set_sub_steps(lib().bd_ext_get_sub_steps(s._h))
    o.set_link_iterations(lib().bd_ext_get_link_iterations(s._h))
    if candidate_mode:
        o.set_candidate_mode(True)
    P = s.get_particles()
    if len(P):
        o.add_particles(P[:, 0:2], P[:, 2:4])
        pins = np.zeros(len(P), np.uint8)
        lib().bd_ext_get_pinned_range(s._h, 0, pins.ctypes.data_as(ctypes.c_void_p), len(P))
        for i in np.nonzero(pins)[0]:
            o.pin(int(i))
    a, b, t = s.get_particle_links()
    if len(a):
        o.add_particle_links(a, b, t)
    C = s.get_circles()
    if len(C):
        o.add_circles(C)
    a, b, t = s.get_circle_links()
    if len(a):
        o.add_circle_links(a, b, t)
    for i in range(s.get_polygons_len()):
        g = s.get_polygon(i)
        o.add_polygon(g["particles"], g["links"], g["is_static"], g["center"], g["scale"])
    ncol = lib().bd_ext_get_static_collider_len(s._h)
    for i in range(ncol):
        c = Vec2()
        r = ctypes.c_float()
        nv = ctypes.c_uint64()
        verts = np.zeros((64, 2), np.float32)
        kind = lib().bd_ext_get_static_collider(s._h, i, ctypes.byref(c), ctypes.byref(r),
                                                verts.ctypes.data_as(ctypes.c_void_p), 64, ctypes.byref(nv))
        if kind == 0:
            o.add_static_circle((c.x, c.y), r.value)
        else:
            assert nv.value <= 64
            o.add_static_polygon(verts[:nv.value])
    return o


def state(s):
    """All body state of a product or oracle solver as one float64 array."""
    parts = [s.get_particles()[:, :4].astype(np.float64).ravel(),
             s.get_circles()[:, :4].astype(np.float64).ravel()]
    n = s.get_polygons_len()
    for i in range(n):
        g = s.get_polygon(i)
        parts.append(np.asarray(g["particles"], np.float64)[:, :4].ravel())
        parts.append(np.asarray(g["center"], np.float64))
    return np.concatenate(parts) if parts else np.zeros(0)


def _canon(x):
    """f32 bit patterns with every NaN mapped to one pattern.  IEEE 754 leaves
    the sign and payload of a NaN produced by an invalid operation open (x86
    SSE yields 0xFFC00000, gfx950 0x7FC00000); NaN-ness itself must match."""
    a = np.array(np.asarray(x, np.float32), copy=True).view(np.uint32)
    a[np.isnan(a.view(np.float32))] = 0x7FC00000
    return a


def bits_equal(x, y):
    """Exact f32 equality (+0 != -0); all NaNs compare equal to each other."""
    xa, ya = _canon(x), _canon(y)
    return xa.shape == ya.shape and bool(np.all(xa == ya))


def first_diff(x, y):
    xa, ya = _canon(x), _canon(y)
    d = np.nonzero(xa != ya)[0]
    if len(d) == 0:
        return None
    i = int(d[0])
    return i, float(np.asarray(x)[i]), float(np.asarray(y)[i]), len(d)


def state_acc(s):
    """state(s) plus the acc of every circle and polygon point (static polygon
    points keep adding gravity; everything else is spent by the update)."""
    parts = [state(s), s.get_circles()[:, 4:6].astype(np.float64).ravel()]
    for i in range(s.get_polygons_len()):
        parts.append(np.asarray(s.get_polygon(i)["particles"], np.float64)[:, 4:6].ravel())
    return np.concatenate(parts)


def polygon_scene(seed, n_polys=6, circles=True, cloth=True, sub_steps=2, link_iters=2):
    """Soft polygons (3 to 19 points, some with an initial acc) falling onto a
    static polygon floor and onto each other, with a static polygon collider,
    a static circle and, optionally, free circles and a pinned cloth."""
    from bendy2d_amd.solver import polygon_circle

    rng = np.random.default_rng(seed)
    s = Solver()
    s.gravity = (0.0, 60.0)
    s.bounds = ((0.0, 0.0), (100.0, 100.0))
    for k in range(n_polys):
        n = int(rng.integers(3, 20))
        pts, links, c = polygon_circle(float(rng.uniform(3.0, 7.0)),
                                       (float(rng.uniform(15.0, 85.0)), float(rng.uniform(10.0, 70.0))), n, False)
        if k % 3 == 1:
            pts[:, 4:6] = rng.normal(0.0, 30.0, (n, 2)).astype(np.float32)
        s.add_polygon(pts, links, False, c, 1.0)
    s.add_polygon_new([(5.0, 88.0), (95.0, 88.0), (95.0, 96.0), (5.0, 96.0)], True)
    s.add_polygon_circle(4.0, (50.0, 80.0), 9, True)
    if circles:
        for k in range(24):
            s.add_circle((float(rng.uniform(10.0, 90.0)), float(rng.uniform(5.0, 60.0))),
                         float(rng.uniform(0.8, 1.6)))
        s.add_circle_link(0, 1, 3.0)
    if cloth:
        from bendy2d_amd import workloads
        workloads.add_grid(s, 6, 5, x0=70.0, y0=20.0, pin_top=True)
    s.add_static_circle((30.0, 75.0), 5.0)
    s.add_static_polygon([(60.0, 60.0), (70.0, 60.0), (65.0, 68.0)])
    s.set_sub_steps(sub_steps)
    s.set_link_iterations(link_iters)
    return s
