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
