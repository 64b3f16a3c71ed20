"""The benchmark workloads, built through the C ABI's scene generators.

Each builder fills a Solver and returns (dt, steps): the update step and the
number of updates the workload runs.  Sizes can be reduced for tests; the
defaults are the full workloads.

  cloth_pinned   32x32 particle cloth, top row pinned, one static circle R=8
                 20 below the centre, g=(0,-9.81), dt=1/60, 8 substeps,
                 8 link sweeps, 1000 steps
  cloth_large    1024x1024 particle cloth, jitter U(-0.01,0.01) (seed 1), top
                 row pinned, 4 substeps, 8 link sweeps, no colliders, 500 steps
  gas_box        262144 circles r=0.5 in a 1024 box walled by 4 static
                 polygons, 64 static circles R~U(4,16), v~N(0,1), 4 substeps,
                 300 steps (seed 2)
  softbody(seed) one scene of the 4096-scene batched workload
"""
import ctypes
import math

from . import _lib
from ._lib import GridDesc, check, lib

HUGE_BOUNDS = ((-1.0e5, -1.0e5), (2.0e5, 2.0e5))


def add_grid(solver, nx, ny, spacing=1.0, x0=0.0, y0=0.0, jitter=0.0, seed=0, order=_lib.BD_ORDER_COLOUR,
             pin_top=False, as_circles=False, radius=0.0):
    g = GridDesc(int(nx), int(ny), float(spacing), float(x0), float(y0), float(jitter), int(seed), int(order),
                 int(bool(pin_top)), int(bool(as_circles)), float(radius))
    return check(lib().bd_scene_add_grid(solver._h, ctypes.byref(g)), "bd_scene_add_grid")


def cloth_pinned(solver, n=32, order=_lib.BD_ORDER_COLOUR, sub_steps=8, link_iters=8, steps=1000):
    """32x32 pinned cloth over a static circle (the reference's CPU-sized case)."""
    add_grid(solver, n, n, order=order, pin_top=True)
    solver.gravity = (0.0, -9.81)
    solver.bounds = HUGE_BOUNDS
    solver.set_sub_steps(sub_steps)
    solver.set_link_iterations(link_iters)
    c = 0.5 * (n - 1)
    solver.add_static_circle((c, -c - 20.0), 8.0)
    return 1.0 / 60.0, steps


def cloth_large(solver, n=1024, order=_lib.BD_ORDER_COLOUR, sub_steps=4, link_iters=8, steps=500, seed=1):
    """1024x1024 jittered pinned cloth (throughput scale, no colliders)."""
    add_grid(solver, n, n, jitter=0.01, seed=seed, order=order, pin_top=True)
    solver.gravity = (0.0, -9.81)
    solver.bounds = HUGE_BOUNDS
    solver.set_sub_steps(sub_steps)
    solver.set_link_iterations(link_iters)
    return 1.0 / 60.0, steps


def gas_box(solver, n=262144, radius=0.5, box=1024.0, n_static=64, seed=2, steps=300):
    """Collision-heavy circles in a walled box."""
    check(lib().bd_scene_gas_box(solver._h, int(n), float(radius), float(box), int(n_static), int(seed)),
          "bd_scene_gas_box")
    return 1.0 / 60.0, steps


def softbody(solver, seed, max_side=0, steps=100):
    """One scene of the batched workload; returns (dt, steps, n_circles)."""
    n = check(lib().bd_scene_softbody(solver._h, int(seed), int(max_side)), "bd_scene_softbody")
    return 1.0 / 60.0, steps, n


def softbody_side(seed):
    """Side of softbody(seed) before any cap: round(4 * 32^u) with the scene's first draw."""
    s = (int(seed) + 0x9e3779b97f4a7c15) & 0xFFFFFFFFFFFFFFFF
    z = s
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & 0xFFFFFFFFFFFFFFFF
    z = z ^ (z >> 31)
    u = (z >> 40) / 16777216.0
    return max(4, min(128, int(round(4.0 * math.pow(32.0, u)))))
