"""Known-answer tests that pin the f32 oracle without the reference crate.

The crate ships no tests, fixtures or golden vectors, and cannot be built here
(no rustc, no nalgebra source), so each oracle function is pinned by what the
mathematics fixes: exact closed forms chosen so f32 arithmetic is exact,
invariants (midpoint, r^2-weighted centroid, reflected velocity), textbook
cases of segment intersection, and brute-force agreement between two
independent implementations (candidate-list sweep vs the reference's all-pairs
loop).  A plausible slip (a dropped term, a swapped index or sign, acc*dt
instead of acc*dt*dt, a reciprocal instead of a division) fails at least one.
"""
import math

import numpy as np
import pytest

from helpers import bits_equal
from oracle import OracleSolver

BIG = ((-1e6, -1e6), (2e6, 2e6))


def solver(bits=32, g=(0.0, 0.0), bounds=BIG):
    o = OracleSolver(bits)
    o.set_gravity(g)
    o.set_bounds(*bounds)
    return o


def test_defaults_match_solver_new():
    # Solver::new (src/solver.rs:34-50): gravity (0, 98.2), bounds 100x100.
    o = OracleSolver(32)
    o.add_particle((50.0, 50.0))
    o.update(1.0)
    p = o.get_particles()[0]
    # acc = 98.2 (f32), one step: pos = pos + 0 + acc*1*1
    assert p[1] == np.float32(50.0) + np.float32(98.2) or p[1] == 100.0  # then clamped below
    # the particle crossed the lower-right bound (y > 100) only after integration,
    # so it is not clamped in this update (bounds run before integrate)
    assert p[1] == np.float32(np.float32(50.0) + np.float32(98.2))
    assert p[3] == 50.0


def test_single_link_known_answer():
    # one link (0,0)-(2,0), target 1: each end moves 0.5 inward, exactly
    o = solver()
    o.add_particle((0.0, 0.0))
    o.add_particle((2.0, 0.0))
    o.add_particle_link(0, 1, 1.0)
    o.apply_particle_links_in_order([0])
    p = o.get_particles()
    assert p[0, 0] == 0.5 and p[1, 0] == 1.5 and p[0, 1] == 0 and p[1, 1] == 0


def test_link_preserves_midpoint_and_reaches_target():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = rng.uniform(-10, 10, 2), rng.uniform(-10, 10, 2)
        t = float(rng.uniform(0.1, 5))
        o = solver(64)
        o.add_particle(a)
        o.add_particle(b)
        o.add_particle_link(0, 1, t)
        o.apply_particle_links_in_order([0])
        p = o.get_particles()
        mid0 = (np.float32(a).astype(np.float64) + np.float32(b)) / 2
        assert np.allclose((p[0, :2] + p[1, :2]) / 2, mid0, atol=1e-12)
        assert abs(np.linalg.norm(p[0, :2] - p[1, :2]) - np.float32(t)) < 1e-9


def test_pinned_link_moves_only_free_end():
    o = solver()
    o.add_particle((0.0, 0.0))
    o.add_particle((3.0, 4.0))
    o.pin(0)
    o.add_particle_link(0, 1, 1.0)
    o.apply_particle_links_in_order([0])
    p = o.get_particles()
    assert p[0, 0] == 0 and p[0, 1] == 0
    # free end lands on the target circle around the pinned end: (0.6, 0.8)
    assert abs(p[1, 0] - 0.6) < 1e-6 and abs(p[1, 1] - 0.8) < 1e-6


def test_free_fall_exact_closed_form():
    # g = -8, dt = 1/16 (1 substep): acc*dt*dt = -1/32 exactly, so
    # y_n = y_0 - n(n+1)/64 exactly in f32 (Verlet from rest, src/particle.rs:20-25)
    o = solver(g=(0.0, -8.0))
    o.add_particle((0.0, 0.0))
    for n in range(1, 101):
        o.update(1.0 / 16.0)
        assert o.get_particles()[0, 1] == -n * (n + 1) / 64.0
    # a term dropped from pos + vel + acc*dt*dt, or acc*dt instead of acc*dt*dt,
    # breaks this sequence at n = 1 or 2


def test_substeps_split_dt():
    # EXT1: delta = dt * (1/n) (src/solver.rs:107-108); 4 substeps of 1/16 under
    # g = -8 from rest: y = -(1/256)*8 * (1+2+3+4) = -10/32... compute exactly
    o = solver(g=(0.0, -8.0))
    o.set_sub_steps(4)
    o.add_particle((0.0, 0.0))
    o.update(0.25)
    # each substep: acc*dt*dt = -8/256 = -1/32; y_4 = -(4*5/2)/32 = -10/32
    assert o.get_particles()[0, 1] == -10.0 / 32.0


def test_bounds_reflect_velocity_exactly():
    # particle moving left through x = 0 (src/particle.rs:27-46)
    o = solver(bounds=((0.0, 0.0), (100.0, 100.0)))
    o.add_particle((-1.0, 50.0), prev=(-0.5, 50.0))
    o.update(1.0)
    p = o.get_particles()[0]
    # bounds: v = prev - pos = 0.5 -> prev = 0 - 0.5, pos = 0; then integrate:
    # vel = pos - prev = 0.5, prev = 0, pos = 0.5
    assert (p[0], p[2]) == (0.5, 0.0)


def test_circle_bounds_use_radius():
    o = solver(bounds=((0.0, 0.0), (10.0, 10.0)))
    o.add_circle((9.5, 5.0), 1.0, prev_pos=(9.5, 5.0))
    o.update(1.0)
    c = o.get_circles()[0]
    # hi - r = 9 -> pos 9, prev = 9 - (9.5 - 9.5 ... ) = 9 - 0 -> then integrate
    # v = prev - pos = 0 at clamp time: prev = 9, pos = 9; Verlet keeps 9
    assert (c[0], c[2]) == (9.0, 9.0)


def test_circle_pair_known_answer_and_centroid():
    # equal radii 1 at distance 1.5: overlap 0.5, each moves 0.25 (src/circle.rs:32-45)
    o = solver()
    o.add_circle((0.0, 0.0), 1.0)
    o.add_circle((1.5, 0.0), 1.0)
    o.update(0.0)
    c = o.get_circles()
    # after update, prev_pos holds the positions collisions left (integrate runs last)
    assert c[0, 2] == -0.25 and c[1, 2] == 1.75
    # unequal radii: r^2-weighted centroid preserved, final distance r1 + r2
    rng = np.random.default_rng(4)
    for _ in range(100):
        r1, r2 = rng.uniform(0.5, 3, 2)
        p1 = rng.uniform(-5, 5, 2)
        p2 = p1 + rng.uniform(-1, 1, 2)
        o = solver(64)
        o.add_circle(p1, r1)
        o.add_circle(p2, r2)
        o.update(0.0)
        c = o.get_circles()
        w1, w2 = np.float32(r1) ** 2, np.float32(r2) ** 2
        before = (w1 * np.float32(p1).astype(np.float64) + w2 * np.float32(p2)) / (w1 + w2)
        after = (w1 * c[0, 2:4] + w2 * c[1, 2:4]) / (w1 + w2)
        assert np.allclose(before, after, atol=1e-9)
        assert abs(np.linalg.norm(c[0, 2:4] - c[1, 2:4]) - (float(np.float32(r1)) + float(np.float32(r2)))) < 1e-9


def test_circle_link_weights():
    # CircleLink (src/link.rs:36-48): a moves by rb^2/(ra^2+rb^2) of the error
    o = solver()
    o.add_circle((0.0, 0.0), 1.0)
    o.add_circle((5.0, 0.0), 3.0)
    o.add_circle_link(0, 1, 4.0)
    o.update(0.0)
    c = o.get_circles()
    # error 1; a moves +0.9, b moves -0.1 (weights 9/10, 1/10)
    assert abs(c[0, 2] - 0.9) < 1e-6 and abs(c[1, 2] - 4.9) < 1e-6


def test_static_circle_collider():
    o = solver()
    o.add_static_circle((0.0, 0.0), 2.0)
    o.add_particle((1.0, 0.0))
    o.update(0.0)
    assert o.get_particles()[0, 2] == 2.0


def test_static_polygon_collider_min_penetration():
    # unit square [0,4]^2; a point at (3.5, 1) leaves through x = 4 (least depth)
    o = solver()
    o.add_static_polygon([(0, 0), (4, 0), (4, 4), (0, 4)])
    o.add_particle((3.5, 1.0))
    o.add_circle((0.2, 2.0), 0.5)
    o.update(0.0)
    assert tuple(o.get_particles()[0, 2:4]) == (4.0, 1.0)
    assert tuple(o.get_circles()[0, 2:4]) == (-0.5, 2.0)
    # clockwise input gives the same planes
    o2 = solver()
    o2.add_static_polygon([(0, 4), (4, 4), (4, 0), (0, 0)])
    o2.add_particle((3.5, 1.0))
    o2.update(0.0)
    assert tuple(o2.get_particles()[0, 2:4]) == (4.0, 1.0)


def test_polygon_edge_push_known_answer():
    # A point of another polygon inside edge (0,0)-(2,0) of a square, with the
    # other polygon's centre below the edge: the segment point->centre crosses
    # the edge at (1,0) (s = 1/3, t = 1/2), the edge ends move 1/6 inward along
    # the inward normal (0,1) and the point is projected onto the edge at (1,0)
    # (src/polygon.rs:164-216, src/common.rs:4-26).
    o = solver()
    sq = np.array([[0, 0, 0, 0, 0, 0], [2, 0, 2, 0, 0, 0], [2, 2, 2, 2, 0, 0], [0, 2, 0, 2, 0, 0]], np.float32)
    o.add_polygon(sq, ([0, 1, 2, 0], [1, 2, 3, 3], [2, 2, 2, 2]), True, (1.0, 1.0))
    tri = np.array([[1, 0.5, 1, 0.5, 0, 0], [1.3, -3.25, 1.3, -3.25, 0, 0], [0.7, -3.25, 0.7, -3.25, 0, 0]],
                   np.float32)
    # links with their current lengths so link solving does not move anything
    def d(i, j):
        return float(np.float32(np.hypot(*(tri[i, :2] - tri[j, :2]))))
    o.add_polygon(tri, ([0, 1, 0], [1, 2, 2], [d(0, 1), d(1, 2), d(0, 2)]), True, (1.0, -2.0))
    o.update(0.0)
    sqo = o.get_polygon(0)["particles"]
    tro = o.get_polygon(1)["particles"]
    assert abs(sqo[0, 1] - 1 / 6) < 1e-6 and abs(sqo[1, 1] - 1 / 6) < 1e-6
    assert sqo[0, 0] == 0 and sqo[1, 0] == 2
    assert abs(tro[0, 0] - 1.0) < 1e-6 and abs(tro[0, 1]) < 1e-6


def test_polygon_ctor_circle_geometry():
    # Polygon::circle (src/polygon.rs:17-82): n points on the circle, links to
    # i+2n/3 and i+n/3 with their current lengths, a < b
    o = solver()
    o.add_polygon_circle(5.0, (10.0, 20.0), 9, False)
    g = o.get_polygon(0)
    pts = g["particles"][:, :2]
    assert np.allclose(np.hypot(pts[:, 0] - 10, pts[:, 1] - 20), 5.0, atol=1e-5)
    assert abs(g["center"][0] - 10) < 1e-5 and abs(g["center"][1] - 20) < 1e-5


def test_candidate_sweep_equals_all_pairs():
    # two independent implementations of A7's lexicographic sweep must agree
    # bit for bit (the candidate list is how large scenes are checked)
    for seed in range(3):
        rng = np.random.default_rng(seed)
        a, b = solver(g=(0.0, -9.81), bounds=((0, 0), (30, 30))), None
        b = solver(g=(0.0, -9.81), bounds=((0, 0), (30, 30)))
        b.set_candidate_mode(True, 0.05)
        for _ in range(300):
            p = rng.uniform(1, 29, 2)
            r = float(rng.uniform(0.3, 0.9))
            v = rng.normal(0, 0.05, 2)
            a.add_circle(p, r, prev_pos=p - v)
            b.add_circle(p, r, prev_pos=p - v)
        for _ in range(40):
            a.update(1 / 60)
            b.update(1 / 60)
        assert bits_equal(a.get_circles(), b.get_circles())
        assert b.cc_redos() > 0  # the margin check did trigger redos


def test_order_of_link_sweep_matters():
    # sequential Gauss-Seidel: a chain solved forward differs from backward,
    # so a schedule that reorders dependent links would be caught
    def run(order):
        o = solver()
        for i in range(4):
            o.add_particle((1.5 * i, 0.3 * (i % 2)))
        for i in range(3):
            o.add_particle_link(i, i + 1, 1.0)
        o.apply_particle_links_in_order(order)
        return o.get_particles()
    assert not bits_equal(run([0, 1, 2]), run([2, 1, 0]))


def test_f64_twin_close_to_f32():
    o32, o64 = solver(32, g=(0, -9.81)), solver(64, g=(0, -9.81))
    for o in (o32, o64):
        o.set_sub_steps(4)
        for i in range(10):
            o.add_particle((i * 1.0, 0.0))
        for i in range(9):
            o.add_particle_link(i, i + 1, 1.0)
        o.pin(0)
        for _ in range(50):
            o.update(1 / 60)
    d = np.abs(o32.get_particles() - o64.get_particles()).max()
    assert 0 < d < 1e-3
