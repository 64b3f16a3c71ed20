"""The f32 oracle against a second restatement written from the Rust source.

tests/rust_model.py transcribes each expression of src/*.rs into binary32
operations with nalgebra's semantics, independently of oracle/oracle_impl.h.
Here both run on the same seeded inputs and must agree bit for bit, law by law
(with inputs spread so that a reassociation, a reciprocal in place of a
division, a sign or an index slip changes some result) and over whole small
scenes that use every phase and extension.  tests/test_oracle_mutations.py
plants such slips in copies of the oracle and checks that these tests, or the
known-answer tests, catch every one.
"""
import numpy as np
import pytest

import rust_model as M
from helpers import bits_equal, state
from oracle import (OracleSolver, law_bounds, law_circle_link, law_circle_pair, law_line_intersection,
                    law_particle_link, law_particle_update, law_resolve_line)

f32 = np.float32


@pytest.fixture(autouse=True)
def _quiet_fp():
    with np.errstate(all="ignore"):
        yield


def _eq(a, b):
    return bits_equal(np.asarray(a, np.float32).ravel(), np.asarray(b, np.float32).ravel())


def _scaled(rng, n, lo, hi):
    """n f32 values, uniform in [lo, hi) times a power of two from 2^-6 to 2^6,
    so results land in many binades and rounding differences show."""
    return (rng.uniform(lo, hi, n) * 2.0 ** rng.integers(-6, 7, n)).astype(np.float32)


def test_particle_link_law_matches_model():
    # ParticleLink::solve, src/link.rs:78-87, and the EXT3 pin rule
    rng = np.random.default_rng(11)
    for k in range(600):
        a, b = _scaled(rng, 2, -4, 4), _scaled(rng, 2, -4, 4)
        if k % 7 == 0:   # nearly vertical, x near the bottom of the normal range: n.x * e subnormal
            a[0] = f32(rng.uniform(1, 8) * 2.0 ** -int(rng.integers(100, 126)))
            b[0] = f32(0.0)
        d = float(np.hypot(*(a.astype(np.float64) - b)))
        t = f32(d * rng.uniform(0.3, 1.7))
        pa, pb = (k % 5 == 1), (k % 5 == 2)
        ga, gb = law_particle_link(a, b, t, pa, pb)
        A, B = M.Particle(a), M.Particle(b)
        M.particle_link_solve(A, B, t, pa, pb)
        assert _eq(ga, A.pos.tup()) and _eq(gb, B.pos.tup()), (k, a, b, t)


def test_circle_link_law_matches_model():
    # CircleLink::solve, src/link.rs:96-108
    rng = np.random.default_rng(12)
    for k in range(600):
        a, b = _scaled(rng, 2, -4, 4), _scaled(rng, 2, -4, 4)
        ra, rb = _scaled(rng, 2, 0.1, 3)
        t = f32(float(np.hypot(*(a.astype(np.float64) - b))) * rng.uniform(0.3, 1.7))
        ga, gb = law_circle_link(a, b, ra, rb, t)
        A, B = M.Circle(a, ra), M.Circle(b, rb)
        M.circle_link_solve(A, B, t)
        assert _eq(ga, A.point.pos.tup()) and _eq(gb, B.point.pos.tup()), (k, a, b, ra, rb, t)


def test_circle_pair_law_matches_model():
    # Circle::solve_circle, src/circle.rs:141-154 (overlapping and not)
    rng = np.random.default_rng(13)
    hits = 0
    for k in range(600):
        ra, rb = _scaled(rng, 2, 0.2, 2)
        a = _scaled(rng, 2, -4, 4)
        ang = rng.uniform(0, 2 * np.pi)
        dist = (float(ra) + float(rb)) * rng.uniform(0.05, 1.2)
        b = (a + dist * np.array([np.cos(ang), np.sin(ang)])).astype(np.float32)
        hit, ga, gb = law_circle_pair(a, b, ra, rb)
        A, B = M.Circle(a, ra), M.Circle(b, rb)
        mhit = A.solve_circle(B)
        hits += hit
        assert hit == mhit and _eq(ga, A.point.pos.tup()) and _eq(gb, B.point.pos.tup()), (k, a, b, ra, rb)
    assert 300 < hits < 600


@pytest.mark.parametrize("circle", [False, True])
def test_bounds_laws_match_model(circle):
    # Particle::solve_bounds (src/particle.rs:27-46), Circle::solve_bounds (src/circle.rs:120-139)
    rng = np.random.default_rng(14 + circle)
    seen = set()
    for k in range(800):
        bp = _scaled(rng, 2, -8, 8)
        bs = _scaled(rng, 2, 1, 30)
        r = _scaled(rng, 1, 0.05, 0.4)[0] * min(bs) if circle else None
        pos = (bp + bs * rng.uniform(-0.3, 1.3, 2)).astype(np.float32)
        prev = (pos + _scaled(rng, 2, -1, 1)).astype(np.float32)
        gp, gq = law_bounds(pos, prev, bp, bs, r)
        if circle:
            C = M.Circle(pos, r, prev)
            C.solve_bounds((M.V(*bp), M.V(*bs)))
            P = C.point
        else:
            P = M.Particle(pos, prev)
            P.solve_bounds((M.V(*bp), M.V(*bs)))
        assert _eq(gp, P.pos.tup()) and _eq(gq, P.prev_pos.tup()), (k, pos, prev, bp, bs, r)
        lo = bp if not circle else bp + r
        hi = bp + bs if not circle else bp + bs - r
        seen |= {(ax, "lo") for ax in range(2) if pos[ax] < lo[ax]} | {(ax, "hi") for ax in range(2) if pos[ax] > hi[ax]}
    assert len(seen) == 4  # every branch ran


def test_particle_update_law_matches_model():
    # Particle::update, src/particle.rs:20-25
    rng = np.random.default_rng(16)
    for k in range(600):
        pos = _scaled(rng, 2, -8, 8)
        prev = (pos + _scaled(rng, 2, -1, 1)).astype(np.float32)
        acc = _scaled(rng, 2, -100, 100)
        dt = _scaled(rng, 1, 0.001, 0.1)[0]
        gp, gq, ga = law_particle_update(pos, prev, acc, dt)
        P = M.Particle(pos, prev, acc)
        P.update(dt)
        assert _eq(gp, P.pos.tup()) and _eq(gq, P.prev_pos.tup()) and _eq(ga, (0, 0)), (k, pos, prev, acc, dt)


def test_line_intersection_law_matches_model():
    # common::line_intersection, src/common.rs:4-26
    rng = np.random.default_rng(17)
    hits = 0
    for k in range(600):
        p = [_scaled(rng, 2, -4, 4) for _ in range(4)]
        if k % 2:   # make most of them cross: p3-p4 through a point of p1-p2
            m = p[0] + rng.uniform(0, 1) * (p[1] - p[0])
            p[3] = (p[2] + (m - p[2]) * rng.uniform(1.0, 2.0)).astype(np.float32)
        g = law_line_intersection(*p)
        m = M.line_intersection((M.V(*p[0]), M.V(*p[1])), (M.V(*p[2]), M.V(*p[3])))
        assert (g is None) == (m is None), (k, p)
        if g is not None:
            hits += 1
            assert _eq(g, m.tup()), (k, p)
    assert hits > 250


def test_resolve_line_law_matches_model():
    # Polygon::resolve_line_intersection, src/polygon.rs:164-216: an edge a-b of
    # a polygon with centre c, a point o of another polygon with centre oc; o
    # placed so the segment o-oc crosses the edge in most cases
    rng = np.random.default_rng(18)
    hits = 0
    for k in range(600):
        a, b = _scaled(rng, 2, -4, 4), _scaled(rng, 2, -4, 4)
        d = b.astype(np.float64) - a
        side = np.array([-d[1], d[0]])
        m = a + rng.uniform(0.05, 0.95) * d
        if k % 2:   # a sliver: the centre far along the edge's line, where the projection is long
            c = (m + d * rng.uniform(3, 40) + side * rng.uniform(0.01, 0.2)).astype(np.float32)
        else:
            c = (m + side * rng.uniform(0.2, 2.0)).astype(np.float32)
        o = (m + side * rng.uniform(0.01, 0.3)).astype(np.float32)
        oc = (m - side * rng.uniform(0.2, 2.0)).astype(np.float32)
        g = law_resolve_line(c, a, b, o, oc)
        P = M.Polygon([M.Particle(a), M.Particle(b)], [], False, c)
        q = M.Particle(o)
        r = P.resolve_line_intersection(M.V(*a), M.V(*b), q, M.V(*oc))
        assert (g is None) == (r is None), (k, c, a, b, o, oc)
        if g is not None:
            hits += 1
            assert _eq(np.concatenate(g), [*r[0].tup(), *r[1].tup(), *r[2].tup()]), (k, c, a, b, o, oc)
    assert hits > 400


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 16, 31, 40])
def test_polygon_constructors_match_model(n):
    # Polygon::circle (src/polygon.rs:17-82) and Polygon::new (src/polygon.rs:84-123)
    o = OracleSolver(32)
    o.add_polygon_circle(3.7, (1.25, -2.5), n, False)
    g = o.get_polygon(0)
    m = M.Polygon.circle(3.7, (1.25, -2.5), n, False)
    assert _eq(g["particles"][:, :2], [p.pos.tup() for p in m.particles])
    assert _eq(g["center"], m.center.tup())
    a, b, t = g["links"]
    assert a.tolist() == [l.particle_a for l in m.particle_links]
    assert b.tolist() == [l.particle_b for l in m.particle_links]
    assert _eq(t, [l.target_distance for l in m.particle_links])
    rng = np.random.default_rng(n)
    pts = _scaled(rng, 2 * n, -5, 5).reshape(n, 2)
    o2 = OracleSolver(32)
    o2.add_polygon_new(pts, True)
    g2 = o2.get_polygon(0)
    m2 = M.Polygon.new(pts, True)
    assert _eq(g2["center"], m2.center.tup())
    a, b, t = g2["links"]
    assert a.tolist() == [l.particle_a for l in m2.particle_links]
    assert b.tolist() == [l.particle_b for l in m2.particle_links]
    assert _eq(t, [l.target_distance for l in m2.particle_links])


def _scene(seed, polygons=True, colliders=True):
    """A small random scene using every phase: particles with links (one
    pinned), circles with links and overlaps, soft polygons that collide, static
    colliders, substeps and link iterations.  Returns (oracle, model, dt)."""
    rng = np.random.default_rng(seed)
    o, m = OracleSolver(32), M.Solver()
    g = (f32(rng.uniform(-3, 3)), f32(rng.uniform(-20, 20)))
    bp, bs = _scaled(rng, 2, -2, 2), f32(rng.uniform(18, 26, 2))
    o.set_gravity(g)
    m.gravity = M.V(*g)
    o.set_bounds(bp, bs)
    m.bounds = (M.V(*bp), M.V(*bs))
    sub, iters = int(rng.integers(1, 4)), int(rng.integers(1, 3))
    o.set_sub_steps(sub)
    o.set_link_iterations(iters)
    m.sub_steps, m.link_iterations = sub, iters
    lo, hi = bp + 1, bp + bs - 1
    np_ = int(rng.integers(3, 9))
    for i in range(np_):
        p = rng.uniform(lo, hi).astype(np.float32)
        q = (p + rng.normal(0, 0.05, 2)).astype(np.float32)
        o.add_particle(p, q)
        m.add_particle(p, q)
    pin = int(rng.integers(0, np_))
    o.pin(pin)
    m.pinned[pin] = True
    for _ in range(2 * np_):
        a, b = sorted(rng.choice(np_, 2, replace=False).tolist())
        t = f32(rng.uniform(0.5, 4))
        o.add_particle_link(a, b, t)
        m.particle_links.append(M.Link(a, b, t))
    nc = int(rng.integers(4, 12))
    for i in range(nc):
        p = rng.uniform(lo, hi).astype(np.float32)
        if i and i % 3 == 0:   # overlapping its predecessor
            p = (np.array(m.circles[-1].point.pos.tup(), np.float32) + rng.uniform(-0.6, 0.6, 2)).astype(np.float32)
        q = (p + rng.normal(0, 0.05, 2)).astype(np.float32)
        r = f32(rng.uniform(0.3, 1.2))
        acc = (f32(rng.uniform(-5, 5)), f32(rng.uniform(-5, 5))) if i % 4 == 1 else (0.0, 0.0)
        o.add_circle(p, r, prev_pos=q, acc=acc)
        m.circles.append(M.Circle(p, r, q, acc))
    for _ in range(nc // 2):
        a, b = sorted(rng.choice(nc, 2, replace=False).tolist())
        t = f32(rng.uniform(0.8, 3))
        o.add_circle_link(a, b, t)
        m.circle_links.append(M.Link(a, b, t))
    if polygons:
        c0 = rng.uniform(lo + 3, hi - 3).astype(np.float32)
        for k in range(3):
            c = (c0 + rng.uniform(-1.5, 1.5, 2)).astype(np.float32)
            n = int(rng.integers(3, 8))
            poly = M.Polygon.circle(f32(rng.uniform(1, 2.5)), c, n, k == 2)
            parts = np.array([[*p.pos.tup(), *p.prev_pos.tup(), 0, 0] for p in poly.particles], np.float32)
            links = ([l.particle_a for l in poly.particle_links], [l.particle_b for l in poly.particle_links],
                     [l.target_distance for l in poly.particle_links])
            o.add_polygon(parts, links, poly.is_static, poly.center.tup())
            m.polygons.append(poly)
    if colliders:
        c = rng.uniform(lo + 2, hi - 2).astype(np.float32)
        r = f32(rng.uniform(1, 3))
        o.add_static_circle(c, r)
        m.colliders.append(("circle", c[0], c[1], r))
        v = (rng.uniform(lo + 2, hi - 2) + np.array([[0, 0], [3, 0.4], [2.5, 2.8], [-0.3, 2]])).astype(np.float32)
        o.add_static_polygon(v[::-1] if seed % 2 else v)
        m.colliders.append(("polygon", M.static_polygon_planes(v[::-1] if seed % 2 else v)))
    return o, m, f32(rng.uniform(0.01, 0.05))


@pytest.mark.parametrize("seed", range(6))
def test_whole_scenes_match_model(seed):
    # Solver::update, src/solver.rs:106-188: phase order, loop orders, every law
    o, m, dt = _scene(seed)
    assert _eq(state(o), m.state())
    for k in range(25):
        assert o.update(dt) == 0
        m.update(dt)
        assert _eq(state(o), m.state()), (seed, k)


def test_model_reaches_collisions():
    # the scenes above do exercise contacts: circle pairs and polygon edges move bodies
    o, m, dt = _scene(3, polygons=True, colliders=False)
    hits = 0
    for _ in range(10):
        for i in range(len(m.circles)):
            for j in range(i + 1, len(m.circles)):
                a, b = m.circles[i], m.circles[j]
                d = (a.point.pos - b.point.pos).magnitude()
                hits += bool(d < a.radius + b.radius)
        m.update(dt)
    assert hits > 0
