"""A second restatement of the reference's arithmetic, written from the Rust
source of Murrent/bendy2d 0.0.1 (src/*.rs), for pinning the C oracle.

TEST INFRASTRUCTURE ONLY.  It is independent of oracle/oracle_impl.h: every
function below transcribes the Rust expression it cites, left to right, into
IEEE-754 binary32 operations on numpy float32 scalars (each +, -, *, / and
sqrt is correctly rounded, ties to even, denormals kept), with nalgebra 0.32
Vector2<f32> semantics:

    a + b, a - b, a * s, s * a, a / s   componentwise
    a.dot(&b)        = a.x * b.x + a.y * b.y
    a.norm_squared() = 0 + a.dot(&a)
    a.magnitude()    = sqrt(a.norm_squared())
    a.normalize()    = a.unscale(a.magnitude()) = (a.x / m, a.y / m)
    a += b           = a = a + b componentwise

Python's operator precedence and left-to-right association of + - * / match
Rust's for these expressions (unary minus binds tighter than *), so each
transcription keeps the source's grouping.  The C oracle and this model agree
bit for bit only when both perform the same operations in the same order; a
slip in either (a dropped term, a sign, a reassociation, a reciprocal in place
of a division, a wrong index) shows up as a mismatch on seeded inputs.

The extensions the reference lacks (pins, static colliders, substeps, link
iterations, initial velocity) are transcribed from their specification in
include/bendy2d.h.

Polygon::circle calls f32::cos / f32::sin, which on x86_64-linux-gnu are
glibc's cosf / sinf; the model calls those through ctypes.
"""
import ctypes
import ctypes.util

import numpy as np

f32 = np.float32
ZERO = f32(0.0)

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.cosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]


def cosf(x):
    return f32(_libm.cosf(float(x)))


def sinf(x):
    return f32(_libm.sinf(float(x)))


class V:
    """nalgebra::Vector2<f32>."""
    __slots__ = ("x", "y")

    def __init__(self, x, y):
        self.x = f32(x)
        self.y = f32(y)

    def __add__(self, o):
        return V(self.x + o.x, self.y + o.y)

    def __sub__(self, o):
        return V(self.x - o.x, self.y - o.y)

    def __mul__(self, s):
        s = f32(s)
        return V(self.x * s, self.y * s)

    def __rmul__(self, s):
        s = f32(s)
        return V(s * self.x, s * self.y)

    def __truediv__(self, s):
        s = f32(s)
        return V(self.x / s, self.y / s)

    def dot(self, o):
        return self.x * o.x + self.y * o.y

    def norm_squared(self):
        return ZERO + self.dot(self)

    def magnitude(self):
        return np.sqrt(self.norm_squared())

    def normalize(self):
        m = self.magnitude()
        return V(self.x / m, self.y / m)

    def copy(self):
        return V(self.x, self.y)

    def tup(self):
        return (self.x, self.y)


class Particle:
    """src/particle.rs:4-18."""

    def __init__(self, pos, prev=None, acc=(0.0, 0.0)):
        self.pos = V(*pos)
        self.prev_pos = V(*(pos if prev is None else prev))
        self.acc = V(*acc)

    def update(self, dt):
        # src/particle.rs:20-25
        dt = f32(dt)
        vel = self.pos - self.prev_pos
        self.prev_pos = self.pos
        self.pos = self.pos + vel + self.acc * dt * dt
        self.acc = V(0.0, 0.0)

    def solve_bounds(self, bounds):
        # src/particle.rs:27-46
        bp, bs = bounds
        if self.pos.x < bp.x:
            vel_x = self.prev_pos.x - self.pos.x
            self.prev_pos.x = bp.x - vel_x
            self.pos.x = bp.x
        elif self.pos.x > bp.x + bs.x:
            vel_x = self.prev_pos.x - self.pos.x
            self.prev_pos.x = bp.x + bs.x - vel_x
            self.pos.x = bp.x + bs.x
        if self.pos.y < bp.y:
            vel_y = self.prev_pos.y - self.pos.y
            self.prev_pos.y = bp.y - vel_y
            self.pos.y = bp.y
        elif self.pos.y > bp.y + bs.y:
            vel_y = self.prev_pos.y - self.pos.y
            self.prev_pos.y = bp.y + bs.y - vel_y
            self.pos.y = bp.y + bs.y

    def add_force_v2(self, force):
        # src/particle.rs:52-54
        self.acc = self.acc + force


class Circle:
    """src/circle.rs:4-8."""

    def __init__(self, pos, radius, prev=None, acc=(0.0, 0.0)):
        self.point = Particle(pos, prev, acc)
        self.radius = f32(radius)

    def solve_bounds(self, bounds):
        # src/circle.rs:120-139
        bp, bs = bounds
        p, r = self.point, self.radius
        if p.pos.x < bp.x + r:
            vel_x = p.prev_pos.x - p.pos.x
            p.prev_pos.x = bp.x + r - vel_x
            p.pos.x = bp.x + r
        elif p.pos.x > bp.x + bs.x - r:
            vel_x = p.prev_pos.x - p.pos.x
            p.prev_pos.x = bp.x + bs.x - r - vel_x
            p.pos.x = bp.x + bs.x - r
        if p.pos.y < bp.y + r:
            vel_y = p.prev_pos.y - p.pos.y
            p.prev_pos.y = bp.y + r - vel_y
            p.pos.y = bp.y + r
        elif p.pos.y > bp.y + bs.y - r:
            vel_y = p.prev_pos.y - p.pos.y
            p.prev_pos.y = bp.y + bs.y - r - vel_y
            p.pos.y = bp.y + bs.y - r

    def solve_circle(self, circle):
        # src/circle.rs:141-154
        dist = self.point.pos - circle.point.pos
        dist_sqr = dist.norm_squared()
        radius_sum = self.radius + circle.radius
        if dist_sqr < radius_sum * radius_sum:
            normal = dist.normalize()
            overlap = radius_sum - np.sqrt(dist_sqr)
            self_rad_sqr = self.radius * self.radius
            circle_rad_sqr = circle.radius * circle.radius
            scale = f32(1.0) / (self_rad_sqr + circle_rad_sqr)
            self.point.pos = self.point.pos + normal * scale * overlap * circle_rad_sqr
            circle.point.pos = circle.point.pos - normal * scale * overlap * self_rad_sqr
            return True
        return False


def particle_link_solve(pa, pb, target, pinned_a=False, pinned_b=False):
    """ParticleLink::solve (src/link.rs:78-87) on two Particles.  With a pinned
    end (EXT3) the free end takes the whole correction; two pinned ends stay."""
    target = f32(target)
    dist_vec = pa.pos - pb.pos
    dist = dist_vec.magnitude()
    normal = dist_vec.normalize()
    if not pinned_a and not pinned_b:
        pa.pos = pa.pos - normal * (dist - target) * f32(0.5)
        pb.pos = pb.pos + normal * (dist - target) * f32(0.5)
    elif not pinned_a:
        pa.pos = pa.pos - normal * (dist - target)
    elif not pinned_b:
        pb.pos = pb.pos + normal * (dist - target)


def circle_link_solve(ca, cb, target):
    """CircleLink::solve (src/link.rs:96-108)."""
    target = f32(target)
    dist_vec = ca.point.pos - cb.point.pos
    dist = dist_vec.magnitude()
    normal = dist_vec.normalize()
    c_a_rad_sqr = ca.radius * ca.radius
    c_b_rad_sqr = cb.radius * cb.radius
    scale = f32(1.0) / (c_a_rad_sqr + c_b_rad_sqr)
    ca.point.pos = ca.point.pos - normal * (dist - target) * scale * c_b_rad_sqr
    cb.point.pos = cb.point.pos + normal * (dist - target) * scale * c_a_rad_sqr


def line_intersection(line1, line2):
    """common::line_intersection (src/common.rs:158-181 of the listing,
    src/common.rs:4-26): Some(V) or None."""
    p1, p2 = line1
    p3, p4 = line2
    s1_x = p2.x - p1.x
    s1_y = p2.y - p1.y
    s2_x = p4.x - p3.x
    s2_y = p4.y - p3.y
    s = (-s1_y * (p1.x - p3.x) + s1_x * (p1.y - p3.y)) / (-s2_x * s1_y + s1_x * s2_y)
    t = (s2_x * (p1.y - p3.y) - s2_y * (p1.x - p3.x)) / (-s2_x * s1_y + s1_x * s2_y)
    if s >= 0.0 and s <= 1.0 and t >= 0.0 and t <= 1.0:
        i_x = p1.x + (t * s1_x)
        i_y = p1.y + (t * s1_y)
        return V(i_x, i_y)
    return None


class Link:
    def __init__(self, a, b, t):
        self.particle_a, self.particle_b, self.target_distance = int(a), int(b), f32(t)


class Polygon:
    """src/polygon.rs:7-14."""

    def __init__(self, particles, links, is_static, center, scale=1.0):
        self.particles = particles
        self.particle_links = links
        self.is_static = bool(is_static)
        self.center = V(*center)
        self.scale = f32(scale)

    @staticmethod
    def circle(radius, pos, point_count, is_static):
        # src/polygon.rs:17-82
        radius = f32(radius)
        pos = V(*pos)
        particles, particle_links = [], []
        center = V(0.0, 0.0)
        angle = f32(0.0)
        for _ in range(point_count):
            x = radius * cosf(angle)
            y = radius * sinf(angle)
            point = Particle((pos + V(x, y)).tup())
            particles.append(point)
            center = center + point.pos
            angle = angle + f32(2.0) * f32(np.pi) / f32(point_count)
        center = center / f32(point_count)
        for i in range(point_count):
            for b_id in ((i + 2 * point_count // 3) % point_count, (i + point_count // 3) % point_count):
                a_id = i
                if a_id > b_id:
                    a_id, b_id = b_id, a_id
                dist_vec = particles[a_id].pos - particles[b_id].pos
                particle_links.append(Link(a_id, b_id, dist_vec.magnitude()))
        return Polygon(particles, particle_links, is_static, center.tup(), 1.0)

    @staticmethod
    def new(points, is_static):
        # src/polygon.rs:84-123
        particles, particle_links = [], []
        center = V(0.0, 0.0)
        for p in points:
            particle = Particle((f32(p[0]), f32(p[1])))
            particles.append(particle)
            center = center + particle.pos
        center = center / f32(len(points))
        for i in range(len(points)):
            a_id, b_id = i, (i + 1) % len(particles)
            if a_id > b_id:
                a_id, b_id = b_id, a_id
            dist_vec = particles[a_id].pos - particles[b_id].pos
            particle_links.append(Link(a_id, b_id, dist_vec.magnitude()))
        return Polygon(particles, particle_links, is_static, center.tup(), 1.0)

    def update(self, dt):
        # src/polygon.rs:125-134
        if self.is_static:
            return
        self.calc_center()
        for point in self.particles:
            point.update(dt)

    def solve_bounds(self, bounds):
        # src/polygon.rs:136-140
        for point in self.particles:
            point.solve_bounds(bounds)

    def solve_polygon(self, other):
        # src/polygon.rs:142-145
        self.solve_polygon_single(other)
        other.solve_polygon_single(self)

    def solve_polygon_single(self, other):
        # src/polygon.rs:147-162: point_a / point_b are copies taken per edge
        n = len(self.particles)
        for i in range(n):
            point_a = self.particles[i].pos.copy()
            b_id = (i + 1) % n
            point_b = self.particles[b_id].pos.copy()
            for others_point in other.particles:
                r = self.resolve_line_intersection(point_a, point_b, others_point, other.center)
                if r is not None:
                    self.particles[i].pos = r[0]
                    self.particles[b_id].pos = r[1]
                    others_point.pos = r[2]

    def resolve_line_intersection(self, point_a, point_b, others_point, other_center):
        # src/polygon.rs:164-216
        result = line_intersection((point_a, point_b), (others_point.pos, other_center))
        if result is None:
            return None
        intersection = result
        normal_line = (point_b - point_a).normalize()
        center_proj = (normal_line.dot(self.center - intersection) / normal_line.dot(normal_line)) * normal_line
        normal_in = (self.center - (intersection + center_proj)).normalize()
        dist_to_a = (intersection - point_a).magnitude()
        dist_to_b = (intersection - point_b).magnitude()
        dist_a_to_b = dist_to_a + dist_to_b
        influence_a = dist_to_b / dist_a_to_b
        influence_b = dist_to_a / dist_a_to_b
        diff_int_to_point = intersection - others_point.pos
        intersection_on_normal = (normal_in.dot(diff_int_to_point) / normal_in.dot(normal_in)) * normal_in
        displace_third = intersection_on_normal / f32(3.0)
        displace_line = displace_third * f32(2.0)
        displacement_a = influence_a * displace_line
        displacement_b = influence_b * displace_line
        new_a = point_a - displacement_a
        new_b = point_b - displacement_b
        intersection_result = line_intersection((point_a, point_b),
                                                (others_point.pos, others_point.pos - normal_in * f32(10000.0)))
        if intersection_result is not None:
            return new_a, new_b, intersection_result
        return None

    def solve_links(self):
        # src/polygon.rs:218-223
        self.calc_center()
        for link in self.particle_links:
            particle_link_solve(self.particles[link.particle_a], self.particles[link.particle_b],
                                link.target_distance)

    def add_force_v2(self, force):
        # src/polygon.rs:225-229
        for point in self.particles:
            point.add_force_v2(force)

    def calc_center(self):
        # src/polygon.rs:231-237
        self.center = V(0.0, 0.0)
        for point in self.particles:
            self.center = self.center + point.pos
        self.center = self.center / f32(len(self.particles))


def static_circle_push(p, radius, cx, cy, cr):
    """EXT4 static circle (include/bendy2d.h): a body of radius r (0 for a
    particle) overlapping the circle is moved out along the centre line by the
    whole overlap: d = p - c, q = d.x*d.x + d.y*d.y, rs = r + R; if q < rs*rs
    and sqrt(q) > 0: p += (d / sqrt(q)) * (rs - sqrt(q))."""
    dx, dy = p.x - f32(cx), p.y - f32(cy)
    q = dx * dx + dy * dy
    rs = f32(radius) + f32(cr)
    if q < rs * rs:
        d = np.sqrt(q)
        if d > 0:
            nx, ny, ov = dx / d, dy / d, rs - d
            p.x = p.x + nx * ov
            p.y = p.y + ny * ov


def static_polygon_planes(verts):
    """EXT4 static convex polygon: counter-clockwise edges (a clockwise input,
    by the sign of its f64 shoelace area, is reversed); edge v_i -> v_i+1 has
    e = v_i+1 - v_i, len = sqrt(e.x*e.x + e.y*e.y), n = (e.y / len, -e.x / len),
    c = n.x*v_i.x + n.y*v_i.y, all in f32."""
    v = [(f32(a), f32(b)) for a, b in verts]
    n = len(v)
    area2 = sum(float(v[i][0]) * float(v[(i + 1) % n][1]) - float(v[(i + 1) % n][0]) * float(v[i][1])
                for i in range(n))
    if area2 < 0:
        v = v[::-1]
    planes = []
    for i in range(n):
        vx, vy = v[i]
        ex, ey = v[(i + 1) % n][0] - vx, v[(i + 1) % n][1] - vy
        ln = np.sqrt(ex * ex + ey * ey)
        nx, ny = ey / ln, (-ex) / ln
        planes.append((nx, ny, nx * vx + ny * vy))
    return planes


def static_polygon_push(p, radius, planes):
    """EXT4 static convex polygon: the plane of largest signed distance
    sd = (n.x*p.x + n.y*p.y) - c (the first on ties) decides; if sd < r the
    body moves along that plane's normal by r - sd."""
    smax, arg = None, None
    for k, (nx, ny, c) in enumerate(planes):
        sd = (nx * p.x + ny * p.y) - c
        if arg is None or sd > smax:
            smax, arg = sd, k
    if smax < f32(radius):
        push = f32(radius) - smax
        nx, ny, _ = planes[arg]
        p.x = p.x + nx * push
        p.y = p.y + ny * push


class Solver:
    """src/solver.rs:19-189, with the extensions of include/bendy2d.h."""

    def __init__(self):
        # Solver::new, src/solver.rs:34-50
        self.gravity = V(0.0, 98.2)
        self.bounds = (V(0.0, 0.0), V(100.0, 100.0))
        self.particles, self.particle_links = [], []
        self.circles, self.circle_links = [], []
        self.polygons = []
        self.sub_steps = 1
        self.sub_steps_multiplier = f32(0.0)
        self.link_iterations = 1      # EXT2
        self.pinned = []              # EXT3
        self.colliders = []           # EXT4: ("circle", cx, cy, r) or ("polygon", planes)

    def add_particle(self, pos, prev=None):
        self.particles.append(Particle(pos, prev))
        self.pinned.append(False)

    def update(self, dt):
        # src/solver.rs:106-116 (EXT1: sub_steps may be above 1)
        self.sub_steps_multiplier = f32(1.0) / f32(self.sub_steps)
        delta = f32(dt) * self.sub_steps_multiplier
        for _ in range(self.sub_steps):
            self.apply_gravity()
            for _ in range(self.link_iterations):
                self.apply_links()
            self.solve_dynamic_collisions()
            self.solve_boundary_collisions()
            self.update_positions(delta)

    def update_positions(self, dt):
        # src/solver.rs:118-128; a pinned particle does not move
        for k, particle in enumerate(self.particles):
            if self.pinned[k]:
                particle.acc = V(0.0, 0.0)
                continue
            particle.update(dt)
        for circle in self.circles:
            circle.point.update(dt)
        for polygon in self.polygons:
            polygon.update(dt)

    def apply_gravity(self):
        # src/solver.rs:130-141
        for particle in self.particles:
            particle.add_force_v2(self.gravity)
        for circle in self.circles:
            circle.point.add_force_v2(self.gravity)
        for polygon in self.polygons:
            polygon.add_force_v2(self.gravity)

    def apply_links(self):
        # src/solver.rs:143-153
        for link in self.particle_links:
            particle_link_solve(self.particles[link.particle_a], self.particles[link.particle_b],
                                link.target_distance, self.pinned[link.particle_a], self.pinned[link.particle_b])
        for link in self.circle_links:
            circle_link_solve(self.circles[link.particle_a], self.circles[link.particle_b], link.target_distance)
        for polygon in self.polygons:
            polygon.solve_links()

    def solve_boundary_collisions(self):
        # src/solver.rs:155-165; a pinned particle is not clamped
        for k, particle in enumerate(self.particles):
            if not self.pinned[k]:
                particle.solve_bounds(self.bounds)
        for circle in self.circles:
            circle.solve_bounds(self.bounds)
        for polygon in self.polygons:
            polygon.solve_bounds(self.bounds)

    def solve_dynamic_collisions(self):
        # src/solver.rs:167-188, then the EXT4 static colliders
        n = len(self.circles)
        for i in range(n):
            for j in range(i + 1, n):
                self.circles[i].solve_circle(self.circles[j])
        n = len(self.polygons)
        for i in range(n):
            for j in range(i + 1, n):
                self.polygons[i].solve_polygon(self.polygons[j])
        if self.colliders:
            bodies = [(p.pos, f32(0.0)) for k, p in enumerate(self.particles) if not self.pinned[k]]
            bodies += [(c.point.pos, c.radius) for c in self.circles]
            bodies += [(q.pos, f32(0.0)) for g in self.polygons for q in g.particles]
            for pos, r in bodies:
                for col in self.colliders:
                    if col[0] == "circle":
                        static_circle_push(pos, r, col[1], col[2], col[3])
                    else:
                        static_polygon_push(pos, r, col[1])

    def state(self):
        """pos.xy and prev.xy of every particle, circle and polygon point, then
        each polygon's centre: the layout of tests/helpers.state()."""
        out = []
        for p in self.particles:
            out += [p.pos.x, p.pos.y, p.prev_pos.x, p.prev_pos.y]
        for c in self.circles:
            out += [c.point.pos.x, c.point.pos.y, c.point.prev_pos.x, c.point.prev_pos.y]
        for g in self.polygons:
            for p in g.particles:
                out += [p.pos.x, p.pos.y, p.prev_pos.x, p.prev_pos.y]
            out += [g.center.x, g.center.y]
        return np.array(out, np.float32)
