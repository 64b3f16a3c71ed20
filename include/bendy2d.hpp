// bendy2d.hpp -- C++ host interface over the C ABI (bendy2d.h), mirroring the
// crate Murrent/bendy2d's public types and Solver methods (src/lib.rs:1-6)
// so that code written against the crate reads the same: Vector2, Particle,
// Link / ParticleLink / CircleLink, Circle, Polygon, Bounds and Solver with
// add_* / get_* / update.
//
// Differences a caller sees:
//   - getters return copies (the crate lends &Vec<T> / Option<&T>); the
//     device holds the state between updates, and a getter downloads it once;
//   - where the crate panics (an invalid link inside update,
//     src/link.rs:19-21), update throws std::runtime_error carrying the
//     library's message, before anything has moved;
//   - Particle::new and Polygon::new are constructors (new is a C++ keyword).
// Header-only; link with -lbendy2d.
#pragma once

#include <cstddef>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bendy2d.h"

namespace bendy2d {

// nalgebra::Vector2<f32>
struct Vector2 {
    float x = 0.0f, y = 0.0f;
};

namespace detail {
inline bd_vec2 v(Vector2 a) { return bd_vec2{a.x, a.y}; }
inline Vector2 v(bd_vec2 a) { return Vector2{a.x, a.y}; }
[[noreturn]] inline void fail(const char *what, int64_t code) {
    throw std::runtime_error(std::string(what) + ": " + bd_error_string((int32_t)code) + " (" + bd_last_error() + ")");
}
inline int64_t check(const char *what, int64_t code) {
    if (code < 0) fail(what, code);
    return code;
}
}  // namespace detail

// Particle {pos, prev_pos, acc}: src/particle.rs:5-18.  acc is private in the
// crate; here it is readable for completeness.
struct Particle {
    Vector2 pos, prev_pos, acc;
    Particle() = default;
    // Particle::new(pos): prev_pos = pos, acc = 0 (src/particle.rs:12-18)
    explicit Particle(Vector2 p) : pos(p), prev_pos(p), acc{} {}
};

// Link {particle_a, particle_b, target_distance}: src/link.rs:5-10
struct Link {
    std::size_t particle_a = 0, particle_b = 0;
    float target_distance = 0.0f;
};
struct ParticleLink { Link link; };  // src/link.rs:13-15
struct CircleLink { Link link; };    // src/link.rs:31-33

// Circle {point, radius}: src/circle.rs:5-8
struct Circle {
    Particle point;
    float radius = 0.0f;
};

// Bounds {pos, size}: src/solver.rs:14-17
struct Bounds {
    Vector2 pos, size;
};

// Polygon {particles, particle_links, is_static, center, scale}:
// src/polygon.rs:8-14, with its two constructors computed by the library
// (bit-identical to the crate's f32 arithmetic).
struct Polygon {
    std::vector<Particle> particles;
    std::vector<ParticleLink> particle_links;
    bool is_static = false;
    Vector2 center;
    float scale = 1.0f;

    Polygon() = default;
    // Polygon::new(points, is_static): src/polygon.rs:84-123
    Polygon(const std::vector<Vector2> &points, bool is_static_) {
        std::vector<bd_vec2> pts(points.size());
        for (std::size_t i = 0; i < points.size(); i++) pts[i] = detail::v(points[i]);
        std::vector<bd_particle> ps(points.size());
        std::vector<bd_link> ls(points.size());
        bd_vec2 c{};
        detail::check("Polygon::new", bd_polygon_new(pts.data(), pts.size(), is_static_ ? 1 : 0, ps.data(), ls.data(), &c));
        assign(ps, ls, is_static_, c);
    }
    // Polygon::circle(radius, pos, point_count, is_static): src/polygon.rs:17-82
    static Polygon circle(float radius, Vector2 pos, std::size_t point_count, bool is_static_) {
        std::vector<bd_particle> ps(point_count);
        std::vector<bd_link> ls(2 * point_count);
        bd_vec2 c{};
        detail::check("Polygon::circle", bd_polygon_circle(radius, detail::v(pos), point_count, is_static_ ? 1 : 0,
                                                           ps.data(), ls.data(), &c));
        Polygon p;
        p.assign(ps, ls, is_static_, c);
        return p;
    }

private:
    void assign(const std::vector<bd_particle> &ps, const std::vector<bd_link> &ls, bool st, bd_vec2 c) {
        particles.clear();
        for (const bd_particle &q : ps) {
            Particle r;
            r.pos = detail::v(q.pos); r.prev_pos = detail::v(q.prev_pos); r.acc = detail::v(q.acc);
            particles.push_back(r);
        }
        particle_links.clear();
        for (const bd_link &l : ls)
            particle_links.push_back(ParticleLink{Link{(std::size_t)l.particle_a, (std::size_t)l.particle_b, l.target_distance}});
        is_static = st;
        center = detail::v(c);
        scale = 1.0f;
    }
};

// Solver: src/solver.rs:20-188.  gravity, bounds and bounds_active are public
// fields as in the crate; update() hands them to the device before it runs.
class Solver {
public:
    Vector2 gravity;
    Bounds bounds;
    bool bounds_active = true;

    // Solver::new (src/solver.rs:34-50): gravity (0, 98.2), bounds (0,0)+(100,100)
    Solver() : s_(bd_solver_new()) {
        if (!s_) throw std::runtime_error(std::string("Solver::new: ") + bd_last_error());
        pull_fields();
    }
    // #[derive(Clone)] (src/solver.rs:19)
    Solver(const Solver &o) : gravity(o.gravity), bounds(o.bounds), bounds_active(o.bounds_active),
                              s_(bd_solver_clone(o.s_)) {
        if (!s_) throw std::runtime_error(std::string("Solver::clone: ") + bd_last_error());
    }
    Solver &operator=(const Solver &o) {
        if (this != &o) { Solver t(o); swap(t); }
        return *this;
    }
    Solver(Solver &&o) noexcept : gravity(o.gravity), bounds(o.bounds), bounds_active(o.bounds_active), s_(o.s_) {
        o.s_ = nullptr;
    }
    Solver &operator=(Solver &&o) noexcept {
        if (this != &o) swap(o);
        return *this;
    }
    ~Solver() { if (s_) bd_solver_free(s_); }

    // add_*: src/solver.rs:52-67
    void add_particle(Vector2 pos) { detail::check("add_particle", bd_add_particle(s_, detail::v(pos))); }
    void add_circle(const Circle &c) {
        const bd_circle r = to_c(c);
        detail::check("add_circle", bd_add_circle(s_, &r));
    }
    void add_polygon(const Polygon &p) {
        std::vector<bd_particle> ps;
        for (const Particle &q : p.particles) ps.push_back(to_c(q));
        std::vector<bd_link> ls;
        for (const ParticleLink &l : p.particle_links) ls.push_back(to_c(l.link));
        bd_polygon_in in{ps.data(), ps.size(), ls.data(), ls.size(), p.is_static ? 1 : 0, detail::v(p.center), p.scale};
        detail::check("add_polygon", bd_add_polygon(s_, &in));
    }
    void add_particle_link(const ParticleLink &l) {
        const bd_link r = to_c(l.link);
        detail::check("add_particle_link", bd_add_particle_link(s_, &r));
    }
    void add_circle_link(const CircleLink &l) {
        const bd_link r = to_c(l.link);
        detail::check("add_circle_link", bd_add_circle_link(s_, &r));
    }

    // get_*: src/solver.rs:69-104 (copies)
    std::size_t get_particle_len() const { return (std::size_t)bd_get_particle_len(s_); }
    std::size_t get_circles_len() const { return (std::size_t)bd_get_circles_len(s_); }
    std::size_t get_polygons_len() const { return (std::size_t)bd_get_polygons_len(s_); }
    std::vector<ParticleLink> get_particle_links() const {
        std::vector<ParticleLink> out;
        for (const bd_link &l : links(bd_get_particle_links)) out.push_back(ParticleLink{from_c(l)});
        return out;
    }
    std::vector<CircleLink> get_circle_links() const {
        std::vector<CircleLink> out;
        for (const bd_link &l : links(bd_get_circle_links)) out.push_back(CircleLink{from_c(l)});
        return out;
    }
    std::vector<Particle> get_particles() const {
        std::vector<bd_particle> raw(get_particle_len());
        if (bd_get_particles(s_, raw.data(), raw.size()) != raw.size()) detail::fail("get_particles", BD_ERR_STATE);
        std::vector<Particle> out;
        for (const bd_particle &p : raw) out.push_back(from_c(p));
        return out;
    }
    std::vector<Circle> get_circles() const {
        std::vector<bd_circle> raw(get_circles_len());
        if (bd_get_circles(s_, raw.data(), raw.size()) != raw.size()) detail::fail("get_circles", BD_ERR_STATE);
        std::vector<Circle> out;
        for (const bd_circle &c : raw) out.push_back(Circle{from_c(c.point), c.radius});
        return out;
    }
    std::vector<Polygon> get_polygons() const {
        std::vector<Polygon> out;
        for (std::size_t i = 0; i < get_polygons_len(); i++) out.push_back(*get_polygon(i));
        return out;
    }
    std::optional<Particle> get_particle(std::size_t i) const {
        bd_particle p{};
        if (!bd_get_particle(s_, i, &p)) return std::nullopt;
        return from_c(p);
    }
    std::optional<Circle> get_circle(std::size_t i) const {
        bd_circle c{};
        if (!bd_get_circle(s_, i, &c)) return std::nullopt;
        return Circle{from_c(c.point), c.radius};
    }
    std::optional<Polygon> get_polygon(std::size_t i) const {
        bd_polygon_info info{};
        if (!bd_get_polygon(s_, i, &info)) return std::nullopt;
        std::vector<bd_particle> ps(info.n_particles);
        std::vector<bd_link> ls(info.n_links);
        bd_get_polygon_particles(s_, i, ps.data(), ps.size());
        bd_get_polygon_links(s_, i, ls.data(), ls.size());
        Polygon p;
        for (const bd_particle &q : ps) p.particles.push_back(from_c(q));
        for (const bd_link &l : ls) p.particle_links.push_back(ParticleLink{from_c(l)});
        p.is_static = info.is_static != 0;
        p.center = detail::v(info.center);
        p.scale = info.scale;
        return p;
    }

    // Solver::update(dt): src/solver.rs:106-116, on the GPU
    void update(float dt) {
        push_fields();
        detail::check("update", bd_update(s_, dt));
    }

    // The C handle, for the bd_ext_* extensions (substeps, pins, colliders, ...).
    bd_solver *raw() { return s_; }
    const bd_solver *raw() const { return s_; }

private:
    bd_solver *s_ = nullptr;

    void swap(Solver &o) noexcept {
        std::swap(gravity, o.gravity);
        std::swap(bounds, o.bounds);
        std::swap(bounds_active, o.bounds_active);
        std::swap(s_, o.s_);
    }
    void pull_fields() {
        gravity = detail::v(bd_get_gravity(s_));
        const bd_bounds b = bd_get_bounds(s_);
        bounds = Bounds{detail::v(b.pos), detail::v(b.size)};
        bounds_active = bd_get_bounds_active(s_) != 0;
    }
    void push_fields() {
        bd_set_gravity(s_, detail::v(gravity));
        bd_set_bounds(s_, bd_bounds{detail::v(bounds.pos), detail::v(bounds.size)});
        bd_set_bounds_active(s_, bounds_active ? 1 : 0);
    }
    template <class F>
    std::vector<bd_link> links(F get) const {
        std::vector<bd_link> raw(get(s_, nullptr, 0));
        get(s_, raw.data(), raw.size());
        return raw;
    }
    static bd_particle to_c(const Particle &p) { return bd_particle{detail::v(p.pos), detail::v(p.prev_pos), detail::v(p.acc)}; }
    static bd_circle to_c(const Circle &c) { return bd_circle{to_c(c.point), c.radius}; }
    static bd_link to_c(const Link &l) { return bd_link{(uint64_t)l.particle_a, (uint64_t)l.particle_b, l.target_distance}; }
    static Particle from_c(const bd_particle &p) {
        Particle r;
        r.pos = detail::v(p.pos); r.prev_pos = detail::v(p.prev_pos); r.acc = detail::v(p.acc);
        return r;
    }
    static Link from_c(const bd_link &l) { return Link{(std::size_t)l.particle_a, (std::size_t)l.particle_b, l.target_distance}; }
};

}  // namespace bendy2d
