// The C++ host interface (include/bendy2d.hpp) used the way code written
// against the crate uses Solver: a scene built from the crate's own types with
// Solver::new's defaults (gravity (0, 98.2), bounds (0,0)+(100,100), one
// substep), advanced by update() on the GPU, and compared bit for bit with the
// f32 oracle (oracle/liboracle.so, test infrastructure) given the same scene.
//
// Exit status: 0 = bit-identical, 1 = a difference (printed), 2 = no HIP device.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/bendy2d.hpp"

extern "C" {
typedef struct or32_solver or32_solver;
or32_solver *or32_new(void);
void or32_free(or32_solver *s);
int64_t or32_add_particle(or32_solver *s, float x, float y);
int64_t or32_add_particle_link(or32_solver *s, int64_t a, int64_t b, float t);
int64_t or32_add_circle(or32_solver *s, float x, float y, float px, float py, float ax, float ay, float r);
int64_t or32_add_circle_link(or32_solver *s, int64_t a, int64_t b, float t);
int64_t or32_add_polygon_circle(or32_solver *s, float radius, float x, float y, int64_t n, int is_static);
int64_t or32_add_polygon_new(or32_solver *s, int64_t n, const float *xy, int is_static);
int or32_update(or32_solver *s, float dt);
int64_t or32_get_particle_len(const or32_solver *s);
int64_t or32_get_circles_len(const or32_solver *s);
int64_t or32_get_polygons_len(const or32_solver *s);
int64_t or32_get_polygon_point_len(const or32_solver *s, int64_t p);
void or32_get_particles(const or32_solver *s, double *out);
void or32_get_circles(const or32_solver *s, double *out);
int or32_get_polygon(const or32_solver *s, int64_t p, double *out, double *center);
}

using bendy2d::Circle;
using bendy2d::CircleLink;
using bendy2d::Link;
using bendy2d::Particle;
using bendy2d::ParticleLink;
using bendy2d::Polygon;
using bendy2d::Solver;
using bendy2d::Vector2;

static bool same(double a, float b) {
    if (std::isnan(a) && std::isnan(b)) return true;
    return (float)a == b && std::signbit(a) == std::signbit(b);
}

int main() {
    if (bd_device_count() < 1) {
        std::printf("no HIP device visible\n");
        return 2;
    }
    const float dt = 1.0f / 60.0f;
    const int updates = 120;
    Solver s;                      // Solver::new
    or32_solver *o = or32_new();

    // a chain of 20 particles, linked in index order
    for (int i = 0; i < 20; i++) {
        s.add_particle(Vector2{10.0f + 2.0f * (float)i, 20.0f});
        or32_add_particle(o, 10.0f + 2.0f * (float)i, 20.0f);
    }
    for (int i = 0; i + 1 < 20; i++) {
        s.add_particle_link(ParticleLink{Link{(size_t)i, (size_t)i + 1, 2.0f}});
        or32_add_particle_link(o, i, i + 1, 2.0f);
    }
    // 3 rows of 10 circles, some with an initial velocity, linked along rows
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 10; c++) {
            Circle ci;
            ci.point = Particle(Vector2{15.0f + 3.1f * (float)c, 40.0f + 3.3f * (float)r});
            ci.point.prev_pos = Vector2{ci.point.pos.x - 0.05f * (float)(c % 3), ci.point.pos.y + 0.02f * (float)r};
            ci.radius = 1.5f;
            s.add_circle(ci);
            or32_add_circle(o, ci.point.pos.x, ci.point.pos.y, ci.point.prev_pos.x, ci.point.prev_pos.y, 0.0f, 0.0f, 1.5f);
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c + 1 < 10; c += 2) {
            const size_t a = (size_t)(10 * r + c);
            s.add_circle_link(CircleLink{Link{a, a + 1, 3.0f}});
            or32_add_circle_link(o, (int64_t)a, (int64_t)a + 1, 3.0f);
        }
    // polygons from both constructors, one static
    s.add_polygon(Polygon::circle(5.0f, Vector2{70.0f, 30.0f}, 8, false));
    or32_add_polygon_circle(o, 5.0f, 70.0f, 30.0f, 8, 0);
    const std::vector<Vector2> square = {{26.0f, 66.0f}, {34.0f, 66.0f}, {34.0f, 74.0f}, {26.0f, 74.0f}};
    s.add_polygon(Polygon(square, false));
    const float sq[8] = {26.0f, 66.0f, 34.0f, 66.0f, 34.0f, 74.0f, 26.0f, 74.0f};
    or32_add_polygon_new(o, 4, sq, 0);
    s.add_polygon(Polygon::circle(6.0f, Vector2{50.0f, 85.0f}, 6, true));
    or32_add_polygon_circle(o, 6.0f, 50.0f, 85.0f, 6, 1);

    for (int k = 0; k < updates; k++) {
        try {
            s.update(dt);
        } catch (const std::exception &e) {
            std::printf("update %d threw: %s\n", k, e.what());
            return 1;
        }
        or32_update(o, dt);
    }

    size_t checked = 0;
    auto cmp = [&](const char *what, size_t i, const double *want, const Particle &got) -> bool {
        const float g[6] = {got.pos.x, got.pos.y, got.prev_pos.x, got.prev_pos.y, got.acc.x, got.acc.y};
        for (int j = 0; j < 4; j++) {  // acc is internal bookkeeping of the step; pos and prev are its result
            checked++;
            if (!same(want[j], g[j])) {
                std::printf("%s %zu value %d: GPU %.9g oracle %.9g\n", what, i, j, (double)g[j], want[j]);
                return false;
            }
        }
        return true;
    };
    const std::vector<Particle> ps = s.get_particles();
    std::vector<double> ref(6 * ps.size());
    or32_get_particles(o, ref.data());
    if ((int64_t)ps.size() != or32_get_particle_len(o)) { std::printf("particle count differs\n"); return 1; }
    for (size_t i = 0; i < ps.size(); i++)
        if (!cmp("particle", i, &ref[6 * i], ps[i])) return 1;
    const std::vector<Circle> cs = s.get_circles();
    std::vector<double> rc(7 * cs.size());
    or32_get_circles(o, rc.data());
    for (size_t i = 0; i < cs.size(); i++)
        if (!cmp("circle", i, &rc[7 * i], cs[i].point)) return 1;
    const std::vector<Polygon> gs = s.get_polygons();
    if ((int64_t)gs.size() != or32_get_polygons_len(o)) { std::printf("polygon count differs\n"); return 1; }
    for (size_t p = 0; p < gs.size(); p++) {
        std::vector<double> rp(6 * (size_t)or32_get_polygon_point_len(o, (int64_t)p));
        double center[2];
        or32_get_polygon(o, (int64_t)p, rp.data(), center);
        for (size_t i = 0; i < gs[p].particles.size(); i++)
            if (!cmp("polygon point", i, &rp[6 * i], gs[p].particles[i])) return 1;
    }
    or32_free(o);
    std::printf("facade parity OK: %zu values after %d updates\n", checked, updates);
    return 0;
}
