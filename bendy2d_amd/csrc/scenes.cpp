// scenes.cpp -- deterministic synthetic scenes for the benchmark workloads.
//
// All randomness comes from splitmix64 seeded per scene, so the same call
// produces the same bits on any host.  Scenes are built through the public
// adders, exactly as a user of the C ABI would build them.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/bendy2d.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1) with 24 random bits (exact in f32)
    float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    float uniform(float lo, float hi) { return lo + (hi - lo) * uniform(); }
    double normal() {
        double u1 = ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        double u2 = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct LinkList {
    std::vector<uint32_t> a, b;
    std::vector<float> t;
    void add(uint32_t x, uint32_t y, float r) { a.push_back(x); b.push_back(y); t.push_back(r); }
};

// Grid links of an nx * ny body grid in the requested emission order.
void grid_links(uint32_t nx, uint32_t ny, float spacing, int order, uint32_t first, LinkList &L) {
    const float rs = spacing, rd = sqrtf(2.0f) * spacing;
    auto id = [&](uint32_t r, uint32_t c) { return first + r * nx + c; };
    if (order == BD_ORDER_PARTICLE) {
        for (uint32_t r = 0; r < ny; r++)
            for (uint32_t c = 0; c < nx; c++) {
                if (c + 1 < nx) L.add(id(r, c), id(r, c + 1), rs);
                if (r + 1 < ny) L.add(id(r, c), id(r + 1, c), rs);
                if (r + 1 < ny && c + 1 < nx) L.add(id(r, c), id(r + 1, c + 1), rd);
                if (r + 1 < ny && c + 1 < nx) L.add(id(r, c + 1), id(r + 1, c), rd);
            }
        return;
    }
    for (uint32_t par = 0; par < 2; par++)  // H, even then odd columns
        for (uint32_t r = 0; r < ny; r++)
            for (uint32_t c = par; c + 1 < nx; c += 2) L.add(id(r, c), id(r, c + 1), rs);
    for (uint32_t par = 0; par < 2; par++)  // V, even then odd rows
        for (uint32_t r = par; r + 1 < ny; r += 2)
            for (uint32_t c = 0; c < nx; c++) L.add(id(r, c), id(r + 1, c), rs);
    for (uint32_t par = 0; par < 2; par++)  // D
        for (uint32_t r = par; r + 1 < ny; r += 2)
            for (uint32_t c = 0; c + 1 < nx; c++) L.add(id(r, c), id(r + 1, c + 1), rd);
    for (uint32_t par = 0; par < 2; par++)  // AD
        for (uint32_t r = par; r + 1 < ny; r += 2)
            for (uint32_t c = 0; c + 1 < nx; c++) L.add(id(r, c + 1), id(r + 1, c), rd);
}

}  // namespace

extern "C" {

int64_t bd_scene_add_grid(bd_solver *s, const bd_grid_desc *g) {
    if (!s || !g || g->nx == 0 || g->ny == 0) return BD_ERR_ARG;
    Rng rng(g->seed);
    const uint64_t n = (uint64_t)g->nx * g->ny;
    std::vector<bd_vec2> pos(n);
    for (uint32_t r = 0; r < g->ny; r++)
        for (uint32_t c = 0; c < g->nx; c++) {
            float x = g->x0 + (float)c * g->spacing, y = g->y0 - (float)r * g->spacing;
            if (g->jitter > 0.0f) {
                x = x + rng.uniform(-g->jitter, g->jitter);
                y = y + rng.uniform(-g->jitter, g->jitter);
            }
            pos[(uint64_t)r * g->nx + c] = bd_vec2{x, y};
        }
    int64_t first;
    if (g->as_circles) {
        std::vector<bd_circle> cs(n);
        for (uint64_t i = 0; i < n; i++) cs[i] = bd_circle{{pos[i], pos[i], {0.0f, 0.0f}}, g->radius};
        first = bd_ext_add_circles(s, cs.data(), n);
    } else {
        first = bd_ext_add_particles(s, pos.data(), nullptr, n);
    }
    if (first < 0) return first;
    LinkList L;
    grid_links(g->nx, g->ny, g->spacing, g->order, (uint32_t)first, L);
    int64_t rc = g->as_circles ? bd_ext_add_circle_links(s, L.a.data(), L.b.data(), L.t.data(), L.a.size())
                               : bd_ext_add_particle_links(s, L.a.data(), L.b.data(), L.t.data(), L.a.size());
    if (rc < 0) return rc;
    if (g->pin_top && !g->as_circles) {
        std::vector<uint8_t> pin(g->nx, 1);
        int32_t prc = bd_ext_set_pinned_range(s, (uint64_t)first, pin.data(), g->nx);
        if (prc < 0) return prc;
    }
    return first;
}

int64_t bd_scene_add_gas(bd_solver *s, uint64_t n, float radius, float lo, float hi, float vsigma, float dt_sub,
                         uint64_t seed) {
    if (!s) return BD_ERR_ARG;
    Rng rng(seed);
    std::vector<bd_circle> cs(n);
    for (uint64_t i = 0; i < n; i++) {
        float x = rng.uniform(lo, hi), y = rng.uniform(lo, hi);
        float vx = (float)(vsigma * rng.normal()), vy = (float)(vsigma * rng.normal());
        bd_vec2 p{x, y}, q{x - vx * dt_sub, y - vy * dt_sub};
        cs[i] = bd_circle{{p, q, {0.0f, 0.0f}}, radius};
    }
    return bd_ext_add_circles(s, cs.data(), n);
}

int32_t bd_scene_gas_box(bd_solver *s, uint64_t n, float radius, float box, uint32_t n_static, uint64_t seed) {
    if (!s || !(box > 4.0f * radius)) return BD_ERR_ARG;
    const float dt_sub = (1.0f / 60.0f) * (1.0f / 4.0f);
    bd_set_gravity(s, bd_vec2{0.0f, -9.81f});
    bd_set_bounds(s, bd_bounds{{0.0f, 0.0f}, {box, box}});
    int32_t rc = bd_ext_set_sub_steps(s, 4);
    if (rc < 0) return rc;
    int64_t r = bd_scene_add_gas(s, n, radius, radius, box - radius, 1.0f, dt_sub, seed);
    if (r < 0) return (int32_t)r;
    Rng rng(seed ^ 0x5f3759dfull);
    for (uint32_t k = 0; k < n_static; k++) {
        float R = rng.uniform(4.0f, 16.0f);
        float x = rng.uniform(0.0f, box), y = rng.uniform(0.0f, box);
        r = bd_ext_add_static_circle(s, bd_vec2{x, y}, R);
        if (r < 0) return (int32_t)r;
    }
    // four walls: thick rectangles just outside the box, counter-clockwise
    const float w = 64.0f;
    const bd_vec2 walls[4][4] = {
        {{-w, -w}, {0.0f, -w}, {0.0f, box + w}, {-w, box + w}},          // left
        {{box, -w}, {box + w, -w}, {box + w, box + w}, {box, box + w}},  // right
        {{-w, -w}, {box + w, -w}, {box + w, 0.0f}, {-w, 0.0f}},          // bottom
        {{-w, box}, {box + w, box}, {box + w, box + w}, {-w, box + w}},  // top
    };
    for (int k = 0; k < 4; k++) {
        r = bd_ext_add_static_polygon(s, walls[k], 4);
        if (r < 0) return (int32_t)r;
    }
    return BD_OK;
}

int64_t bd_scene_softbody(bd_solver *s, uint64_t seed, uint32_t max_side) {
    if (!s) return BD_ERR_ARG;
    Rng rng(seed);
    float u = rng.uniform();
    uint32_t side = (uint32_t)lrintf(4.0f * powf(32.0f, u));
    if (side < 4) side = 4;
    if (side > 128) side = 128;
    if (max_side && side > max_side) side = max_side;
    const float W = 3.0f * (float)side + 40.0f;
    bd_set_gravity(s, bd_vec2{0.0f, -9.81f});
    bd_set_bounds(s, bd_bounds{{0.0f, 0.0f}, {W, W}});
    int32_t rc = bd_ext_set_sub_steps(s, 8);
    if (rc < 0) return rc;
    rc = bd_ext_set_link_iterations(s, 1);
    if (rc < 0) return rc;
    bd_grid_desc g;
    g.nx = side; g.ny = side; g.spacing = 1.0f;
    g.x0 = 0.5f * W - 0.5f * (float)(side - 1);
    g.y0 = W - 2.0f;
    g.jitter = 0.0f; g.seed = seed; g.order = BD_ORDER_COLOUR; g.pin_top = 0;
    g.as_circles = 1; g.radius = 0.35f;
    int64_t first = bd_scene_add_grid(s, &g);
    if (first < 0) return first;
    uint32_t ncol = 1 + (uint32_t)(rng.next() % 8);
    uint32_t kind = (uint32_t)(rng.next() & 1);
    for (uint32_t k = 0; k < ncol; k++, kind ^= 1u) {
        float R = rng.uniform(2.0f, 10.0f);
        float cx = rng.uniform(0.1f * W, 0.9f * W), cy = rng.uniform(0.1f * W, 0.6f * W);
        int64_t r;
        if (kind == 0) {
            r = bd_ext_add_static_circle(s, bd_vec2{cx, cy}, R);
        } else {
            uint32_t m = 3 + (uint32_t)(rng.next() % 6);
            bd_vec2 v[8];
            for (uint32_t j = 0; j < m; j++) {
                float a = (6.2831853f * (float)j + rng.uniform(-0.3f, 0.3f) * 3.14159265f) / (float)m;
                v[j] = bd_vec2{cx + R * cosf(a), cy + R * sinf(a)};
            }
            r = bd_ext_add_static_polygon(s, v, m);
        }
        if (r < 0) return r;
    }
    return (int64_t)side * side;
}

}  // extern "C"
