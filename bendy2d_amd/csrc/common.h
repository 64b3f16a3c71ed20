// common.h -- records and arithmetic shared by the host runtime and every
// engine of the MI355X bendy2d solver step.
//
// The arithmetic helpers restate the reference's f32 expressions operation for
// operation (Rust evaluates left to right and never contracts a*b+c into an
// FMA); every translation unit that includes this file is compiled with
// -ffp-contract=off and without fast-math, and HIP's default correctly
// rounded sqrtf and division are kept.
#pragma once

#include <cstdint>
#include <cstddef>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BD_HD __host__ __device__ __forceinline__
#else
#define BD_HD inline
#endif

namespace bd {

struct V2 { float x, y; };

// Device link record, 12 bytes, sorted by dependency layer.
//   a: bits 0-29 point index, bit 30 = circle link (r^2-weighted law),
//      bit 31 = end a pinned.
//   b: bits 0-29 point index, bit 31 = end b pinned.
struct Link12 { uint32_t a, b; float t; };

constexpr uint32_t kIndexMask = 0x3fffffffu;
constexpr uint32_t kCircleLinkBit = 0x40000000u;
constexpr uint32_t kPinBit = 0x80000000u;
constexpr uint64_t kMaxPoints = (1ull << 30) - 1;

// ParticleLink::solve (src/link.rs:18-27) with the pin extension: a pinned
// end does not move and the free end takes the whole correction; two free
// ends use the reference's 0.5 split.
BD_HD void solve_particle_link(float &ax, float &ay, float &bx, float &by, float t, bool pin_a, bool pin_b) {
    float dx = ax - bx, dy = ay - by;
    float dist = sqrtf(dx * dx + dy * dy);
    float nx = dx / dist, ny = dy / dist;
    float e = dist - t;
    if (!pin_a && !pin_b) {
        float cx = (nx * e) * 0.5f, cy = (ny * e) * 0.5f;
        ax = ax - cx; ay = ay - cy;
        bx = bx + cx; by = by + cy;
    } else if (!pin_a) {
        ax = ax - nx * e; ay = ay - ny * e;
    } else if (!pin_b) {
        bx = bx + nx * e; by = by + ny * e;
    }
}

// Reference link law for two free ends (src/link.rs:22-26).
BD_HD void solve_particle_link_free(float &ax, float &ay, float &bx, float &by, float t) {
    float dx = ax - bx, dy = ay - by;
    float dist = sqrtf(dx * dx + dy * dy);
    float nx = dx / dist, ny = dy / dist;
    float e = dist - t;
    float cx = (nx * e) * 0.5f, cy = (ny * e) * 0.5f;
    ax = ax - cx; ay = ay - cy;
    bx = bx + cx; by = by + cy;
}

// CircleLink::solve (src/link.rs:36-48) with ra*ra, rb*rb and
// 1.0f / (ra*ra + rb*rb) computed once by the caller: the same f32 values, so
// the same results as solve_circle_link.
BD_HD void solve_circle_link_w(float &ax, float &ay, float &bx, float &by, float t, float ra2, float rb2,
                               float scale) {
    float dx = ax - bx, dy = ay - by;
    float dist = sqrtf(dx * dx + dy * dy);
    float nx = dx / dist, ny = dy / dist;
    float e = dist - t;
    ax = ax - ((nx * e) * scale) * rb2; ay = ay - ((ny * e) * scale) * rb2;
    bx = bx + ((nx * e) * scale) * ra2; by = by + ((ny * e) * scale) * ra2;
}

// CircleLink::solve (src/link.rs:36-48).
BD_HD void solve_circle_link(float &ax, float &ay, float &bx, float &by, float t, float ra, float rb) {
    float dx = ax - bx, dy = ay - by;
    float dist = sqrtf(dx * dx + dy * dy);
    float nx = dx / dist, ny = dy / dist;
    float ra2 = ra * ra, rb2 = rb * rb;
    float scale = 1.0f / (ra2 + rb2);
    float e = dist - t;
    ax = ax - ((nx * e) * scale) * rb2; ay = ay - ((ny * e) * scale) * rb2;
    bx = bx + ((nx * e) * scale) * ra2; by = by + ((ny * e) * scale) * ra2;
}

// Circle::solve_circle (src/circle.rs:32-45).  Returns true when the pair
// overlapped; push_i / push_j receive the length of each circle's move.
BD_HD bool solve_circle_pair(float &ix, float &iy, float &jx, float &jy, float ri, float rj,
                             float &push_i, float &push_j) {
    float dx = ix - jx, dy = iy - jy;
    float q = dx * dx + dy * dy;
    float rs = ri + rj;
    if (q < rs * rs) {
        float n = sqrtf(q);
        float nx = dx / n, ny = dy / n;
        float ov = rs - sqrtf(q);
        float ri2 = ri * ri, rj2 = rj * rj;
        float scale = 1.0f / (ri2 + rj2);
        ix = ix + ((nx * scale) * ov) * rj2; iy = iy + ((ny * scale) * ov) * rj2;
        jx = jx - ((nx * scale) * ov) * ri2; jy = jy - ((ny * scale) * ov) * ri2;
        push_i = fabsf((scale * ov) * rj2);
        push_j = fabsf((scale * ov) * ri2);
        return true;
    }
    return false;
}

// Largest distance from a sweep's start position (sx, sy) reached so far,
// after a push moved the circle to (px, py).  The distance is rounded up by a
// relative 2^-20, far above float rounding, so the bound never falls below the
// truth.  A circle pushed back and forth keeps a small excursion where the sum
// of its pushes would grow without end.  A non-finite position leaves the
// bound unchanged: such a circle takes part in no pair.
BD_HD float excursion(float cur, float px, float py, float sx, float sy) {
    float dx = px - sx, dy = py - sy;
    float d = sqrtf(dx * dx + dy * dy) * (1.0f + 0x1p-20f);
    return fmaxf(cur, d);
}

// Particle::update (src/particle.rs:20-25): pos + vel + acc*dt*dt, acc = g
// (plus any initial acc on the first substep).
BD_HD void verlet(float &px, float &py, float &qx, float &qy, float ax, float ay, float dt) {
    float vx = px - qx, vy = py - qy;
    qx = px; qy = py;
    px = (px + vx) + ((ax * dt) * dt);
    py = (py + vy) + ((ay * dt) * dt);
}

// Particle::solve_bounds (src/particle.rs:27-46): lo/hi are bounds.pos and
// bounds.pos + bounds.size, per axis, if / else-if.
BD_HD void bounds_axis(float &p, float &q, float lo, float hi) {
    if (p < lo) {
        float v = q - p; q = lo - v; p = lo;
    } else if (p > hi) {
        float v = q - p; q = hi - v; p = hi;
    }
}

// Circle::solve_bounds (src/circle.rs:11-30): limits shifted by the radius,
// lo + r and (lo + size) - r.
BD_HD void circle_bounds_axis(float &p, float &q, float lo, float hi, float r) {
    float l = lo + r, h = hi - r;
    if (p < l) {
        float v = q - p; q = l - v; p = l;
    } else if (p > h) {
        float v = q - p; q = h - v; p = h;
    }
}

// EXT4 static circle collider: a body of radius r is pushed out of a fixed
// circle (cx, cy, R) by the whole overlap.  A body exactly at the centre is
// left in place.
BD_HD void static_circle(float &px, float &py, float r, float cx, float cy, float R) {
    float dx = px - cx, dy = py - cy;
    float q = dx * dx + dy * dy;
    float rs = r + R;
    if (q < rs * rs) {
        float d = sqrtf(q);
        if (d > 0.0f) {
            float nx = dx / d, ny = dy / d, ov = rs - d;
            px = px + nx * ov; py = py + ny * ov;
        }
    }
}

// Cell hashing for the circle-pair broadphase.
BD_HD uint32_t cell_hash(int32_t cx, int32_t cy) {
    return ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u);
}

}  // namespace bd
