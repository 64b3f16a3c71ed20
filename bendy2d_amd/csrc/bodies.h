// bodies.h -- the per-body phase shared by the engines: EXT4 static
// colliders, then bounds (src/solver.rs:155-165), then Verlet
// (src/solver.rs:118-128); and the edge-against-point push of soft polygons.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace bd {

struct BodyParams {
    float dt, gex, gey, gx, gy;
    float lox, loy, hix, hiy;
    uint32_t N, Np, Nc;
    uint32_t first_acc;
    uint32_t ncol, nplanes;
};

// Static colliders in insertion order against a body of radius r (0 for
// particles and polygon points).  hdr: (kind, plane offset, planes, -);
// geo: circle (cx, cy, R, -); pl: polygon edge planes (nx, ny, c, -).
__device__ __forceinline__ void static_colliders(float2 &p, float r, uint32_t ncol, const int4 *hdr,
                                                 const float4 *geo, const float4 *pl) {
    for (uint32_t k = 0; k < ncol; k++) {
        int4 h = hdr[k];
        if (h.x == 0) {
            float4 g = geo[k];
            static_circle(p.x, p.y, r, g.x, g.y, g.z);
        } else {
            // the push uses the plane of largest signed distance, and only when
            // that distance is below r; a plane at r or beyond already rules
            // the push out, so the scan stops there (a NaN distance, first or
            // later, never pushes either way)
            float smax = 0.0f;
            int arg = -1;
            for (int e = 0; e < h.z; e++) {
                float4 w = pl[h.y + e];
                float sd = (w.x * p.x + w.y * p.y) - w.z;
                if (sd >= r) { arg = -1; break; }
                if (arg < 0 || sd > smax) { smax = sd; arg = e; }
            }
            if (arg >= 0 && smax < r) {
                float4 w = pl[h.y + arg];
                float push = r - smax;
                p.x = p.x + w.x * push; p.y = p.y + w.y * push;
            }
        }
    }
}

// A particle (not pinned): colliders, bounds, Verlet with acc = 0 + g.
__device__ __forceinline__ void particle_body(float2 &p, float2 &q, const BodyParams &P, const int4 *hdr,
                                              const float4 *geo, const float4 *pl) {
    static_colliders(p, 0.0f, P.ncol, hdr, geo, pl);
    bounds_axis(p.x, q.x, P.lox, P.hix);
    bounds_axis(p.y, q.y, P.loy, P.hiy);
    verlet(p.x, p.y, q.x, q.y, P.gex, P.gey, P.dt);
}

// ---- soft polygons ----------------------------------------------------------

// common::line_intersection (src/common.rs:4-26)
__device__ __forceinline__ bool line_intersection(float2 p1, float2 p2, float2 p3, float2 p4, float2 &out) {
    float s1x = p2.x - p1.x, s1y = p2.y - p1.y;
    float s2x = p4.x - p3.x, s2y = p4.y - p3.y;
    float den = (-s2x) * s1y + s1x * s2y;
    float s = ((-s1y) * (p1.x - p3.x) + s1x * (p1.y - p3.y)) / den;
    float t = (s2x * (p1.y - p3.y) - s2y * (p1.x - p3.x)) / den;
    if (s >= 0.0f && s <= 1.0f && t >= 0.0f && t <= 1.0f) {
        out = make_float2(p1.x + (t * s1x), p1.y + (t * s1y));
        return true;
    }
    return false;
}

__device__ __forceinline__ float2 f2sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 f2add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 f2mul(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float f2dot(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float2 f2normalize(float2 a) {
    float n = sqrtf(f2dot(a, a));
    return make_float2(a.x / n, a.y / n);
}
__device__ __forceinline__ float f2mag(float2 a) { return sqrtf(f2dot(a, a)); }

// Polygon::resolve_line_intersection (src/polygon.rs:164-216)
__device__ __forceinline__ bool resolve_line(float2 self_center, float2 pa, float2 pb, float2 op, float2 other_center,
                             float2 &na, float2 &nb, float2 &no) {
    float2 inter;
    if (!line_intersection(pa, pb, op, other_center, inter)) return false;
    float2 nl = f2normalize(f2sub(pb, pa));                                    // :175
    float2 cp = f2mul(nl, f2dot(nl, f2sub(self_center, inter)) / f2dot(nl, nl));  // :177-179
    float2 nin = f2normalize(f2sub(self_center, f2add(inter, cp)));           // :181
    float da = f2mag(f2sub(inter, pa));                                       // :183
    float db = f2mag(f2sub(inter, pb));                                       // :184
    float dab = da + db;                                                      // :186
    float ia = db / dab, ib = da / dab;                                       // :188-189
    float2 diff = f2sub(inter, op);                                           // :191
    float2 ion = f2mul(nin, f2dot(nin, diff) / f2dot(nin, nin));             // :193-194
    float2 third = make_float2(ion.x / 3.0f, ion.y / 3.0f);                   // :196
    float2 line = f2mul(third, 2.0f);                                         // :198
    float2 da_v = f2mul(line, ia), db_v = f2mul(line, ib);                    // :200-201
    float2 ra = f2sub(pa, da_v), rb = f2sub(pb, db_v);                        // :203-204
    float2 ray_end = f2sub(op, f2mul(nin, 10000.0f));                         // :208
    float2 np;
    if (!line_intersection(pa, pb, op, ray_end, np)) return false;            // :206-213
    na = ra; nb = rb; no = np;
    return true;
}

}  // namespace bd
