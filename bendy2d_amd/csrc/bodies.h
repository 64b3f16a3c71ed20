// bodies.h -- the per-body phase shared by the engines: EXT4 static
// colliders, then bounds (src/solver.rs:155-165), then Verlet
// (src/solver.rs:118-128).
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

}  // namespace bd
