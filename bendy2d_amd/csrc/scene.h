// scene.h -- host-side model of one solver's state and the device plan built
// from it.
//
// HostScene mirrors the fields of the reference Solver (src/solver.rs:19-31)
// plus the extensions; it is the authoritative copy until a plan is uploaded,
// after which the device copy is authoritative and the host copy is refreshed
// lazily on readback.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"

namespace bd {

// A link as the user gave it (src/link.rs:4-10).  Indices that do not fit in
// 32 bits are stored as kBadIndex, which no scene can hold, so validation
// rejects them.
struct LinkH { uint32_t a, b; float t; };
constexpr uint32_t kBadIndex = 0xffffffffu;

// One reference Polygon (src/polygon.rs:7-14): points [off, off+n) of the
// g_* arrays, links [loff, loff+nl) of glinks with polygon-local indices.
struct Poly {
    uint32_t off, n, loff, nl;
    int32_t is_static;
    V2 center;
    float scale;
};

// EXT4 static collider.  kind 0: circle (cx, cy, r).  kind 1: convex polygon
// whose nv outward edge planes (nx, ny, c) start at planes[3*plane_off].
struct Collider {
    int32_t kind;
    float cx, cy, r;
    uint32_t plane_off, nv;
};

struct HostScene {
    V2 gravity{0.0f, 98.2f};                 // src/solver.rs:36
    V2 bpos{0.0f, 0.0f}, bsize{100.0f, 100.0f};  // src/solver.rs:37-40
    int32_t bounds_active = 1;               // src/solver.rs:41
    uint32_t sub_steps = 1;                  // src/solver.rs:47 (EXT1)
    uint32_t link_iters = 1;                 // EXT2

    // particles (Particle::new: prev = pos, acc = 0)
    std::vector<V2> p_pos, p_prev;
    std::vector<uint8_t> p_pin;              // EXT3
    std::vector<LinkH> plinks;

    // circles
    std::vector<V2> c_pos, c_prev, c_acc;
    std::vector<float> c_r;
    std::vector<LinkH> clinks;

    // polygons
    std::vector<Poly> polys;
    std::vector<V2> g_pos, g_prev, g_acc;
    std::vector<LinkH> glinks;

    // EXT4 static colliders
    std::vector<Collider> cols;
    std::vector<float> planes;  // 3 floats (nx, ny, c) per polygon edge
    std::vector<float> verts;   // 2 floats per polygon vertex as given, same offsets

    uint64_t n_points() const { return p_pos.size() + c_pos.size() + g_pos.size(); }
    uint64_t n_links() const { return plinks.size() + clinks.size() + glinks.size(); }
    bool any_pinned() const;
    bool any_initial_acc() const;  // circle or polygon point with acc != 0
};

// Device-ready plan: unified point indexing
//   [0, Np) particles, [Np, Np+Nc) circle points, [Np+Nc, N) polygon points,
// and every link of one sweep in dependency-layer order.
struct Plan {
    uint32_t Np = 0, Nc = 0, Ng = 0, N = 0;
    std::vector<Link12> links;         // sorted by (layer, sequence index)
    std::vector<uint32_t> layer_off;   // depth + 1 offsets into links
    uint32_t depth = 0;
    // index of the first layer that holds a polygon link (polygon centres are
    // computed before it in the last sweep); depth if there are none
    uint32_t first_poly_layer = 0;
};

// Validates every link (a < b < group size, src/link.rs:19-21,37-39) and
// builds the plan.  Returns BD_OK or BD_ERR_LINK / BD_ERR_CAPACITY, with a
// message in err.
int build_plan(const HostScene &s, Plan &plan, std::string &err);

// Order-preserving dependency layering of a link sequence.  layer[i] is
// 1 + the largest layer among earlier links sharing an end (0-based), so
// links of one layer share no point and applying the layers in order repeats
// the sequential sweep exactly.  order receives the link indices sorted by
// (layer, index).  Returns the number of layers.
uint32_t plan_layers(const uint32_t *a, const uint32_t *b, uint64_t n_links, uint64_t n_points,
                     uint32_t *layer, uint64_t *order);

// Static polygon collider planes: counter-clockwise edges (clockwise input is
// reversed), outward normal n = (e.y/|e|, -e.x/|e|), offset c = n . v_i.
void polygon_planes(const V2 *v, uint64_t n, std::vector<float> &out);

// Polygon::circle (src/polygon.rs:17-82) and Polygon::new (src/polygon.rs:84-123).
void polygon_circle(float radius, V2 pos, uint64_t count, std::vector<V2> &pts,
                    std::vector<LinkH> &links, V2 &center);
void polygon_new(const V2 *points, uint64_t n, std::vector<V2> &pts, std::vector<LinkH> &links,
                 V2 &center);

}  // namespace bd
