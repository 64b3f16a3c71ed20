// solver.cpp -- the C ABI of include/bendy2d.h over the host scene and the
// device engines.
//
// State model: HostScene is authoritative until the first bd_update after a
// structural change (an add, a pin); the plan is then rebuilt and uploaded and
// the device copy becomes authoritative.  Getters and structural changes first
// copy device positions back (lazy readback).  Scalar fields (gravity, bounds,
// sub_steps, link iterations) are passed with every update and need no upload.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bendy2d.h"
#include "engine.h"
#include "scene.h"
#include "tiled_plan.h"

using namespace bd;

namespace bd {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
static const std::string &bd_last_error_str() { return g_last_error; }

StepParams make_params(const HostScene &s, float dt, bool acc_pending) {
    StepParams p;
    float mult = 1.0f / (float)s.sub_steps;   // src/solver.rs:107
    p.delta = dt * mult;                      // src/solver.rs:108
    p.sub_steps = s.sub_steps;
    p.link_iters = s.link_iters;
    p.gx = s.gravity.x; p.gy = s.gravity.y;
    p.gex = 0.0f + s.gravity.x;               // acc = (0,0) + g (src/particle.rs:52-54)
    p.gey = 0.0f + s.gravity.y;
    p.lox = s.bpos.x; p.loy = s.bpos.y;
    p.hix = s.bpos.x + s.bsize.x;             // bounds.pos.x + bounds.size.x (src/particle.rs:221)
    p.hiy = s.bpos.y + s.bsize.y;
    p.first_acc = acc_pending;
    return p;
}

}  // namespace bd

struct bd_solver {
    HostScene host;
    Plan plan;
    bool plan_valid = false;
    bool dev_newer = false;    // device positions newer than host copy
    bool acc_pending = false;  // initial circle / polygon-point acc not yet consumed
    std::vector<uint8_t> center_from_prev;  // per polygon: centre = sum(prev)/n after an update
    std::unique_ptr<Engine> eng;
    int32_t engine_pref = BD_ENGINE_AUTO;
    int32_t engine_used = 0;
    // TILED shape: 2,048-point tiles, two workgroups of 1,024 per CU on the
    // 1024x1024 cloth (measured fastest of 1,024..8,192 points x 256..1,024 lanes)
    uint32_t tile_points = 2048, tile_wg = 1024;
    hipStream_t stream = nullptr;
    int device = -1;
    Stats stats;
    // Readback (bd_ext_set_readback): after each update, a device-side copy of
    // the state into one of two staging buffers on the solver's stream, then a
    // copy to pinned host memory on a second stream, which overlaps the next
    // update's kernels.  `slot` holds the latest update's copy when `valid`.
    struct Readback {
        bool on = false, valid = false;
        int slot = -1;
        DevBuf<float2> dev[2];
        float2 *host[2] = {nullptr, nullptr};
        size_t host_cap[2] = {0, 0};
        hipEvent_t snap[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
        bool used[2] = {false, false};
        hipStream_t copy = nullptr;
    } rb;
    ~bd_solver() {
        for (int k = 0; k < 2; k++) {
            if (rb.done[k]) (void)hipEventSynchronize(rb.done[k]);
            if (rb.host[k]) (void)hipHostFree(rb.host[k]);
            if (rb.snap[k]) (void)hipEventDestroy(rb.snap[k]);
            if (rb.done[k]) (void)hipEventDestroy(rb.done[k]);
        }
        if (rb.copy) (void)hipStreamDestroy(rb.copy);
    }
};

static int fail(int code, const std::string &msg) {
    set_last_error(msg);
    return code;
}

// Polygon::update recomputes the centre from the positions it is about to
// integrate (src/polygon.rs:130), which are prev afterwards.
static void centre_from_prev(HostScene &h, size_t p) {
    Poly &pg = h.polys[p];
    float cx = 0.0f, cy = 0.0f;
    for (uint32_t i = 0; i < pg.n; i++) {
        cx = cx + h.g_prev[pg.off + i].x;
        cy = cy + h.g_prev[pg.off + i].y;
    }
    pg.center = V2{cx / (float)pg.n, cy / (float)pg.n};
}

// Copy device state back into the host scene if the device is newer: from
// the readback copy of the latest update when there is one, else directly.
static int ensure_host(const bd_solver *cs) {
    bd_solver *s = const_cast<bd_solver *>(cs);
    if (!s->dev_newer || !s->eng) return BD_OK;
    std::string err;
    int rc;
    if (s->rb.valid) {
        const int k = s->rb.slot;
        if (hipEventSynchronize(s->rb.done[k]) != hipSuccess) return fail(BD_ERR_DEVICE, "readback copy failed");
        rc = s->eng->unpack(s->rb.host[k], s->host, err);
    } else {
        rc = s->eng->download(s->host, err);
    }
    if (rc) return fail(rc, err);
    for (size_t p = 0; p < s->host.polys.size(); p++)
        if (!s->host.polys[p].is_static && p < s->center_from_prev.size() && s->center_from_prev[p])
            centre_from_prev(s->host, p);
    s->dev_newer = false;
    return BD_OK;
}

// Structural change: host must be current and the plan is rebuilt next update.
// When the device copy cannot be brought back, nothing changes and the error
// is returned (the host copy would be stale, and the device copy would no
// longer match the host scene's shape).
static int begin_change(bd_solver *s) {
    int rc = ensure_host(s);
    if (rc) return rc;
    s->plan_valid = false;
    return BD_OK;
}

#define BD_BEGIN_CHANGE(s)                   \
    do {                                     \
        int rc_ = begin_change(s);           \
        if (rc_) return rc_;                 \
    } while (0)

static uint32_t idx32(uint64_t v) { return v >= 0xffffffffull ? kBadIndex : (uint32_t)v; }

namespace bd {

// Internal accessors for the batched form and the server.
int solver_export(const bd_solver *s, HostScene &out, std::string &err) {
    int rc = ensure_host(s);
    if (rc) { err = bd_last_error_str(); return rc; }
    out = s->host;
    return BD_OK;
}

// Overwrite the body state of `s` with that of `src` (same scene shape), as
// left by `updates` updates run elsewhere with src.gravity and src.sub_steps
// from the acc in src.c_acc / src.g_acc.  src.polys holds each polygon's
// last centre.
int solver_import_state(bd_solver *s, const HostScene &src, uint64_t updates, std::string &err) {
    int rc = ensure_host(s);
    if (rc) { err = bd_last_error_str(); return rc; }
    HostScene &h = s->host;
    if (h.p_pos.size() != src.p_pos.size() || h.c_pos.size() != src.c_pos.size() ||
        h.g_pos.size() != src.g_pos.size() || h.polys.size() != src.polys.size()) {
        err = "scene shape differs from the batch scene";
        return BD_ERR_ARG;
    }
    h.p_pos = src.p_pos; h.p_prev = src.p_prev;
    h.c_pos = src.c_pos; h.c_prev = src.c_prev;
    h.g_pos = src.g_pos; h.g_prev = src.g_prev;
    if (updates) {
        // what settle_acc does after a local update: acc spent, except static
        // polygon points, which gain gravity every substep
        const uint64_t substeps = updates * src.sub_steps;
        for (V2 &a : h.c_acc) a = V2{0.0f, 0.0f};
        for (size_t p = 0; p < h.polys.size(); p++) {
            const Poly &pg = h.polys[p];
            for (uint32_t i = 0; i < pg.n; i++) {
                V2 &a = h.g_acc[pg.off + i];
                if (pg.is_static) {
                    a = src.g_acc[pg.off + i];
                    for (uint64_t k = 0; k < substeps; k++) a = V2{a.x + src.gravity.x, a.y + src.gravity.y};
                } else {
                    a = V2{0.0f, 0.0f};
                }
            }
            if (pg.is_static) h.polys[p].center = src.polys[p].center;
            else centre_from_prev(h, p);
            if (p < s->center_from_prev.size()) s->center_from_prev[p] = pg.is_static ? 0 : 1;
        }
        s->acc_pending = false;
    }
    s->plan_valid = false;  // device copy (if any) is stale now
    s->dev_newer = false;
    return BD_OK;
}

const HostScene &solver_host(const bd_solver *s) { return s->host; }

}  // namespace bd

extern "C" {

const char *bd_error_string(int32_t code) {
    switch (code) {
    case BD_OK: return "ok";
    case BD_ERR_ARG: return "invalid argument";
    case BD_ERR_LINK: return "invalid link (particle_a must be < particle_b < count)";
    case BD_ERR_INDEX: return "index out of range";
    case BD_ERR_DEVICE: return "HIP device error";
    case BD_ERR_NOMEM: return "out of memory";
    case BD_ERR_CAPACITY: return "engine capacity exceeded";
    case BD_ERR_COLLISION: return "circle pair schedule did not converge";
    case BD_ERR_NODEVICE: return "no HIP device visible";
    case BD_ERR_BUDGET: return "memory budget exceeded";
    case BD_ERR_STATE: return "invalid state";
    default: return "unknown error";
    }
}

const char *bd_last_error(void) { return g_last_error.c_str(); }

const char *bd_version(void) { return "bendy2d_amd 0.1 (gfx950)"; }

int32_t bd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

bd_solver *bd_solver_new(void) {
    try {
        return new bd_solver();
    } catch (...) {
        set_last_error("allocation failed");
        return nullptr;
    }
}

void bd_solver_free(bd_solver *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->eng.reset();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

bd_solver *bd_solver_clone(const bd_solver *s) {
    if (!s) return nullptr;
    if (ensure_host(s)) return nullptr;
    bd_solver *c = bd_solver_new();
    if (!c) return nullptr;
    c->host = s->host;
    c->acc_pending = false;
    c->center_from_prev = s->center_from_prev;
    c->engine_pref = s->engine_pref;
    return c;
}

void bd_set_gravity(bd_solver *s, bd_vec2 g) { if (s) s->host.gravity = V2{g.x, g.y}; }
bd_vec2 bd_get_gravity(const bd_solver *s) { return s ? bd_vec2{s->host.gravity.x, s->host.gravity.y} : bd_vec2{0, 0}; }
void bd_set_bounds(bd_solver *s, bd_bounds b) {
    if (!s) return;
    s->host.bpos = V2{b.pos.x, b.pos.y};
    s->host.bsize = V2{b.size.x, b.size.y};
}
bd_bounds bd_get_bounds(const bd_solver *s) {
    bd_bounds b{{0, 0}, {0, 0}};
    if (s) { b.pos = bd_vec2{s->host.bpos.x, s->host.bpos.y}; b.size = bd_vec2{s->host.bsize.x, s->host.bsize.y}; }
    return b;
}
void bd_set_bounds_active(bd_solver *s, int32_t a) { if (s) s->host.bounds_active = a ? 1 : 0; }
int32_t bd_get_bounds_active(const bd_solver *s) { return s ? s->host.bounds_active : 0; }

int64_t bd_add_particle(bd_solver *s, bd_vec2 pos) {
    if (!s) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->host.p_pos.push_back(V2{pos.x, pos.y});
    s->host.p_prev.push_back(V2{pos.x, pos.y});
    s->host.p_pin.push_back(0);
    return (int64_t)s->host.p_pos.size() - 1;
}

int64_t bd_ext_add_particle_with_prev(bd_solver *s, bd_vec2 pos, bd_vec2 prev) {
    if (!s) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->host.p_pos.push_back(V2{pos.x, pos.y});
    s->host.p_prev.push_back(V2{prev.x, prev.y});
    s->host.p_pin.push_back(0);
    return (int64_t)s->host.p_pos.size() - 1;
}

int64_t bd_ext_add_particles(bd_solver *s, const bd_vec2 *pos, const bd_vec2 *prev, uint64_t n) {
    if (!s || (!pos && n)) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    int64_t first = (int64_t)s->host.p_pos.size();
    s->host.p_pos.reserve(first + n); s->host.p_prev.reserve(first + n);
    for (uint64_t i = 0; i < n; i++) {
        s->host.p_pos.push_back(V2{pos[i].x, pos[i].y});
        const bd_vec2 &q = prev ? prev[i] : pos[i];
        s->host.p_prev.push_back(V2{q.x, q.y});
    }
    s->host.p_pin.resize(first + n, 0);
    return first;
}

int64_t bd_add_circle(bd_solver *s, const bd_circle *c) {
    if (!s || !c) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->host.c_pos.push_back(V2{c->point.pos.x, c->point.pos.y});
    s->host.c_prev.push_back(V2{c->point.prev_pos.x, c->point.prev_pos.y});
    s->host.c_acc.push_back(V2{c->point.acc.x, c->point.acc.y});
    s->host.c_r.push_back(c->radius);
    return (int64_t)s->host.c_pos.size() - 1;
}

int64_t bd_ext_add_circles(bd_solver *s, const bd_circle *c, uint64_t n) {
    if (!s || (!c && n)) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    int64_t first = (int64_t)s->host.c_pos.size();
    for (uint64_t i = 0; i < n; i++) {
        s->host.c_pos.push_back(V2{c[i].point.pos.x, c[i].point.pos.y});
        s->host.c_prev.push_back(V2{c[i].point.prev_pos.x, c[i].point.prev_pos.y});
        s->host.c_acc.push_back(V2{c[i].point.acc.x, c[i].point.acc.y});
        s->host.c_r.push_back(c[i].radius);
    }
    return first;
}

int64_t bd_add_particle_link(bd_solver *s, const bd_link *l) {
    if (!s || !l) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->host.plinks.push_back(LinkH{idx32(l->particle_a), idx32(l->particle_b), l->target_distance});
    return (int64_t)s->host.plinks.size() - 1;
}

int64_t bd_add_circle_link(bd_solver *s, const bd_link *l) {
    if (!s || !l) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->host.clinks.push_back(LinkH{idx32(l->particle_a), idx32(l->particle_b), l->target_distance});
    return (int64_t)s->host.clinks.size() - 1;
}

static int64_t add_links_bulk(bd_solver *s, std::vector<LinkH> &dst, const uint32_t *a, const uint32_t *b,
                              const float *t, uint64_t n) {
    if (!s || ((!a || !b || !t) && n)) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    int64_t first = (int64_t)dst.size();
    dst.reserve(first + n);
    for (uint64_t i = 0; i < n; i++) dst.push_back(LinkH{a[i], b[i], t[i]});
    return first;
}

int64_t bd_ext_add_particle_links(bd_solver *s, const uint32_t *a, const uint32_t *b, const float *t, uint64_t n) {
    return s ? add_links_bulk(s, s->host.plinks, a, b, t, n) : BD_ERR_ARG;
}

int64_t bd_ext_add_circle_links(bd_solver *s, const uint32_t *a, const uint32_t *b, const float *t, uint64_t n) {
    return s ? add_links_bulk(s, s->host.clinks, a, b, t, n) : BD_ERR_ARG;
}

int64_t bd_add_polygon(bd_solver *s, const bd_polygon_in *p) {
    if (!s || !p || (!p->particles && p->n_particles) || (!p->particle_links && p->n_links)) return BD_ERR_ARG;
    if (p->n_particles >= 0xffffffffull || p->n_links >= 0xffffffffull) return fail(BD_ERR_CAPACITY, "polygon too large");
    BD_BEGIN_CHANGE(s);
    HostScene &h = s->host;
    Poly pg;
    pg.off = (uint32_t)h.g_pos.size();
    pg.n = (uint32_t)p->n_particles;
    pg.loff = (uint32_t)h.glinks.size();
    pg.nl = (uint32_t)p->n_links;
    pg.is_static = p->is_static ? 1 : 0;
    pg.center = V2{p->center.x, p->center.y};
    pg.scale = p->scale;
    for (uint64_t i = 0; i < p->n_particles; i++) {
        const bd_particle &q = p->particles[i];
        h.g_pos.push_back(V2{q.pos.x, q.pos.y});
        h.g_prev.push_back(V2{q.prev_pos.x, q.prev_pos.y});
        h.g_acc.push_back(V2{q.acc.x, q.acc.y});
    }
    for (uint64_t i = 0; i < p->n_links; i++) {
        const bd_link &l = p->particle_links[i];
        h.glinks.push_back(LinkH{idx32(l.particle_a), idx32(l.particle_b), l.target_distance});
    }
    h.polys.push_back(pg);
    s->center_from_prev.push_back(0);
    return (int64_t)h.polys.size() - 1;
}

int32_t bd_polygon_circle(float radius, bd_vec2 pos, uint64_t point_count, int32_t is_static,
                          bd_particle *out_particles, bd_link *out_links, bd_vec2 *out_center) {
    (void)is_static;
    if ((!out_particles || !out_links) && point_count) return BD_ERR_ARG;
    std::vector<V2> pts;
    std::vector<LinkH> links;
    V2 c;
    polygon_circle(radius, V2{pos.x, pos.y}, point_count, pts, links, c);
    for (uint64_t i = 0; i < pts.size(); i++)
        out_particles[i] = bd_particle{{pts[i].x, pts[i].y}, {pts[i].x, pts[i].y}, {0.0f, 0.0f}};
    for (uint64_t i = 0; i < links.size(); i++) out_links[i] = bd_link{links[i].a, links[i].b, links[i].t};
    if (out_center) *out_center = bd_vec2{c.x, c.y};
    return BD_OK;
}

int32_t bd_polygon_new(const bd_vec2 *points, uint64_t n, int32_t is_static, bd_particle *out_particles,
                       bd_link *out_links, bd_vec2 *out_center) {
    (void)is_static;
    if ((!points || !out_particles || !out_links) && n) return BD_ERR_ARG;
    std::vector<V2> in(n), pts;
    for (uint64_t i = 0; i < n; i++) in[i] = V2{points[i].x, points[i].y};
    std::vector<LinkH> links;
    V2 c;
    polygon_new(in.data(), n, pts, links, c);
    for (uint64_t i = 0; i < pts.size(); i++)
        out_particles[i] = bd_particle{{pts[i].x, pts[i].y}, {pts[i].x, pts[i].y}, {0.0f, 0.0f}};
    for (uint64_t i = 0; i < links.size(); i++) out_links[i] = bd_link{links[i].a, links[i].b, links[i].t};
    if (out_center) *out_center = bd_vec2{c.x, c.y};
    return BD_OK;
}

uint64_t bd_get_particle_len(const bd_solver *s) { return s ? s->host.p_pos.size() : 0; }
uint64_t bd_get_circles_len(const bd_solver *s) { return s ? s->host.c_pos.size() : 0; }
uint64_t bd_get_polygons_len(const bd_solver *s) { return s ? s->host.polys.size() : 0; }

static uint64_t copy_links(const std::vector<LinkH> &v, uint64_t begin, uint64_t count, bd_link *out, uint64_t cap) {
    if (out)
        for (uint64_t i = 0; i < count && i < cap; i++) {
            const LinkH &l = v[begin + i];
            out[i] = bd_link{l.a == kBadIndex ? ~0ull : l.a, l.b == kBadIndex ? ~0ull : l.b, l.t};
        }
    return count;
}

uint64_t bd_get_particle_links(const bd_solver *s, bd_link *out, uint64_t cap) {
    return s ? copy_links(s->host.plinks, 0, s->host.plinks.size(), out, cap) : 0;
}
uint64_t bd_get_circle_links(const bd_solver *s, bd_link *out, uint64_t cap) {
    return s ? copy_links(s->host.clinks, 0, s->host.clinks.size(), out, cap) : 0;
}

// acc of a body after updates: 0 except static polygon points, which keep
// accumulating gravity (src/polygon.rs:126-128, 225-229).
uint64_t bd_get_particles(const bd_solver *s, bd_particle *out, uint64_t cap) {
    if (!s) return 0;
    if (ensure_host(s)) return 0;
    const HostScene &h = s->host;
    if (out)
        for (uint64_t i = 0; i < h.p_pos.size() && i < cap; i++)
            out[i] = bd_particle{{h.p_pos[i].x, h.p_pos[i].y}, {h.p_prev[i].x, h.p_prev[i].y}, {0.0f, 0.0f}};
    return h.p_pos.size();
}

uint64_t bd_get_circles(const bd_solver *s, bd_circle *out, uint64_t cap) {
    if (!s) return 0;
    if (ensure_host(s)) return 0;
    const HostScene &h = s->host;
    if (out)
        for (uint64_t i = 0; i < h.c_pos.size() && i < cap; i++)
            out[i] = bd_circle{{{h.c_pos[i].x, h.c_pos[i].y}, {h.c_prev[i].x, h.c_prev[i].y}, {h.c_acc[i].x, h.c_acc[i].y}},
                               h.c_r[i]};
    return h.c_pos.size();
}

int32_t bd_get_particle(const bd_solver *s, uint64_t i, bd_particle *out) {
    if (!s || i >= s->host.p_pos.size()) return 0;
    if (ensure_host(s)) return 0;
    const HostScene &h = s->host;
    if (out) *out = bd_particle{{h.p_pos[i].x, h.p_pos[i].y}, {h.p_prev[i].x, h.p_prev[i].y}, {0.0f, 0.0f}};
    return 1;
}

int32_t bd_get_circle(const bd_solver *s, uint64_t i, bd_circle *out) {
    if (!s || i >= s->host.c_pos.size()) return 0;
    if (ensure_host(s)) return 0;
    const HostScene &h = s->host;
    if (out)
        *out = bd_circle{{{h.c_pos[i].x, h.c_pos[i].y}, {h.c_prev[i].x, h.c_prev[i].y}, {h.c_acc[i].x, h.c_acc[i].y}},
                         h.c_r[i]};
    return 1;
}

int32_t bd_get_polygon(const bd_solver *s, uint64_t i, bd_polygon_info *info) {
    if (!s || i >= s->host.polys.size()) return 0;
    if (ensure_host(s)) return 0;
    const Poly &p = s->host.polys[i];
    if (info) *info = bd_polygon_info{p.n, p.nl, p.is_static, {p.center.x, p.center.y}, p.scale};
    return 1;
}

uint64_t bd_get_polygon_particles(const bd_solver *s, uint64_t i, bd_particle *out, uint64_t cap) {
    if (!s || i >= s->host.polys.size()) return 0;
    if (ensure_host(s)) return 0;
    const HostScene &h = s->host;
    const Poly &p = h.polys[i];
    if (out)
        for (uint64_t k = 0; k < p.n && k < cap; k++) {
            uint32_t g = p.off + (uint32_t)k;
            out[k] = bd_particle{{h.g_pos[g].x, h.g_pos[g].y}, {h.g_prev[g].x, h.g_prev[g].y}, {h.g_acc[g].x, h.g_acc[g].y}};
        }
    return p.n;
}

uint64_t bd_get_polygon_links(const bd_solver *s, uint64_t i, bd_link *out, uint64_t cap) {
    if (!s || i >= s->host.polys.size()) return 0;
    const Poly &p = s->host.polys[i];
    return copy_links(s->host.glinks, p.loff, p.nl, out, cap);
}

int32_t bd_ext_set_sub_steps(bd_solver *s, uint32_t n) {
    if (!s || n < 1 || n > 65535) return BD_ERR_ARG;  // sub_steps is a u16 (src/solver.rs:29)
    s->host.sub_steps = n;
    return BD_OK;
}
uint32_t bd_ext_get_sub_steps(const bd_solver *s) { return s ? s->host.sub_steps : 0; }

int32_t bd_ext_set_link_iterations(bd_solver *s, uint32_t k) {
    if (!s || k < 1) return BD_ERR_ARG;
    s->host.link_iters = k;
    return BD_OK;
}
uint32_t bd_ext_get_link_iterations(const bd_solver *s) { return s ? s->host.link_iters : 0; }

int32_t bd_ext_set_pinned(bd_solver *s, uint64_t i, int32_t pinned) {
    if (!s) return BD_ERR_ARG;
    if (i >= s->host.p_pos.size()) return BD_ERR_INDEX;
    BD_BEGIN_CHANGE(s);
    s->host.p_pin[i] = pinned ? 1 : 0;
    return BD_OK;
}

int32_t bd_ext_set_pinned_range(bd_solver *s, uint64_t first, const uint8_t *pinned, uint64_t n) {
    if (!s || (!pinned && n)) return BD_ERR_ARG;
    if (first + n > s->host.p_pos.size()) return BD_ERR_INDEX;
    BD_BEGIN_CHANGE(s);
    for (uint64_t i = 0; i < n; i++) s->host.p_pin[first + i] = pinned[i] ? 1 : 0;
    return BD_OK;
}

int64_t bd_ext_add_static_circle(bd_solver *s, bd_vec2 c, float r) {
    if (!s) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->host.cols.push_back(Collider{0, c.x, c.y, r, 0, 0});
    return (int64_t)s->host.cols.size() - 1;
}

int64_t bd_ext_add_static_polygon(bd_solver *s, const bd_vec2 *v, uint64_t n) {
    if (!s || !v || n < 3) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    std::vector<V2> pts(n);
    for (uint64_t i = 0; i < n; i++) pts[i] = V2{v[i].x, v[i].y};
    uint32_t off = (uint32_t)(s->host.planes.size() / 3);
    polygon_planes(pts.data(), n, s->host.planes);
    for (uint64_t i = 0; i < n; i++) { s->host.verts.push_back(v[i].x); s->host.verts.push_back(v[i].y); }
    s->host.cols.push_back(Collider{1, 0.0f, 0.0f, 0.0f, off, (uint32_t)n});
    return (int64_t)s->host.cols.size() - 1;
}

uint64_t bd_ext_get_static_collider_len(const bd_solver *s) { return s ? s->host.cols.size() : 0; }

int32_t bd_ext_get_static_collider(const bd_solver *s, uint64_t i, bd_vec2 *center, float *radius,
                                   bd_vec2 *vertices, uint64_t cap, uint64_t *n_vertices) {
    if (!s) return BD_ERR_ARG;
    if (i >= s->host.cols.size()) return BD_ERR_INDEX;
    const Collider &c = s->host.cols[i];
    if (center) *center = bd_vec2{c.cx, c.cy};
    if (radius) *radius = c.r;
    if (n_vertices) *n_vertices = c.kind == 1 ? c.nv : 0;
    if (c.kind == 1 && vertices)
        for (uint32_t k = 0; k < c.nv && k < cap; k++)
            vertices[k] = bd_vec2{s->host.verts[2 * (c.plane_off + k)], s->host.verts[2 * (c.plane_off + k) + 1]};
    return c.kind;
}

int32_t bd_ext_get_pinned_range(const bd_solver *s, uint64_t first, uint8_t *out, uint64_t n) {
    if (!s || (!out && n)) return BD_ERR_ARG;
    if (first + n > s->host.p_pin.size()) return BD_ERR_INDEX;
    for (uint64_t k = 0; k < n; k++) out[k] = s->host.p_pin[first + k];
    return BD_OK;
}

int32_t bd_ext_set_engine(bd_solver *s, int32_t engine) {
    if (!s || engine < BD_ENGINE_AUTO || engine > BD_ENGINE_TILED) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->engine_pref = engine;
    return BD_OK;
}

int32_t bd_ext_get_engine_used(const bd_solver *s) { return s ? s->engine_used : 0; }

int32_t bd_ext_set_tiling(bd_solver *s, uint32_t tile_points, uint32_t workgroup) {
    if (!s || tile_points < 64 || tile_points > 16384) return BD_ERR_ARG;
    if (workgroup != 256 && workgroup != 512 && workgroup != 1024) return BD_ERR_ARG;
    BD_BEGIN_CHANGE(s);
    s->tile_points = tile_points;
    s->tile_wg = workgroup;
    return BD_OK;
}

int64_t bd_ext_tiled_selfcheck(const bd_solver *s, uint32_t tile_points, uint32_t region_cap, uint32_t *passes) {
    if (!s) return BD_ERR_ARG;
    int rc = ensure_host(s);
    if (rc) return rc;
    Plan p;
    std::string err;
    rc = build_plan(s->host, p, err);
    if (rc) return fail(rc, err);
    TilePlan tp;
    rc = build_tile_plan(s->host, p, tile_points, region_cap, tp, err);
    if (rc) return fail(rc, err);
    if (passes) *passes = (uint32_t)tp.passes.size();
    return (int64_t)tile_plan_selfcheck(s->host, p, tp);
}

int64_t bd_plan_layers(const uint32_t *a, const uint32_t *b, uint64_t n_links, uint64_t n_points, uint32_t *layer,
                       uint64_t *order) {
    if ((!a || !b || !layer) && n_links) return BD_ERR_ARG;
    for (uint64_t i = 0; i < n_links; i++)
        if (a[i] >= n_points || b[i] >= n_points) return BD_ERR_INDEX;
    return (int64_t)plan_layers(a, b, n_links, n_points, layer, order);
}

static int ensure_device(bd_solver *s) {
    if (s->stream) return BD_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(BD_ERR_NODEVICE, "no HIP device visible");
    if (hipGetDevice(&s->device) != hipSuccess) return fail(BD_ERR_DEVICE, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(BD_ERR_DEVICE, "hipStreamCreate failed");
    return BD_OK;
}

static void settle_acc(bd_solver *s, uint64_t substeps);

// Queue the readback of the state the last update left (see bd_solver::rb).
static int readback_begin(bd_solver *s) {
    bd_solver::Readback &rb = s->rb;
    const size_t n = s->eng->state_count();
    const int k = rb.slot < 0 ? 0 : (rb.slot ^ 1);
    auto hip_ok = [](hipError_t e) { return e == hipSuccess; };
    if (!rb.copy) {
        if (!hip_ok(hipStreamCreateWithFlags(&rb.copy, hipStreamNonBlocking))) return fail(BD_ERR_DEVICE, "readback stream");
        for (int j = 0; j < 2; j++)
            if (!hip_ok(hipEventCreateWithFlags(&rb.snap[j], hipEventDisableTiming)) ||
                !hip_ok(hipEventCreateWithFlags(&rb.done[j], hipEventDisableTiming)))
                return fail(BD_ERR_DEVICE, "readback events");
    }
    // slot k's previous host copy must be finished before its device
    // staging is overwritten (on the solver's stream) or its host buffer grows
    if (rb.used[k] && !hip_ok(hipStreamWaitEvent(s->stream, rb.done[k], 0))) return fail(BD_ERR_DEVICE, "readback wait");
    if (rb.dev[k].alloc(std::max<size_t>(n, 1)) != hipSuccess) return fail(BD_ERR_NOMEM, "readback staging");
    if (rb.host_cap[k] < n) {
        if (rb.used[k] && !hip_ok(hipEventSynchronize(rb.done[k]))) return fail(BD_ERR_DEVICE, "readback wait");
        if (rb.host[k]) (void)hipHostFree(rb.host[k]);
        rb.host[k] = nullptr;
        rb.host_cap[k] = 0;
        if (!hip_ok(hipHostMalloc((void **)&rb.host[k], n * sizeof(float2), hipHostMallocDefault)))
            return fail(BD_ERR_NOMEM, "readback pinned buffer");
        rb.host_cap[k] = n;
    }
    std::string err;
    int rc = s->eng->copy_state(rb.dev[k].p, hipMemcpyDeviceToDevice, s->stream, err);
    if (rc) return fail(rc, err);
    if (!hip_ok(hipEventRecord(rb.snap[k], s->stream)) || !hip_ok(hipStreamWaitEvent(rb.copy, rb.snap[k], 0)) ||
        (n && !hip_ok(hipMemcpyAsync(rb.host[k], rb.dev[k].p, n * sizeof(float2), hipMemcpyDeviceToHost, rb.copy))) ||
        !hip_ok(hipEventRecord(rb.done[k], rb.copy)))
        return fail(BD_ERR_DEVICE, "readback copy");
    rb.used[k] = true;
    rb.slot = k;
    rb.valid = true;
    return BD_OK;
}

int32_t bd_update(bd_solver *s, float dt) {
    if (!s) return BD_ERR_ARG;
    std::string err;
    if (!s->plan_valid) {
        int rc = ensure_host(s);
        if (rc) return rc;
        rc = build_plan(s->host, s->plan, err);
        if (rc) return fail(rc, err);
        rc = ensure_device(s);
        if (rc) return rc;
        int want = s->engine_pref;
        if (want == BD_ENGINE_AUTO) {
            if (scene_engine_fits(s->host, s->plan)) want = BD_ENGINE_SCENE;
            else if (tiled_engine_fits(s->host, s->plan)) want = BD_ENGINE_TILED;
            else want = BD_ENGINE_GLOBAL;
        }
        if (want == BD_ENGINE_SCENE && !scene_engine_fits(s->host, s->plan))
            return fail(BD_ERR_CAPACITY, "scene does not fit the SCENE engine");
        if (want == BD_ENGINE_TILED && !tiled_engine_fits(s->host, s->plan))
            return fail(BD_ERR_CAPACITY, "link graph does not suit the TILED engine");
        s->eng.reset();
        if (want == BD_ENGINE_SCENE) s->eng.reset(make_scene_engine(s->stream));
        else if (want == BD_ENGINE_TILED) s->eng.reset(make_tiled_engine(s->stream, s->tile_points, s->tile_wg));
        else s->eng.reset(make_global_engine(s->stream));
        rc = s->eng->upload(s->host, s->plan, err);
        if (rc == BD_ERR_CAPACITY && want == BD_ENGINE_TILED && s->engine_pref == BD_ENGINE_AUTO) {
            // the link graph does not tile within the LDS budget: one launch per layer
            want = BD_ENGINE_GLOBAL;
            s->eng.reset(make_global_engine(s->stream));
            rc = s->eng->upload(s->host, s->plan, err);
        }
        if (rc) return fail(rc, err);
        s->plan_valid = true;
        s->acc_pending = s->host.any_initial_acc();
        s->engine_used = want;
    }
    StepParams sp = make_params(s->host, dt, s->acc_pending);
    s->rb.valid = false;  // a readback copy, if any, is of the state before this update
    const uint64_t substeps_before = s->stats.substeps;
    int rc = s->eng->update(sp, s->stats, err);
    if (rc) {
        // The substeps that completed stay applied on the device (getters show
        // them); the initial circle / polygon-point acc is spent once one ran.
        if (s->stats.substeps > substeps_before) settle_acc(s, s->stats.substeps - substeps_before);
        s->dev_newer = true;
        return fail(rc, err);
    }
    s->stats.updates++;
    settle_acc(s, s->host.sub_steps);
    s->dev_newer = true;
    if (s->rb.on) return readback_begin(s);
    return BD_OK;
}

int32_t bd_ext_set_readback(bd_solver *s, int32_t on) {
    if (!s) return BD_ERR_ARG;
    s->rb.on = on != 0;
    if (!s->rb.on) s->rb.valid = false;
    return BD_OK;
}

// acc after update_positions: zero, except static polygons, which skip
// Polygon::update and keep adding gravity (src/polygon.rs:125-128); a static
// polygon's points gain one gravity per substep that ran.
static void settle_acc(bd_solver *s, uint64_t substeps) {
    HostScene &h = s->host;
    for (V2 &a : h.c_acc) a = V2{0.0f, 0.0f};
    for (size_t p = 0; p < h.polys.size(); p++) {
        const Poly &pg = h.polys[p];
        for (uint32_t i = 0; i < pg.n; i++) {
            V2 &a = h.g_acc[pg.off + i];
            if (pg.is_static) {
                for (uint64_t k = 0; k < substeps; k++) a = V2{a.x + h.gravity.x, a.y + h.gravity.y};
            } else {
                a = V2{0.0f, 0.0f};
            }
        }
        if (p < s->center_from_prev.size()) s->center_from_prev[p] = pg.is_static ? 0 : 1;
    }
    s->acc_pending = false;
}

}  // extern "C"

// ---- snapshot format --------------------------------------------------------
//
// A versioned, self-contained byte image of a solver (little-endian):
//   "BD2DSNAP" | u32 version | u32 header words | header (u32/f32 words) |
//   sections: particles (pos, prev, pin), particle links, circles (pos, prev,
//   acc, radius), circle links, polygons (records, points, links), static
//   colliders (records, planes, vertices), per-polygon centre flags |
//   u64 FNV-1a of every preceding byte.
// Only the host scene is saved: the device copy is brought back first, and a
// loaded solver plans and uploads on its next update, so a run continued from
// a snapshot equals the uninterrupted run bit for bit.

namespace {

constexpr char kSnapMagic[8] = {'B', 'D', '2', 'D', 'S', 'N', 'A', 'P'};
constexpr uint32_t kSnapVersion = 1;

struct SnapWriter {
    std::vector<uint8_t> b;
    template <typename T> void put(const T &v) {
        const uint8_t *p = (const uint8_t *)&v;
        b.insert(b.end(), p, p + sizeof(T));
    }
    template <typename T> void arr(const std::vector<T> &v) {
        const uint8_t *p = (const uint8_t *)v.data();
        b.insert(b.end(), p, p + sizeof(T) * v.size());
    }
};

struct SnapReader {
    const uint8_t *p, *end;
    bool ok = true;
    template <typename T> T get() {
        T v{};
        if ((size_t)(end - p) < sizeof(T)) { ok = false; return v; }
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    template <typename T> void arr(std::vector<T> &v, uint64_t n) {
        if (!ok || n > (uint64_t)(end - p) / sizeof(T)) { ok = false; return; }
        v.resize(n);
        std::memcpy(v.data(), p, sizeof(T) * n);
        p += sizeof(T) * n;
    }
};

uint64_t fnv1a(const uint8_t *p, size_t n) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; i++) { h ^= p[i]; h *= 0x100000001b3ull; }
    return h;
}

}  // namespace

extern "C" int64_t bd_solver_save(const bd_solver *s, void *buf, uint64_t cap) {
    if (!s) return BD_ERR_ARG;
    int rc = ensure_host(s);
    if (rc) return rc;
    const HostScene &h = s->host;
    SnapWriter w;
    w.b.insert(w.b.end(), kSnapMagic, kSnapMagic + 8);
    w.put(kSnapVersion);
    const uint32_t header_words = 21;
    w.put(header_words);
    w.put(h.gravity); w.put(h.bpos); w.put(h.bsize);                        // 6 words
    w.put(h.bounds_active); w.put(h.sub_steps); w.put(h.link_iters);        // 3
    w.put(s->engine_pref); w.put((uint32_t)s->acc_pending);                 // 2
    w.put((uint32_t)h.p_pos.size()); w.put((uint32_t)h.plinks.size());      // 2
    w.put((uint32_t)h.c_pos.size()); w.put((uint32_t)h.clinks.size());      // 2
    w.put((uint32_t)h.polys.size()); w.put((uint32_t)h.g_pos.size());      // 2
    w.put((uint32_t)h.glinks.size()); w.put((uint32_t)h.cols.size());      // 2
    w.put((uint32_t)h.planes.size()); w.put((uint32_t)h.verts.size());     // 2
    w.arr(h.p_pos); w.arr(h.p_prev); w.arr(h.p_pin); w.arr(h.plinks);
    w.arr(h.c_pos); w.arr(h.c_prev); w.arr(h.c_acc); w.arr(h.c_r); w.arr(h.clinks);
    w.arr(h.polys); w.arr(h.g_pos); w.arr(h.g_prev); w.arr(h.g_acc); w.arr(h.glinks);
    w.arr(h.cols); w.arr(h.planes); w.arr(h.verts);
    std::vector<uint8_t> cfp(h.polys.size(), 0);
    for (size_t p = 0; p < cfp.size() && p < s->center_from_prev.size(); p++) cfp[p] = s->center_from_prev[p];
    w.arr(cfp);
    w.put(fnv1a(w.b.data(), w.b.size()));
    if (buf && cap >= w.b.size()) std::memcpy(buf, w.b.data(), w.b.size());
    else if (buf) return fail(BD_ERR_ARG, "buffer smaller than the snapshot");
    return (int64_t)w.b.size();
}

extern "C" bd_solver *bd_solver_load(const void *buf, uint64_t bytes) {
    if (!buf || bytes < 8 + 8 + 8) { set_last_error("not a solver snapshot"); return nullptr; }
    const uint8_t *b = (const uint8_t *)buf;
    if (std::memcmp(b, kSnapMagic, 8) != 0) { set_last_error("not a solver snapshot"); return nullptr; }
    uint64_t sum;
    std::memcpy(&sum, b + bytes - 8, 8);
    if (fnv1a(b, bytes - 8) != sum) { set_last_error("snapshot checksum mismatch"); return nullptr; }
    SnapReader r{b + 8, b + bytes - 8};
    const uint32_t version = r.get<uint32_t>(), words = r.get<uint32_t>();
    if (version != kSnapVersion || words != 21) { set_last_error("unsupported snapshot version"); return nullptr; }
    std::unique_ptr<bd_solver> s(bd_solver_new());
    if (!s) return nullptr;
    HostScene &h = s->host;
    h.gravity = r.get<V2>(); h.bpos = r.get<V2>(); h.bsize = r.get<V2>();
    h.bounds_active = r.get<int32_t>(); h.sub_steps = r.get<uint32_t>(); h.link_iters = r.get<uint32_t>();
    s->engine_pref = r.get<int32_t>();
    s->acc_pending = r.get<uint32_t>() != 0;
    const uint32_t np = r.get<uint32_t>(), npl = r.get<uint32_t>(), nc = r.get<uint32_t>(), ncl = r.get<uint32_t>();
    const uint32_t npoly = r.get<uint32_t>(), ng = r.get<uint32_t>(), ngl = r.get<uint32_t>(), ncol = r.get<uint32_t>();
    const uint32_t nplanes = r.get<uint32_t>(), nverts = r.get<uint32_t>();
    r.arr(h.p_pos, np); r.arr(h.p_prev, np); r.arr(h.p_pin, np); r.arr(h.plinks, npl);
    r.arr(h.c_pos, nc); r.arr(h.c_prev, nc); r.arr(h.c_acc, nc); r.arr(h.c_r, nc); r.arr(h.clinks, ncl);
    r.arr(h.polys, npoly); r.arr(h.g_pos, ng); r.arr(h.g_prev, ng); r.arr(h.g_acc, ng); r.arr(h.glinks, ngl);
    r.arr(h.cols, ncol); r.arr(h.planes, nplanes); r.arr(h.verts, nverts);
    r.arr(s->center_from_prev, npoly);
    bool ok = r.ok && r.p == r.end && h.sub_steps >= 1 && h.sub_steps <= 65535 && h.link_iters >= 1 &&
              s->engine_pref >= BD_ENGINE_AUTO && s->engine_pref <= BD_ENGINE_TILED && nplanes % 3 == 0;
    for (const Poly &pg : h.polys)
        ok = ok && (uint64_t)pg.off + pg.n <= ng && (uint64_t)pg.loff + pg.nl <= ngl;
    for (const Collider &c : h.cols)
        ok = ok && (c.kind == 0 || (c.kind == 1 && (uint64_t)c.plane_off + c.nv <= nplanes / 3 &&
                                    2ull * ((uint64_t)c.plane_off + c.nv) <= nverts));
    if (!ok) { set_last_error("malformed snapshot"); return nullptr; }
    // the first update after loading plans and uploads; the initial acc of
    // circles and polygon points is pending only if no update has run
    s->acc_pending = false;
    return s.release();
}

extern "C" {

int32_t bd_ext_sync(bd_solver *s) {
    if (!s) return BD_ERR_ARG;
    if (!s->eng) return BD_OK;
    std::string err;
    int rc = s->eng->sync(err);
    return rc ? fail(rc, err) : BD_OK;
}

int32_t bd_ext_get_stats(const bd_solver *s, bd_stats *out) {
    if (!s || !out) return BD_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->n_points = s->host.n_points();
    out->n_links = s->host.n_links();
    out->link_layers = s->plan_valid ? s->plan.depth : 0;
    out->updates = s->stats.updates;
    out->substeps = s->stats.substeps;
    out->kernel_launches = s->stats.launches;
    out->cc_rounds = s->stats.cc_rounds;
    out->cc_redos = s->stats.cc_redos;
    out->cc_pairs = s->stats.cc_pairs;
    out->cc_margin = s->stats.cc_margin;
    out->engine = s->engine_used;
    out->tile_passes = s->stats.tile_passes;
    return BD_OK;
}

}  // extern "C"
