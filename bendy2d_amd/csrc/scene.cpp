// scene.cpp -- validation, dependency-layer planning and host constructors.
#include "scene.h"

#include <cmath>
#include <cstring>

#include "../../include/bendy2d.h"

namespace bd {

bool HostScene::any_pinned() const {
    for (uint8_t v : p_pin)
        if (v) return true;
    return false;
}

bool HostScene::any_initial_acc() const {
    for (const V2 &a : c_acc)
        if (a.x != 0.0f || a.y != 0.0f || std::signbit(a.x) || std::signbit(a.y)) return true;
    for (size_t p = 0; p < polys.size(); p++) {
        if (polys[p].is_static) continue;  // a static polygon never integrates
        for (uint32_t i = 0; i < polys[p].n; i++) {
            const V2 &a = g_acc[polys[p].off + i];
            if (a.x != 0.0f || a.y != 0.0f || std::signbit(a.x) || std::signbit(a.y)) return true;
        }
    }
    return false;
}

uint32_t plan_layers(const uint32_t *a, const uint32_t *b, uint64_t n_links, uint64_t n_points,
                     uint32_t *layer, uint64_t *order) {
    // last[p] = 1 + layer of the latest link touching p (0 = none yet)
    std::vector<uint32_t> last(n_points, 0u);
    uint32_t depth = 0;
    for (uint64_t i = 0; i < n_links; i++) {
        uint32_t la = last[a[i]], lb = last[b[i]];
        uint32_t L = la > lb ? la : lb;
        layer[i] = L;
        last[a[i]] = L + 1;
        last[b[i]] = L + 1;
        if (L + 1 > depth) depth = L + 1;
    }
    if (order) {
        std::vector<uint64_t> off(depth + 1, 0);
        for (uint64_t i = 0; i < n_links; i++) off[layer[i] + 1]++;
        for (uint32_t d = 0; d < depth; d++) off[d + 1] += off[d];
        for (uint64_t i = 0; i < n_links; i++) order[off[layer[i]]++] = i;  // stable
    }
    return depth;
}

static bool check_links(const std::vector<LinkH> &v, uint64_t begin, uint64_t count, uint64_t n,
                        const char *what, std::string &err) {
    for (uint64_t i = begin; i < begin + count; i++) {
        const LinkH &l = v[i];
        if (l.a == kBadIndex || l.b == kBadIndex || l.b >= n || l.a >= l.b) {
            err = std::string(what) + " link " + std::to_string(i - begin) + " (" +
                  std::to_string(l.a) + ", " + std::to_string(l.b) +
                  ") needs particle_a < particle_b < " + std::to_string(n);
            return false;
        }
    }
    return true;
}

int build_plan(const HostScene &s, Plan &plan, std::string &err) {
    uint64_t Np = s.p_pos.size(), Nc = s.c_pos.size(), Ng = s.g_pos.size();
    uint64_t N = Np + Nc + Ng;
    if (N > kMaxPoints) {
        err = "scene holds more than 2^30-1 points";
        return BD_ERR_CAPACITY;
    }
    if (!check_links(s.plinks, 0, s.plinks.size(), Np, "particle", err)) return BD_ERR_LINK;
    if (!check_links(s.clinks, 0, s.clinks.size(), Nc, "circle", err)) return BD_ERR_LINK;
    for (const Poly &p : s.polys)
        if (!check_links(s.glinks, p.loff, p.nl, p.n, "polygon", err)) return BD_ERR_LINK;

    plan.Np = (uint32_t)Np; plan.Nc = (uint32_t)Nc; plan.Ng = (uint32_t)Ng; plan.N = (uint32_t)N;
    uint64_t L = s.n_links();
    std::vector<uint32_t> ua(L), ub(L);
    std::vector<Link12> seq(L);
    uint64_t k = 0;
    // Sequence order of apply_links (src/solver.rs:143-153): particle links,
    // circle links, then each polygon's links.
    for (const LinkH &l : s.plinks) {
        ua[k] = l.a; ub[k] = l.b;
        uint32_t fa = (s.p_pin[l.a] ? kPinBit : 0u), fb = (s.p_pin[l.b] ? kPinBit : 0u);
        seq[k] = Link12{l.a | fa, l.b | fb, l.t};
        k++;
    }
    for (const LinkH &l : s.clinks) {
        uint32_t a = (uint32_t)Np + l.a, b = (uint32_t)Np + l.b;
        ua[k] = a; ub[k] = b;
        seq[k] = Link12{a | kCircleLinkBit, b, l.t};
        k++;
    }
    uint64_t first_poly_seq = k;
    for (const Poly &p : s.polys) {
        uint32_t base = (uint32_t)(Np + Nc) + p.off;
        for (uint32_t i = 0; i < p.nl; i++) {
            const LinkH &l = s.glinks[p.loff + i];
            ua[k] = base + l.a; ub[k] = base + l.b;
            seq[k] = Link12{base + l.a, base + l.b, l.t};
            k++;
        }
    }
    std::vector<uint32_t> layer(L);
    std::vector<uint64_t> order(L);
    plan.depth = plan_layers(ua.data(), ub.data(), L, N, layer.data(), order.data());
    plan.links.resize(L);
    plan.layer_off.assign(plan.depth + 1, 0);
    plan.first_poly_layer = plan.depth;
    for (uint64_t i = 0; i < L; i++) {
        plan.links[i] = seq[order[i]];
        plan.layer_off[layer[order[i]] + 1]++;
        if (order[i] >= first_poly_seq && layer[order[i]] < plan.first_poly_layer)
            plan.first_poly_layer = layer[order[i]];
    }
    for (uint32_t d = 0; d < plan.depth; d++) plan.layer_off[d + 1] += plan.layer_off[d];
    return BD_OK;
}

void polygon_planes(const V2 *v, uint64_t n, std::vector<float> &out) {
    double area2 = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = (i + 1) % n;
        area2 += (double)v[i].x * v[j].y - (double)v[j].x * v[i].y;
    }
    for (uint64_t k = 0; k < n; k++) {
        uint64_t i = area2 >= 0 ? k : n - 1 - k;
        uint64_t j = area2 >= 0 ? (k + 1) % n : (n - 1 - ((k + 1) % n));
        float vx = v[i].x, vy = v[i].y;
        float ex = v[j].x - vx, ey = v[j].y - vy;
        float len = sqrtf(ex * ex + ey * ey);
        float nx = ey / len, ny = (-ex) / len;
        out.push_back(nx);
        out.push_back(ny);
        out.push_back(nx * vx + ny * vy);
    }
}

void polygon_circle(float radius, V2 pos, uint64_t count, std::vector<V2> &pts,
                    std::vector<LinkH> &links, V2 &center) {
    pts.clear(); links.clear();
    float cx = 0.0f, cy = 0.0f, angle = 0.0f;
    for (uint64_t i = 0; i < count; i++) {
        float x = radius * cosf(angle);                 // src/polygon.rs:23
        float y = radius * sinf(angle);                 // src/polygon.rs:24
        V2 p{pos.x + x, pos.y + y};                     // src/polygon.rs:25
        pts.push_back(p);
        cx = cx + p.x; cy = cy + p.y;                   // src/polygon.rs:27
        angle = angle + (2.0f * 3.14159265358979323846f) / (float)count;  // src/polygon.rs:28
    }
    center = V2{cx / (float)count, cy / (float)count};  // src/polygon.rs:30
    for (uint64_t i = 0; i < count; i++) {
        for (int part = 0; part < 2; part++) {          // src/polygon.rs:33-52, 53-72
            uint64_t a = i;
            uint64_t b = part == 0 ? (i + 2 * count / 3) % count : (i + count / 3) % count;
            if (a > b) { uint64_t t = a; a = b; b = t; }
            float dx = pts[a].x - pts[b].x, dy = pts[a].y - pts[b].y;
            links.push_back(LinkH{(uint32_t)a, (uint32_t)b, sqrtf(dx * dx + dy * dy)});
        }
    }
}

void polygon_new(const V2 *points, uint64_t n, std::vector<V2> &pts, std::vector<LinkH> &links,
                 V2 &center) {
    pts.assign(points, points + n); links.clear();
    float cx = 0.0f, cy = 0.0f;
    for (uint64_t i = 0; i < n; i++) { cx = cx + points[i].x; cy = cy + points[i].y; }  // :88-92
    center = V2{cx / (float)n, cy / (float)n};                                          // :93
    for (uint64_t i = 0; i < n; i++) {                                                  // :95-114
        uint64_t a = i, b = (i + 1) % n;
        if (a > b) { uint64_t t = a; a = b; b = t; }
        float dx = points[a].x - points[b].x, dy = points[a].y - points[b].y;
        links.push_back(LinkH{(uint32_t)a, (uint32_t)b, sqrtf(dx * dx + dy * dy)});
    }
}

}  // namespace bd
