// engine_global.hip -- GLOBAL engine: the whole device works on one scene.
//
// Per substep (src/solver.rs:110-114):
//   links      one launch per dependency layer (k_link_layer), k sweeps; the
//              layers reproduce the sequential Gauss-Seidel order exactly
//              (src/solver.rs:143-153)
//   circles    candidate pairs from a hashed cell broadphase (counting sort
//              with wave prefix scans), then an exact replay of the reference's
//              (i, j) lexicographic sweep (src/solver.rs:168-177): pair (i, j)
//              runs once every earlier pair touching i or j has run
//   polygons   one workgroup replays the polygon-pair loop (src/solver.rs:178-187)
//   per body   EXT4 static colliders (staged in LDS), bounds, Verlet, fused in
//              one streaming kernel (src/solver.rs:155-165, 118-128)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bendy2d.h"
#include "common.h"
#include "engine.h"
#include "bodies.h"
#include "kernels.h"

namespace bd {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// One dependency layer of links: no two links of the layer share a point.
// Algorithmic traffic: 12 B record + 2 x 8 B reads + 2 x 8 B writes per link.
__global__ void __launch_bounds__(256) k_link_layer(float2 *__restrict__ pos, const Link12 *__restrict__ links,
                                                    uint32_t n, const float *__restrict__ crad, uint32_t cbase) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    Link12 l = links[i];
    uint32_t a = l.a & kIndexMask, b = l.b & kIndexMask;
    float2 A = pos[a], B = pos[b];
    if (l.a & kCircleLinkBit) {
        solve_circle_link(A.x, A.y, B.x, B.y, l.t, crad[a - cbase], crad[b - cbase]);
    } else {
        solve_particle_link(A.x, A.y, B.x, B.y, l.t, (l.a & kPinBit) != 0, (l.b & kPinBit) != 0);
    }
    pos[a] = A;
    pos[b] = B;
}

// Polygon centres (Polygon::calc_center, src/polygon.rs:231-237): sequential
// f32 sum in point order, one thread per polygon.
__global__ void k_poly_centers(const float2 *__restrict__ pos, const uint2 *__restrict__ ptab, uint32_t P,
                               float2 *__restrict__ centers) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    uint2 t = ptab[p];
    float sx = 0.0f, sy = 0.0f;
    for (uint32_t i = 0; i < t.y; i++) {
        float2 v = pos[t.x + i];
        sx = sx + v.x; sy = sy + v.y;
    }
    float n = (float)t.y;
    centers[p] = make_float2(sx / n, sy / n);
}

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
__device__ bool resolve_line(float2 self_center, float2 pa, float2 pb, float2 op, float2 other_center,
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

// Polygon pair loop (src/solver.rs:178-187 -> Polygon::solve_polygon
// src/polygon.rs:142-162) in one workgroup.  Edges run in order; within an
// edge the other polygon's points are independent (each reads only the stale
// edge copies, its own position and the two centres), so they run across the
// workgroup, and the hit with the highest point index writes the edge ends
// ("last hit wins").
__global__ void __launch_bounds__(256) k_poly_pairs(float2 *pos, const uint2 *__restrict__ ptab, uint32_t P,
                                                    const float2 *__restrict__ centers) {
    __shared__ int s_last;
    __shared__ float2 s_pa, s_pb;
    for (uint32_t i = 0; i < P; i++) {
        for (uint32_t j = i + 1; j < P; j++) {
            for (int dir = 0; dir < 2; dir++) {
                uint32_t si = dir == 0 ? i : j, oi = dir == 0 ? j : i;
                uint2 st = ptab[si], ot = ptab[oi];
                float2 sc = centers[si], oc = centers[oi];
                for (uint32_t e = 0; e < st.y; e++) {
                    uint32_t bid = (e + 1) % st.y;
                    if (threadIdx.x == 0) {
                        s_last = -1;
                        s_pa = pos[st.x + e];
                        s_pb = pos[st.x + bid];
                    }
                    __syncthreads();
                    float2 pa = s_pa, pb = s_pb;
                    int mine = -1;
                    float2 na = make_float2(0.0f, 0.0f), nb = na;
                    for (uint32_t k = threadIdx.x; k < ot.y; k += blockDim.x) {
                        float2 op = pos[ot.x + k];
                        float2 a2, b2, o2;
                        if (resolve_line(sc, pa, pb, op, oc, a2, b2, o2)) {
                            pos[ot.x + k] = o2;
                            na = a2; nb = b2; mine = (int)k;
                        }
                    }
                    if (mine >= 0) atomicMax(&s_last, mine);
                    __syncthreads();
                    if (mine >= 0 && mine == s_last) {
                        pos[st.x + e] = na;
                        pos[st.x + bid] = nb;
                    }
                    __threadfence_block();
                    __syncthreads();
                }
            }
        }
    }
}

// Fused per-body phase: EXT4 static colliders -> bounds -> Verlet.
constexpr int kMaxStagedColliders = 256;
constexpr int kMaxStagedPlanes = 1024;

template <bool STAGED>
__global__ void __launch_bounds__(256) k_bodies(float2 *__restrict__ pos, float2 *__restrict__ prev,
                                                const uint32_t *__restrict__ pinmask,
                                                const float *__restrict__ crad,
                                                const float2 *__restrict__ acc0,
                                                const uint8_t *__restrict__ gflags,
                                                const int4 *__restrict__ colhdr,
                                                const float4 *__restrict__ colgeo,
                                                const float4 *__restrict__ planes, BodyParams P) {
    __shared__ int4 s_hdr[STAGED ? kMaxStagedColliders : 1];
    __shared__ float4 s_geo[STAGED ? kMaxStagedColliders : 1];
    __shared__ float4 s_pl[STAGED ? kMaxStagedPlanes : 1];
    const int4 *hdr = colhdr;
    const float4 *geo = colgeo;
    const float4 *pl = planes;
    if (STAGED) {
        for (uint32_t k = threadIdx.x; k < P.ncol; k += blockDim.x) { s_hdr[k] = colhdr[k]; s_geo[k] = colgeo[k]; }
        for (uint32_t k = threadIdx.x; k < P.nplanes; k += blockDim.x) s_pl[k] = planes[k];
        __syncthreads();
        hdr = s_hdr; geo = s_geo; pl = s_pl;
    }
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < P.N; u += gridDim.x * blockDim.x) {
        float r = 0.0f;
        float ax = P.gex, ay = P.gey;
        bool circle = false, integrate = true;
        if (u < P.Np) {
            if (pinmask && ((pinmask[u >> 5] >> (u & 31)) & 1u)) continue;  // EXT3: pinned never moves
        } else if (u < P.Np + P.Nc) {
            circle = true;
            r = crad[u - P.Np];
            if (P.first_acc) { float2 a = acc0[u - P.Np]; ax = a.x + P.gx; ay = a.y + P.gy; }
        } else {
            integrate = (gflags[u - P.Np - P.Nc] & 1u) == 0;  // Polygon::update returns if static
            if (P.first_acc && integrate) { float2 a = acc0[u - P.Np]; ax = a.x + P.gx; ay = a.y + P.gy; }
        }
        float2 p = pos[u], q = prev[u];
        static_colliders(p, r, P.ncol, hdr, geo, pl);
        if (circle) {
            circle_bounds_axis(p.x, q.x, P.lox, P.hix, r);
            circle_bounds_axis(p.y, q.y, P.loy, P.hiy, r);
        } else {
            bounds_axis(p.x, q.x, P.lox, P.hix);
            bounds_axis(p.y, q.y, P.loy, P.hiy);
        }
        if (integrate) verlet(p.x, p.y, q.x, q.y, ax, ay, P.dt);
        pos[u] = p;
        prev[u] = q;
    }
}

__global__ void k_zero_acc0(float2 *acc0, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc0[i] = make_float2(0.0f, 0.0f);
}

// ---------------------------------------------------------------------------
// Circle-pair schedule kernels
// ---------------------------------------------------------------------------

// Cell of a point for cell size h (conservatively above the candidate
// distance), clamped to +-2^30 so huge or non-finite coordinates stay finite.
__device__ __forceinline__ int cell_coord(float x, double inv_h) {
    double c = floor((double)x * inv_h);
    if (!(c > -1073741824.0)) c = -1073741824.0;
    if (c > 1073741824.0) c = 1073741824.0;
    return (int)c;
}

__global__ void k_cc_prep(const float2 *__restrict__ pos, uint32_t cbase, uint32_t nc, double inv_h, uint32_t tmask,
                          int2 *__restrict__ cell, uint32_t *__restrict__ key, uint32_t *__restrict__ count,
                          float2 *__restrict__ save, float *__restrict__ path) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    float2 p = pos[cbase + i];
    save[i] = p;
    path[i] = 0.0f;
    int cx = cell_coord(p.x, inv_h), cy = cell_coord(p.y, inv_h);
    cell[i] = make_int2(cx, cy);
    uint32_t k = cell_hash(cx, cy) & tmask;
    key[i] = k;
    atomicAdd(&count[k], 1u);
}

__global__ void k_cc_scatter(const uint32_t *__restrict__ key, uint32_t nc, const uint32_t *__restrict__ start,
                             uint32_t *__restrict__ fill, uint32_t *__restrict__ items) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    uint32_t k = key[i];
    uint32_t slot = start[k] + atomicAdd(&fill[k], 1u);
    items[slot] = i;
}

// Candidates of circle i: every j != i whose distance at sweep start is below
// r_i + r_j + margin.  MODE 0 counts, MODE 1 writes and sorts ascending.
template <int MODE>
__global__ void k_cc_cand(const float2 *__restrict__ save, const float *__restrict__ crad, const int2 *__restrict__ cell,
                          uint32_t nc, uint32_t tmask, const uint32_t *__restrict__ start,
                          const uint32_t *__restrict__ items, float margin, uint32_t *__restrict__ cnt,
                          const uint32_t *__restrict__ off, uint32_t *__restrict__ cand) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    float2 p = save[i];
    float ri = crad[i];
    int2 c = cell[i];
    uint32_t n = 0;
    uint32_t base = MODE == 1 ? off[i] : 0u;
    for (int dy = -1; dy <= 1; dy++) {
        for (int dx = -1; dx <= 1; dx++) {
            int cx = c.x + dx, cy = c.y + dy;
            uint32_t k = cell_hash(cx, cy) & tmask;
            // a bucket can hold several cells; two neighbour cells can share a
            // bucket, so each j is taken only from its own cell
            for (uint32_t s = start[k], e = start[k + 1]; s < e; s++) {
                uint32_t j = items[s];
                if (j == i) continue;
                int2 cj = cell[j];
                if (cj.x != cx || cj.y != cy) continue;
                float2 pj = save[j];
                float ddx = p.x - pj.x, ddy = p.y - pj.y;
                float lim = (ri + crad[j]) + margin;
                if (ddx * ddx + ddy * ddy < lim * lim) {
                    if (MODE == 1) cand[base + n] = j;
                    n++;
                }
            }
        }
    }
    if (MODE == 0) {
        cnt[i] = n;
    } else {
        // insertion sort of this circle's list (short)
        for (uint32_t a = 1; a < n; a++) {
            uint32_t v = cand[base + a];
            uint32_t b = a;
            while (b > 0 && cand[base + b - 1] > v) { cand[base + b] = cand[base + b - 1]; b--; }
            cand[base + b] = v;
        }
    }
}

// One round of the exact pair schedule.  Pair (i, j), i < j, is ready when
// j is the smallest unprocessed neighbour of i and i the smallest unprocessed
// neighbour of j: every earlier pair touching i or j (in lexicographic order)
// has then run.  Ready pairs share no circle.  Heads ping-pong between two
// buffers so every thread reads the previous round's state.
__global__ void k_cc_round(float2 *__restrict__ pos, uint32_t cbase, const float *__restrict__ crad,
                           uint32_t nc, const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                           const uint32_t *__restrict__ cand, const uint32_t *__restrict__ head_in,
                           uint32_t *__restrict__ head_out, float *__restrict__ path,
                           uint32_t *__restrict__ remaining, uint32_t r, uint32_t *__restrict__ pmax_bits) {
    if (r > 0 && remaining[r - 1] == 0u) return;  // schedule finished in an earlier round
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    uint32_t h = head_in[i], ci = cnt[i];
    uint32_t nh = h;
    if (h < ci) {
        uint32_t j = cand[off[i] + h];
        uint32_t hj = head_in[j];
        bool ready = hj < cnt[j] && cand[off[j] + hj] == i;
        if (ready) {
            if (i < j) {
                float2 pi = pos[cbase + i], pj = pos[cbase + j];
                float qi = 0.0f, qj = 0.0f;
                if (solve_circle_pair(pi.x, pi.y, pj.x, pj.y, crad[i], crad[j], qi, qj)) {
                    pos[cbase + i] = pi;
                    pos[cbase + j] = pj;
                    float a = path[i] + qi, b = path[j] + qj;
                    path[i] = a; path[j] = b;
                    float m = fmaxf(a, b);
                    if (!(m == m)) m = INFINITY;
                    atomicMax(pmax_bits, __float_as_uint(m));
                }
            }
            nh = h + 1;
        }
    }
    head_out[i] = nh;
    if (nh < ci) atomicAdd(&remaining[r], 1u);
}

__global__ void k_cc_restore(float2 *__restrict__ pos, uint32_t cbase, const float2 *__restrict__ save, uint32_t nc) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) pos[cbase + i] = save[i];
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------

class GlobalEngine : public Engine {
public:
    explicit GlobalEngine(hipStream_t s) : stream_(s) {}
    ~GlobalEngine() override {
        if (host_ctl_) (void)hipHostFree(host_ctl_);
    }
    int kind() const override { return BD_ENGINE_GLOBAL; }

    int upload(const HostScene &s, const Plan &p, std::string &err) override {
        Np_ = p.Np; Nc_ = p.Nc; Ng_ = p.Ng; N_ = p.N;
        P_ = (uint32_t)s.polys.size();
        layer_off_ = p.layer_off;
        depth_ = p.depth;
        std::vector<float2> pos(N_), prev(N_);
        auto put = [&](const std::vector<V2> &a, const std::vector<V2> &b, uint32_t base) {
            for (size_t i = 0; i < a.size(); i++) {
                pos[base + i] = make_float2(a[i].x, a[i].y);
                prev[base + i] = make_float2(b[i].x, b[i].y);
            }
        };
        put(s.p_pos, s.p_prev, 0);
        put(s.c_pos, s.c_prev, Np_);
        put(s.g_pos, s.g_prev, Np_ + Nc_);
        BD_HIP(pos_.alloc(std::max<size_t>(N_, 1)));
        BD_HIP(prev_.alloc(std::max<size_t>(N_, 1)));
        if (N_) {
            BD_HIP(hipMemcpyAsync(pos_.p, pos.data(), N_ * sizeof(float2), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(prev_.p, prev.data(), N_ * sizeof(float2), hipMemcpyHostToDevice, stream_));
        }
        // links
        BD_HIP(links_.alloc(std::max<size_t>(p.links.size(), 1)));
        if (!p.links.empty())
            BD_HIP(hipMemcpyAsync(links_.p, p.links.data(), p.links.size() * sizeof(Link12), hipMemcpyHostToDevice, stream_));
        // pins
        has_pins_ = s.any_pinned();
        if (has_pins_) {
            std::vector<uint32_t> mask((Np_ + 31) / 32, 0u);
            for (uint32_t i = 0; i < Np_; i++)
                if (s.p_pin[i]) mask[i >> 5] |= 1u << (i & 31);
            BD_HIP(pinmask_.alloc(mask.size()));
            BD_HIP(hipMemcpyAsync(pinmask_.p, mask.data(), mask.size() * 4, hipMemcpyHostToDevice, stream_));
        }
        // circle radii
        BD_HIP(crad_.alloc(std::max<size_t>(Nc_, 1)));
        if (Nc_) BD_HIP(hipMemcpyAsync(crad_.p, s.c_r.data(), Nc_ * 4, hipMemcpyHostToDevice, stream_));
        // initial acc of circles and polygon points
        if (Nc_ + Ng_) {
            std::vector<float2> acc(Nc_ + Ng_);
            for (uint32_t i = 0; i < Nc_; i++) acc[i] = make_float2(s.c_acc[i].x, s.c_acc[i].y);
            for (uint32_t i = 0; i < Ng_; i++) acc[Nc_ + i] = make_float2(s.g_acc[i].x, s.g_acc[i].y);
            BD_HIP(acc0_.alloc(acc.size()));
            BD_HIP(hipMemcpyAsync(acc0_.p, acc.data(), acc.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
        }
        // polygons
        if (P_) {
            std::vector<uint2> tab(P_);
            std::vector<uint8_t> flags(std::max<uint32_t>(Ng_, 1), 0);
            for (uint32_t i = 0; i < P_; i++) {
                tab[i] = make_uint2(Np_ + Nc_ + s.polys[i].off, s.polys[i].n);
                for (uint32_t k = 0; k < s.polys[i].n; k++) flags[s.polys[i].off + k] = s.polys[i].is_static ? 1 : 0;
            }
            BD_HIP(ptab_.alloc(P_));
            BD_HIP(centers_.alloc(P_));
            BD_HIP(gflags_.alloc(flags.size()));
            BD_HIP(hipMemcpyAsync(ptab_.p, tab.data(), P_ * sizeof(uint2), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(gflags_.p, flags.data(), flags.size(), hipMemcpyHostToDevice, stream_));
            std::vector<float2> c(P_);
            for (uint32_t i = 0; i < P_; i++) c[i] = make_float2(s.polys[i].center.x, s.polys[i].center.y);
            BD_HIP(hipMemcpyAsync(centers_.p, c.data(), P_ * sizeof(float2), hipMemcpyHostToDevice, stream_));
            static_poly_.clear();
            for (uint32_t i = 0; i < P_; i++)
                if (s.polys[i].is_static) static_poly_.push_back(i);
        }
        // static colliders
        ncol_ = (uint32_t)s.cols.size();
        nplanes_ = (uint32_t)(s.planes.size() / 3);
        if (ncol_) {
            std::vector<int4> hdr(ncol_);
            std::vector<float4> geo(ncol_);
            for (uint32_t k = 0; k < ncol_; k++) {
                const Collider &c = s.cols[k];
                hdr[k] = make_int4(c.kind, (int)c.plane_off, (int)c.nv, 0);
                geo[k] = make_float4(c.cx, c.cy, c.r, 0.0f);
            }
            std::vector<float4> pl(std::max<uint32_t>(nplanes_, 1));
            for (uint32_t k = 0; k < nplanes_; k++)
                pl[k] = make_float4(s.planes[3 * k], s.planes[3 * k + 1], s.planes[3 * k + 2], 0.0f);
            BD_HIP(colhdr_.alloc(ncol_));
            BD_HIP(colgeo_.alloc(ncol_));
            BD_HIP(planes_.alloc(pl.size()));
            BD_HIP(hipMemcpyAsync(colhdr_.p, hdr.data(), ncol_ * sizeof(int4), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(colgeo_.p, geo.data(), ncol_ * sizeof(float4), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(planes_.p, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice, stream_));
        }
        if (!host_ctl_) BD_HIP(hipHostMalloc((void **)&host_ctl_, 64 * sizeof(uint32_t), hipHostMallocDefault));
        rmax_ = 0.0f;
        for (float r : s.c_r) rmax_ = std::max(rmax_, r);
        if (margin_ <= 0.0f) margin_ = std::max(0.25f * rmax_, 0.05f);
        // the upload copies read host vectors that die at return
        BD_HIP(hipStreamSynchronize(stream_));
        return BD_OK;
    }

    int update(const StepParams &sp, Stats &st, std::string &err) override {
        BodyParams bp;
        bp.dt = sp.delta; bp.gex = sp.gex; bp.gey = sp.gey; bp.gx = sp.gx; bp.gy = sp.gy;
        bp.lox = sp.lox; bp.loy = sp.loy; bp.hix = sp.hix; bp.hiy = sp.hiy;
        bp.N = N_; bp.Np = Np_; bp.Nc = Nc_;
        bp.ncol = ncol_; bp.nplanes = nplanes_;
        for (uint32_t step = 0; step < sp.sub_steps; step++) {
            // links, k sweeps; polygon centres before the last sweep
            for (uint32_t it = 0; it < sp.link_iters; it++) {
                if (P_ && it + 1 == sp.link_iters) {
                    hipLaunchKernelGGL(k_poly_centers, dim3((P_ + 255) / 256), dim3(256), 0, stream_, pos_.p, ptab_.p, P_, centers_.p);
                    st.launches++;
                }
                for (uint32_t d = 0; d < depth_; d++) {
                    uint32_t o = layer_off_[d], n = layer_off_[d + 1] - o;
                    hipLaunchKernelGGL(k_link_layer, dim3((n + 255) / 256), dim3(256), 0, stream_, pos_.p, links_.p + o, n, crad_.p, Np_);
                    st.launches++;
                }
            }
            if (P_ && sp.link_iters == 0) {
                hipLaunchKernelGGL(k_poly_centers, dim3((P_ + 255) / 256), dim3(256), 0, stream_, pos_.p, ptab_.p, P_, centers_.p);
                st.launches++;
            }
            if (Nc_ >= 2) {
                int rc = circles(st, err);
                if (rc) return rc;
            }
            if (P_ >= 2) {
                hipLaunchKernelGGL(k_poly_pairs, dim3(1), dim3(256), 0, stream_, pos_.p, ptab_.p, P_, centers_.p);
                st.launches++;
            }
            bp.first_acc = (sp.first_acc && step == 0 && acc0_.p) ? 1u : 0u;
            if (N_) {
                uint32_t blocks = std::min<uint32_t>((N_ + 255) / 256, 65535u * 4);
                if (ncol_ <= (uint32_t)kMaxStagedColliders && nplanes_ <= (uint32_t)kMaxStagedPlanes)
                    hipLaunchKernelGGL(k_bodies<true>, dim3(blocks), dim3(256), 0, stream_, pos_.p, prev_.p,
                                       has_pins_ ? pinmask_.p : nullptr, crad_.p, acc0_.p, gflags_.p, colhdr_.p,
                                       colgeo_.p, planes_.p, bp);
                else
                    hipLaunchKernelGGL(k_bodies<false>, dim3(blocks), dim3(256), 0, stream_, pos_.p, prev_.p,
                                       has_pins_ ? pinmask_.p : nullptr, crad_.p, acc0_.p, gflags_.p, colhdr_.p,
                                       colgeo_.p, planes_.p, bp);
                st.launches++;
            }
            st.substeps++;
        }
        if (sp.first_acc && acc0_.p) {
            uint32_t n = Nc_ + Ng_;
            hipLaunchKernelGGL(k_zero_acc0, dim3((n + 255) / 256), dim3(256), 0, stream_, acc0_.p, n);
            st.launches++;
        }
        BD_HIP(hipGetLastError());
        return BD_OK;
    }

    int download(HostScene &s, std::string &err) override {
        if (N_) {
            std::vector<float2> pos(N_), prev(N_);
            BD_HIP(hipMemcpyAsync(pos.data(), pos_.p, N_ * sizeof(float2), hipMemcpyDeviceToHost, stream_));
            BD_HIP(hipMemcpyAsync(prev.data(), prev_.p, N_ * sizeof(float2), hipMemcpyDeviceToHost, stream_));
            std::vector<float2> centers(P_);
            if (P_) BD_HIP(hipMemcpyAsync(centers.data(), centers_.p, P_ * sizeof(float2), hipMemcpyDeviceToHost, stream_));
            BD_HIP(hipStreamSynchronize(stream_));
            auto get = [&](std::vector<V2> &a, std::vector<V2> &b, uint32_t base) {
                for (size_t i = 0; i < a.size(); i++) {
                    a[i] = V2{pos[base + i].x, pos[base + i].y};
                    b[i] = V2{prev[base + i].x, prev[base + i].y};
                }
            };
            get(s.p_pos, s.p_prev, 0);
            get(s.c_pos, s.c_prev, Np_);
            get(s.g_pos, s.g_prev, Np_ + Nc_);
            for (uint32_t p : static_poly_) s.polys[p].center = V2{centers[p].x, centers[p].y};
        }
        return BD_OK;
    }

    int sync(std::string &err) override {
        BD_HIP(hipStreamSynchronize(stream_));
        return BD_OK;
    }

private:
    // Exact replay of the circle-pair sweep (see k_cc_round).  The sweep is
    // redone with a larger margin if any circle moved by 0.45 * margin or more,
    // since then a pair left out of the candidate list could have met.
    int circles(Stats &st, std::string &err) {
        const uint32_t nc = Nc_;
        const uint32_t nb = (nc + 255) / 256;
        uint32_t T = 1;
        while (T < 2 * nc) T <<= 1;
        const uint32_t tmask = T - 1;
        BD_HIP(cc_cell_.alloc(nc)); BD_HIP(cc_key_.alloc(nc)); BD_HIP(cc_items_.alloc(nc));
        BD_HIP(cc_save_.alloc(nc)); BD_HIP(cc_path_.alloc(nc));
        BD_HIP(cc_cnt_.alloc(nc + 1)); BD_HIP(cc_off_.alloc(nc + 1));
        BD_HIP(cc_head_.alloc(2 * (size_t)nc));
        BD_HIP(cc_count_.alloc(T + 1)); BD_HIP(cc_start_.alloc(T + 1)); BD_HIP(cc_fill_.alloc(T));
        BD_HIP(cc_ctl_.alloc(kBurst + 2));
        for (int attempt = 0; attempt < 12; attempt++) {
            float margin = margin_;
            double h = ((double)(2.0f * rmax_) + (double)margin) * 1.001 + 1e-30;
            double inv_h = 1.0 / h;
            fill_u32(cc_count_.p, T + 1, 0u, st);
            fill_u32(cc_fill_.p, T, 0u, st);
            hipLaunchKernelGGL(k_cc_prep, dim3(nb), dim3(256), 0, stream_, pos_.p, Np_, nc, inv_h, tmask, cc_cell_.p,
                               cc_key_.p, cc_count_.p, cc_save_.p, cc_path_.p);
            st.launches++;
            int rc = exclusive_scan_u32(cc_count_.p, cc_start_.p, T + 1, stream_, scan_tmp_, st.launches, err);
            if (rc) return rc;
            hipLaunchKernelGGL(k_cc_scatter, dim3(nb), dim3(256), 0, stream_, cc_key_.p, nc, cc_start_.p, cc_fill_.p, cc_items_.p);
            hipLaunchKernelGGL(k_cc_cand<0>, dim3(nb), dim3(256), 0, stream_, cc_save_.p, crad_.p, cc_cell_.p, nc, tmask,
                               cc_start_.p, cc_items_.p, margin, cc_cnt_.p, cc_off_.p, nullptr);
            st.launches += 2;
            fill_u32(cc_cnt_.p + nc, 1, 0u, st);
            rc = exclusive_scan_u32(cc_cnt_.p, cc_off_.p, nc + 1, stream_, scan_tmp_, st.launches, err);
            if (rc) return rc;
            // total candidate entries -> allocate the list
            BD_HIP(hipMemcpyAsync(host_ctl_, cc_off_.p + nc, 4, hipMemcpyDeviceToHost, stream_));
            BD_HIP(hipStreamSynchronize(stream_));
            uint32_t total = host_ctl_[0];
            BD_HIP(cc_cand_.alloc(std::max<uint32_t>(total, 1)));
            hipLaunchKernelGGL(k_cc_cand<1>, dim3(nb), dim3(256), 0, stream_, cc_save_.p, crad_.p, cc_cell_.p, nc, tmask,
                               cc_start_.p, cc_items_.p, margin, cc_cnt_.p, cc_off_.p, cc_cand_.p);
            st.launches++;
            st.cc_pairs = total / 2;
            // rounds
            fill_u32(cc_head_.p, nc, 0u, st);
            uint32_t *hin = cc_head_.p, *hout = cc_head_.p + nc;
            fill_u32(cc_ctl_.p + kBurst, 1, 0u, st);  // pmax bits
            bool done = total == 0;
            uint32_t rounds = 0;
            while (!done) {
                fill_u32(cc_ctl_.p, kBurst, 0u, st);
                for (uint32_t r = 0; r < kBurst; r++) {
                    hipLaunchKernelGGL(k_cc_round, dim3(nb), dim3(256), 0, stream_, pos_.p, Np_, crad_.p, nc, cc_cnt_.p,
                                       cc_off_.p, cc_cand_.p, hin, hout, cc_path_.p, cc_ctl_.p, r, cc_ctl_.p + kBurst);
                    std::swap(hin, hout);
                    st.launches++;
                }
                rounds += kBurst;
                BD_HIP(hipMemcpyAsync(host_ctl_, cc_ctl_.p + kBurst - 1, 8, hipMemcpyDeviceToHost, stream_));
                BD_HIP(hipStreamSynchronize(stream_));
                done = host_ctl_[0] == 0;
                if (rounds > 4 * total + 64) {
                    err = "circle pair schedule made no progress";
                    return BD_ERR_COLLISION;
                }
            }
            st.cc_rounds += rounds;
            float pmax = 0.0f;
            if (total) std::memcpy(&pmax, &host_ctl_[1], 4);
            if ((double)pmax <= 0.45 * (double)margin) {
                if ((double)pmax < 0.1 * (double)margin && margin > 0.05f * std::max(rmax_, 1e-3f)) margin_ = margin * 0.5f;
                st.cc_margin = margin_;
                return BD_OK;
            }
            // a circle moved too far for the candidate margin: undo and widen
            st.cc_redos++;
            hipLaunchKernelGGL(k_cc_restore, dim3(nb), dim3(256), 0, stream_, pos_.p, Np_, cc_save_.p, nc);
            st.launches++;
            if (!(pmax < 1e30f)) {
                err = "circle pair sweep produced a non-finite push";
                return BD_ERR_COLLISION;
            }
            margin_ = 2.5f * pmax + margin;
        }
        err = "circle pair margin did not converge";
        return BD_ERR_COLLISION;
    }

    void fill_u32(uint32_t *p, size_t n, uint32_t v, Stats &st) {
        launch_fill_u32(p, n, v, stream_);
        st.launches++;
    }

    static constexpr uint32_t kBurst = 16;
    hipStream_t stream_;
    uint32_t Np_ = 0, Nc_ = 0, Ng_ = 0, N_ = 0, P_ = 0, depth_ = 0, ncol_ = 0, nplanes_ = 0;
    bool has_pins_ = false;
    float rmax_ = 0.0f, margin_ = 0.0f;
    std::vector<uint32_t> layer_off_;
    std::vector<uint32_t> static_poly_;
    DevBuf<float2> pos_, prev_, acc0_, centers_;
    DevBuf<Link12> links_;
    DevBuf<uint32_t> pinmask_;
    DevBuf<float> crad_;
    DevBuf<uint2> ptab_;
    DevBuf<uint8_t> gflags_;
    DevBuf<int4> colhdr_;
    DevBuf<float4> colgeo_, planes_;
    // circle schedule workspace
    DevBuf<int2> cc_cell_;
    DevBuf<uint32_t> cc_key_, cc_items_, cc_cnt_, cc_off_, cc_head_, cc_count_, cc_start_, cc_fill_, cc_ctl_, cc_cand_;
    DevBuf<float2> cc_save_;
    DevBuf<float> cc_path_;
    ScanTemp scan_tmp_;
    uint32_t *host_ctl_ = nullptr;
};

Engine *make_global_engine(hipStream_t stream) { return new GlobalEngine(stream); }

}  // namespace bd
