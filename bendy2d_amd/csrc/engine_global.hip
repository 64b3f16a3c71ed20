// engine_global.hip -- GLOBAL engine: the whole device works on one scene.
//
// Per substep (src/solver.rs:110-114):
//   links      one launch per dependency layer (k_link_layer), k sweeps; the
//              layers reproduce the sequential Gauss-Seidel order exactly
//              (src/solver.rs:143-153)
//   circles    candidate pairs from a hierarchical hashed cell broadphase
//              (per-circle margins, counting sort with wave prefix scans), the
//              connected components of the candidate graph (union-find), and an
//              exact replay of the reference's (i, j) lexicographic sweep
//              (src/solver.rs:168-177) per component: by one lane, by one
//              wavefront in LDS, or by a workgroup in global memory, by size;
//              a sweep whose circles left their margins is checked pair by pair
//              at each skipped pair's turn and redone only if one met
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
//
// Every sweep builds candidate lists from the circles' positions at its
// start (save).  Circle i gets a margin m_i = max(s0, need_i), where need_i
// follows how far the circle's pushes took it from save in recent sweeps; its
// reach is r_i + m_i, and (i, j) is a candidate pair when
// |save_i - save_j| < reach_i + reach_j.  A pair left out can only meet during
// the sweep if e_i + e_j > m_i + m_j, e being each circle's largest distance
// from save during the sweep (excursion), so the sweep stands when every
// e_i <= 0.9 m_i; otherwise it is undone and run again with the larger needs.
//
// Margins can differ by orders of magnitude (a jammed ring of circles around a
// collider beside quiet gas), so the grid is hierarchical: circle i sits at
// level L_i, the smallest L with m_i <= lim_L = s0 4^L, in a grid of cell size
// h_L = 2 (rmax + lim_L) (x1.001); all levels share one hash table keyed by
// (L, cell).  Circle i scans the levels L <= L_i over the cells that its reach
// plus level L's largest reach covers, and the levels above L_i over the 3x3
// cells around its own cell there.  Both sides of a pair find it, so each
// circle writes its own list without atomics.
//
// The sweep then runs per connected component of the candidate graph: pairs
// of different components share no circle, so they commute, and within one
// component the reference's (i, j) order (src/solver.rs:168-177) is kept.
// Components of at most kLaneMax circles are replayed pair by pair by one
// lane; larger ones by a workgroup with the round protocol: pair (i, j) runs
// once j is the next unprocessed partner of i and i the next of j.

constexpr int kLevels = 12;
constexpr uint32_t kSlotG = 8;  // candidates kept by the counting pass
constexpr uint32_t kLvlShift = 26;  // bucket record .w: index | level << 26
constexpr uint32_t kIdxMaskG = (1u << kLvlShift) - 1u;
constexpr uint32_t kLaneMax = 4;
constexpr uint32_t kLdsMembers = 128, kLdsEntries = 1024;
constexpr int64_t kMaxScanCells = 1024;
constexpr uint32_t kNoneG = 0xffffffffu;
constexpr int kCompWG = 256;

// Per-sweep grid of every level, in device memory.
struct CGrid {
    double inv_h[kLevels];
    float lim[kLevels];        // largest margin of level L: s0 * 4^L
    float reach_max[kLevels];  // largest reach of level L: rmax + lim
};

// Control words of a sweep (zeroed with the bucket counters).
enum : uint32_t {
    kCtlMask = 0, kCtlOver, kCtlBad, kCtlRounds, kCtlNbig, kCtlQueue, kCtlInf, kCtlMissed, kCtlQuiet, kCtlNmid,
    kCtlEmax = 16, kCtlWords = kCtlEmax + kLevels
};

__device__ __forceinline__ int cell_coord_d(double x, double inv_h) {
    double c = floor(x * inv_h);
    if (!(c > -1073741824.0)) c = -1073741824.0;
    if (c > 1073741824.0) c = 1073741824.0;
    return (int)c;
}

// Cell of a point for a cell size 1 / inv_h, clamped to +-2^30 so that huge
// or non-finite coordinates stay finite.
__device__ __forceinline__ int cell_coord(float x, double inv_h) { return cell_coord_d((double)x, inv_h); }

__device__ __forceinline__ uint32_t level_hash(uint32_t L, int cx, int cy) {
    return cell_hash(cx, cy) ^ (L * 0x9e3779b9u);
}

__global__ void k_cg_grid(float s0, float rmax, CGrid *__restrict__ g) {
    const int L = threadIdx.x;
    if (L >= kLevels) return;
    const float lim = ldexpf(s0, 2 * L);
    const float rm = rmax + lim;
    g->lim[L] = lim;
    g->reach_max[L] = rm;
    g->inv_h[L] = 1.0 / ((2.0 * (double)rm) * 1.001 + 1e-30);
}

__global__ void __launch_bounds__(256) k_cg_prep(const float2 *__restrict__ pos, uint32_t cbase,
                                                 const float *__restrict__ crad, const float *__restrict__ need,
                                                 const CGrid *__restrict__ g, uint32_t nc, uint32_t tmask,
                                                 float2 *__restrict__ save, float *__restrict__ exc,
                                                 float *__restrict__ marg, float *__restrict__ reach,
                                                 uint8_t *__restrict__ lvl, uint32_t *__restrict__ key,
                                                 uint32_t *__restrict__ count, uint32_t *__restrict__ par,
                                                 uint32_t *__restrict__ ctl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bit = 0;
    if (i < nc) {
        const float2 p = pos[cbase + i];
        save[i] = p;
        exc[i] = 0.0f;
        par[i] = i;
        float m = fmaxf(g->lim[0], need[i]);
        uint32_t L = 0;
        while (L + 1 < (uint32_t)kLevels && m > g->lim[L]) L++;
        m = fminf(m, g->lim[L]);
        marg[i] = m;
        reach[i] = crad[i] + m;
        lvl[i] = (uint8_t)L;
        const double ih = g->inv_h[L];
        const uint32_t k = level_hash(L, cell_coord(p.x, ih), cell_coord(p.y, ih)) & tmask;
        key[i] = k;
        atomicAdd(&count[k], 1u);
        bit = 1u << L;
    }
    // occupied levels above 0 (level 0 counts as occupied): one atomic per
    // wavefront that holds such a level, so the common all-level-0 wavefront
    // takes none
    for (int o = 32; o > 0; o >>= 1) bit |= __shfl_xor(bit, o);
    if ((threadIdx.x & 63u) == 0u && (bit & ~1u)) atomicOr(&ctl[kCtlMask], bit);
}

__global__ void __launch_bounds__(256) k_cg_scatter(const float2 *__restrict__ save, const float *__restrict__ reach,
                                                    const uint8_t *__restrict__ lvl, const uint32_t *__restrict__ key,
                                                    uint32_t nc, const uint32_t *__restrict__ start,
                                                    uint32_t *__restrict__ fill, float4 *__restrict__ rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const uint32_t k = key[i];
    const uint32_t s = start[k] + atomicAdd(&fill[k], 1u);
    const float2 p = save[i];
    rec[s] = make_float4(p.x, p.y, reach[i], __uint_as_float(i | ((uint32_t)lvl[i] << kLvlShift)));
}

__device__ __forceinline__ void cswap_u32(uint32_t &a, uint32_t &b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo; b = hi;
}

// Ascending sort of a list in global memory: insertion for short lists,
// heapsort for long ones.
__device__ void sort_list(uint32_t *a, uint32_t n) {
    if (n <= 32u) {
        for (uint32_t x = 1; x < n; x++) {
            const uint32_t v = a[x];
            uint32_t y = x;
            while (y > 0 && a[y - 1] > v) { a[y] = a[y - 1]; y--; }
            a[y] = v;
        }
        return;
    }
    auto sift = [&](uint32_t root, uint32_t end) {
        for (;;) {
            uint32_t child = 2u * root + 1u;
            if (child >= end) return;
            if (child + 1u < end && a[child] < a[child + 1u]) child++;
            const uint32_t vr = a[root], vc = a[child];
            if (!(vr < vc)) return;
            a[root] = vc; a[child] = vr;
            root = child;
        }
    };
    for (uint32_t s = n / 2u; s-- > 0;) sift(s, n);
    for (uint32_t e = n - 1u; e > 0; e--) {
        const uint32_t t = a[0]; a[0] = a[e]; a[e] = t;
        sift(0, e);
    }
}

// Candidates of circle i (see above).  MODE 0 counts them and keeps the
// first kSlotG in slot; MODE 1 writes the list at off[i], from slot when it
// fits there and by a second gather otherwise, and sorts it ascending.
template <int MODE>
__global__ void __launch_bounds__(256) k_cg_cand(const float2 *__restrict__ save, const float *__restrict__ reach,
                                                 const uint8_t *__restrict__ lvl, const CGrid *__restrict__ g,
                                                 const uint32_t *__restrict__ ctl, const uint32_t *__restrict__ start,
                                                 const float4 *__restrict__ rec, uint32_t nc, uint32_t tmask,
                                                 uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                                                 uint32_t *__restrict__ cand, uint32_t *__restrict__ slot) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (MODE == 0 && i == 0) cnt[nc] = 0u;  // the scan's last entry
    if (i >= nc) return;
    uint32_t *const keep = slot + (size_t)i * kSlotG;
    uint32_t *out = MODE == 1 ? cand + off[i] : nullptr;
    if (MODE == 1) {
        const uint32_t n0 = cnt[i];
        if (n0 <= kSlotG) {
            uint32_t v[kSlotG];
#pragma unroll
            for (uint32_t k = 0; k < kSlotG; k++) v[k] = k < n0 ? keep[k] : kNoneG;
#pragma unroll
            for (uint32_t k = 2; k <= kSlotG; k <<= 1)
#pragma unroll
                for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                    for (uint32_t x = 0; x < kSlotG; x++) {
                        const uint32_t l = x ^ j;
                        if (l > x) {
                            if ((x & k) == 0u) cswap_u32(v[x], v[l]);
                            else cswap_u32(v[l], v[x]);
                        }
                    }
#pragma unroll
            for (uint32_t k = 0; k < kSlotG; k++)
                if (k < n0) out[k] = v[k];
            return;
        }
    }
    const float2 p = save[i];
    uint32_t n = 0;
    // a circle at a non-finite position takes part in no pair
    if (fabsf(p.x) <= 3.0e38f && fabsf(p.y) <= 3.0e38f) {
        const float ri = reach[i];
        const uint32_t Li = lvl[i];
        const uint32_t mask = ctl[kCtlMask] | 1u;
        const uint32_t below = mask & ((2u << Li) - 1u);  // levels <= Li
        auto test = [&](uint32_t j, float qx, float qy, float rj) {
            const float dx = p.x - qx, dy = p.y - qy, lim = ri + rj;
            if (dx * dx + dy * dy < lim * lim) {
                if (MODE == 1) out[n] = j;
                else if (n < kSlotG) keep[n] = j;
                n++;
            }
        };
        auto visit = [&](uint32_t L, int cx, int cy, double ih) {
            const uint32_t k = level_hash(L, cx, cy) & tmask;
            for (uint32_t s = start[k], e = start[k + 1]; s < e; s++) {
                const float4 q = rec[s];
                const uint32_t w = __float_as_uint(q.w);
                const uint32_t j = w & kIdxMaskG;
                if ((w >> kLvlShift) != L || j == i) continue;
                if (cell_coord(q.x, ih) != cx || cell_coord(q.y, ih) != cy) continue;
                test(j, q.x, q.y, q.z);
            }
        };
        // a reach spanning too many cells of a finer level scans every circle
        bool brute = false;
        for (uint32_t m = below; m; m &= m - 1u) {
            const uint32_t L = (uint32_t)__ffs(m) - 1u;
            const double R = ((double)ri + (double)g->reach_max[L]) * (1.0 + 1e-6);
            const double ih = g->inv_h[L];
            const int64_t nx = (int64_t)cell_coord_d((double)p.x + R, ih) - cell_coord_d((double)p.x - R, ih) + 1;
            const int64_t ny = (int64_t)cell_coord_d((double)p.y + R, ih) - cell_coord_d((double)p.y - R, ih) + 1;
            if (nx * ny > kMaxScanCells) brute = true;
        }
        if (brute) {
            for (uint32_t j = 0; j < nc; j++) {
                if (j == i) continue;
                const float2 q = save[j];
                test(j, q.x, q.y, reach[j]);
            }
        } else {
            for (uint32_t m = below; m; m &= m - 1u) {
                const uint32_t L = (uint32_t)__ffs(m) - 1u;
                const double R = ((double)ri + (double)g->reach_max[L]) * (1.0 + 1e-6);
                const double ih = g->inv_h[L];
                const int x0 = cell_coord_d((double)p.x - R, ih), x1 = cell_coord_d((double)p.x + R, ih);
                const int y0 = cell_coord_d((double)p.y - R, ih), y1 = cell_coord_d((double)p.y + R, ih);
                for (int cy = y0; cy <= y1; cy++)
                    for (int cx = x0; cx <= x1; cx++) visit(L, cx, cy, ih);
            }
            for (uint32_t m = mask & ~((2u << Li) - 1u); m; m &= m - 1u) {
                const uint32_t L = (uint32_t)__ffs(m) - 1u;
                const double ih = g->inv_h[L];
                const int cx = cell_coord(p.x, ih), cy = cell_coord(p.y, ih);
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) visit(L, cx + dx, cy + dy, ih);
            }
        }
    }
    if (MODE == 0) cnt[i] = n;
    else sort_list(out, n);
}

// Union-find over the candidate graph.  parent[x] <= x always (a root is
// hooked under the smaller root), so every value ever read is an ancestor.
__device__ __forceinline__ uint32_t uf_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t uf_find(uint32_t *par, uint32_t x) {
    for (;;) {
        const uint32_t p = uf_load(par + x);
        if (p == x) return x;
        const uint32_t g = uf_load(par + p);
        if (g == p) return p;
        __hip_atomic_store(par + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // path halving
        x = g;
    }
}

__global__ void __launch_bounds__(256) k_cg_union(const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ cand, uint32_t nc, uint32_t *par) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    for (uint32_t s = off[i], e = s + cnt[i]; s < e; s++) {
        uint32_t a = i, b = cand[s];
        if (b < a) continue;  // each edge once, from its smaller end
        for (;;) {
            a = uf_find(par, a);
            b = uf_find(par, b);
            if (a == b) break;
            if (a > b) { const uint32_t t = a; a = b; b = t; }
            const uint32_t old = atomicCAS(par + b, b, a);
            if (old == b) break;
            b = old;  // b was hooked meanwhile
        }
    }
}

// Component of each listed circle (its smallest index) and component sizes.
__global__ void __launch_bounds__(256) k_cg_comp(const uint32_t *__restrict__ cnt, uint32_t nc, uint32_t *par,
                                                 uint32_t *__restrict__ root, uint32_t *__restrict__ csize) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc || cnt[i] == 0u) return;
    const uint32_t r = uf_find(par, i);
    root[i] = r;
    atomicAdd(&csize[r], 1u);
}

// Members of each component, grouped at cstart[root]; large components queue
// for a workgroup.
__global__ void __launch_bounds__(256) k_cg_members(const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ root,
                                                    const uint32_t *__restrict__ csize,
                                                    const uint32_t *__restrict__ cstart, uint32_t nc,
                                                    uint32_t *__restrict__ cfill, uint32_t *__restrict__ members,
                                                    uint32_t *__restrict__ midq, uint32_t *__restrict__ bigq,
                                                    uint32_t *__restrict__ ctl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    bool mid = false, big = false;
    if (i < nc && cnt[i] != 0u) {
        const uint32_t r = root[i];
        members[cstart[r] + atomicAdd(&cfill[r], 1u)] = i;
        if (r == i && csize[r] > kLaneMax) {
            mid = csize[r] <= kLdsMembers;
            big = !mid;
        }
    }
    // queue appends: one atomic per wavefront and queue
    const uint64_t bm = __ballot(mid), bb = __ballot(big);
    const uint64_t below = (1ull << lane) - 1ull;
    if (bm) {
        uint32_t base = 0;
        if (lane == 0u) base = atomicAdd(&ctl[kCtlNmid], (uint32_t)__popcll(bm));
        base = __shfl(base, 0);
        if (mid) midq[base + (uint32_t)__popcll(bm & below)] = i;
    }
    if (bb) {
        uint32_t base = 0;
        if (lane == 0u) base = atomicAdd(&ctl[kCtlNbig], (uint32_t)__popcll(bb));
        base = __shfl(base, 0);
        if (big) bigq[base + (uint32_t)__popcll(bb & below)] = i;
    }
}

// Pairs (c, j), j > c, of circle c's list in order, with live positions; the
// calling lane owns every circle of the component.
// Both circles' positions after each pair go to traj at the pair's entry in
// their lists; head[j] counts j's pairs done so far, which is the index of
// the current pair in j's list since each circle meets its pairs in list order.
__device__ __forceinline__ void replay_list(uint32_t c, float2 *__restrict__ P, const float *__restrict__ crad,
                                            const float2 *__restrict__ save, float *__restrict__ exc,
                                            const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                                            const uint32_t *__restrict__ cand, uint32_t *__restrict__ head,
                                            float2 *__restrict__ traj) {
    float2 pc = P[c];
    const float rc = crad[c];
    const float2 sc = save[c];
    float ec = exc[c];
    for (uint32_t s = off[c], e = s + cnt[c]; s < e; s++) {
        const uint32_t j = cand[s];
        if (j <= c) continue;  // done at j's turn
        float2 pj = P[j];
        float qi = 0.0f, qj = 0.0f;
        if (solve_circle_pair(pc.x, pc.y, pj.x, pj.y, rc, crad[j], qi, qj)) {
            P[j] = pj;
            const float2 sj = save[j];
            exc[j] = excursion(exc[j], pj.x, pj.y, sj.x, sj.y);
            ec = excursion(ec, pc.x, pc.y, sc.x, sc.y);
        }
        traj[s] = pc;
        const uint32_t hj = head[j];
        if (hj < cnt[j]) traj[off[j] + hj] = pj;  // always true for symmetric lists
        head[j] = hj + 1u;
    }
    P[c] = pc;
    exc[c] = ec;
}

// Components of at most kLaneMax circles: the lane of the component's root
// sorts its members and replays them in ascending order.
__global__ void __launch_bounds__(256) k_cg_lane(float2 *__restrict__ pos, uint32_t cbase, const float *__restrict__ crad,
                                                 const float2 *__restrict__ save, float *__restrict__ exc,
                                                 const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                                                 const uint32_t *__restrict__ cand, const uint32_t *__restrict__ root,
                                                 const uint32_t *__restrict__ csize,
                                                 const uint32_t *__restrict__ cstart, uint32_t *__restrict__ members,
                                                 uint32_t nc, uint32_t *__restrict__ head, float2 *__restrict__ traj) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc || cnt[i] == 0u || root[i] != i) return;
    const uint32_t n = csize[i];
    if (n > kLaneMax) return;
    uint32_t *M = members + cstart[i];
    uint32_t v[kLaneMax];
#pragma unroll
    for (uint32_t k = 0; k < kLaneMax; k++) v[k] = k < n ? M[k] : kNoneG;
    // bitonic network, ascending
#pragma unroll
    for (uint32_t k = 2; k <= kLaneMax; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (uint32_t x = 0; x < kLaneMax; x++) {
                const uint32_t l = x ^ j;
                if (l > x) {
                    if ((x & k) == 0u) cswap_u32(v[x], v[l]);
                    else cswap_u32(v[l], v[x]);
                }
            }
#pragma unroll
    for (uint32_t k = 0; k < kLaneMax; k++)
        if (k < n) {
            M[k] = v[k];
            head[v[k]] = 0u;
        }
    float2 *P = pos + cbase;
    for (uint32_t k = 0; k < n; k++) replay_list(M[k], P, crad, save, exc, cnt, off, cand, head, traj);
}

// Components of kLaneMax+1 .. kLdsMembers circles: one wavefront each (a
// grid stride over the queue), with the round protocol in LDS.  The members' positions, start
// positions, radii and excursions are staged into LDS and the lists remapped to
// local indices (slot[c] = c's place among the members); the lists keep their
// order, which is all the protocol depends on.  Positions after each pair go to
// traj in global memory as in the other replays.  A component whose lists do
// not fit joins the queue of the global-memory workgroups.
__global__ void __launch_bounds__(64) k_cg_lds(float2 *__restrict__ pos, uint32_t cbase,
                                               const float *__restrict__ crad, float *__restrict__ exc,
                                               const float2 *__restrict__ save, const uint32_t *__restrict__ cnt,
                                               const uint32_t *__restrict__ off, const uint32_t *__restrict__ cand,
                                               const uint32_t *__restrict__ csize, const uint32_t *__restrict__ cstart,
                                               const uint32_t *__restrict__ members, const uint32_t *__restrict__ midq,
                                               uint32_t *__restrict__ bigq, uint32_t *__restrict__ slot,
                                               float2 *__restrict__ traj, uint32_t *__restrict__ ctl) {
    __shared__ float2 s_p[kLdsMembers], s_s[kLdsMembers];
    __shared__ float s_r[kLdsMembers], s_e[kLdsMembers];
    __shared__ uint32_t s_off[kLdsMembers];
    __shared__ uint16_t s_lb[kLdsMembers + 1], s_cnt[kLdsMembers], s_head[kLdsMembers], s_nxp[kLdsMembers],
        s_stamp[kLdsMembers];
    __shared__ uint16_t s_list[kLdsEntries];
    __shared__ uint32_t s_wl[2][kLdsMembers / 2 + 1];
    __shared__ uint32_t s_n[3], s_bad;
    constexpr uint16_t kNone16 = 0xffffu;
    const uint32_t tid = threadIdx.x;
    const uint32_t nmid = ctl[kCtlNmid];
    float2 *P = pos + cbase;
    for (uint32_t q = blockIdx.x; q < nmid; q += gridDim.x) {  // uniform per workgroup
        const uint32_t r = midq[q], n = csize[r];
        const uint32_t *M = members + cstart[r];
        // stage; each lane takes kLdsMembers / 64 consecutive members for the scan
        constexpr uint32_t kPer = kLdsMembers / 64u;
        uint32_t sum = 0;
        for (uint32_t u = 0; u < kPer; u++) {
            const uint32_t k = tid * kPer + u;
            if (k < n) {
                const uint32_t c = M[k];
                slot[c] = k;
                s_p[k] = P[c];
                s_s[k] = save[c];
                s_r[k] = crad[c];
                s_e[k] = exc[c];
                s_off[k] = off[c];
                const uint32_t m = cnt[c];
                s_cnt[k] = (uint16_t)min(m, 0xffffu);
                sum += m;
            }
        }
        // exclusive list starts: a wavefront scan of the lanes' sums
        uint32_t incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if ((int)tid >= o) incl += t;
        }
        const uint32_t total = __shfl(incl, 63);
        if (total > kLdsEntries) {  // uniform: the lists do not fit
            if (tid == 0) bigq[atomicAdd(&ctl[kCtlNbig], 1u)] = r;
            __syncthreads();
            continue;
        }
        uint32_t run = incl - sum;
        for (uint32_t u = 0; u < kPer; u++) {
            const uint32_t k = tid * kPer + u;
            if (k < n) {
                s_lb[k] = (uint16_t)run;
                run += s_cnt[k];
            }
        }
        if (tid == 0) { s_n[0] = 0u; s_n[1] = 0u; s_n[2] = 0u; s_bad = 0u; }
        __syncthreads();
        // local lists
        for (uint32_t k = tid; k < n; k += 64u) {
            const uint32_t *src = cand + s_off[k];
            uint16_t *dst = s_list + s_lb[k];
            for (uint32_t e = 0, m = s_cnt[k]; e < m; e++) dst[e] = (uint16_t)slot[src[e]];
        }
        __syncthreads();
        for (uint32_t k = tid; k < n; k += 64u) {
            s_head[k] = 0;
            s_nxp[k] = s_cnt[k] ? s_list[s_lb[k]] : kNone16;
            s_stamp[k] = kNone16;
        }
        __syncthreads();
        for (uint32_t k = tid; k < n; k += 64u) {
            const uint32_t o = s_nxp[k];
            if (o != kNone16 && k < o && s_nxp[o] == k) s_wl[0][atomicAdd(&s_n[0], 1u)] = k | (o << 16);
        }
        __syncthreads();
        uint32_t cur = s_n[0], rounds = 0;
        for (; cur > 0; rounds++) {
            const uint32_t cb = rounds & 1u, nb = cb ^ 1u;
            for (uint32_t k = tid; k < cur; k += 64u) {
                const uint32_t pr = s_wl[cb][k], c = pr & 0xffffu, o = pr >> 16;
                float2 pi = s_p[c], pj = s_p[o];
                const uint32_t hc = s_head[c], ho = s_head[o];
                const uint32_t xc = hc + 1u < s_cnt[c] ? s_list[s_lb[c] + hc + 1u] : kNone16;
                const uint32_t xo = ho + 1u < s_cnt[o] ? s_list[s_lb[o] + ho + 1u] : kNone16;
                float qi = 0.0f, qj = 0.0f;
                if (solve_circle_pair(pi.x, pi.y, pj.x, pj.y, s_r[c], s_r[o], qi, qj)) {
                    s_p[c] = pi;
                    s_p[o] = pj;
                    s_e[c] = excursion(s_e[c], pi.x, pi.y, s_s[c].x, s_s[c].y);
                    s_e[o] = excursion(s_e[o], pj.x, pj.y, s_s[o].x, s_s[o].y);
                }
                traj[s_off[c] + hc] = pi;
                traj[s_off[o] + ho] = pj;
                s_head[c] = (uint16_t)(hc + 1u);
                s_head[o] = (uint16_t)(ho + 1u);
                s_nxp[c] = (uint16_t)xc;
                s_nxp[o] = (uint16_t)xo;
                s_stamp[c] = (uint16_t)rounds;
                s_stamp[o] = (uint16_t)rounds;
            }
            if (tid == 0) s_n[(rounds + 2u) % 3u] = 0u;
            __syncthreads();
            for (uint32_t k = tid; k < cur; k += 64u) {
                const uint32_t pr = s_wl[cb][k];
                for (int e = 0; e < 2; e++) {
                    const uint32_t x = e ? pr >> 16 : pr & 0xffffu, y = s_nxp[x];
                    if (y == kNone16 || s_nxp[y] != x) continue;
                    if (s_stamp[y] == (uint16_t)rounds && y < x) continue;
                    s_wl[nb][atomicAdd(&s_n[(rounds + 1u) % 3u], 1u)] = min(x, y) | (max(x, y) << 16);
                }
            }
            __syncthreads();
            cur = s_n[(rounds + 1u) % 3u];
        }
        for (uint32_t k = tid; k < n; k += 64u) {
            const uint32_t c = M[k];
            P[c] = s_p[k];
            exc[c] = s_e[k];
            if (s_head[k] != s_cnt[k]) s_bad = 1u;
        }
        __syncthreads();
        if (tid == 0) {
            if (s_bad) atomicOr(&ctl[kCtlBad], 1u);
            atomicMax(&ctl[kCtlRounds], rounds);
        }
        __syncthreads();
    }
}

// Components above kLdsMembers circles (or whose lists overflow LDS): one
// workgroup each, taken from a queue,
// with the round protocol.  head[c] counts c's processed partners and nxp[c]
// is the next one (kNoneG when done), so (x, y) is ready exactly when
// nxp[x] == y and nxp[y] == x.  Every round applies at least one pair, so the
// loop ends; a list left unfinished (a schedule that stalls) is reported.
__global__ void __launch_bounds__(kCompWG) k_cg_comps(float2 *__restrict__ pos, uint32_t cbase,
                                                      const float *__restrict__ crad, const float2 *__restrict__ save,
                                                      float *__restrict__ exc, const uint32_t *__restrict__ cnt,
                                                      const uint32_t *__restrict__ off,
                                                      const uint32_t *__restrict__ cand,
                                                      const uint32_t *__restrict__ csize,
                                                      const uint32_t *__restrict__ cstart,
                                                      const uint32_t *__restrict__ members,
                                                      const uint32_t *__restrict__ bigq, uint32_t nc,
                                                      uint32_t *__restrict__ head, uint32_t *__restrict__ nxp,
                                                      uint32_t *__restrict__ stamp, uint2 *__restrict__ wl,
                                                      float2 *__restrict__ traj, uint32_t *__restrict__ ctl) {
    __shared__ uint32_t s_q, s_n[3], s_bad;
    const uint32_t tid = threadIdx.x;
    const uint32_t nbig = ctl[kCtlNbig];
    float2 *P = pos + cbase;
    for (;;) {
        if (tid == 0) s_q = atomicAdd(&ctl[kCtlQueue], 1u);
        __syncthreads();
        const uint32_t q = s_q;
        if (q >= nbig) break;
        const uint32_t r = bigq[q], base = cstart[r], n = csize[r];
        const uint32_t *M = members + base;
        uint2 *const W[2] = {wl + base, wl + nc + base};
        for (uint32_t k = tid; k < n; k += kCompWG) {
            const uint32_t c = M[k];
            head[c] = 0u;
            nxp[c] = cnt[c] ? cand[off[c]] : kNoneG;
            stamp[c] = kNoneG;
        }
        if (tid == 0) { s_n[0] = 0u; s_n[1] = 0u; s_n[2] = 0u; s_bad = 0u; }
        __syncthreads();
        for (uint32_t k = tid; k < n; k += kCompWG) {
            const uint32_t c = M[k], o = nxp[c];
            if (o != kNoneG && c < o && nxp[o] == c) W[0][atomicAdd(&s_n[0], 1u)] = make_uint2(c, o);
        }
        __syncthreads();
        uint32_t cur = s_n[0], rounds = 0;
        for (; cur > 0; rounds++) {
            const uint32_t cb = rounds & 1u, nb = cb ^ 1u;
            for (uint32_t k = tid; k < cur; k += kCompWG) {
                const uint2 pr = W[cb][k];
                const uint32_t c = pr.x, o = pr.y;
                float2 pi = P[c], pj = P[o];
                const uint32_t hc = head[c], ho = head[o];
                const uint32_t oc = off[c], oo = off[o];
                const uint32_t xc = hc + 1u < cnt[c] ? cand[oc + hc + 1u] : kNoneG;
                const uint32_t xo = ho + 1u < cnt[o] ? cand[oo + ho + 1u] : kNoneG;
                float qi = 0.0f, qj = 0.0f;
                if (solve_circle_pair(pi.x, pi.y, pj.x, pj.y, crad[c], crad[o], qi, qj)) {
                    P[c] = pi;
                    P[o] = pj;
                    const float2 sc = save[c], so = save[o];
                    exc[c] = excursion(exc[c], pi.x, pi.y, sc.x, sc.y);
                    exc[o] = excursion(exc[o], pj.x, pj.y, so.x, so.y);
                }
                traj[oc + hc] = pi;
                traj[oo + ho] = pj;
                head[c] = hc + 1u;
                head[o] = ho + 1u;
                nxp[c] = xc;
                nxp[o] = xo;
                stamp[c] = rounds;
                stamp[o] = rounds;
            }
            if (tid == 0) s_n[(rounds + 2u) % 3u] = 0u;  // read two barriers ago
            __syncthreads();
            // pairs completed by a moved head; when both ends moved this round
            // the smaller one adds the pair
            for (uint32_t k = tid; k < cur; k += kCompWG) {
                const uint2 pr = W[cb][k];
                for (int e = 0; e < 2; e++) {
                    const uint32_t x = e ? pr.y : pr.x, y = nxp[x];
                    if (y == kNoneG || nxp[y] != x) continue;
                    if (stamp[y] == rounds && y < x) continue;
                    W[nb][atomicAdd(&s_n[(rounds + 1u) % 3u], 1u)] = make_uint2(min(x, y), max(x, y));
                }
            }
            __syncthreads();
            cur = s_n[(rounds + 1u) % 3u];
        }
        for (uint32_t k = tid; k < n; k += kCompWG)
            if (head[M[k]] != cnt[M[k]]) s_bad = 1u;
        __syncthreads();
        if (tid == 0) {
            if (s_bad) atomicOr(&ctl[kCtlBad], 1u);
            atomicMax(&ctl[kCtlRounds], rounds);
        }
        __syncthreads();  // s_q, s_n and s_bad are rewritten for the next component
    }
}

// Validity of the sweep and the circles' needs for the next one.  Circles
// that went further than 0.9 of their margin are queued in overq, and for each
// level the largest r + e among them goes to ctl[kCtlEmax + L] (as bits:
// positive floats order like their bit patterns).  Atomics are taken once per
// wavefront.
__global__ void __launch_bounds__(256) k_cg_check(const float *__restrict__ exc, const float *__restrict__ marg,
                                                  const float *__restrict__ crad, const uint8_t *__restrict__ lvl,
                                                  float *__restrict__ need, float s0, uint32_t nc,
                                                  uint32_t *__restrict__ overq, uint32_t *__restrict__ ctl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    bool over = false, inf = false, quiet = false;
    uint32_t L = 0, ebits = 0;
    if (i < nc) {
        const float e = exc[i];
        over = !(e <= 0.9f * marg[i]);
        inf = !(e <= 3.0e38f);
        // jammed circles swing from sweep to sweep: the need keeps twice the
        // excursion and forgets slowly
        if (!inf) need[i] = fmaxf(2.0f * e, 0.9f * need[i]);
        if (over && !inf) {
            L = lvl[i];
            ebits = __float_as_uint(crad[i] + e);
            quiet = marg[i] <= s0;  // a circle at the base margin overran
        }
    }
    const uint64_t b = __ballot(over);
    if (!b) return;  // uniform: the wavefront has no overrun
    const uint64_t binf = __ballot(inf), bquiet = __ballot(quiet);
    if (lane == 0u && binf) atomicOr(&ctl[kCtlInf], 1u);
    if (lane == 0u && bquiet) atomicOr(&ctl[kCtlQuiet], 1u);
    uint32_t base = 0;
    if (lane == 0u) base = atomicAdd(&ctl[kCtlOver], (uint32_t)__popcll(b));
    base = __shfl(base, 0);
    if (over) overq[base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull))] = i;
    // per-level maxima: one atomic per (wavefront, level present)
    uint64_t todo = __ballot(ebits != 0u);
    while (todo) {
        const uint32_t src = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
        const uint32_t Ls = __shfl(L, (int)src);
        const bool mine = ebits != 0u && L == Ls;
        uint32_t m = mine ? ebits : 0u;
        for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o));
        if (lane == src) atomicMax(&ctl[kCtlEmax + Ls], m);
        todo &= ~__ballot(mine);
    }
}

// Exact check of a sweep with overruns, one wavefront per overrunning circle.
// Pair (i, j) was skipped if j is not in i's list.  It cannot have met during
// the sweep if |save_i - save_j| >= (r_i + e_i) + (r_j + e_j), e being the
// excursions; a circle within its margin has r_j + e_j <= reach_j, so only
// pairs with an overrunning end can have met, and each overrunning circle
// scans every level L over (r_i + e_i) + max(largest reach of L, largest
// r + e of the overruns at L).  A skipped pair (a, b), a < b, that passes that
// test is tested at its own turn, as the reference would: a has then done its
// pairs with partners below b, b its pairs with partners below a, and traj
// holds each circle's position after each of its pairs.  If it overlapped
// there, the sweep is marked for a redo; otherwise skipping it changed nothing.
__global__ void __launch_bounds__(256) k_cg_verify(const float2 *__restrict__ save, const float *__restrict__ crad,
                                                   const float *__restrict__ exc, const CGrid *__restrict__ g,
                                                   const uint32_t *__restrict__ start, const float4 *__restrict__ rec,
                                                   uint32_t nc, uint32_t tmask, const uint32_t *__restrict__ cnt,
                                                   const uint32_t *__restrict__ off, const uint32_t *__restrict__ cand,
                                                   const uint32_t *__restrict__ overq, uint32_t n_over,
                                                   const float2 *__restrict__ traj, uint32_t *__restrict__ ctl) {
    const uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (k >= n_over) return;  // uniform per wavefront
    const uint32_t i = overq[k];
    const float2 p = save[i];
    if (!(fabsf(p.x) <= 3.0e38f && fabsf(p.y) <= 3.0e38f)) return;
    const float ai = crad[i] + exc[i];
    const uint32_t mask = ctl[kCtlMask] | 1u;
    const uint32_t *const list = cand + off[i];
    const uint32_t nl = cnt[i];
    bool missed = false;
    // entries of a sorted list below v
    auto rank = [&](const uint32_t *l, uint32_t n, uint32_t v) {
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (l[mid] < v) lo = mid + 1u; else hi = mid;
        }
        return lo;
    };
    auto test = [&](uint32_t j, float qx, float qy) {
        const float lim = (ai + (crad[j] + exc[j])) * (1.0f + 0x1p-16f);
        const float dx = p.x - qx, dy = p.y - qy;
        if (!(dx * dx + dy * dy < lim * lim)) return;
        const uint32_t ri = rank(list, nl, j);
        if (ri < nl && list[ri] == j) return;  // listed: replayed in order
        const uint32_t a = min(i, j), b = max(i, j);
        const uint32_t ka = a == i ? ri : rank(cand + off[a], cnt[a], b);
        const uint32_t kb = b == i ? ri : rank(cand + off[b], cnt[b], a);
        const float2 pa = ka ? traj[off[a] + ka - 1u] : save[a];
        const float2 pb = kb ? traj[off[b] + kb - 1u] : save[b];
        const float ex = pa.x - pb.x, ey = pa.y - pb.y;
        const float rs = crad[a] + crad[b];
        if (ex * ex + ey * ey < rs * rs) missed = true;
    };
    auto reach_at = [&](uint32_t L) {
        return ((double)ai + (double)fmaxf(g->reach_max[L], __uint_as_float(ctl[kCtlEmax + L]))) * (1.0 + 1e-4);
    };
    bool brute = false;
    for (uint32_t m = mask; m; m &= m - 1u) {
        const uint32_t L = (uint32_t)__ffs(m) - 1u;
        const double R = reach_at(L), ih = g->inv_h[L];
        const int64_t nx = (int64_t)cell_coord_d((double)p.x + R, ih) - cell_coord_d((double)p.x - R, ih) + 1;
        const int64_t ny = (int64_t)cell_coord_d((double)p.y + R, ih) - cell_coord_d((double)p.y - R, ih) + 1;
        if (nx * ny > kMaxScanCells) brute = true;
    }
    if (brute) {
        for (uint32_t j = lane; j < nc; j += 64u) {
            if (j == i) continue;
            const float2 q = save[j];
            test(j, q.x, q.y);
        }
    } else {
        for (uint32_t m = mask; m; m &= m - 1u) {
            const uint32_t L = (uint32_t)__ffs(m) - 1u;
            const double R = reach_at(L), ih = g->inv_h[L];
            const int x0 = cell_coord_d((double)p.x - R, ih), x1 = cell_coord_d((double)p.x + R, ih);
            const int y0 = cell_coord_d((double)p.y - R, ih), y1 = cell_coord_d((double)p.y + R, ih);
            const uint32_t w = (uint32_t)(x1 - x0 + 1), cells = w * (uint32_t)(y1 - y0 + 1);
            for (uint32_t c = lane; c < cells; c += 64u) {
                const int cx = x0 + (int)(c % w), cy = y0 + (int)(c / w);
                const uint32_t kb = level_hash(L, cx, cy) & tmask;
                for (uint32_t s = start[kb], e = start[kb + 1]; s < e; s++) {
                    const float4 q = rec[s];
                    const uint32_t wq = __float_as_uint(q.w);
                    const uint32_t j = wq & kIdxMaskG;
                    if ((wq >> kLvlShift) != L || j == i) continue;
                    if (cell_coord(q.x, ih) != cx || cell_coord(q.y, ih) != cy) continue;
                    test(j, q.x, q.y);
                }
            }
        }
    }
    if (missed) atomicOr(&ctl[kCtlMissed], 1u);
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
        if (Nc_ > kIdxMaskG) {
            err = "too many circles for one scene";
            return BD_ERR_ARG;
        }
        // per-circle margins start at the base margin
        s0_ = std::max(0.125f * rmax_, s0_min());
        clean_ = 0;
        BD_HIP(cg_need_.alloc(std::max<uint32_t>(Nc_, 1)));
        if (Nc_) BD_HIP(hipMemsetAsync(cg_need_.p, 0, Nc_ * sizeof(float), stream_));
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

    size_t state_count() const override { return 2ull * N_ + P_; }
    hipStream_t stream() const override { return stream_; }

    int copy_state(void *dst, hipMemcpyKind kind, hipStream_t st, std::string &err) override {
        float2 *d = (float2 *)dst;
        if (N_) {
            BD_HIP(hipMemcpyAsync(d, pos_.p, N_ * sizeof(float2), kind, st));
            BD_HIP(hipMemcpyAsync(d + N_, prev_.p, N_ * sizeof(float2), kind, st));
        }
        if (P_) BD_HIP(hipMemcpyAsync(d + 2 * N_, centers_.p, P_ * sizeof(float2), kind, st));
        return BD_OK;
    }

    int unpack(const float2 *src, HostScene &s, std::string &err) override {
        if (!host_shape_matches(s, Np_, Nc_, Ng_, err)) return BD_ERR_STATE;
        const float2 *pos = src, *prev = src + N_, *centers = src + 2 * N_;
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
        return BD_OK;
    }

    int sync(std::string &err) override {
        BD_HIP(hipStreamSynchronize(stream_));
        return BD_OK;
    }

private:
    // Exact replay of the circle-pair sweep (see the circle-pair kernels): lists
    // from per-circle margins on the level grid, components, then the lane and
    // workgroup replays.  A sweep in which some circle went further than 0.9 of
    // its margin is undone and run again with the needs it showed.
    int circles(Stats &st, std::string &err) {
        const uint32_t nc = Nc_;
        const uint32_t nb = (nc + 255) / 256;
        uint32_t T = 1;
        while (T < 2 * nc) T <<= 1;
        const uint32_t tmask = T - 1;
        // counters and control words, zeroed by one fill per attempt
        const size_t zn = (size_t)(T + 1) + T + (nc + 1) + nc + kCtlWords;
        BD_HIP(cg_zero_.alloc(zn));
        uint32_t *const count = cg_zero_.p, *const fill = count + T + 1, *const csize = fill + T;
        uint32_t *const cfill = csize + nc + 1, *const ctl = cfill + nc;
        BD_HIP(cg_grid_.alloc(1));
        BD_HIP(cc_save_.alloc(nc)); BD_HIP(cg_exc_.alloc(nc)); BD_HIP(cg_marg_.alloc(nc)); BD_HIP(cg_reach_.alloc(nc));
        BD_HIP(cg_lvl_.alloc(nc)); BD_HIP(cc_key_.alloc(nc)); BD_HIP(cg_par_.alloc(nc)); BD_HIP(cg_root_.alloc(nc));
        BD_HIP(cc_start_.alloc(T + 1)); BD_HIP(cg_rec_.alloc(nc));
        BD_HIP(cc_cnt_.alloc(nc + 1)); BD_HIP(cc_off_.alloc(nc + 1));
        BD_HIP(cg_cstart_.alloc(nc + 1)); BD_HIP(cg_members_.alloc(nc)); BD_HIP(cg_bigq_.alloc(nc));
        BD_HIP(cc_head_.alloc(nc)); BD_HIP(cg_nxp_.alloc(nc)); BD_HIP(cg_stamp_.alloc(nc));
        BD_HIP(cg_wl_.alloc(2 * (size_t)nc)); BD_HIP(cg_overq_.alloc(nc));
        BD_HIP(cg_midq_.alloc(nc)); BD_HIP(cg_slot_.alloc(nc)); BD_HIP(cg_keep_.alloc((size_t)nc * kSlotG));
        for (int attempt = 0; attempt < 12; attempt++) {
            fill_u32(cg_zero_.p, zn, 0u, st);
            hipLaunchKernelGGL(k_cg_grid, dim3(1), dim3(32), 0, stream_, s0_, rmax_, cg_grid_.p);
            hipLaunchKernelGGL(k_cg_prep, dim3(nb), dim3(256), 0, stream_, pos_.p, Np_, crad_.p, cg_need_.p, cg_grid_.p,
                               nc, tmask, cc_save_.p, cg_exc_.p, cg_marg_.p, cg_reach_.p, cg_lvl_.p, cc_key_.p, count,
                               cg_par_.p, ctl);
            st.launches += 2;
            int rc = exclusive_scan_u32(count, cc_start_.p, T + 1, stream_, scan_tmp_, st.launches, err);
            if (rc) return rc;
            hipLaunchKernelGGL(k_cg_scatter, dim3(nb), dim3(256), 0, stream_, cc_save_.p, cg_reach_.p, cg_lvl_.p,
                               cc_key_.p, nc, cc_start_.p, fill, cg_rec_.p);
            hipLaunchKernelGGL(k_cg_cand<0>, dim3(nb), dim3(256), 0, stream_, cc_save_.p, cg_reach_.p, cg_lvl_.p,
                               cg_grid_.p, ctl, cc_start_.p, cg_rec_.p, nc, tmask, cc_cnt_.p, nullptr, nullptr,
                               cg_keep_.p);
            st.launches += 2;
            rc = exclusive_scan_u32(cc_cnt_.p, cc_off_.p, nc + 1, stream_, scan_tmp_, st.launches, err);
            if (rc) return rc;
            // total candidate entries -> size the list
            BD_HIP(hipMemcpyAsync(host_ctl_, cc_off_.p + nc, 4, hipMemcpyDeviceToHost, stream_));
            BD_HIP(hipStreamSynchronize(stream_));
            const uint32_t total = host_ctl_[0];
            BD_HIP(cc_cand_.alloc(std::max<uint32_t>(total, 1)));
            BD_HIP(cg_traj_.alloc(std::max<uint32_t>(total, 1)));
            hipLaunchKernelGGL(k_cg_cand<1>, dim3(nb), dim3(256), 0, stream_, cc_save_.p, cg_reach_.p, cg_lvl_.p,
                               cg_grid_.p, ctl, cc_start_.p, cg_rec_.p, nc, tmask, cc_cnt_.p, cc_off_.p, cc_cand_.p,
                               cg_keep_.p);
            hipLaunchKernelGGL(k_cg_union, dim3(nb), dim3(256), 0, stream_, cc_cnt_.p, cc_off_.p, cc_cand_.p, nc,
                               cg_par_.p);
            hipLaunchKernelGGL(k_cg_comp, dim3(nb), dim3(256), 0, stream_, cc_cnt_.p, nc, cg_par_.p, cg_root_.p, csize);
            st.launches += 3;
            rc = exclusive_scan_u32(csize, cg_cstart_.p, nc + 1, stream_, scan_tmp_, st.launches, err);
            if (rc) return rc;
            hipLaunchKernelGGL(k_cg_members, dim3(nb), dim3(256), 0, stream_, cc_cnt_.p, cg_root_.p, csize,
                               cg_cstart_.p, nc, cfill, cg_members_.p, cg_midq_.p, cg_bigq_.p, ctl);
            hipLaunchKernelGGL(k_cg_lane, dim3(nb), dim3(256), 0, stream_, pos_.p, Np_, crad_.p, cc_save_.p, cg_exc_.p,
                               cc_cnt_.p, cc_off_.p, cc_cand_.p, cg_root_.p, csize, cg_cstart_.p, cg_members_.p, nc,
                               cc_head_.p, cg_traj_.p);
            hipLaunchKernelGGL(k_cg_lds, dim3(kLdsGrid), dim3(64), 0, stream_, pos_.p, Np_, crad_.p, cg_exc_.p,
                               cc_save_.p, cc_cnt_.p, cc_off_.p, cc_cand_.p, csize, cg_cstart_.p, cg_members_.p,
                               cg_midq_.p, cg_bigq_.p, cg_slot_.p, cg_traj_.p, ctl);
            hipLaunchKernelGGL(k_cg_comps, dim3(kCompGrid), dim3(kCompWG), 0, stream_, pos_.p, Np_, crad_.p,
                               cc_save_.p, cg_exc_.p, cc_cnt_.p, cc_off_.p, cc_cand_.p, csize, cg_cstart_.p,
                               cg_members_.p, cg_bigq_.p, nc, cc_head_.p, cg_nxp_.p, cg_stamp_.p, cg_wl_.p,
                               cg_traj_.p, ctl);
            hipLaunchKernelGGL(k_cg_check, dim3(nb), dim3(256), 0, stream_, cg_exc_.p, cg_marg_.p, crad_.p, cg_lvl_.p,
                               cg_need_.p, s0_, nc, cg_overq_.p, ctl);
            st.launches += 5;
            BD_HIP(hipMemcpyAsync(host_ctl_, ctl, kCtlWords * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
            BD_HIP(hipStreamSynchronize(stream_));
            st.cc_pairs = total / 2;
            st.cc_rounds += host_ctl_[kCtlRounds];
            if (host_ctl_[kCtlBad]) {
                err = "circle pair schedule stalled";
                return BD_ERR_COLLISION;
            }
            if (host_ctl_[kCtlInf]) {
                err = "circle pair sweep produced a non-finite push";
                return BD_ERR_COLLISION;
            }
            const uint32_t n_over = host_ctl_[kCtlOver];
            bool missed = false;
            if (n_over) {
                hipLaunchKernelGGL(k_cg_verify, dim3((n_over + 3) / 4), dim3(256), 0, stream_, cc_save_.p, crad_.p,
                                   cg_exc_.p, cg_grid_.p, cc_start_.p, cg_rec_.p, nc, tmask, cc_cnt_.p, cc_off_.p,
                                   cc_cand_.p, cg_overq_.p, n_over, cg_traj_.p, ctl);
                st.launches++;
                BD_HIP(hipMemcpyAsync(host_ctl_ + kCtlMissed, ctl + kCtlMissed, sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, stream_));
                BD_HIP(hipStreamSynchronize(stream_));
                missed = host_ctl_[kCtlMissed] != 0u;
            }
            if (!missed) {
                // quiet sweeps let the base margin shrink back
                if (++clean_ >= 16) {
                    s0_ = std::max(0.8f * s0_, s0_min());
                    clean_ = 0;
                }
                st.cc_margin = s0_;
                return BD_OK;
            }
            // a skipped pair may have met: undo, widen, redo
            st.cc_redos++;
            hipLaunchKernelGGL(k_cc_restore, dim3(nb), dim3(256), 0, stream_, pos_.p, Np_, cc_save_.p, nc);
            st.launches++;
            clean_ = 0;
            // capped at 0.75 r_max: measured best over the gas box's run (0.25-1.0)
            if (host_ctl_[kCtlQuiet]) s0_ = std::min(1.5f * s0_, std::max(0.75f * rmax_, s0_min()));
        }
        err = "circle pair margin did not converge";
        return BD_ERR_COLLISION;
    }

    float s0_min() const { return 0.05f * std::max(rmax_, 1e-3f); }

    void fill_u32(uint32_t *p, size_t n, uint32_t v, Stats &st) {
        launch_fill_u32(p, n, v, stream_);
        st.launches++;
    }

    static constexpr uint32_t kCompGrid = 256, kLdsGrid = 16384;
    hipStream_t stream_;
    uint32_t Np_ = 0, Nc_ = 0, Ng_ = 0, N_ = 0, P_ = 0, depth_ = 0, ncol_ = 0, nplanes_ = 0;
    bool has_pins_ = false;
    float rmax_ = 0.0f, s0_ = 0.0f;
    uint32_t clean_ = 0;
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
    DevBuf<CGrid> cg_grid_;
    DevBuf<uint32_t> cg_zero_, cc_key_, cc_cnt_, cc_off_, cc_head_, cc_start_, cc_cand_;
    DevBuf<uint32_t> cg_par_, cg_root_, cg_cstart_, cg_members_, cg_bigq_, cg_nxp_, cg_stamp_, cg_overq_, cg_midq_, cg_slot_, cg_keep_;
    DevBuf<float2> cc_save_, cg_traj_;
    DevBuf<float> cg_need_, cg_exc_, cg_marg_, cg_reach_;
    DevBuf<uint8_t> cg_lvl_;
    DevBuf<float4> cg_rec_;
    DevBuf<uint2> cg_wl_;
    ScanTemp scan_tmp_;
    uint32_t *host_ctl_ = nullptr;
};

Engine *make_global_engine(hipStream_t stream) { return new GlobalEngine(stream); }

}  // namespace bd
