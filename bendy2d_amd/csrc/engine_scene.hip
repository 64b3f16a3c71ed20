// engine_scene.hip -- SCENE engine: one workgroup runs a whole scene.
//
// A persistent kernel takes scenes from a work queue (largest first).  For
// each scene the workgroup keeps positions and previous positions in LDS
// when both fit the LDS budget, positions alone when only they fit, or works
// from global memory; it runs all updates and substeps of the launch with
// workgroup barriers between dependent phases (src/solver.rs:106-116):
//
//   links     dependency layers streamed from global memory, a barrier
//             between layers (src/solver.rs:143-153)
//   circles   exact replay of the (i, j) lexicographic pair sweep
//             (src/solver.rs:168-177) from per-circle candidate lists: a
//             round runs every pair whose earlier pairs (sharing a circle)
//             are done; a sweep stands when no pair missing from the lists
//             can have touched at its turn, else it is undone and rerun with
//             new lists (circle_sweep)
//   bodies    EXT4 static colliders (staged in LDS), bounds, Verlet
//             (src/solver.rs:155-165, 118-128)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/bendy2d.h"
#include "common.h"
#include "engine.h"
#include "bodies.h"
#include "kernels.h"
#include "pack.h"
#include "scene_pack.h"

namespace bd {

constexpr int kStageCol = 16;    // colliders staged in LDS per scene
constexpr int kStagePlanes = 64; // polygon collider planes staged in LDS

struct ColliderStage {
    int4 hdr[kStageCol];
    float4 geo[kStageCol];
    float4 pl[kStagePlanes];
};

// A value every lane of the workgroup holds alike, moved to a scalar register
// (the compiler cannot see that values read back from LDS are uniform).
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) { return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v))); }

// Orders one wavefront's memory operations between its lanes: a wavefront's
// LDS and vector memory instructions execute in order, so keeping the
// compiler from moving them across this point is all that is needed.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A uniform pointer into global memory, typed as such, with its base read
// from the first lane: loads through it take a 32-bit offset beside a scalar
// base instead of a 64-bit flat address.
#define BD_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ BD_GLOBAL T *to_global(T *p) {
    const uint64_t a = (uint64_t)p;
    // each half through uni(), which returns it unsigned: the builtin's int,
    // widened directly, would carry bit 31 of the low half into the high one
    const uint64_t lo = uni((uint32_t)a);
    const uint64_t hi = uni((uint32_t)(a >> 32));
    return (BD_GLOBAL T *)((hi << 32) | lo);
}

// The same for read-only data read at uniform indices: typed as constant, it
// is read by scalar loads into scalar registers.
#define BD_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ BD_CONST const T *to_const(const T *p) {
    const uint64_t a = (uint64_t)p;
    const uint64_t lo = uni((uint32_t)a);
    const uint64_t hi = uni((uint32_t)(a >> 32));
    return (BD_CONST const T *)((hi << 32) | lo);
}

// A pointer into LDS, typed as such: the low 32 bits of a generic LDS address
// are its LDS offset.  Accesses through it are LDS instructions, where a
// generic pointer takes flat ones (slower, and counted against both the LDS
// and the vector-memory wait counters).
#define BD_LDS __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ BD_LDS T *to_lds(T *p) {
    return (BD_LDS T *)(uint32_t)(uint64_t)p;
}


template <int WG>
__device__ __forceinline__ float wg_max(float v, float *s_red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[wave] = v;
    __syncthreads();
    float m = s_red[0];
#pragma unroll 2
    for (int w = 1; w < WG / 64; w++) m = fmaxf(m, s_red[w]);
    return uni(m);
}

// inclusive prefix sum over the lanes of the wavefront, all lanes active
// (ROCm's device library: a row-shift and row-broadcast sequence)
extern "C" __device__ uint32_t __ockl_wfscan_add_u32(uint32_t, bool);

// Exclusive scan in place over a[0..n) in global memory by one workgroup.
template <int WG>
__device__ void wg_scan_global(uint32_t *a, uint32_t n, uint32_t *s_wave) {
    uint32_t carry = 0;
    // The wave index is uniform and the lane scan takes no lane masks: the
    // loops over the waves below are scalar loops, and no per-lane compare
    // stays live for the whole kernel.
    const int lane = threadIdx.x & 63;
    const uint32_t wave = uni(threadIdx.x >> 6);
    for (uint32_t base = 0; base < n; base += WG * 4) {
        uint32_t i0 = base + threadIdx.x * 4;
        uint32_t v[4], local = 0;
        for (int k = 0; k < 4; k++) { v[k] = (i0 + k < n) ? a[i0 + k] : 0u; local += v[k]; }
        const uint32_t inc = __ockl_wfscan_add_u32(local, true);
        __syncthreads();
        if (lane == 63) s_wave[wave] = inc;
        __syncthreads();
        uint32_t woff = 0, total = 0;
#pragma unroll 1
        for (uint32_t w = 0; w < wave; w++) woff += s_wave[w];
        total = woff;
#pragma unroll 1
        for (uint32_t w = wave; w < WG / 64; w++) total += s_wave[w];
        uint32_t run = carry + woff + inc - local;
        for (int k = 0; k < 4; k++) {
            if (i0 + k < n) a[i0 + k] = run;
            run += v[k];
        }
        carry += total;
    }
    __syncthreads();
}

// Upper bound on how far a circle can have been from its list-build position
// `r0` during a sweep that started at `p` and took it at most `path` away from
// `p` (its largest excursion, see below).  A circle at a non-finite position
// takes part in no pair (every comparison with it is false), so it cannot
// invalidate the list; one whose list-build position was non-finite but is
// finite now forces a rebuild.
//
// The bound is taken relative to a common shift u of the whole scene: for a
// pair, |(p_i - r_i) - (p_j - r_j)| <= |p_i - r_i - u| + |p_j - r_j - u| for
// any fixed u, so bulk motion (a falling body) does not count.
__device__ __forceinline__ float moved_bound(float2 p, float2 r0, float2 u, float path) {
    if (!(fabsf(p.x) <= 3.0e38f) || !(fabsf(p.y) <= 3.0e38f)) return 0.0f;
    if (!(fabsf(r0.x) <= 3.0e38f) || !(fabsf(r0.y) <= 3.0e38f)) return INFINITY;
    float dx = (p.x - r0.x) - u.x, dy = (p.y - r0.y) - u.y;
    float v = sqrtf(dx * dx + dy * dy) + path;
    return v == v ? v : INFINITY;
}


__device__ __forceinline__ int scell(float x, double inv_h) {
    double c = floor((double)x * inv_h);
    if (!(c > -1073741824.0)) c = -1073741824.0;
    if (c > 1073741824.0) c = 1073741824.0;
    return (int)c;
}

// The circle-pair workspace of a scene, in its chunk (layout: cc_layout).
struct CCWork {
    float2 *ref, *save;
    float *path, *skin, *need;
    uint32_t *cnt, *off;
    uint16_t *cand;
    float2 *traj;  // per list entry: the circle's position after that pair ran
    uint32_t *head, *stamp, *wl, *key;
    float4 *rec;  // circles in bucket order: x, y, radius + skin, index (| kOutlierBit)
    uint32_t *start, *fill, *items;
};

constexpr uint32_t kOutlierBit = 0x80000000u;

constexpr uint32_t kOutCap = 512;   // circles with an outlying skin, per list build
constexpr uint32_t kMidList = 256;  // longer candidate lists are sorted by the whole workgroup
constexpr uint32_t kSlot = 24;      // normal partners kept from the counting pass (more are gathered again)
static_assert(2 * kSlot <= 8 * kCandCap, "count-pass slots must fit in the trajectory array");
constexpr uint32_t kNone = 0xffffffffu;
// sweeps with at least this many candidate pairs park previous positions (or
// positions) in global memory to hold their round state in LDS
constexpr uint32_t kParkPairs = 32;
// Failed attempts at a sweep after which it runs sequentially instead.
constexpr int kFallbackAttempt = 4;
// Sweeps after such a fallback that run sequentially without trying lists.
constexpr uint32_t kFallbackStreak = 16;
// A list is rebuilt before a sweep once some circle drifted this fraction of
// its skin from its list-build position (relative to the scene's mean shift).
constexpr float kDriftRebuild = 1.0f;

// FbGrid: bucket heads (u16, tmask + 1 of them, 0xffff for an empty bucket),
// per-circle next links (u16) and a candidate buffer of cap entries (u16): in
// LDS when the scene leaves room, else in the workspace.  A circle's bucket
// is always that of its current position, so it is not stored.
struct FbGrid {
    uint16_t *head;
    uint16_t *next, *cand;
    uint32_t tmask, cap;
};

struct CCShared {
    CCWork work;
    uint16_t out[kOutCap];   // outlier circles of the current list
    uint32_t hist[33];       // skins by log2(skin / base)
    uint32_t nout;
    uint32_t nmid, nlong;    // candidate lists sorted by a wavefront / by the workgroup
    uint32_t halt;           // the scene's time budget for this launch is spent
    float red9[9 * 16];      // workgroup drift summary (see drift_summary)
    uint32_t cnt3[3];        // worklist lengths, rotating (see the round loop)
    uint32_t poly_last[2];   // polygon pairs: 1 + the last point of the other polygon that
                             // hit the current edge, two slots used alternately
    // LDS the sequential fallback may use for its grid: the scene's previous
    // positions when they are held in LDS (parked in q_global meanwhile), or
    // what the positions leave of the launch's LDS
    struct Fallback {
        uint8_t *lds;
        uint32_t lds_bytes, nq;
        float2 *q_lds, *q_global;
        float2 *p_global;  // positions' global copy when they are in LDS without prev (MODE 1)
    } fb;
    FbGrid grid;             // the fallback's grid (see sequential_sweep)
    uint32_t fb_streak;      // sweeps left that go straight to the fallback
    // statistics of the scene during a launch, written by lane 0 only (in LDS
    // rather than in every lane's registers)
    struct Counts {
        uint32_t rounds, redos, rebuilds, kept, unchecked, n_out, n_far, max_list, fallbacks;
        float smax;
        uint64_t t_build, t_count, t_lists, t_sort, t_links, t_sweep, t_bodies;
    } cnt;
};

// Drift of the circles from their list-build positions, in one pass: the sum
// and count of the finite displacements, their bounding box, the smallest
// skin, and whether some circle has a finite position but a non-finite
// list-build one.  Every thread gets the workgroup's result.
struct Drift {
    float sx, sy, n, xmin, xmax, ymin, ymax, smin, nonfin;
};

template <int WG>
__device__ __forceinline__ Drift drift_summary(const float2 *P, uint32_t Np, uint32_t Nc, const CCWork &w,
                                               float *s_red9, float smin) {
    // skins change only when the list is built, which keeps their minimum
    Drift d{0.0f, 0.0f, 0.0f, INFINITY, -INFINITY, INFINITY, -INFINITY, smin, 0.0f};
#pragma unroll 1
    for (uint32_t c = threadIdx.x; c < Nc; c += WG) {
        const float2 p = P[Np + c], r0 = w.ref[c];
        if (!(fabsf(p.x) <= 3.0e38f) || !(fabsf(p.y) <= 3.0e38f)) continue;  // touches nothing
        const float dx = p.x - r0.x, dy = p.y - r0.y;
        if (fabsf(dx) <= 3.0e38f && fabsf(dy) <= 3.0e38f) {
            d.sx += dx; d.sy += dy; d.n += 1.0f;
            d.xmin = fminf(d.xmin, dx); d.xmax = fmaxf(d.xmax, dx);
            d.ymin = fminf(d.ymin, dy); d.ymax = fmaxf(d.ymax, dy);
        } else {
            d.nonfin = 1.0f;
        }
    }
    float v[9] = {d.sx, d.sy, d.n, d.xmin, d.xmax, d.ymin, d.ymax, d.smin, d.nonfin};
    auto comb = [](int i, float a, float b) {
        return i < 3 ? a + b : (i == 3 || i == 5 || i == 7) ? fminf(a, b) : fmaxf(a, b);
    };
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int i = 0; i < 9; i++) v[i] = comb(i, v[i], __shfl_xor(v[i], o, 64));
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < 9; i++) s_red9[9 * wave + i] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 9; i++) v[i] = s_red9[i];
#pragma unroll 1
    for (int wv = 1; wv < WG / 64; wv++)
#pragma unroll
        for (int i = 0; i < 9; i++) v[i] = comb(i, v[i], s_red9[9 * wv + i]);
    return Drift{uni(v[0]), uni(v[1]), uni(v[2]), uni(v[3]), uni(v[4]), uni(v[5]), uni(v[6]), uni(v[7]), uni(v[8])};
}

// Candidate pairs are the pairs closer than (r_c + skin_c) + (r_o + skin_o)
// when the list is built.  A normal circle (skin <= s_cell) finds its normal
// partners in the 3x3 cells of size h = 2 (r_max + s_cell) around it: the
// circles sit in hash buckets as 16-byte records, and each distinct bucket of
// the 9 is scanned once (a record from another cell that hashes there is
// still a true partner when it passes the distance test, and it is in no
// other scanned bucket).  An outlier (larger skin) finds all its partners
// itself and writes itself into its normal partners' lists too, so normal
// circles never search far.  The first `cap` partners go to out; the count
// is returned in full.
__device__ __noinline__ uint32_t gather_normal(uint32_t c, const CCWork &w, float2 p, float ec, double inv_h,
                                               uint32_t tmask, uint16_t *out, uint32_t cap) {
    const int cx0 = scell(p.x, inv_h), cy0 = scell(p.y, inv_h);
    uint32_t kk[9];
    bool dup[9];
#pragma unroll
    for (int i = 0; i < 9; i++) kk[i] = cell_hash(cx0 + (i % 3) - 1, cy0 + (i / 3) - 1) & tmask;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        dup[i] = false;
#pragma unroll
        for (int j = 0; j < i; j++) dup[i] = dup[i] || kk[j] == kk[i];
    }
    // the buckets' record ranges, loaded together, then their records as one
    // run, kU at a time (a record sits at start + rank in its bucket's range)
    uint32_t s0[9], pre[10];
    pre[0] = 0u;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t a = w.start[kk[i]], b = w.start[kk[i] + 1];
        s0[i] = a;
        pre[i + 1] = pre[i] + (dup[i] ? 0u : b - a);
    }
    const uint32_t total = pre[9];
    constexpr uint32_t kU = 4;
    uint32_t n = 0;
#pragma unroll 1
    for (uint32_t t0 = 0; t0 < total; t0 += kU) {
        float4 r[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint32_t t = t0 + u;
            uint32_t s = 0;
#pragma unroll
            for (int i = 0; i < 9; i++)
                if (t >= pre[i]) s = s0[i] + (t - pre[i]);
            r[u] = t < total ? w.rec[s] : make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(c));
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint32_t o = __float_as_uint(r[u].w);
            if (o == c || (o & kOutlierBit)) continue;  // an outlier lists its pairs itself; padding is c
            float ex = p.x - r[u].x, ey = p.y - r[u].y;
            float lim = ec + r[u].z;
            if (ex * ex + ey * ey < lim * lim) {
                if (n < cap) out[n] = (uint16_t)o;
                n++;
            }
        }
    }
    return n;
}

// Outliers' partners.  A near outlier (its normal partners lie within 8
// cells) has those cells scanned: the (outlier, cell) items of all near
// outliers are spread over the workgroup's threads when the outliers fit one
// stage, else a wavefront takes one outlier at a time.  A far one is searched
// by the whole workgroup over every circle, and pairs of two outliers by all pairs among
// the outliers.  An outlier-normal pair is found once, by the outlier, and
// goes into both lists; an outlier-outlier pair is found once per direction
// (the test is symmetric) and goes into the first one's list.  MODE 0 counts
// into cnt, MODE 1 appends at the lists' cursors.  stage: WG / 2 entries of
// LDS scratch.  Returns the far outliers.
template <int WG, int MODE>
__device__ __noinline__ uint32_t gather_outliers(uint32_t Nc, const CCWork &w, const float *crad, const uint16_t *outs,
                                    uint32_t n_out, float rmax, float s_cell, double h, double inv_h,
                                    uint32_t tmask, uint32_t *cursor, float4 *stage, uint32_t *s_wave) {
    auto add = [&](uint32_t c, uint32_t o) {
        if (MODE == 0) atomicAdd(&w.cnt[c], 1u);
        else w.cand[atomicAdd(&cursor[c], 1u)] = (uint16_t)o;
    };
    auto reach_of = [&](float ec) { return ((double)ec + (double)rmax + (double)s_cell) / h; };
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // the outliers' positions and reaches in LDS (stage), WG / 2 at a time
    constexpr uint32_t kStage = WG / 2;
    auto stage_tile = [&](uint32_t t0, uint32_t tn) {
        __syncthreads();  // the previous tile is no longer read
        for (uint32_t k = threadIdx.x; k < tn; k += WG) {
            const uint32_t o = outs[t0 + k];
            const float2 po = w.ref[o];
            stage[k] = make_float4(po.x, po.y, crad[o] + w.skin[o], 0.0f);
        }
        __syncthreads();
    };
    const bool one_tile = n_out <= kStage;
    // the cells a near outlier scans: (2R + 1)^2 around its own, R = ceil(reach)
    auto near_side = [&](float ec) -> int {
        const double reach = reach_of(ec);
        return reach <= 8.0 ? 2 * (int)ceil(reach) + 1 : 0;
    };
    auto scan_cell = [&](uint32_t c, float2 p, float ec, int cx, int cy) {
        const uint32_t kk = cell_hash(cx, cy) & tmask;
#pragma unroll 4
        for (uint32_t s = w.start[kk], s1 = w.start[kk + 1]; s < s1; s++) {
            const float4 r = w.rec[s];
            const uint32_t o = __float_as_uint(r.w);
            if (o & kOutlierBit) continue;  // outliers: below
            if (scell(r.x, inv_h) != cx || scell(r.y, inv_h) != cy) continue;  // another cell, same bucket
            float ex = p.x - r.x, ey = p.y - r.y;
            float lim = ec + r.z;
            if (ex * ex + ey * ey < lim * lim) { add(c, o); add(o, c); }
        }
    };
    if (one_tile && n_out) {
        stage_tile(0u, n_out);
        // every (near outlier, cell) item across the workgroup: each outlier's
        // first item (an exclusive prefix over the outliers' cell counts) goes
        // in its stage entry's w, and a thread takes a contiguous run of items,
        // finding its first outlier by bisection
        const uint32_t t = threadIdx.x;
        uint32_t cells = 0;
        if (t < n_out) {
            const int side = near_side(stage[t].z);
            cells = (uint32_t)(side * side);
        }
        const uint32_t inc = __ockl_wfscan_add_u32(cells, true);
        if (lane == 63u) s_wave[wave] = inc;
        __syncthreads();
        const uint32_t uw = uni(wave);
        uint32_t below = 0, items = 0;
#pragma unroll 1
        for (uint32_t v = 0; v < uw; v++) below += s_wave[v];
        items = below;
#pragma unroll 1
        for (uint32_t v = uw; v < WG / 64; v++) items += s_wave[v];
        if (t < n_out) stage[t].w = __uint_as_float(below + inc - cells);
        __syncthreads();
        const uint32_t per = (items + WG - 1) / WG;
        uint32_t i = t * per;
        const uint32_t i1 = min(items, i + per);
        if (i < i1) {
            // the last outlier whose first item is at or before i (one with no
            // cells shares its first item with the next, and is passed over)
            uint32_t lo = 0, hi = n_out;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (__float_as_uint(stage[mid].w) <= i) lo = mid; else hi = mid;
            }
            uint32_t k = lo;
            float4 e = stage[k];
            int side = near_side(e.z);
            uint32_t q = i - __float_as_uint(e.w);
#pragma unroll 1
            for (; i < i1; i++) {
                if (q == (uint32_t)(side * side)) {
                    // the next outlier with cells
                    do {
                        e = stage[++k];
                        side = near_side(e.z);
                    } while (side == 0 && k + 1u < n_out);
                    q = 0;
                }
                const float2 p = make_float2(e.x, e.y);
                const int R = side >> 1;
                const int cx = scell(p.x, inv_h) + (int)(q % (uint32_t)side) - R;
                const int cy = scell(p.y, inv_h) + (int)(q / (uint32_t)side) - R;
                scan_cell(outs[k], p, e.z, cx, cy);
                q++;
            }
        }
    }
    // more outliers than one stage holds: a wavefront per outlier, its lanes
    // over the cells
    for (uint32_t k = wave; k < n_out && !one_tile; k += WG / 64) {
        const uint32_t c = outs[k];
        const float2 p = w.ref[c];
        const float ec = crad[c] + w.skin[c];
        const int side = near_side(ec);
        if (side == 0) continue;
        const int R = side >> 1;
        const int cx0 = scell(p.x, inv_h), cy0 = scell(p.y, inv_h);
        for (int q = (int)lane; q < side * side; q += 64)
            scan_cell(c, p, ec, cx0 + (q % side) - R, cy0 + (q / side) - R);
    }
    uint32_t n_far = 0;
    for (uint32_t k = 0; k < n_out; k++) {
        const uint32_t c = outs[k];
        const float ec = crad[c] + w.skin[c];
        if (reach_of(ec) <= 8.0) continue;
        n_far++;
        const float2 p = w.ref[c];
        for (uint32_t s = threadIdx.x; s < Nc; s += WG) {
            const float4 r = w.rec[s];
            const uint32_t o = __float_as_uint(r.w);
            if (o & kOutlierBit) continue;
            float ex = p.x - r.x, ey = p.y - r.y;
            float lim = ec + r.z;
            if (ex * ex + ey * ey < lim * lim) { add(c, o); add(o, c); }
        }
    }
    // outlier pairs: a wavefront takes outlier a, its lanes the staged ones
    for (uint32_t t0 = 0; t0 < n_out; t0 += kStage) {
        const uint32_t tn = min(kStage, n_out - t0);
        if (!one_tile) stage_tile(t0, tn);
        for (uint32_t a = wave; a < n_out; a += WG / 64) {
            const uint32_t c = outs[a];
            float2 p;
            float ec;
            if (one_tile) {
                const float4 e = stage[a];
                p = make_float2(e.x, e.y);
                ec = e.z;
            } else {
                p = w.ref[c];
                ec = crad[c] + w.skin[c];
            }
            for (uint32_t k = lane; k < tn; k += 64) {
                if (t0 + k == a) continue;
                const float4 q = stage[k];
                float ex = p.x - q.x, ey = p.y - q.y;
                float lim = ec + q.z;
                if (ex * ex + ey * ey < lim * lim) add(c, outs[t0 + k]);
            }
        }
    }
    return n_far;
}

// Ascending sort of a list of at most K < 65536 entries by one thread: a
// bitonic network over registers (K a power of two), padding sorting last.
template <int K>
__device__ __forceinline__ void sort_in_registers(uint16_t *my, uint32_t n) {
    uint32_t v[K];
#pragma unroll
    for (int i = 0; i < K; i++) v[i] = i < (int)n ? (uint32_t)my[i] : 0xffffffffu;
#pragma unroll
    for (int k = 2; k <= K; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < K; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t a = v[i], b = v[l];
                    const bool up = (i & k) == 0;
                    v[i] = up ? min(a, b) : max(a, b);
                    v[l] = up ? max(a, b) : min(a, b);
                }
            }
#pragma unroll
    for (int i = 0; i < K; i++)
        if (i < (int)n) my[i] = (uint16_t)v[i];
}

// Ascending sort of a list of at most kMidList distinct entries by one
// wavefront: each lane ranks its entries against the whole list, then writes
// them to their ranks.  Every lane runs the rank loop to the end (the trip
// count is uniform) before any lane stores, so no entry is overwritten while
// it is still being read.
__device__ __forceinline__ void sort_by_wave(uint16_t *my, uint32_t n) {
    constexpr int M = kMidList / 64;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t v[M], r[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        const uint32_t k = lane + 64u * m;
        v[m] = k < n ? (uint32_t)my[k] : 0xffffffffu;
        r[m] = 0u;
    }
#pragma unroll 8
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t x = my[j];
#pragma unroll
        for (int m = 0; m < M; m++) r[m] += x < v[m] ? 1u : 0u;
    }
#pragma unroll
    for (int m = 0; m < M; m++)
        if (lane + 64u * m < n) my[r[m]] = (uint16_t)v[m];
}

// Cell size of the candidate grid: twice the largest radius plus the skin of
// normal circles, so a normal pair lies in neighbouring cells.
__device__ __forceinline__ double cell_size(float rmax, float s_cell) {
    return 2.0 * ((double)rmax + (double)s_cell) * 1.001 + 1e-30;
}

// Position of x in c's sorted candidate list, or where it would go; found
// tells whether the list holds it.
__device__ __forceinline__ uint32_t list_rank(const CCWork &w, uint32_t c, uint32_t x, bool &found) {
    const uint16_t *my = w.cand + w.off[c];
    uint32_t lo = 0, hi = w.cnt[c];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t v = my[mid];
        if (v == x) { found = true; return mid; }
        if (v < x) lo = mid + 1; else hi = mid;
    }
    found = false;
    return lo;
}

// Whether the pair (c, x), missing from the candidate lists, would have
// overlapped when the reference tests it.  Its turn comes after the pairs of
// c with smaller partners and those of x with smaller partners (the
// reference's (i, j) order), so each circle is where the sweep left it after
// that prefix of its own list: exactly, as long as no other missing pair
// touched first, which this same test covers for every missing pair.  The
// overlap test is Circle::solve_circle's (src/circle.rs:32-35) on those
// positions, so a pair it clears was a no-op in the reference too.
__device__ __forceinline__ bool missed_pair_touches(const CCWork &w, const float *crad, uint32_t c, uint32_t x) {
    bool found = false;
    const uint32_t kc = list_rank(w, c, x, found);
    if (found) return false;  // on the list: run in order by the sweep
    const uint32_t kx = list_rank(w, x, c, found);
    const float2 pc = kc ? w.traj[w.off[c] + kc - 1] : w.save[c];
    const float2 px = kx ? w.traj[w.off[x] + kx - 1] : w.save[x];
    const bool c_first = c < x;
    const float2 pi = c_first ? pc : px, pj = c_first ? px : pc;
    const float ri = crad[c_first ? c : x], rj = crad[c_first ? x : c];
    float dx = pi.x - pj.x, dy = pi.y - pj.y;
    float q = dx * dx + dy * dy;
    float rs = ri + rj;
    return q < rs * rs;
}

// A sweep in which some circles moved further than 0.8 of their skins is
// still exact when no pair missing from the list touched at its turn.  Pairs
// of two circles within their skins stayed apart by construction; this looks
// at every pair with a circle beyond its skin (`over`, with -v in vb; the
// others hold v <= vmax, v being a circle's bound on its distance from its
// list-build position, relative to the scene shift, over the whole sweep).
// A pair whose list-build positions are more than r_c + r_x + v_c + v_x apart
// stayed apart; the rest get the exact test above.
//
// With few overruns (at most kOutCap, listed in LDS) pairs of two overrun
// circles are tested all against all and the grid scan skips them; with more
// (listed in the workspace) the grid scan covers them too, vmax then being
// the largest bound of any circle, and a pair of two overruns is tested from
// both ends.
template <int WG, typename Idx>
__device__ int overruns_check(const CCWork &w, const float *crad, const float *vb, const Idx *over,
                              uint32_t n_over, float vmax, float rmax, double h, double inv_h, uint32_t tmask,
                              float *s_red, bool scan_overruns) {
    // rounding margin: a fifth of the overrun circle's skin (the margin the
    // skin test leaves too) plus about 80 ulps of the coordinates
    auto margin = [&](float2 p, float sk) { return 0.2f * sk + 1e-5f * (fabsf(p.x) + fabsf(p.y) + 1.0f); };
    auto apart = [&](float2 p, float2 q, float lim) {
        float dx = p.x - q.x, dy = p.y - q.y;
        return !(sqrtf(dx * dx + dy * dy) <= lim);  // NaN: a non-finite circle touches nothing
    };
    float bad = 0.0f, far = 0.0f;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t k = wave; k < n_over; k += WG / 64) {
        const uint32_t c = over[k];
        const float2 p = w.ref[c];
        const float rc = crad[c], vc = -vb[c], mc = margin(p, w.skin[c]);
        const double reach = ((double)rc + (double)vc + (double)mc + (double)rmax + (double)vmax) / h;
        if (!(reach <= 8.0)) { far = 1.0f; continue; }
        const int R = (int)ceil(reach), side = 2 * R + 1;
        const int cx0 = scell(p.x, inv_h), cy0 = scell(p.y, inv_h);
        for (int q = (int)lane; q < side * side; q += 64) {
            const uint32_t kk = cell_hash(cx0 + (q % side) - R, cy0 + (q / side) - R) & tmask;
            for (uint32_t s = w.start[kk], s1 = w.start[kk + 1]; s < s1; s++) {
                const float4 r = w.rec[s];
                const uint32_t x = __float_as_uint(r.w) & ~kOutlierBit;
                float vx = vb[x];
                if (x == c) continue;
                if (!(vx >= 0.0f)) {  // beyond its skin: below, or here when overruns are scanned
                    if (!scan_overruns) continue;
                    vx = -vx;
                }
                if (apart(p, make_float2(r.x, r.y), ((rc + crad[x]) + (vc + vx)) + mc)) continue;
                if (missed_pair_touches(w, crad, c, x)) bad = 1.0f;
            }
        }
    }
    if (!scan_overruns)
    for (uint32_t q = threadIdx.x; q < n_over * n_over; q += WG) {
        const uint32_t a = q / n_over, b = q - a * n_over;
        if (a >= b) continue;
        const uint32_t c = over[a], o = over[b];
        const float2 p = w.ref[c];
        const float lim = ((crad[c] + crad[o]) + (-vb[c] - vb[o])) + margin(p, w.skin[c] + w.skin[o]);
        if (apart(p, w.ref[o], lim)) continue;
        if (missed_pair_touches(w, crad, c, o)) bad = 1.0f;
    }
    // 0: every missing pair stayed apart; 1: one touched; 2: an overrun too
    // far to search (the sweep is redone without knowing)
    return (int)wg_max<WG>(fmaxf(bad, 2.0f * far), s_red);
}

// Per-scene state of the circle-pair schedule that steers it, kept in
// registers during a launch (uniform across the workgroup); its statistics
// are in CCShared::cnt.
struct CCStats {
    uint32_t pairs, since_build;
    float s_cell, smin;
};

// Sequential fallback of the circle-pair sweep, for a sweep whose candidate
// lists keep failing (circles pushed far within the sweep, lists beyond the
// workspace): the reference's (i, j) loop (src/solver.rs:168-177) row by
// row, which needs no list made in advance.  Run by the first wavefront
// alone; the caller holds the others at a barrier.
//
// The circles sit in a hashed grid of cell size h = 2 r_max + M, each in
// the bucket of the cell its current position is in (a push that moves a
// circle to another bucket moves it in the grid).  Row i gathers every
// j > i closer than r_i + r_j + M to circle i (nine lanes walk the nine
// buckets of the 3x3 cells around it).  The lanes hold the candidates and
// their positions in registers (a candidate does not move before its turn);
// the next pair of the row is the smallest candidate above the last one
// that overlaps circle i where it is now (a wavefront minimum), since the
// reference's test is false, and the pair a no-op, for every j in between.
// One lane runs that pair.  Only circle i moves within its row
// besides each partner at its own turn, so the gather is complete while
// circle i stays within M/2 of where it was gathered from; when it moves
// further the rest of the row is gathered again, with a margin one cell
// wider each time (up to two cells, over 5x5 and then 7x7 cells), so a
// circle pushed far within its row is gathered again only a few times.
// The halved margin leaves room for f32 rounding of the distance tests.
//
constexpr uint32_t kNone16 = 0xffffu;

// minimum over the active lanes of the wavefront (ROCm's device library)
extern "C" __device__ uint32_t __ockl_wfred_min_u32(uint32_t);

template <int WG>
__device__ __noinline__ void sequential_sweep(float2 *P, uint32_t Np, uint32_t Nc, const float *crad, bool uni_r,
                                              float runi, float rmax, const FbGrid *gp, uint32_t *s_fb) {
    const FbGrid g = *gp;
    const uint32_t lane = threadIdx.x & 63u;
    const float rm = fmaxf(rmax, 1e-3f);
    const float M = 0.5f * rm;
    const double hd = 2.0 * (double)rm + (double)M;
    const double inv_h = 1.0 / hd;
    const float hstep = 0.999f * (float)hd;
    uint16_t *const head = g.head;
    uint16_t *const next = g.next, *const cand = g.cand;
    const uint32_t tmask = g.tmask;
    auto radius = [=](uint32_t c) { return uni_r ? runi : crad[c]; };
    auto finite2 = [](float2 p) { return fabsf(p.x) <= 3.0e38f && fabsf(p.y) <= 3.0e38f; };
    auto bucket = [=](float2 p) {
        return finite2(p) ? (cell_hash(scell(p.x, inv_h), scell(p.y, inv_h)) & tmask) : kNone;
    };
    auto wave_sync = []() { __threadfence_block(); __builtin_amdgcn_wave_barrier(); };
    auto link = [=](uint32_t c, uint32_t k) {  // c into bucket k (lane 0 alone)
        next[c] = head[k];
        head[k] = (uint16_t)c;
    };
    auto unlink = [=](uint32_t c, uint32_t k) {  // c out of bucket k (lane 0 alone; lists are acyclic)
        uint32_t x = head[k];
        if (x == c) {
            head[k] = next[c];
        } else {
#pragma unroll 1
            for (uint32_t guard = 0; x != kNone16 && guard < Nc; guard++) {
                const uint32_t y = next[x];
                if (y == c) { next[x] = next[c]; break; }
                x = y;
            }
        }
    };
    // heads are u16, two to a 32-bit word: a circle goes in by compare-and-swap
    // on its bucket's word (each attempt that fails saw another lane succeed)
    uint32_t *const head32 = (uint32_t *)head;
#pragma unroll 1
    for (uint32_t b = lane; b <= tmask / 2u; b += 64) head32[b] = 0xffffffffu;
    wave_sync();
#pragma unroll 1
    for (uint32_t c = lane; c < Nc; c += 64) {
        const uint32_t k = bucket(P[Np + c]);
        if (k != kNone) {
            const uint32_t sh = 16u * (k & 1u);
            uint32_t old = head32[k >> 1], seen;
            do {
                seen = old;
                next[c] = (uint16_t)(seen >> sh);
                old = atomicCAS(&head32[k >> 1], seen, (seen & ~(0xffffu << sh)) | (c << sh));
            } while (old != seen);
        }
    }
    wave_sync();
#pragma unroll 1
    for (uint32_t i = 0; i + 1 < Nc; i++) {
        if (!finite2(P[Np + i])) continue;  // touches nothing (every comparison with it is false)
        const float ri = radius(i);
        uint32_t jlast = i, level = 0, lcap = 2;
        // a pass runs a pair or ends the row, except a wider gather that
        // found too many candidates, which the next pass redoes one cell
        // narrower (the row's widest from then on)
#pragma unroll 1
        for (uint32_t pass = 0; pass < 2u * Nc + 2u; pass++) {
            const float2 base = P[Np + i];
            // margin of this gather: M, and one cell more for each gather of
            // the row that circle i outran; the cells within R of its own
            // hold every circle within reach (R cells apart is more than R h)
            const float Ml = M + (float)level * hstep;
            const int R = (int)level + 1, side = 2 * R + 1;
            if (lane == 0) s_fb[0] = 0u;
            wave_sync();
            // gather: lane q walks the bucket of cell q of the side x side
            // block (a bucket two cells share is walked twice; its circles are
            // then listed twice, which changes no pair's turn)
            const int cx0 = scell(base.x, inv_h), cy0 = scell(base.y, inv_h);
            if (lane < (uint32_t)(side * side)) {
                const uint32_t kk =
                    cell_hash(cx0 + (int)(lane % (uint32_t)side) - R, cy0 + (int)(lane / (uint32_t)side) - R) & tmask;
                uint32_t j = head[kk];
#pragma unroll 1
                for (uint32_t guard = 0; j != kNone16 && guard < Nc; guard++) {
                    const uint32_t nx = next[j];
                    if (j > jlast) {
                        const float2 pj = P[Np + j];
                        const float dx = base.x - pj.x, dy = base.y - pj.y, lim = (ri + radius(j)) + Ml;
                        if (dx * dx + dy * dy < lim * lim) {
                            const uint32_t at = atomicAdd(&s_fb[0], 1u);
                            if (at < g.cap) cand[at] = (uint16_t)j;
                        }
                    }
                    j = nx;
                }
            }
            wave_sync();
            const uint32_t n = s_fb[0];
            const bool fits = n <= g.cap && n <= kMidList;
            if (!fits && level > 0u) {
                lcap = level - 1u;
                level = lcap;
                continue;
            }
            if (fits) {
                constexpr int KM = kMidList / 64;
                uint32_t v[KM];
                float2 pv[KM];
                float rv[KM];
#pragma unroll
                for (int m = 0; m < KM; m++) {
                    const uint32_t k = lane + 64u * m;
                    v[m] = k < n ? (uint32_t)cand[k] : kNone;
                    pv[m] = make_float2(0.0f, 0.0f);
                    rv[m] = 0.0f;
                    if (v[m] != kNone) { pv[m] = P[Np + v[m]]; rv[m] = ri + radius(v[m]); }
                }
                uint32_t again = 0u;
#pragma unroll 1
                for (;;) {
                    const float2 a = P[Np + i];
                    uint32_t best = kNone;
#pragma unroll
                    for (int m = 0; m < KM; m++) {
                        const float dx = a.x - pv[m].x, dy = a.y - pv[m].y;
                        if (v[m] != kNone && v[m] > jlast && dx * dx + dy * dy < rv[m] * rv[m]) best = min(best, v[m]);
                    }
                    const uint32_t j = __ockl_wfred_min_u32(best);
                    if (j == kNone) break;  // the row is done
                    if (lane == 0) {
                        float2 ai = a;
                        const float2 b0 = P[Np + j];
                        float2 b = b0;
                        float qi, qj;
                        uint32_t moved = 0u;
                        if (solve_circle_pair(ai.x, ai.y, b.x, b.y, ri, radius(j), qi, qj)) {
                            P[Np + i] = ai;
                            P[Np + j] = b;
                            const uint32_t ob = bucket(b0), nb = bucket(b);  // b0 is finite: it was gathered
                            if (nb != ob) {
                                unlink(j, ob);
                                if (nb != kNone) link(j, nb);
                            }
                            const float gx = ai.x - base.x, gy = ai.y - base.y;
                            moved = !(gx * gx + gy * gy <= 0.25f * Ml * Ml) ? 1u : 0u;
                        }
                        s_fb[1] = moved;
                    }
                    wave_sync();
                    jlast = j;
                    if (s_fb[1]) { again = 1u; break; }
                }
                if (!again) break;  // the row is done
                level = min(level + 1u, lcap);
                continue;           // gather the rest of the row again
            }
            wave_sync();
            if (lane == 0) {
                uint32_t again = 0u;
                float2 a = P[Np + i];
#pragma unroll 1
                for (uint32_t k = 0; k < n; k++) {
                    uint32_t j;
                    {
                        // too many to rank: the next j above jlast within reach,
                        // by walking the block's buckets again (lane 0 alone)
                        j = kNone;
#pragma unroll 1
                        for (int q = 0; q < side * side; q++) {
                            const uint32_t kq = cell_hash(cx0 + q % side - R, cy0 + q / side - R) & tmask;
                            uint32_t x = head[kq];
#pragma unroll 1
                            for (uint32_t guard = 0; x != kNone16 && guard < Nc; guard++) {
                                const uint32_t nx = next[x];
                                if (x > jlast && x < j) {
                                    const float2 px = P[Np + x];
                                    const float dx = base.x - px.x, dy = base.y - px.y;
                                    const float lim = (ri + radius(x)) + Ml;
                                    if (dx * dx + dy * dy < lim * lim) j = x;
                                }
                                x = nx;
                            }
                        }
                        if (j == kNone) break;
                    }
                    jlast = j;
                    const float2 b0 = P[Np + j];
                    float2 b = b0;
                    float qi, qj;
                    if (solve_circle_pair(a.x, a.y, b.x, b.y, ri, radius(j), qi, qj)) {
                        P[Np + i] = a;
                        P[Np + j] = b;
                        const uint32_t ob = bucket(b0), nb = bucket(b);  // b0 is finite: it was gathered
                        if (nb != ob) {
                            unlink(j, ob);
                            if (nb != kNone) link(j, nb);
                        }
                        const float gx = a.x - base.x, gy = a.y - base.y;
                        if (!(gx * gx + gy * gy <= 0.25f * Ml * Ml)) { again = 1u; break; }
                    }
                }
                s_fb[1] = again;
                s_fb[2] = jlast;
            }
            wave_sync();
            jlast = s_fb[2];
            if (!s_fb[1]) break;
            level = min(level + 1u, lcap);
        }
    }
    wave_sync();
}

// The sequential sweep for the whole workgroup: the first wavefront runs it,
// then every circle's skin need follows from how far the sweep moved it, and
// the lists are marked stale.  Its grid goes into LDS when the scene leaves
// room there (previous positions held in LDS are parked in global memory
// for the sweep), else into the circle-pair workspace.
template <int WG>
__device__ __forceinline__ void fallback_sweep(float2 *P, uint32_t Np, uint32_t Nc, const float *crad, bool uni_r,
                                               float runi, float rmax, uint32_t T, const CCWork &w, CCShared *sh,
                                               float &sbase, bool &list_valid) {
    const int tid = threadIdx.x;
    const CCShared::Fallback fb = sh->fb;
    for (uint32_t c = tid; c < Nc; c += WG) w.save[c] = P[Np + c];
    if (fb.q_lds)
        for (uint32_t j = tid; j < fb.nq; j += WG) fb.q_global[j] = fb.q_lds[j];
    __syncthreads();
    // heads 2 B per bucket, next links 2 B per circle, 256 candidates
    FbGrid &g = sh->grid;
    const uint64_t fixed = 2ull * Nc + 2ull * kMidList;
    uint32_t tb = 0;
    if (fixed + 2ull * 64 <= fb.lds_bytes) {
        tb = 64;
        while (tb < 32768u && tb < 2 * Nc && fixed + 2ull * 2 * tb <= fb.lds_bytes) tb <<= 1;
    }
    if (tb) {
        uint8_t *at = fb.lds;
        g.head = (uint16_t *)at; at += 2ull * tb;
        g.next = (uint16_t *)at; at += 2ull * Nc;
        g.cand = (uint16_t *)at;
        g.tmask = tb - 1;
        g.cap = kMidList;
    } else {
        // workspace: 2T heads (in start's T + 1 words), next links in key,
        // candidates in stamp
        g.head = (uint16_t *)w.start;
        g.next = (uint16_t *)w.key;
        g.cand = (uint16_t *)w.stamp;
        g.tmask = 2 * T - 1;
        g.cap = 2 * Nc;
    }
    __syncthreads();  // the grid's description, written by every thread alike, before the first wavefront reads it
    if (tid < 64) sequential_sweep<WG>(P, Np, Nc, crad, uni_r, runi, rmax, &sh->grid, sh->cnt3);
    __syncthreads();
    if (fb.q_lds)
        for (uint32_t j = tid; j < fb.nq; j += WG) fb.q_lds[j] = fb.q_global[j];
    for (uint32_t c = tid; c < Nc; c += WG) {
        const float2 p = P[Np + c], q = w.save[c];
        const float dx = p.x - q.x, dy = p.y - q.y;
        const float d = sqrtf(dx * dx + dy * dy) * (1.0f + 0x1p-20f);
        w.need[c] = (d <= 3.0e38f) ? fmaxf(1.5f * d, 0.75f * w.need[c]) : 0.0f;
    }
    if (tid == 0) sh->cnt.fallbacks++;
    list_valid = false;
    // a base skin driven below its floor by lists that did not fit starts
    // over from the usual one
    sbase = fmaxf(sbase, 0.05f * fmaxf(rmax, 1e-3f));
    __syncthreads();
}

// The rounds themselves (see sweep_rounds).  Per circle: nxp (u16) its next
// unprocessed partner, nx2 (u16) the partner after that, and cst (u32) the
// entry of nxp in the candidate array (its list offset plus its head, low 24
// bits) with the round in which the circle last moved (high 8 bits, kStamp8
// for none).  nxp, nx2 and cst point into LDS or, typed as global, into the
// workspace.  A thread's first pair of a round takes its new next partners
// from nx2 and loads the entries after them from the candidate array after
// the pair is solved; they are written to nx2 only at the end of the
// readiness pass, so that load overlaps the barrier and the pass instead of
// lengthening the round.
constexpr uint32_t kCurMask = 0x00ffffffu;
constexpr uint32_t kStamp8 = 0xffu;
static_assert(kCandCap * 65536ull <= kCurMask + 1ull, "a list cursor must fit 24 bits");
template <int WG, typename PCP, typename U16P, typename U32P>
__device__ __forceinline__ void rounds_core(PCP Pc, uint32_t Nc, const float *crad_, bool uni_r, float runi,
                                            CCShared *sh, uint32_t *s_wl, U16P nxp, U16P nx2, U32P cst) {
    const int tid = threadIdx.x;
    const CCWork &ww = sh->work;
    struct {
        BD_GLOBAL const uint32_t *off, *cnt;
        BD_GLOBAL const uint16_t *cand;
        BD_GLOBAL float *path;
        BD_GLOBAL uint64_t *save, *traj;  // f32x2 as their bits
        BD_GLOBAL uint32_t *head;
    } w{to_global((const uint32_t *)ww.off), to_global((const uint32_t *)ww.cnt),
        to_global((const uint16_t *)ww.cand), to_global(ww.path), to_global((uint64_t *)ww.save),
        to_global((uint64_t *)ww.traj), to_global(ww.head)};
    auto bits = [](float2 v) { return (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32); };
    auto f32x2 = [](uint64_t b) { return make_float2(__uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32))); };
    BD_GLOBAL const float *const crad = to_global(crad_);
    // worklists: the first WG ready pairs of a round in LDS (s_wl, two
    // buffers of WG), the rest in HBM at the same index
    BD_GLOBAL uint32_t *const gwl[2] = {to_global(ww.wl), to_global(ww.wl) + (Nc / 2 + 1)};
    auto wl_put = [&](uint32_t b, uint32_t i, uint32_t v) {
        if (i < (uint32_t)WG) s_wl[b * WG + i] = v; else gwl[b][i] = v;
    };
    auto wl_get = [&](uint32_t b, uint32_t i) { return i < (uint32_t)WG ? s_wl[b * WG + i] : gwl[b][i]; };
    // the last candidate entry: an entry past a list's end is read before
    // that end is known, and only used below it
    const uint32_t cand_last = kCandCap * Nc - 1u;
#pragma unroll 1
    for (uint32_t c = tid; c < Nc; c += WG) {
        w.save[c] = Pc[c];
        const uint32_t o0 = w.off[c], n0 = w.cnt[c];
        cst[c] = o0 | (kStamp8 << 24);
        nxp[c] = n0 > 0u ? w.cand[o0] : (uint16_t)kNone16;
        nx2[c] = n0 > 1u ? w.cand[o0 + 1u] : (uint16_t)kNone16;
    }
    if (tid == 0) { sh->cnt3[0] = 0u; sh->cnt3[1] = 0u; sh->cnt3[2] = 0u; }
    __syncthreads();
#pragma unroll 1
    for (uint32_t c = tid; c < Nc; c += WG) {
        const uint32_t o = nxp[c];
        if (o != kNone16 && c < o && nxp[o] == c) wl_put(0, atomicAdd(&sh->cnt3[0], 1u), c | (o << 16));
    }
    __syncthreads();
    uint32_t n = uni(sh->cnt3[0]);
    uint32_t r = 0, r8 = 0;
    // one pair: solve it, move both heads to their next partners (from nx2),
    // and between the solve and the stores call far(uc, uo) with the pair's
    // cursors, to issue the loads for the partners after next: vector memory
    // operations complete in order, stores included, so loads issued after
    // the stores would have to wait for them
    auto apply = [&](uint32_t c, uint32_t o, uint32_t &xc, uint32_t &xo, auto &&far) {
        float2 pi = f32x2(Pc[c]), pj = f32x2(Pc[o]);
        const uint32_t uc = cst[c] & kCurMask, uo = cst[o] & kCurMask;
        xc = nx2[c];
        xo = nx2[o];
        float qi = 0.0f, qj = 0.0f;
        const float rc = uni_r ? runi : crad[c], ro = uni_r ? runi : crad[o];
        const bool touch = solve_circle_pair(pi.x, pi.y, pj.x, pj.y, rc, ro, qi, qj);
        __builtin_amdgcn_sched_barrier(0);
        far(uc, uo);
        __builtin_amdgcn_sched_barrier(0);
        if (touch) {
            Pc[c] = bits(pi);
            Pc[o] = bits(pj);
        }
        // each circle's position after each of its pairs, in list order: its
        // excursion from the sweep's start is taken from these at the end
        w.traj[uc] = bits(pi);
        w.traj[uo] = bits(pj);
        cst[c] = (uc + 1u) | (r8 << 24);
        cst[o] = (uo + 1u) | (r8 << 24);
        nxp[c] = (uint16_t)xc;
        nxp[o] = (uint16_t)xo;
    };
    // rounds with few ready pairs: the first wavefront runs them alone, in
    // step, with no workgroup barrier, while the count stays within it (the
    // same steps as a workgroup round below, for lanes < n <= 64); the
    // others wait, then take the round number and count from LDS
    __shared__ uint32_t s_rr[2];
    auto wave_rounds = [&]() {
#pragma unroll 1
        do {
            const uint32_t cb = r & 1u, nb = cb ^ 1u;
            const bool first = (uint32_t)tid < n;
            uint32_t mc = 0, mo = 0, mnc = kNone16, mno = kNone16;
            uint32_t fec, feo, fuc, fuo;
            uint16_t fyc, fyo;
            if (first) {
                const uint32_t pr = s_wl[cb * WG + tid];
                mc = pr & 0xffffu;
                mo = pr >> 16;
                apply(mc, mo, mnc, mno, [&](uint32_t uc, uint32_t uo) {
                    fuc = uc;
                    fuo = uo;
                    fec = w.off[mc + 1];
                    feo = w.off[mo + 1];
                    fyc = w.cand[min(uc + 2u, cand_last)];
                    fyo = w.cand[min(uo + 2u, cand_last)];
                });
            }
            if (tid == 0) sh->cnt3[(r + 2) % 3] = 0u;
            wave_sync();
            if (first) {
#pragma unroll 1
                for (int e = 0; e < 2; e++) {
                    const uint32_t x = e ? mo : mc, y = e ? mno : mnc;
                    if (y == kNone16 || nxp[y] != x) continue;
                    if ((cst[y] >> 24) == r8 && y < x) continue;  // both moved: the smaller one adds it
                    const uint32_t a = x < y ? x : y, b = x < y ? y : x;
                    wl_put(nb, atomicAdd(&sh->cnt3[(r + 1) % 3], 1u), a | (b << 16));
                }
                nx2[mc] = fuc + 2u < fec ? fyc : (uint16_t)kNone16;
                nx2[mo] = fuo + 2u < feo ? fyo : (uint16_t)kNone16;
            }
            wave_sync();
            n = uni(sh->cnt3[(r + 1) % 3]);
            if (tid == 0) sh->cnt.rounds++;
            if (++r8 == kStamp8) {
#pragma unroll 1
                for (uint32_t c = tid; c < Nc; c += 64) cst[c] = cst[c] | (kStamp8 << 24);
                wave_sync();
                r8 = 0u;
            }
            r++;
        } while (n > 0 && n <= 64u);
    };
#pragma unroll 1
    while (n > 0) {
        if (n <= 64u) {
            if (tid < 64) {
                wave_rounds();
                if (tid == 0) { s_rr[0] = r; s_rr[1] = r8; }
            }
            __syncthreads();
            r = uni(s_rr[0]);
            r8 = uni(s_rr[1]);
            n = uni(sh->cnt3[r % 3]);
            __syncthreads();  // s_rr is read before a later pass of the first wavefront writes it
            continue;
        }
        const uint32_t cb = r & 1u, nb = cb ^ 1u;
        // a thread's first pair: its circles, their new next partners, and
        // the loads (list ends and entries) for the partners after them
        uint32_t mc = 0, mo = 0, mnc = kNone16, mno = kNone16;
        // set and read only under `first`: left without a value otherwise, so
        // that no merge of the two paths has to wait for the loads
        uint32_t fec, feo, fuc, fuo;
        uint16_t fyc, fyo;  // as loaded: no widening to wait for
        // further pairs (more ready pairs than threads) complete at once, and
        // before the first pair, whose loads nothing after them may wait for
#pragma unroll 1
        for (uint32_t k = tid + WG; k < n; k += WG) {
            const uint32_t pr = wl_get(cb, k), c = pr & 0xffffu, o = pr >> 16;
            uint32_t xc, xo, uc = 0, uo = 0, ec = 0, eo = 0, yc = 0, yo = 0;
            apply(c, o, xc, xo, [&](uint32_t ucc, uint32_t uoo) {
                uc = ucc;
                uo = uoo;
                ec = w.off[c + 1];
                eo = w.off[o + 1];
                yc = w.cand[min(ucc + 2u, cand_last)];
                yo = w.cand[min(uoo + 2u, cand_last)];
            });
            nx2[c] = uc + 2u < ec ? (uint16_t)yc : (uint16_t)kNone16;
            nx2[o] = uo + 2u < eo ? (uint16_t)yo : (uint16_t)kNone16;
        }
        const bool first = (uint32_t)tid < n;
        if (first) {
            const uint32_t pr = wl_get(cb, tid);
            mc = pr & 0xffffu;
            mo = pr >> 16;
            apply(mc, mo, mnc, mno, [&](uint32_t uc, uint32_t uo) {
                fuc = uc;
                fuo = uo;
                fec = w.off[mc + 1];
                feo = w.off[mo + 1];
                fyc = w.cand[min(uc + 2u, cand_last)];
                fyo = w.cand[min(uo + 2u, cand_last)];
            });
        }
        if (tid == 0) sh->cnt3[(r + 2) % 3] = 0u;  // read at the end of round r-1, two barriers ago
        __syncthreads();
        if (first) {
            // the first pair's partners after next, read across the barrier
            // (no one reads nx2 before the next round); written before the
            // readiness pass, whose loop a wait could not count through, so
            // the wait leaves the pair's stores in flight
            nx2[mc] = fuc + 2u < fec ? fyc : (uint16_t)kNone16;
            nx2[mo] = fuo + 2u < feo ? fyo : (uint16_t)kNone16;
        }
        // pairs completed by a moved head
#pragma unroll 1
        for (uint32_t k = tid; k < n; k += WG) {
            uint32_t c = mc, o = mo, nc = mnc, no = mno;
            if (k != (uint32_t)tid) {
                const uint32_t pr = wl_get(cb, k);
                c = pr & 0xffffu; o = pr >> 16;
                nc = nxp[c]; no = nxp[o];
            }
#pragma unroll 1
            for (int e = 0; e < 2; e++) {
                const uint32_t x = e ? o : c, y = e ? no : nc;
                if (y == kNone16 || nxp[y] != x) continue;
                if ((cst[y] >> 24) == r8 && y < x) continue;  // both moved: the smaller one adds it
                const uint32_t a = x < y ? x : y, b = x < y ? y : x;
                wl_put(nb, atomicAdd(&sh->cnt3[(r + 1) % 3], 1u), a | (b << 16));
            }
        }
        __syncthreads();
        n = uni(sh->cnt3[(r + 1) % 3]);
        if (tid == 0) sh->cnt.rounds++;
        if (++r8 == kStamp8) {
            // every round number is about to recur: no stamp may match one
#pragma unroll 1
            for (uint32_t c = tid; c < Nc; c += WG) cst[c] = cst[c] | (kStamp8 << 24);
            __syncthreads();
            r8 = 0u;
        }
        r++;
    }
    // heads, and each circle's largest excursion from its start position
    // (the running maximum over its positions after its pairs, in their
    // order: a pair that did not touch repeats the position before it)
#pragma unroll 1
    for (uint32_t c = tid; c < Nc; c += WG) {
        const uint32_t o0 = w.off[c], u1 = cst[c] & kCurMask;
        w.head[c] = u1 - o0;
        const float2 sv = f32x2(w.save[c]);
        float pth = 0.0f;
#pragma unroll 4
        for (uint32_t u = o0; u < u1; u++) {
            const float2 q = f32x2(w.traj[u]);
            pth = excursion(pth, q.x, q.y, sv.x, sv.y);
        }
        w.path[c] = pth;
    }
}

// The pair rounds of one circle-pair sweep (see circle_sweep), out of line:
// the sweep's list builds and the kernel's link and body phases keep their
// registers.  The round state is 8 bytes a circle: its next two unprocessed
// partners and its list cursor with the round it last moved in (see
// rounds_core).  With it in LDS a round's list state is an LDS read, and the
// candidate load it needs overlaps the round's barrier.  It sits in the LDS
// the scene leaves free (rs_off bytes
// into the dynamic LDS, kNone for none), or, for a sweep with enough
// candidate pairs to pay for two copies, in place of the circles' previous
// positions (held in LDS, unused by the sweep) or of their positions (the
// rounds then work on the global copy), parked in global memory meanwhile;
// else in the workspace (key, stamp, head).  The rounds are instantiated
// once for each placement, so the LDS one addresses LDS directly.  The heads
// go back to the workspace at the end.
template <int WG>
__device__ __noinline__ void sweep_rounds(float2 *P, uint32_t Np, uint32_t Nc, const float *crad, bool uni_r,
                                          float runi, CCShared *sh, uint32_t *s_wl, uint32_t rs_off,
                                          uint32_t pairs) {
    extern __shared__ float2 s_dyn[];
    const int tid = threadIdx.x;
    const CCWork &w = sh->work;
    const CCShared::Fallback &fb = sh->fb;
    float2 *Pc = P + Np, *pl = nullptr, *pg = nullptr;
    if (rs_off == kNone && pairs >= kParkPairs) {
        if (fb.q_lds) {
            pl = fb.q_lds + Np;
            pg = fb.q_global + Np;
        } else if (fb.p_global) {
            pl = P + Np;
            pg = fb.p_global + Np;
            Pc = pg;
        }
        if (pl) {
#pragma unroll 1
            for (uint32_t c = tid; c < Nc; c += WG) pg[c] = pl[c];
            __syncthreads();
            rs_off = (uint32_t)((uint8_t *)pl - (uint8_t *)s_dyn);
        }
    }
    // the positions as 64-bit words, typed by where they are (LDS, or global
    // memory for a scene without them in LDS or with them parked): a generic
    // LDS address carries the LDS aperture in its high half
    const bool pc_lds = ((uint64_t)Pc >> 32) == ((uint64_t)s_dyn >> 32);
    if (rs_off != kNone) {
        uint16_t *const rs = (uint16_t *)((uint8_t *)s_dyn + rs_off);
        if (pc_lds)
            rounds_core<WG>(to_lds((uint64_t *)Pc), Nc, crad, uni_r, runi, sh, s_wl, rs, rs + Nc,
                            (uint32_t *)(rs + 2 * Nc));
        else
            rounds_core<WG>(to_global((uint64_t *)Pc), Nc, crad, uni_r, runi, sh, s_wl, rs, rs + Nc,
                            (uint32_t *)(rs + 2 * Nc));
    } else {
        BD_GLOBAL uint16_t *const key = to_global((uint16_t *)w.key);
        BD_GLOBAL uint16_t *const stamp = to_global((uint16_t *)w.stamp);
        BD_GLOBAL uint32_t *const head = to_global(w.head);
        if (pc_lds)
            rounds_core<WG>(to_lds((uint64_t *)Pc), Nc, crad, uni_r, runi, sh, s_wl, key, stamp, head);
        else
            rounds_core<WG>(to_global((uint64_t *)Pc), Nc, crad, uni_r, runi, sh, s_wl, key, stamp, head);
    }
    if (pl) {
        // every cursor has been read (the state sits where the positions go back)
        __syncthreads();
#pragma unroll 1
        for (uint32_t c = tid; c < Nc; c += WG) pl[c] = pg[c];
    }
    __syncthreads();
}

// One circle-pair sweep of a scene (src/solver.rs:168-177), exact: the
// candidate pairs run in rounds; pair (c, o), c < o, is ready when o is the
// smallest unprocessed candidate of c and c that of o, so every earlier pair
// (in the reference's (i, j) order) sharing c or o has run.  Ready pairs share
// no circle.  After a round only circles whose head moved can complete a new
// ready pair, so each round looks at those alone (a worklist).  The sweep is
// valid when no circle moved more than 0.8 of its skin since the list was
// built (relative to the scene's mean shift): a pair not on the list started
// further apart than the sum of the two radii and skins, so it cannot have
// overlapped.  When some circles went further, overruns_check checks every
// missing pair they belong to; only if one of them touched is the sweep
// undone and rerun with a new list and larger skins for the circles that
// moved.  The list is rebuilt before a sweep when a circle drifted more than
// kDriftRebuild of its skin since it was built.
// Returns the new error state (0 = fine).
template <int WG>
__device__ uint32_t circle_sweep(float2 *P, uint32_t Np, uint32_t Nc, const float *crad, bool uni_r, float runi,
                                 float rmax, uint32_t T,
                                 const CCWork &w, CCShared *sh, float *s_red, uint32_t *s_u32, uint32_t *s_wl,
                                 float &sbase, bool &list_valid, CCStats &cs, uint32_t rs_off) {
    const int tid = threadIdx.x;
    const uint32_t tmask = T - 1;
    // a scene whose lists just failed goes on without them for a few sweeps
    bool fallback = false;
    for (int attempt = uni(sh->fb_streak) != 0u ? kFallbackAttempt : 0;; attempt++) {
        if (fallback || attempt >= kFallbackAttempt) {
            // the lists keep failing: the sequential sweep, from the sweep's
            // start positions (P holds them: every redo restores them)
            fallback_sweep<WG>(P, Np, Nc, crad, uni_r, runi, rmax, T, w, sh, sbase, list_valid);
            if (tid == 0) sh->fb_streak = sh->fb_streak ? sh->fb_streak - 1u : kFallbackStreak;
            __syncthreads();
            return 0u;
        }
        bool rebuild = !list_valid;
        float2 u = make_float2(0.0f, 0.0f);
        if (!rebuild) {
            // drift since the list was built, relative to the mean shift u, as a
            // fraction of each circle's skin (at most box / smallest skin, with
            // a margin for rounding, when the displacements' bounding box around
            // u is small; else each circle is measured)
            const Drift dr = drift_summary<WG>(P, Np, Nc, w, sh->red9, cs.smin);
            if (dr.n > 0.0f) u = make_float2(dr.sx / dr.n, dr.sy / dr.n);
            const float bx = fmaxf(fabsf(dr.xmin - u.x), fabsf(dr.xmax - u.x));
            const float by = fmaxf(fabsf(dr.ymin - u.y), fabsf(dr.ymax - u.y));
            const float box = dr.n > 0.0f ? sqrtf(bx * bx + by * by) : 0.0f;
            float drift = (box * 1.0001f) / (dr.smin * 0.9999f);
            if (dr.nonfin > 0.0f) {
                drift = INFINITY;  // a circle came back from a non-finite list-build position
            } else if (!(drift <= kDriftRebuild)) {
                float m = 0.0f;
#pragma unroll 4
                for (uint32_t c = tid; c < Nc; c += WG)
                    m = fmaxf(m, moved_bound(P[Np + c], w.ref[c], u, 0.0f) / w.skin[c]);
                drift = wg_max<WG>(m, s_red);
            }
            rebuild = !(drift <= kDriftRebuild);
            // no candidate pair at all and every circle within 0.8 of its skin:
            // no pair can overlap during the sweep, which therefore changes nothing
            if (!rebuild && cs.pairs == 0u && drift <= 0.8f) {
                cs.since_build++;
                return 0u;
            }
        }
        if (rebuild) {
            if (tid == 0) sh->cnt.rebuilds++;
            const uint64_t tb0 = wall_clock64();
            // per-circle skins: the base, or what the circle's recent sweeps pushed it;
            // s_cell: the skin at most kOutCap circles exceed (log2 histogram)
            if (tid <= 32) sh->hist[tid] = 0u;
            __syncthreads();
            float smax = 0.0f, nsmin = -INFINITY;
            for (uint32_t c = tid; c < Nc; c += WG) {
                float sk = fmaxf(sbase, w.need[c]);
                w.skin[c] = sk;
                smax = fmaxf(smax, sk);
                nsmin = fmaxf(nsmin, -sk);
                int b = 0;
                if (sk > sbase) b = sk <= 3.0e38f ? min(32, 1 + (int)floorf(log2f(sk / sbase))) : 32;
                atomicAdd(&sh->hist[b], 1u);
            }
            smax = wg_max<WG>(smax, s_red);
            cs.smin = -wg_max<WG>(nsmin, s_red);
            int B = 0;
            {
                uint32_t above = 0;
                for (int b = 32; b >= 1; b--) {
                    if (above + sh->hist[b] > kOutCap) { B = b; break; }
                    above += sh->hist[b];
                }
            }
            const float s_cell = uni(fminf(smax, sbase * exp2f((float)B)));
            if (tid == 0) sh->nout = 0u;
            __syncthreads();
            for (uint32_t c = tid; c < Nc; c += WG)
                if (w.skin[c] > s_cell) sh->out[atomicAdd(&sh->nout, 1u)] = (uint16_t)c;
            __syncthreads();
            const uint32_t n_out = uni(sh->nout);
            const double h = cell_size(rmax, s_cell);
            const double inv_h = 1.0 / h;
            for (uint32_t b = tid; b <= T; b += WG) w.start[b] = 0u;
            for (uint32_t b = tid; b < T; b += WG) w.fill[b] = 0u;
            __syncthreads();
            for (uint32_t c = tid; c < Nc; c += WG) {
                float2 p = P[Np + c];
                w.ref[c] = p;
                uint32_t kk = cell_hash(scell(p.x, inv_h), scell(p.y, inv_h)) & tmask;
                w.key[c] = kk;
                atomicAdd(&w.start[kk], 1u);
            }
            __syncthreads();
            wg_scan_global<WG>(w.start, T + 1, s_u32);  // bucket k = [start[k], start[k+1])
            for (uint32_t c = tid; c < Nc; c += WG) {
                uint32_t kk = w.key[c];
                float2 p = w.ref[c];
                float sk = w.skin[c];
                w.rec[w.start[kk] + atomicAdd(&w.fill[kk], 1u)] =
                    make_float4(p.x, p.y, crad[c] + sk, __uint_as_float(c | (sk > s_cell ? kOutlierBit : 0u)));
            }
            __syncthreads();
            // counts: normal circles gather their normal partners into a slot of
            // kSlot entries (in traj, free until the sweep) and keep the count in
            // stamp; then outliers add themselves to their partners' counts
            uint16_t *const slot = (uint16_t *)w.traj;
            for (uint32_t c = tid; c < Nc; c += WG) {
                const float sk = w.skin[c];
                uint32_t n = sk > s_cell ? 0u
                                         : gather_normal(c, w, w.ref[c], crad[c] + sk, inv_h, tmask,
                                                         slot + (size_t)c * kSlot, kSlot);
                w.cnt[c] = n;
                w.stamp[c] = n;
            }
            __syncthreads();
            const uint32_t n_far =
                gather_outliers<WG, 0>(Nc, w, crad, sh->out, n_out, rmax, s_cell, h, inv_h, tmask, nullptr,
                                      (float4 *)s_wl, s_u32);
            __syncthreads();
            const uint64_t tb1 = wall_clock64();
            for (uint32_t c = tid; c < Nc; c += WG) w.off[c] = w.cnt[c];
            if (tid == 0) w.off[Nc] = 0u;
            __syncthreads();
            wg_scan_global<WG>(w.off, Nc + 1, s_u32);
            const uint32_t total = uni(w.off[Nc]);
            if (total > kCandCap * Nc) {
                // more candidates than the workspace holds: first the outlying
                // skins come down to the outlier threshold (they list the most
                // partners), and only with no outliers left the base skin
                // shrinks; the skins the other circles showed they need stay
                __syncthreads();
                if (n_out) {
                    for (uint32_t c = tid; c < Nc; c += WG) w.need[c] = fminf(w.need[c], 0.5f * s_cell);
                } else {
                    sbase = 0.7f * sbase;
                    for (uint32_t c = tid; c < Nc; c += WG) w.need[c] = fminf(w.need[c], sbase);
                }
                __syncthreads();
                list_valid = false;
                if (sbase < 1e-4f * fmaxf(rmax, 1e-3f)) fallback = true;
                continue;
            }
            // lists: normal partners first (from the slot, or gathered again when
            // they did not fit it), outliers append at a cursor (in head)
            for (uint32_t c = tid; c < Nc; c += WG) {
                const float sk = w.skin[c];
                uint32_t own = 0;
                if (sk <= s_cell) {
                    own = w.stamp[c];
                    uint16_t *dst = w.cand + w.off[c];
                    if (own <= kSlot) {
                        const uint16_t *src = slot + (size_t)c * kSlot;
                        for (uint32_t i = 0; i < own; i++) dst[i] = src[i];
                    } else {
                        gather_normal(c, w, w.ref[c], crad[c] + sk, inv_h, tmask, dst, own);
                    }
                }
                w.head[c] = w.off[c] + own;
            }
            __syncthreads();
            gather_outliers<WG, 1>(Nc, w, crad, sh->out, n_out, rmax, s_cell, h, inv_h, tmask, w.head, (float4 *)s_wl,
                                   s_u32);
            __syncthreads();
            const uint64_t tb2 = wall_clock64();
            // ascending partner order (the reference's pair order): short lists by
            // their thread, longer ones by a wavefront, the longest by the workgroup
            // (queued at the two ends of wl)
            if (tid == 0) { sh->nmid = 0u; sh->nlong = 0u; }
            __syncthreads();
            const uint32_t wl_end = 2u * (Nc / 2u + 1u) - 1u;
            uint32_t longest = 0;
            for (uint32_t c = tid; c < Nc; c += WG) {
                uint16_t *my = w.cand + w.off[c];
                const uint32_t n = w.cnt[c];
                longest = max(longest, n);
                if (n <= 1u) continue;
                if (n <= 8u) sort_in_registers<8>(my, n);
                else if (WG < 1024 && n <= 16u) sort_in_registers<16>(my, n);
                // 32 registers more than a 1,024-lane workgroup's 128 allow
                // without spilling: there such lists go to a wavefront
                else if (WG < 1024 && n <= 32u) sort_in_registers<(WG < 1024 ? 32 : 16)>(my, n);
                else if (n <= kMidList) w.wl[atomicAdd(&sh->nmid, 1u)] = c;
                else w.wl[wl_end - atomicAdd(&sh->nlong, 1u)] = c;
            }
            __syncthreads();
            const uint32_t nmid = uni(sh->nmid), nlong = uni(sh->nlong);
            for (uint32_t q = tid >> 6; q < nmid; q += WG / 64) {
                const uint32_t c = w.wl[q];
                sort_by_wave(w.cand + w.off[c], w.cnt[c]);
            }
            for (uint32_t q = 0; q < nlong; q++) {
                const uint32_t c = w.wl[wl_end - q], n = w.cnt[c];
                uint16_t *my = w.cand + w.off[c];
                for (uint32_t k = tid; k < n; k += WG) w.items[k] = my[k];
                __syncthreads();
                // partners are distinct, so the ranks are a permutation of [0, n)
                for (uint32_t k = tid; k < n; k += WG) {
                    const uint32_t v = w.items[k];
                    uint32_t rank = 0;
                    for (uint32_t j2 = 0; j2 < n; j2++) rank += w.items[j2] < v ? 1u : 0u;
                    my[rank] = (uint16_t)v;
                }
                __syncthreads();
            }
            __syncthreads();
            // base skin: rebuilds every few sweeps cost more than larger lists
            if (cs.since_build < 4) sbase = fminf(1.25f * sbase, 0.5f * fmaxf(rmax, 1e-3f));
            else if (cs.since_build > 32) sbase = fmaxf(0.8f * sbase, 0.05f * fmaxf(rmax, 1e-3f));
            cs.since_build = 0;
            list_valid = true;
            cs.pairs = total / 2;
            cs.s_cell = s_cell;
            u = make_float2(0.0f, 0.0f);
            const uint32_t max_list = (uint32_t)wg_max<WG>((float)longest, s_red);
            const uint64_t tb3 = wall_clock64();
            if (tid == 0) {
                sh->cnt.n_out = n_out;
                sh->cnt.n_far = n_far;
                sh->cnt.smax = smax;
                sh->cnt.max_list = max_list;
                sh->cnt.t_build += tb3 - tb0;
                sh->cnt.t_count += tb1 - tb0;
                sh->cnt.t_lists += tb2 - tb1;
                sh->cnt.t_sort += tb3 - tb2;
            }
        }
        // ---- sweep: worklist rounds.  nxp: each circle's next unprocessed
        // partner, kNone16 when its list is done, so a pair (x, y) is ready
        // exactly when nxp[x] == y and nxp[y] == x; with it each circle keeps
        // the partner after that, its list cursor, and the round (counted
        // modulo 255, with a clear at each wrap) in which it last moved (see
        // rounds_core).
        sweep_rounds<WG>(P, Np, Nc, crad, uni_r, runi, sh, s_wl, rs_off, cs.pairs);
        // every list must be exhausted (a schedule that stalls is an error);
        // circles beyond 0.8 of their skin go to out[] (free after list builds)
        // and vb (stamp, free after the rounds) gets each circle's bound v,
        // negated for those
        // (the worklist array, free after the rounds, lists every overrun
        // circle when there are more than kOutCap)
        float left = 0.0f, vmax = 0.0f, vall = 0.0f;
        float *vb = (float *)w.stamp;
        if (tid == 0) sh->nout = 0u;
        __syncthreads();
        for (uint32_t c = tid; c < Nc; c += WG) {
            if (w.head[c] != w.cnt[c]) left = 1.0f;
            float pth = w.path[c];
            float v = moved_bound(w.save[c], w.ref[c], u, pth);
            // the sweep's own pushes, decaying slowly so circles in lasting contact keep their skin
            w.need[c] = (pth <= 3.0e38f) ? fmaxf(1.5f * pth, 0.75f * w.need[c]) : 0.0f;
            if (v <= 0.8f * w.skin[c]) {
                vmax = fmaxf(vmax, v);
                vb[c] = v;
            } else {
                vb[c] = -v;
                vall = fmaxf(vall, v);
                uint32_t k = atomicAdd(&sh->nout, 1u);
                if (k < kOutCap) sh->out[k] = (uint16_t)c;
                w.wl[k] = c;
            }
        }
        if (wg_max<WG>(left, s_red) > 0.0f) return 3u;
        const uint32_t n_over = uni(sh->nout);
        vmax = wg_max<WG>(vmax, s_red);
        const double h = cell_size(rmax, cs.s_cell);  // the grid of the list build
        int check = 0;
        if (n_over > kOutCap) {
            vall = fmaxf(wg_max<WG>(vall, s_red), vmax);
            check = overruns_check<WG>(w, crad, vb, w.wl, n_over, vall, rmax, h, 1.0 / h, tmask, s_red, true);
        } else if (n_over) {
            check = overruns_check<WG>(w, crad, vb, sh->out, n_over, vmax, rmax, h, 1.0 / h, tmask, s_red, false);
        }
        if (check == 0) {
            cs.since_build++;
            if (n_over && tid == 0) sh->cnt.kept++;
            return 0u;
        }
        if (check == 2 && tid == 0) sh->cnt.unchecked++;
        // undo and rebuild with the skins the circles needed
        for (uint32_t c = tid; c < Nc; c += WG) P[Np + c] = w.save[c];
        __syncthreads();
        if (tid == 0) sh->cnt.redos++;
        list_valid = false;
    }
}

// Polygon pairs of one scene (see run_scene), kept out of line: a rare phase
// whose registers need not weigh on the rest of the kernel.
template <int WG>
__device__ __noinline__ void poly_pairs(float2 *P, const uint4 *ptab, const float2 *ctr, uint32_t npoly,
                                        uint32_t *poly_last) {
    const int tid = threadIdx.x;
    if (tid == 0) { poly_last[0] = 0u; poly_last[1] = 0u; }
    __syncthreads();
    uint32_t turn = 0;
    for (uint32_t i = 0; i < npoly; i++)
        for (uint32_t j = i + 1; j < npoly; j++)
            for (uint32_t dir = 0; dir < 2; dir++) {
                const uint32_t si = dir ? j : i, oi = dir ? i : j;
                const uint4 ts = ptab[si], to = ptab[oi];
                const float2 sc = ctr[si], oc = ctr[oi];
                for (uint32_t e = 0; e < ts.y; e++, turn++) {
                    const uint32_t ia = ts.x + e, ib = ts.x + (e + 1 == ts.y ? 0u : e + 1);
                    const float2 pa = P[ia], pb = P[ib];
                    uint32_t mine = 0;
                    float2 na = pa, nb = pb;
                    for (uint32_t k = tid; k < to.y; k += WG) {
                        float2 a2, b2, o2;
                        if (resolve_line(sc, pa, pb, P[to.x + k], oc, a2, b2, o2)) {
                            P[to.x + k] = o2;
                            na = a2; nb = b2; mine = k + 1;
                        }
                    }
                    uint32_t *slot = &poly_last[turn & 1u];
                    if (mine) atomicMax(slot, mine);
                    __syncthreads();
                    if (mine && mine == *slot) { P[ia] = na; P[ib] = nb; }
                    if (tid == 0) poly_last[(turn & 1u) ^ 1u] = 0u;
                    __syncthreads();
                }
            }
}

// One scene for `steps` updates.  MODE 2 keeps positions and previous
// positions in LDS (s_lds), MODE 1 positions only, MODE 0 works on the scene
// chunk in global memory.
template <int WG, int MODE, bool SEQ>
__device__ __forceinline__ void run_scene(const SceneDesc *__restrict__ dp, uint32_t steps, SceneState *st,
                                          float2 *s_lds, ColliderStage *s_col, float *s_red, uint32_t *s_u32,
                                          CCShared *s_cc, uint32_t *s_wl, uint32_t lds_bytes) {
    // fields are read where used: the descriptor is uniform and read-only, so
    // they come through scalar loads instead of a private copy
    const SceneDesc &d = *dp;
    char *base = (char *)d.base;
    float2 *gpos = (float2 *)(base + d.pos_off);
    float2 *gprev = (float2 *)(base + d.prev_off);
    const float *crad = (const float *)(base + d.crad_off);
    const uint32_t *pinmask = (const uint32_t *)(base + d.pin_off);
    const uint32_t *loff = (const uint32_t *)(base + d.loff_off);
    // layer bounds: every lane reads the same two before each layer, through
    // the scalar cache (a vector load here put HBM latency before every layer)
    BD_CONST const uint32_t *const kloff = to_const(loff);
    const LinkS *links = (const LinkS *)(base + d.links_off);
    float2 *P, *Q;
    if constexpr (MODE >= 1) P = s_lds; else P = gpos;
    if constexpr (MODE == 2) Q = s_lds + d.N; else Q = gprev;
    const uint32_t N = d.N, Np = d.Np, Nc = d.Nc, NpNc = d.Np + d.Nc, npoly = d.P;
    const uint4 *ptab = (const uint4 *)(base + d.poly_off);
    float2 *ctr = (float2 *)(base + d.ctr_off);
    const uint32_t *gstatic = (const uint32_t *)(base + d.gst_off);
    const int tid = threadIdx.x;
    // circle radius: the descriptor's common one when all circles share it
    const bool uni_r = (d.flags & kFlagUniformR) != 0;
    const float runi = d.runi;
    auto radius = [&](uint32_t c) { return uni_r ? runi : crad[c]; };
    const float r2u = runi * runi, scaleu = 1.0f / (r2u + r2u);

    // stage colliders
    const bool staged = d.ncol <= (uint32_t)kStageCol && d.nplanes <= (uint32_t)kStagePlanes;
    const int4 *ghdr = (const int4 *)(base + d.hdr_off);
    const float4 *ggeo = (const float4 *)(base + d.geo_off);
    const float4 *gpl = (const float4 *)(base + d.pl_off);
    if (staged) {
        for (uint32_t k = tid; k < d.ncol; k += WG) { s_col->hdr[k] = ghdr[k]; s_col->geo[k] = ggeo[k]; }
        for (uint32_t k = tid; k < d.nplanes; k += WG) s_col->pl[k] = gpl[k];
    }
    // load state into LDS; with positions and previous positions there, the
    // link records follow them when they fit too
    if constexpr (MODE >= 1) {
        for (uint32_t j = tid; j < N; j += WG) {
            P[j] = gpos[j];
            if constexpr (MODE == 2) Q[j] = gprev[j];
        }
    }
    const bool links_lds = MODE == 2 && 16ull * N + 8ull * d.L <= lds_bytes;
    LinkS *const s_links = (LinkS *)(s_lds + 2 * N);
    // the circle-pair rounds' per-circle state (8 bytes a circle) in the LDS
    // left after what stays resident, when it fits
    const uint64_t rs_at = MODE == 2 ? 16ull * N + (links_lds ? 8ull * d.L : 0ull) : MODE == 1 ? 8ull * N : 0ull;
    const uint32_t rs_off = rs_at + 8ull * Nc <= lds_bytes ? (uint32_t)rs_at : kNone;
    if (links_lds)
        for (uint32_t l = tid; l < d.L; l += WG) s_links[l] = links[l];
    const bool pins = (d.flags & kFlagPins) != 0;
    if (tid == 0) {
        CCShared::Fallback &f = s_cc->fb;
        if constexpr (MODE == 2) {
            f.lds = (uint8_t *)Q; f.lds_bytes = 8u * N; f.q_lds = Q; f.q_global = gprev; f.nq = N;
            f.p_global = nullptr;
        } else {
            const uint32_t used = MODE == 1 ? 8u * N : 0u;
            f.lds = (uint8_t *)(s_lds) + used; f.lds_bytes = lds_bytes > used ? lds_bytes - used : 0u;
            f.q_lds = nullptr; f.q_global = nullptr; f.nq = 0u;
            f.p_global = MODE == 1 ? gpos : nullptr;
        }
    }
    __syncthreads();

    // circle-pair workspace
    char *ws = d.ws_base ? (char *)d.ws_base : base + d.ws_off;
    const CCLayout cl = cc_layout(Nc, d.T);
    CCWork &w = s_cc->work;
    if (tid == 0) {
    w.ref = (float2 *)(ws + cl.ref); w.save = (float2 *)(ws + cl.save);
    w.path = (float *)(ws + cl.path); w.skin = (float *)(ws + cl.skin); w.need = (float *)(ws + cl.need);
    w.cnt = (uint32_t *)(ws + cl.cnt); w.off = (uint32_t *)(ws + cl.off); w.cand = (uint16_t *)(ws + cl.cand);
    w.traj = (float2 *)(ws + cl.traj);
    w.head = (uint32_t *)(ws + cl.head); w.stamp = (uint32_t *)(ws + cl.stamp); w.wl = (uint32_t *)(ws + cl.wl);
    w.key = (uint32_t *)(ws + cl.key); w.rec = (float4 *)(ws + cl.rec); w.start = (uint32_t *)(ws + cl.start);
    w.fill = (uint32_t *)(ws + cl.fill); w.items = (uint32_t *)(ws + cl.items);
    }
    __syncthreads();
    if (Nc >= 2 && !st->ws_ready) {
        for (uint32_t c = tid; c < Nc; c += WG) w.need[c] = 0.0f;
        __syncthreads();
    }

    float skin = uni(st->skin);
    bool list_valid = uni((uint32_t)st->list_valid) != 0;
    uint32_t err = uni(st->err);
    CCStats cs{};
    cs.pairs = uni(st->pairs);
    cs.since_build = uni(st->since_build);
    cs.s_cell = uni(st->s_cell); cs.smin = uni(st->smin);
    if (tid == 0) {
        s_cc->fb_streak = 0u;
        CCShared::Counts &k = s_cc->cnt;
        k = CCShared::Counts{};
        k.n_out = st->n_out; k.n_far = st->n_far; k.max_list = st->max_list; k.smax = st->smax;
    }
    __syncthreads();
    bool acc = (d.flags & kFlagAcc) != 0;
    const float2 *acc0 = (const float2 *)(base + d.acc0_off);

    const uint64_t t_scene0 = wall_clock64();
    uint32_t ran = 0;
    // a scene stopped by its time budget inside an update resumes at the
    // substep after the last one it ran
    uint32_t sub0 = uni((uint32_t)st->sub_phase), resume = 0u;
    for (uint32_t step = 0; step < steps; step++) {
        for (uint32_t sub = sub0; sub < d.sub_steps; sub++) {
            // ---- links: dependency layers, k sweeps
            const uint64_t tl0 = wall_clock64();
            // polygon centres (Polygon::calc_center, src/polygon.rs:231-237) as
            // Polygon::solve_links computes them before the polygon's links in
            // the last sweep; no other link moves a polygon's points, so the
            // start of that sweep gives the same sums
            auto poly_centres = [&]() {
                for (uint32_t q = tid; q < npoly; q += WG) {
                    const uint4 r = ptab[q];
                    float sx = 0.0f, sy = 0.0f;
                    for (uint32_t i = 0; i < r.y; i++) {
                        const float2 v = P[r.x + i];
                        sx = sx + v.x; sy = sy + v.y;
                    }
                    const float n = (float)r.y;
                    ctr[q] = make_float2(sx / n, sy / n);
                }
                __syncthreads();
            };
            auto link_sweeps = [&](const LinkS *lk) {
            for (uint32_t it = 0; it < d.link_iters; it++) {
                if (npoly && it + 1 == d.link_iters) poly_centres();
                // with the records in global memory, each layer's first record
                // for this thread loads before the barrier that ends the layer
                // before it, so the load overlaps the barrier (a barrier waits
                // for outstanding LDS reads, so records in LDS load after it)
                LinkS nx0{0, 0, 0.0f};
                if (!links_lds && d.depth && kloff[0] + tid < kloff[1]) nx0 = lk[kloff[0] + tid];
                for (uint32_t L = 0; L < d.depth; L++) {
                    uint32_t lo = kloff[L], hi = kloff[L + 1];
                    // not unrolled (measured faster than unrolling by 2, 4 or
                    // 8: register pressure of the whole persistent kernel); the
                    // thread's next record loads while this one is solved
                    uint32_t l = lo + tid;
                    LinkS nxt = nx0;
                    if (links_lds && l < hi) nxt = lk[l];
#pragma unroll 1
                    for (; l < hi; l += WG) {
                        const LinkS rec = nxt;
                        if (l + WG < hi) nxt = lk[l + WG];
                        uint32_t a = rec.a & 0x7fffu, b = rec.b & 0x7fffu;
                        float2 A = P[a], B = P[b];
                        if (a >= Np && a < NpNc) {
                            if (uni_r) solve_circle_link_w(A.x, A.y, B.x, B.y, rec.t, r2u, r2u, scaleu);
                            else solve_circle_link(A.x, A.y, B.x, B.y, rec.t, crad[a - Np], crad[b - Np]);
                        } else {
                            solve_particle_link(A.x, A.y, B.x, B.y, rec.t, (rec.a & kScenePinBit) != 0,
                                                (rec.b & kScenePinBit) != 0);
                        }
                        P[a] = A;
                        P[b] = B;
                    }
                    if (!links_lds && L + 1 < d.depth) {
                        const uint32_t l2 = hi + tid;
                        nx0 = l2 < kloff[L + 2] ? lk[l2] : LinkS{0, 0, 0.0f};
                    }
                    __syncthreads();
                }
            }
            };
            if (links_lds) link_sweeps(s_links);
            else link_sweeps(links);
            if (npoly && d.link_iters == 0) poly_centres();
            if (tid == 0) s_cc->cnt.t_links += wall_clock64() - tl0;
            // ---- circle pairs
            if (Nc >= 2 && !err)
            {
                const uint64_t t0 = wall_clock64();
                if constexpr (SEQ)
                    fallback_sweep<WG>(P, Np, Nc, crad, uni_r, runi, d.rmax, d.T, w, s_cc, skin, list_valid);
                else
                    err = circle_sweep<WG>(P, Np, Nc, crad, uni_r, runi, d.rmax, d.T, w, s_cc, s_red, s_u32, s_wl,
                                           skin, list_valid, cs, rs_off);
                if (tid == 0) s_cc->cnt.t_sweep += wall_clock64() - t0;
            }
            // ---- polygon pairs (src/solver.rs:178-187): Polygon::solve_polygon
            // (src/polygon.rs:142-162) pair by pair, edge by edge.  For one edge
            // the other polygon's points are independent (each reads the edge's
            // ends as they were when its turn came, its own position and the two
            // centres), so they run across the workgroup; the last point that
            // hits (highest index) decides where the edge's ends go.
            if (npoly >= 2) poly_pairs<WG>(P, ptab, ctr, npoly, s_cc->poly_last);
            // ---- bodies: colliders -> bounds -> Verlet
            const uint64_t tq0 = wall_clock64();
            const int4 *hdr = staged ? s_col->hdr : ghdr;
            const float4 *geo = staged ? s_col->geo : ggeo;
            const float4 *pl = staged ? s_col->pl : gpl;
            for (uint32_t j = tid; j < N; j += WG) {
                if (!(pins && j < Np && ((pinmask[j >> 5] >> (j & 31)) & 1u))) {
                    float2 p = P[j];
                    const bool circle = j >= Np && j < NpNc;
                    float r = circle ? radius(j - Np) : 0.0f;
                    static_colliders(p, r, d.ncol, hdr, geo, pl);
                    float2 q = Q[j];
                    float ax = d.gex, ay = d.gey;
                    // a static polygon's points skip Polygon::update (src/polygon.rs:125-128)
                    const bool integrate = j < NpNc || !((gstatic[(j - NpNc) >> 5] >> ((j - NpNc) & 31)) & 1u);
                    if (circle) {
                        circle_bounds_axis(p.x, q.x, d.lox, d.hix, r);
                        circle_bounds_axis(p.y, q.y, d.loy, d.hiy, r);
                    } else {
                        bounds_axis(p.x, q.x, d.lox, d.hix);
                        bounds_axis(p.y, q.y, d.loy, d.hiy);
                    }
                    if (acc && j >= Np) { float2 a0 = acc0[j - Np]; ax = a0.x + d.gx; ay = a0.y + d.gy; }
                    if (integrate) verlet(p.x, p.y, q.x, q.y, ax, ay, d.delta);
                    P[j] = p;
                    Q[j] = q;
                }
            }
            acc = false;
            __syncthreads();
            if (tid == 0) s_cc->cnt.t_bodies += wall_clock64() - tq0;
            if (d.budget && sub + 1 < d.sub_steps) {
                // stop after a substep once the budget is spent: the state
                // between substeps is whole, and the next launch resumes
                // there (one clock read, shared, so the workgroup leaves
                // together)
                if (tid == 0) s_cc->halt = (wall_clock64() - t_scene0 > (uint64_t)d.budget) ? 1u : 0u;
                __syncthreads();
                if (uni(s_cc->halt)) { resume = sub + 1; break; }
            }
        }
        if (resume) break;
        sub0 = 0u;
        ran = step + 1;
        if (d.budget && ran < steps) {
            // stop after a whole update once the budget is spent
            if (tid == 0) s_cc->halt = (wall_clock64() - t_scene0 > (uint64_t)d.budget) ? 1u : 0u;
            __syncthreads();
            if (uni(s_cc->halt)) break;
        }
    }
    // write back
    if constexpr (MODE >= 1) {
        for (uint32_t j = tid; j < N; j += WG) {
            gpos[j] = P[j];
            if constexpr (MODE == 2) gprev[j] = Q[j];
        }
    }
    // statistics (uniform across the workgroup)
    __syncthreads();
    if (tid == 0) {
        st->skin = skin;
        st->list_valid = list_valid ? 1u : 0u;
        st->err = err;
        const CCShared::Counts &k = s_cc->cnt;
        st->rounds += k.rounds;
        st->redos += k.redos;
        st->rebuilds += k.rebuilds;
        st->pairs = cs.pairs;
        if (Nc >= 2) st->ws_ready = 1u;
        st->t_links += k.t_links;
        st->t_build += k.t_build;
        st->t_sweep += k.t_sweep - k.t_build;
        st->t_bodies += k.t_bodies;
        st->since_build = cs.since_build;
        st->n_out = k.n_out;
        st->n_far = k.n_far;
        st->max_list = k.max_list;
        st->s_cell = cs.s_cell;
        st->smax = k.smax;
        st->smin = cs.smin;
        st->t_count += k.t_count;
        st->t_lists += k.t_lists;
        st->t_sort += k.t_sort;
        st->kept += k.kept;
        st->unchecked += k.unchecked;
        st->fallbacks += k.fallbacks;
        st->last_steps = ran;
        st->sub_phase = (uint16_t)resume;
    }
    __syncthreads();
}

template <int WG>
__global__ void __launch_bounds__(WG) k_scenes(const SceneDesc *__restrict__ descs, SceneState *states,
                                               const uint32_t *__restrict__ order, uint32_t n,
                                               uint32_t *queue, uint32_t lds_bytes, uint32_t launch_steps) {
    extern __shared__ float2 s_lds[];
    __shared__ ColliderStage s_col;
    __shared__ float s_red[WG / 64];
    __shared__ uint32_t s_u32[WG / 64];
    __shared__ CCShared s_cc;
    __shared__ __align__(16) uint32_t s_wl[2 * WG];  // also float4 scratch of the list builds
    __shared__ uint32_t s_next;
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t k = uni(s_next);
        __syncthreads();
        if (k >= n) break;
        const uint32_t si = uni(order[k]);
        const SceneDesc *__restrict__ d = descs + si;
        const uint32_t steps = d->steps < launch_steps ? d->steps : launch_steps;
        if (steps == 0 || d->N == 0) continue;
        if (d->flags & kFlagSequential)  // diagnostic: every sweep sequential, state in global memory
            run_scene<WG, 0, true>(d, steps, states + d->state_slot, s_lds, &s_col, s_red, s_u32, &s_cc, s_wl,
                                   lds_bytes);
        else if (16ull * d->N <= lds_bytes)
            run_scene<WG, 2, false>(d, steps, states + d->state_slot, s_lds, &s_col, s_red, s_u32, &s_cc, s_wl,
                                    lds_bytes);
        else if (8ull * d->N <= lds_bytes)
            run_scene<WG, 1, false>(d, steps, states + d->state_slot, s_lds, &s_col, s_red, s_u32, &s_cc, s_wl,
                                    lds_bytes);
        else
            run_scene<WG, 0, false>(d, steps, states + d->state_slot, s_lds, &s_col, s_red, s_u32, &s_cc, s_wl,
                                    lds_bytes);
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

template <int WG>
static void launch_variant(dim3 grid, size_t lds, hipStream_t s, const SceneDesc *d, SceneState *st,
                           const uint32_t *order, uint32_t n, uint32_t *queue, uint32_t steps) {
    hipLaunchKernelGGL((k_scenes<WG>), grid, dim3(WG), lds, s, d, st, order, n, queue, (uint32_t)lds, steps);
}

using LaunchFn = void (*)(dim3, size_t, hipStream_t, const SceneDesc *, SceneState *, const uint32_t *, uint32_t,
                          uint32_t *, uint32_t);

static const struct {
    int wg;
    LaunchFn fn;
    const void *kern;
} kVariants[] = {
    {64, launch_variant<64>, (const void *)k_scenes<64>},
    {128, launch_variant<128>, (const void *)k_scenes<128>},
    {256, launch_variant<256>, (const void *)k_scenes<256>},
    {512, launch_variant<512>, (const void *)k_scenes<512>},
    {1024, launch_variant<1024>, (const void *)k_scenes<1024>},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
constexpr size_t kMaxDynLds = 144 * 1024;

int scene_variant_count() { return kNumVariants; }
int scene_variant_wg(int v) { return kVariants[v].wg; }
size_t scene_max_lds() { return kMaxDynLds; }

int launch_scene_variant(int v, uint32_t blocks, size_t lds_bytes, hipStream_t stream, const SceneDesc *d,
                         SceneState *st, const uint32_t *order, uint32_t n, uint32_t *queue, uint32_t steps,
                         std::string &err) {
    if (lds_bytes > kMaxDynLds) lds_bytes = kMaxDynLds;
    kVariants[v].fn(dim3(blocks), lds_bytes, stream, d, st, order, n, queue, steps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("scene kernel launch: ") + hipGetErrorString(e);
        return BD_ERR_DEVICE;
    }
    return BD_OK;
}

// Allow every variant the full dynamic LDS budget (once per process, before
// any launch or graph capture).
int scene_kernels_init(std::string &err) {
    static std::once_flag once;
    static hipError_t result = hipSuccess;
    std::call_once(once, []() {
        for (int v = 0; v < kNumVariants; v++) {
            hipError_t e = hipFuncSetAttribute(kVariants[v].kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxDynLds);
            if (e != hipSuccess) result = e;
        }
    });
    if (result != hipSuccess) {
        err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(result);
        return BD_ERR_DEVICE;
    }
    return BD_OK;
}

int scene_blocks_per_cu(int v, size_t lds_bytes) {
    int nb = 0;
    if (lds_bytes > kMaxDynLds) lds_bytes = kMaxDynLds;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kVariants[v].kern, kVariants[v].wg, lds_bytes) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

int variant_for_wg(int wg) {
    for (int v = 0; v < kNumVariants; v++)
        if (kVariants[v].wg == wg) return v;
    return -1;
}

// ---- scene images ----------------------------------------------------------

int build_scene_image(const HostScene &s, const Plan &p, SceneImage &img, std::string &err) {
    if (p.N > kSceneMaxPoints) { err = "scene too large for the SCENE engine"; return BD_ERR_CAPACITY; }
    SceneDesc &d = img.desc;
    std::memset(&d, 0, sizeof(d));
    d.N = p.N; d.Np = p.Np; d.Nc = p.Nc; d.Ng = p.Ng; d.depth = p.depth;
    d.P = (uint32_t)s.polys.size();
    d.L = (uint32_t)p.links.size();
    d.ncol = (uint32_t)s.cols.size();
    d.nplanes = (uint32_t)(s.planes.size() / 3);
    d.T = 0;
    if (d.Nc >= 2) { d.T = 1; while (d.T < 2 * d.Nc) d.T <<= 1; }
    uint64_t o = 0;
    d.pos_off = (uint32_t)o; o = align16(o + 8ull * d.N);
    d.prev_off = (uint32_t)o; o = align16(o + 8ull * d.N);
    d.crad_off = (uint32_t)o; o = align16(o + 4ull * d.Nc);
    d.acc0_off = (uint32_t)o; o = align16(o + 8ull * (d.Nc + d.Ng));
    d.pin_off = (uint32_t)o; o = align16(o + 4ull * ((d.Np + 31) / 32));
    d.loff_off = (uint32_t)o; o = align16(o + 4ull * (d.depth + 1));
    d.links_off = (uint32_t)o; o = align16(o + 8ull * d.L);
    d.hdr_off = (uint32_t)o; o = align16(o + 16ull * d.ncol);
    d.geo_off = (uint32_t)o; o = align16(o + 16ull * d.ncol);
    d.pl_off = (uint32_t)o; o = align16(o + 16ull * d.nplanes);
    d.poly_off = (uint32_t)o; o = align16(o + 16ull * d.P);
    d.ctr_off = (uint32_t)o; o = align16(o + 8ull * d.P);
    d.gst_off = (uint32_t)o; o = align16(o + 4ull * ((d.Ng + 31) / 32));
    img.init_bytes = o;  // everything above is initialised from the host
    d.ws_off = (uint32_t)o;
    d.ws_bytes = d.Nc >= 2 ? cc_layout(d.Nc, d.T).bytes : 0u;
    o = align16(o + d.ws_bytes);
    img.total_bytes = o;
    img.bytes.assign(img.init_bytes, 0);
    uint8_t *b = img.bytes.data();
    auto put = [&](uint32_t off, const void *src, size_t n) { if (n) std::memcpy(b + off, src, n); };
    std::vector<V2> pos(d.N), prev(d.N);
    for (uint32_t i = 0; i < d.Np; i++) { pos[i] = s.p_pos[i]; prev[i] = s.p_prev[i]; }
    for (uint32_t i = 0; i < d.Nc; i++) { pos[d.Np + i] = s.c_pos[i]; prev[d.Np + i] = s.c_prev[i]; }
    for (uint32_t i = 0; i < d.Ng; i++) { pos[d.Np + d.Nc + i] = s.g_pos[i]; prev[d.Np + d.Nc + i] = s.g_prev[i]; }
    put(d.pos_off, pos.data(), 8ull * d.N);
    put(d.prev_off, prev.data(), 8ull * d.N);
    put(d.crad_off, s.c_r.data(), 4ull * d.Nc);
    put(d.acc0_off, s.c_acc.data(), 8ull * d.Nc);
    put(d.acc0_off + 8ull * d.Nc, s.g_acc.data(), 8ull * d.Ng);
    {
        std::vector<uint4> tab(d.P);
        std::vector<float2> ctr(d.P);
        std::vector<uint32_t> mask((d.Ng + 31) / 32, 0u);
        for (uint32_t i = 0; i < d.P; i++) {
            const Poly &pg = s.polys[i];
            tab[i] = make_uint4(d.Np + d.Nc + pg.off, pg.n, pg.is_static ? 1u : 0u, 0u);
            ctr[i] = make_float2(pg.center.x, pg.center.y);
            if (pg.is_static)
                for (uint32_t k = pg.off; k < pg.off + pg.n; k++) mask[k >> 5] |= 1u << (k & 31);
        }
        put(d.poly_off, tab.data(), 16ull * d.P);
        put(d.ctr_off, ctr.data(), 8ull * d.P);
        put(d.gst_off, mask.data(), 4ull * mask.size());
    }
    bool pins = false;
    {
        std::vector<uint32_t> mask((d.Np + 31) / 32, 0u);
        for (uint32_t i = 0; i < d.Np; i++)
            if (s.p_pin[i]) { mask[i >> 5] |= 1u << (i & 31); pins = true; }
        put(d.pin_off, mask.data(), 4ull * mask.size());
    }
    put(d.loff_off, p.layer_off.data(), 4ull * (d.depth + 1));
    {
        std::vector<LinkS> ls(d.L);
        for (uint32_t i = 0; i < d.L; i++) {
            const Link12 &l = p.links[i];
            uint32_t a = l.a & kIndexMask, bb = l.b & kIndexMask;
            ls[i].a = (uint16_t)(a | ((l.a & kPinBit) ? kScenePinBit : 0u));
            ls[i].b = (uint16_t)(bb | ((l.b & kPinBit) ? kScenePinBit : 0u));
            ls[i].t = l.t;
        }
        put(d.links_off, ls.data(), 8ull * d.L);
    }
    {
        std::vector<int4> hdr(d.ncol);
        std::vector<float4> geo(d.ncol), pl(d.nplanes);
        for (uint32_t k = 0; k < d.ncol; k++) {
            const Collider &c = s.cols[k];
            hdr[k] = make_int4(c.kind, (int)c.plane_off, (int)c.nv, 0);
            geo[k] = make_float4(c.cx, c.cy, c.r, 0.0f);
        }
        for (uint32_t k = 0; k < d.nplanes; k++)
            pl[k] = make_float4(s.planes[3 * k], s.planes[3 * k + 1], s.planes[3 * k + 2], 0.0f);
        put(d.hdr_off, hdr.data(), 16ull * d.ncol);
        put(d.geo_off, geo.data(), 16ull * d.ncol);
        put(d.pl_off, pl.data(), 16ull * d.nplanes);
    }
    d.flags = pins ? kFlagPins : 0u;
    // diagnostic switch: every circle-pair sweep by the sequential fallback
    // (its results are the same; tests use it to run that path on any scene)
    if (const char *e = std::getenv("BD_SCENE_SEQUENTIAL_SWEEP"))
        if (e[0] == '1') d.flags |= kFlagSequential;
    img.acc_pending = s.any_initial_acc();
    d.rmax = 0.0f;
    for (float r : s.c_r) d.rmax = std::max(d.rmax, r);
    // one radius for every circle (same bits): the kernel reads it from the
    // descriptor instead of crad[]
    d.runi = 0.0f;
    if (!s.c_r.empty() &&
        std::all_of(s.c_r.begin(), s.c_r.end(), [&](float r) { return std::memcmp(&r, &s.c_r[0], 4) == 0; })) {
        d.flags |= kFlagUniformR;
        d.runi = s.c_r[0];
    }
    img.skin0 = std::max(0.25f * d.rmax, 0.025f);  // per circle: a pair's margin is the sum of both
    return BD_OK;
}

void scene_chunk_to_host(const SceneDesc &d, const uint8_t *chunk, HostScene &h) {
    const float2 *pos = (const float2 *)(chunk + d.pos_off);
    const float2 *prev = (const float2 *)(chunk + d.prev_off);
    const float2 *ctr = (const float2 *)(chunk + d.ctr_off);
    const uint32_t NpNc = d.Np + d.Nc;
    h.p_pos.resize(d.Np); h.p_prev.resize(d.Np);
    h.c_pos.resize(d.Nc); h.c_prev.resize(d.Nc);
    h.g_pos.resize(d.Ng); h.g_prev.resize(d.Ng);
    for (uint32_t i = 0; i < d.N; i++) {
        V2 p{pos[i].x, pos[i].y}, q{prev[i].x, prev[i].y};
        if (i < d.Np) { h.p_pos[i] = p; h.p_prev[i] = q; }
        else if (i < NpNc) { h.c_pos[i - d.Np] = p; h.c_prev[i - d.Np] = q; }
        else { h.g_pos[i - NpNc] = p; h.g_prev[i - NpNc] = q; }
    }
    // the acc the bodies held when the chunk was built (the kernel reads it
    // on the first substep and never writes it)
    const float2 *acc0 = (const float2 *)(chunk + d.acc0_off);
    h.c_acc.resize(d.Nc);
    h.g_acc.resize(d.Ng);
    for (uint32_t i = 0; i < d.Nc; i++) h.c_acc[i] = V2{acc0[i].x, acc0[i].y};
    for (uint32_t i = 0; i < d.Ng; i++) h.g_acc[i] = V2{acc0[d.Nc + i].x, acc0[d.Nc + i].y};
    h.polys.assign(d.P, Poly{});
    for (uint32_t i = 0; i < d.P; i++) h.polys[i].center = V2{ctr[i].x, ctr[i].y};
    h.gravity = V2{d.gx, d.gy};
    h.sub_steps = d.sub_steps;
}

void set_scene_params(SceneDesc &d, const StepParams &sp, uint32_t steps, bool acc_pending) {
    d.steps = steps;
    d.sub_steps = sp.sub_steps;
    d.link_iters = sp.link_iters;
    d.delta = sp.delta;
    d.gex = sp.gex; d.gey = sp.gey; d.gx = sp.gx; d.gy = sp.gy;
    d.lox = sp.lox; d.loy = sp.loy; d.hix = sp.hix; d.hiy = sp.hiy;
    d.flags = (d.flags & ~kFlagAcc) | (acc_pending ? kFlagAcc : 0u);
}

// ---- single-scene engine ---------------------------------------------------

class SceneEngine : public Engine {
public:
    explicit SceneEngine(hipStream_t s) : stream_(s) {}
    ~SceneEngine() override {
        if (chunk_) (void)hipFree(chunk_);
        if (dev_) (void)hipFree(dev_);
    }
    int kind() const override { return BD_ENGINE_SCENE; }

    int upload(const HostScene &s, const Plan &p, std::string &err) override {
        SceneImage img;
        int rc = build_scene_image(s, p, img, err);
        if (rc) return rc;
        rc = scene_kernels_init(err);
        if (rc) return rc;
        if (chunk_) { (void)hipFree(chunk_); chunk_ = nullptr; }
        BD_HIP(hipMalloc((void **)&chunk_, std::max<uint64_t>(img.total_bytes, 16)));
        BD_HIP(hipMemcpyAsync(chunk_, img.bytes.data(), img.init_bytes, hipMemcpyHostToDevice, stream_));
        desc_ = img.desc;
        desc_.base = (uint64_t)(uintptr_t)chunk_;
        acc_pending_ = img.acc_pending;
        if (!dev_) BD_HIP(hipMalloc((void **)&dev_, 512));
        SceneState st{};
        st.skin = img.skin0;
        BD_HIP(hipMemcpyAsync(state_ptr(), &st, sizeof(st), hipMemcpyHostToDevice, stream_));
        // one scene on an otherwise idle device: about two points per lane
        int want = 64;
        while (want < 1024 && (uint32_t)want * 2 < desc_.N) want <<= 1;
        variant_ = variant_for_wg(want);
        // positions, previous positions and, when they fit, the link records
        const size_t with_links = 16ull * desc_.N + 8ull * desc_.L;
        lds_bytes_ = with_links <= scene_max_lds() ? with_links : std::min<size_t>(16ull * desc_.N, scene_max_lds());
        BD_HIP(hipStreamSynchronize(stream_));
        N_ = p.N; Np_ = p.Np; Nc_ = p.Nc; P_ = desc_.P;
        return BD_OK;
    }

    int update(const StepParams &sp, Stats &st, std::string &err) override {
        set_scene_params(desc_, sp, 1, acc_pending_ && sp.first_acc);
        acc_pending_ = false;
        // descriptor, order and queue live in one small buffer (one copy per update)
        std::memcpy(host_blob_, &desc_, sizeof(desc_));
        uint32_t zero_order[2] = {0u, 0u};
        std::memcpy(host_blob_ + 256, zero_order, 8);
        BD_HIP(hipMemcpyAsync(dev_, host_blob_, 264, hipMemcpyHostToDevice, stream_));
        int rc = launch_scene_variant(variant_, 1, lds_bytes_, stream_, (const SceneDesc *)dev_, state_ptr(),
                                      (const uint32_t *)(dev_ + 256), 1, (uint32_t *)(dev_ + 260), 1u, err);
        if (rc) return rc;
        st.launches++;
        st.substeps += sp.sub_steps;
        return BD_OK;
    }

    // pos[N], prev[N], the scene's SceneState, then the polygon centres
    static constexpr size_t kStateWords = (sizeof(SceneState) + sizeof(float2) - 1) / sizeof(float2);
    size_t state_count() const override { return 2ull * N_ + kStateWords + P_; }
    hipStream_t stream() const override { return stream_; }

    int copy_state(void *dst, hipMemcpyKind kind, hipStream_t st, std::string &err) override {
        float2 *d = (float2 *)dst;
        if (N_) {
            BD_HIP(hipMemcpyAsync(d, chunk_ + desc_.pos_off, 8ull * N_, kind, st));
            BD_HIP(hipMemcpyAsync(d + N_, chunk_ + desc_.prev_off, 8ull * N_, kind, st));
        }
        BD_HIP(hipMemcpyAsync(d + 2 * N_, state_ptr(), sizeof(SceneState), kind, st));
        if (P_) BD_HIP(hipMemcpyAsync(d + 2 * N_ + kStateWords, chunk_ + desc_.ctr_off, 8ull * P_, kind, st));
        return BD_OK;
    }

    int unpack(const float2 *src, HostScene &s, std::string &err) override {
        if (!host_shape_matches(s, Np_, Nc_, N_ - Np_ - Nc_, err) || s.polys.size() != P_) return BD_ERR_STATE;
        const float2 *pos = src, *prev = src + N_, *ctr = src + 2 * N_ + kStateWords;
        SceneState stt;
        std::memcpy(&stt, src + 2 * N_, sizeof(stt));
        const uint32_t NpNc = Np_ + Nc_;
        for (uint32_t i = 0; i < N_; i++) {
            V2 p{pos[i].x, pos[i].y}, q{prev[i].x, prev[i].y};
            if (i < Np_) { s.p_pos[i] = p; s.p_prev[i] = q; }
            else if (i < NpNc) { s.c_pos[i - Np_] = p; s.c_prev[i - Np_] = q; }
            else { s.g_pos[i - NpNc] = p; s.g_prev[i - NpNc] = q; }
        }
        // a static polygon keeps the centre of its last Polygon::solve_links
        // (the others' centres follow from prev on the host)
        for (uint32_t i = 0; i < P_; i++)
            if (s.polys[i].is_static) s.polys[i].center = V2{ctr[i].x, ctr[i].y};
        if (stt.err) {
            err = "circle pair schedule did not converge (kind " + std::to_string(stt.err) + ", skin " +
                  std::to_string(stt.skin) + ", rounds " + std::to_string(stt.rounds) + ", redos " +
                  std::to_string(stt.redos) + ", rebuilds " + std::to_string(stt.rebuilds) + ")";
            return BD_ERR_COLLISION;
        }
        return BD_OK;
    }

    int sync(std::string &err) override {
        BD_HIP(hipStreamSynchronize(stream_));
        return BD_OK;
    }

private:
    SceneState *state_ptr() { return (SceneState *)(dev_ + 384); }
    hipStream_t stream_;
    uint8_t *chunk_ = nullptr;
    uint8_t *dev_ = nullptr;  // [0,256) desc, [256,260) order, [260,264) queue, [384,512) state
    uint8_t host_blob_[264];
    SceneDesc desc_{};
    bool acc_pending_ = false;
    int variant_ = 0;
    size_t lds_bytes_ = 0;
    uint32_t N_ = 0, Np_ = 0, Nc_ = 0, P_ = 0;
};

bool scene_engine_fits(const HostScene &s, const Plan &p) {
    (void)s;
    return p.N <= 16384 && p.N > 0;
}

Engine *make_scene_engine(hipStream_t s) { return new SceneEngine(s); }

}  // namespace bd
