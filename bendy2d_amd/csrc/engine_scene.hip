// engine_scene.hip -- SCENE engine: one workgroup runs a whole scene.
//
// A persistent kernel takes scenes from a work queue (largest first).  For
// each scene the workgroup keeps positions in LDS (or in global memory for
// scenes too large for the LDS budget), keeps each lane's previous positions
// in registers, and runs all updates and substeps of the launch with
// workgroup barriers between dependent phases (src/solver.rs:106-116):
//
//   links     dependency layers streamed from global memory, a barrier
//             between layers (src/solver.rs:143-153)
//   circles   exact replay of the (i, j) lexicographic pair sweep
//             (src/solver.rs:168-177) from a Verlet candidate list: pairs
//             within r_i + r_j + skin of each other when the list was built;
//             a round runs every pair whose earlier pairs (sharing a circle)
//             are done; a sweep whose circles moved 0.45 * skin or more since
//             the list was built is undone and rerun with a new list
//   bodies    EXT4 static colliders (staged in LDS), bounds, Verlet
//             (src/solver.rs:155-165, 118-128)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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

template <int WG>
__device__ __forceinline__ float wg_max(float v, float *s_red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[wave] = v;
    __syncthreads();
    float m = s_red[0];
    for (int w = 1; w < WG / 64; w++) m = fmaxf(m, s_red[w]);
    return m;
}

// Exclusive scan in place over a[0..n) in global memory by one workgroup.
template <int WG>
__device__ void wg_scan_global(uint32_t *a, uint32_t n, uint32_t *s_wave) {
    uint32_t carry = 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t base = 0; base < n; base += WG * 4) {
        uint32_t i0 = base + threadIdx.x * 4;
        uint32_t v[4], local = 0;
        for (int k = 0; k < 4; k++) { v[k] = (i0 + k < n) ? a[i0 + k] : 0u; local += v[k]; }
        uint32_t inc = local;
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t t = __shfl_up(inc, d, 64);
            if (lane >= d) inc += t;
        }
        __syncthreads();
        if (lane == 63) s_wave[wave] = inc;
        __syncthreads();
        uint32_t woff = 0, total = 0;
        for (int w = 0; w < WG / 64; w++) {
            if (w < wave) woff += s_wave[w];
            total += s_wave[w];
        }
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
// `r0` during a sweep that started at `p` and pushed it `path` in total.  A
// circle at a non-finite position takes part in no pair (every comparison
// with it is false), so it cannot invalidate the list; one whose list-build
// position was non-finite but is finite now forces a rebuild.
__device__ __forceinline__ float moved_bound(float2 p, float2 r0, float path) {
    if (!(fabsf(p.x) <= 3.0e38f) || !(fabsf(p.y) <= 3.0e38f)) return 0.0f;
    if (!(fabsf(r0.x) <= 3.0e38f) || !(fabsf(r0.y) <= 3.0e38f)) return INFINITY;
    float dx = p.x - r0.x, dy = p.y - r0.y;
    float v = sqrtf(dx * dx + dy * dy) + path;
    return v == v ? v : INFINITY;
}

// Candidates of circle c from the hashed cells around it (3x3 cells of size
// >= 2 r_max + skin); writes them to out when given, returns the count.
__device__ __noinline__ uint32_t gather_candidates(uint32_t c, const float2 *__restrict__ ref,
                                                   const float *__restrict__ crad, const int2 *__restrict__ cell,
                                                   const uint32_t *__restrict__ start,
                                                   const uint32_t *__restrict__ items, uint32_t tmask, float skin,
                                                   uint16_t *out) {
    float2 p = ref[c];
    float rc = crad[c];
    int2 ce = cell[c];
    uint32_t n = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            int cx = ce.x + dx, cy = ce.y + dy;
            uint32_t kk = cell_hash(cx, cy) & tmask;
            for (uint32_t s = start[kk], s1 = start[kk + 1]; s < s1; s++) {
                uint32_t o = items[s];
                if (o == c) continue;
                int2 co = cell[o];
                if (co.x != cx || co.y != cy) continue;
                float2 po = ref[o];
                float ex = p.x - po.x, ey = p.y - po.y;
                float lim = (rc + crad[o]) + skin;
                if (ex * ex + ey * ey < lim * lim) {
                    if (out) out[n] = (uint16_t)o;
                    n++;
                }
            }
        }
    return n;
}

__device__ __forceinline__ int scell(float x, double inv_h) {
    double c = floor((double)x * inv_h);
    if (!(c > -1073741824.0)) c = -1073741824.0;
    if (c > 1073741824.0) c = 1073741824.0;
    return (int)c;
}

// MODE 2: pos and prev in LDS; MODE 1: pos in LDS, prev in global memory;
// MODE 0: both in global memory (the scene chunk).
template <int WG, int MODE>
__device__ void run_scene(const SceneDesc &d, uint32_t steps, SceneState *st, float2 *s_lds, ColliderStage *s_col,
                          float *s_red, uint32_t *s_u32, uint32_t *s_flag) {
    char *base = (char *)d.base;
    float2 *gpos = (float2 *)(base + d.pos_off);
    float2 *gprev = (float2 *)(base + d.prev_off);
    const float *crad = (const float *)(base + d.crad_off);
    const uint32_t *pinmask = (const uint32_t *)(base + d.pin_off);
    const uint32_t *loff = (const uint32_t *)(base + d.loff_off);
    const LinkS *links = (const LinkS *)(base + d.links_off);
    float2 *P, *Q;
    if constexpr (MODE >= 1) P = s_lds; else P = gpos;
    if constexpr (MODE == 2) Q = s_lds + d.N; else Q = gprev;
    const uint32_t N = d.N, Np = d.Np, Nc = d.Nc;
    const int tid = threadIdx.x;

    // stage colliders
    const bool staged = d.ncol <= (uint32_t)kStageCol && d.nplanes <= (uint32_t)kStagePlanes;
    const int4 *ghdr = (const int4 *)(base + d.hdr_off);
    const float4 *ggeo = (const float4 *)(base + d.geo_off);
    const float4 *gpl = (const float4 *)(base + d.pl_off);
    if (staged) {
        for (uint32_t k = tid; k < d.ncol; k += WG) { s_col->hdr[k] = ghdr[k]; s_col->geo[k] = ggeo[k]; }
        for (uint32_t k = tid; k < d.nplanes; k += WG) s_col->pl[k] = gpl[k];
    }
    // load state into LDS
    if constexpr (MODE >= 1) {
        for (uint32_t j = tid; j < N; j += WG) {
            P[j] = gpos[j];
            if constexpr (MODE == 2) Q[j] = gprev[j];
        }
    }
    const bool pins = (d.flags & kFlagPins) != 0;
    __syncthreads();

    // circle-pair workspace
    char *ws = base + d.ws_off;
    const CCLayout cl = cc_layout(Nc, d.T);
    float2 *ref = (float2 *)(ws + cl.ref);
    float2 *save = (float2 *)(ws + cl.save);
    float *path = (float *)(ws + cl.path);
    uint32_t *cnt = (uint32_t *)(ws + cl.cnt);
    uint32_t *coff = (uint32_t *)(ws + cl.off);
    uint16_t *cand = (uint16_t *)(ws + cl.cand);
    uint32_t *head = (uint32_t *)(ws + cl.head);
    uint32_t *key = (uint32_t *)(ws + cl.key);
    int2 *cell = (int2 *)(ws + cl.cell);
    uint32_t *start = (uint32_t *)(ws + cl.start);
    uint32_t *fill = (uint32_t *)(ws + cl.fill);
    uint32_t *items = (uint32_t *)(ws + cl.items);
    const uint32_t tmask = d.T ? d.T - 1 : 0;

    float skin = st->skin;
    bool list_valid = st->list_valid != 0;
    uint32_t err = st->err;
    uint32_t n_rounds = 0, n_redos = 0, n_rebuilds = 0, n_pairs = st->pairs;
    bool acc = (d.flags & kFlagAcc) != 0;
    const float2 *acc0 = (const float2 *)(base + d.acc0_off);

    for (uint32_t step = 0; step < steps; step++) {
        for (uint32_t sub = 0; sub < d.sub_steps; sub++) {
            // ---- links: dependency layers, k sweeps
            for (uint32_t it = 0; it < d.link_iters; it++) {
                for (uint32_t L = 0; L < d.depth; L++) {
                    uint32_t lo = loff[L], hi = loff[L + 1];
                    for (uint32_t l = lo + tid; l < hi; l += WG) {
                        LinkS rec = links[l];
                        uint32_t a = rec.a & 0x7fffu, b = rec.b & 0x7fffu;
                        float2 A = P[a], B = P[b];
                        if (a >= Np) {
                            solve_circle_link(A.x, A.y, B.x, B.y, rec.t, crad[a - Np], crad[b - Np]);
                        } else {
                            solve_particle_link(A.x, A.y, B.x, B.y, rec.t, (rec.a & kScenePinBit) != 0,
                                                (rec.b & kScenePinBit) != 0);
                        }
                        P[a] = A;
                        P[b] = B;
                    }
                    __syncthreads();
                }
            }
            // ---- circle pairs
            if (Nc >= 2 && !err) {
                for (int attempt = 0;; attempt++) {
                    // displacement since the list was built
                    float dmax = 0.0f;
                    if (list_valid) {
                        for (uint32_t j = Np + tid; j < N; j += WG) {
                            {
                                float2 p = P[j], r0 = ref[j - Np];
                                dmax = fmaxf(dmax, moved_bound(p, r0, 0.0f));
                            }
                        }
                        dmax = wg_max<WG>(dmax, s_red);
                    }
                    if (!list_valid || dmax > 0.25f * skin) {
                        // ---- build the candidate list
                        n_rebuilds++;
                        double inv_h = 1.0 / (((double)(2.0f * d.rmax) + (double)skin) * 1.001 + 1e-30);
                        for (uint32_t b = tid; b <= d.T; b += WG) start[b] = 0u;
                        for (uint32_t b = tid; b < d.T; b += WG) fill[b] = 0u;
                        __syncthreads();
                        for (uint32_t j = Np + tid; j < N; j += WG) {
                            {
                                uint32_t c = j - Np;
                                float2 p = P[j];
                                ref[c] = p;
                                int cx = scell(p.x, inv_h), cy = scell(p.y, inv_h);
                                cell[c] = make_int2(cx, cy);
                                uint32_t kk = cell_hash(cx, cy) & tmask;
                                key[c] = kk;
                                atomicAdd(&start[kk], 1u);
                            }
                        }
                        __syncthreads();
                        // counts -> bucket starts; start[T] = Nc, bucket k = [start[k], start[k+1])
                        wg_scan_global<WG>(start, d.T + 1, s_u32);
                        for (uint32_t j = Np + tid; j < N; j += WG) {
                            {
                                uint32_t c = j - Np;
                                uint32_t kk = key[c];
                                uint32_t slot = start[kk] + atomicAdd(&fill[kk], 1u);
                                items[slot] = c;
                            }
                        }
                        __syncthreads();
                        // candidates of circle c: every o != c within r_c + r_o + skin of it
                        // at list build; counted, scanned into offsets, then written and
                        // sorted ascending (the reference's pair order)
                        for (uint32_t j = Np + tid; j < N; j += WG) {
                            uint32_t c = j - Np;
                            uint32_t n = gather_candidates(c, ref, crad, cell, start, items, tmask, skin, nullptr);
                            cnt[c] = n;
                            coff[c] = n;
                        }
                        if (tid == 0) coff[Nc] = 0u;
                        __syncthreads();
                        wg_scan_global<WG>(coff, Nc + 1, s_u32);
                        if (coff[Nc] <= kCandCap * Nc) {
                            for (uint32_t j = Np + tid; j < N; j += WG) {
                                uint32_t c = j - Np;
                                uint16_t *my = cand + coff[c];
                                uint32_t n = gather_candidates(c, ref, crad, cell, start, items, tmask, skin, my);
                                for (uint32_t x = 1; x < n; x++) {
                                    uint16_t v = my[x];
                                    uint32_t y = x;
                                    while (y > 0 && my[y - 1] > v) { my[y] = my[y - 1]; y--; }
                                    my[y] = v;
                                }
                            }
                        }
                        uint32_t total = coff[Nc];
                        __syncthreads();
                        if (total > kCandCap * Nc) {
                            // more candidates than the workspace holds: shrink the skin
                            list_valid = false;
                            skin = skin * 0.7f;
                            if (skin < 1e-4f * fmaxf(d.rmax, 1e-3f) || attempt > 12) { err = 1; break; }
                            continue;
                        }
                        uint32_t pairs = tid == 0 ? total : 0u;
                        list_valid = true;
                        n_pairs = pairs;  // per-lane partial; summed below for statistics
                    }
                    // ---- sweep
                    for (uint32_t j = Np + tid; j < N; j += WG) {
                        {
                            uint32_t c = j - Np;
                            save[c] = P[j];
                            path[c] = 0.0f;
                            head[c] = 0u;
                        }
                    }
                    if (tid == 0) { s_flag[0] = 0u; s_flag[1] = 0u; s_flag[2] = 0u; }
                    __syncthreads();
                    uint32_t *hin = head, *hout = head + Nc;
                    for (uint32_t r = 0;; r++) {
                        uint32_t any = 0;
                        for (uint32_t j = Np + tid; j < N; j += WG) {
                            {
                                uint32_t c = j - Np;
                                uint32_t h = hin[c], nc_ = cnt[c], nh = h;
                                if (h < nc_) {
                                    uint32_t o = cand[coff[c] + h];
                                    uint32_t ho = hin[o];
                                    if (ho < cnt[o] && cand[coff[o] + ho] == c) {
                                        if (c < o) {
                                            float2 pi = P[Np + c], pj = P[Np + o];
                                            float qi = 0.0f, qj = 0.0f;
                                            if (solve_circle_pair(pi.x, pi.y, pj.x, pj.y, crad[c], crad[o], qi, qj)) {
                                                P[Np + c] = pi;
                                                P[Np + o] = pj;
                                                path[c] = path[c] + qi;
                                                path[o] = path[o] + qj;
                                            }
                                        }
                                        nh = h + 1;
                                    }
                                }
                                hout[c] = nh;
                                if (nh < nc_) any = 1;
                            }
                        }
                        if (tid == 0) s_flag[(r + 1) % 3] = 0u;
                        if (any) s_flag[r % 3] = 1u;
                        __syncthreads();
                        uint32_t *t = hin; hin = hout; hout = t;
                        n_rounds++;
                        if (!s_flag[r % 3]) break;
                    }
                    // ---- check: a pair left out of the list can only have met if a
                    // circle moved 0.45 * skin or more since the list was built
                    float vmax = 0.0f;
                    for (uint32_t j = Np + tid; j < N; j += WG) {
                        {
                            uint32_t c = j - Np;
                            vmax = fmaxf(vmax, moved_bound(save[c], ref[c], path[c]));
                        }
                    }
                    vmax = wg_max<WG>(vmax, s_red);
                    if (vmax <= 0.45f * skin) break;
                    // undo the sweep, widen the skin (a stale list alone just
                    // needs rebuilding), rebuild
                    for (uint32_t j = Np + tid; j < N; j += WG) P[j] = save[j - Np];
                    __syncthreads();
                    n_redos++;
                    list_valid = false;
                    if (attempt > 12) { err = 2; break; }
                    if (vmax < 1e30f) skin = fmaxf(2.0f * skin, 2.5f * vmax);
                }
            }
            // ---- bodies: colliders -> bounds -> Verlet
            const int4 *hdr = staged ? s_col->hdr : ghdr;
            const float4 *geo = staged ? s_col->geo : ggeo;
            const float4 *pl = staged ? s_col->pl : gpl;
            for (uint32_t j = tid; j < N; j += WG) {
                if (!(pins && j < Np && ((pinmask[j >> 5] >> (j & 31)) & 1u))) {
                    float2 p = P[j];
                    float r = j >= Np ? crad[j - Np] : 0.0f;
                    static_colliders(p, r, d.ncol, hdr, geo, pl);
                    float2 q = Q[j];
                    float ax = d.gex, ay = d.gey;
                    if (j >= Np) {
                        circle_bounds_axis(p.x, q.x, d.lox, d.hix, r);
                        circle_bounds_axis(p.y, q.y, d.loy, d.hiy, r);
                        if (acc) { float2 a0 = acc0[j - Np]; ax = a0.x + d.gx; ay = a0.y + d.gy; }
                    } else {
                        bounds_axis(p.x, q.x, d.lox, d.hix);
                        bounds_axis(p.y, q.y, d.loy, d.hiy);
                    }
                    verlet(p.x, p.y, q.x, q.y, ax, ay, d.delta);
                    P[j] = p;
                    Q[j] = q;
                }
            }
            acc = false;
            __syncthreads();
        }
    }
    // write back
    if constexpr (MODE >= 1) {
        for (uint32_t j = tid; j < N; j += WG) {
            gpos[j] = P[j];
            if constexpr (MODE == 2) gprev[j] = Q[j];
        }
    }
    // statistics (lane 0 holds uniform values except pairs)
    uint32_t pairs_total = n_pairs;
    for (int o = 32; o > 0; o >>= 1) pairs_total += __shfl_xor(pairs_total, o, 64);
    __syncthreads();
    if ((tid & 63) == 0) s_u32[tid >> 6] = pairs_total;
    __syncthreads();
    if (tid == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < WG / 64; w++) tot += s_u32[w];
        st->skin = skin;
        st->list_valid = list_valid ? 1u : 0u;
        st->err = err;
        st->rounds += n_rounds;
        st->redos += n_redos;
        st->rebuilds += n_rebuilds;
        if (n_rebuilds) st->pairs = tot / 2;  // each pair is listed by both circles
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
    __shared__ uint32_t s_flag[4];
    __shared__ uint32_t s_next;
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(queue, 1u);
        __syncthreads();
        uint32_t k = s_next;
        __syncthreads();
        if (k >= n) break;
        uint32_t si = order[k];
        const SceneDesc d = descs[si];
        const uint32_t steps = d.steps < launch_steps ? d.steps : launch_steps;
        if (steps == 0 || d.N == 0) continue;
        if (16ull * d.N <= lds_bytes)
            run_scene<WG, 2>(d, steps, states + d.state_slot, s_lds, &s_col, s_red, s_u32, s_flag);
        else if (8ull * d.N <= lds_bytes)
            run_scene<WG, 1>(d, steps, states + d.state_slot, s_lds, &s_col, s_red, s_u32, s_flag);
        else
            run_scene<WG, 0>(d, steps, states + d.state_slot, s_lds, &s_col, s_red, s_u32, s_flag);
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
    if (!s.polys.empty()) { err = "soft polygons are not supported by the SCENE engine"; return BD_ERR_CAPACITY; }
    if (p.N > kSceneMaxPoints) { err = "scene too large for the SCENE engine"; return BD_ERR_CAPACITY; }
    SceneDesc &d = img.desc;
    std::memset(&d, 0, sizeof(d));
    d.N = p.N; d.Np = p.Np; d.Nc = p.Nc; d.depth = p.depth;
    d.L = (uint32_t)p.links.size();
    d.ncol = (uint32_t)s.cols.size();
    d.nplanes = (uint32_t)(s.planes.size() / 3);
    d.T = 0;
    if (d.Nc >= 2) { d.T = 1; while (d.T < 2 * d.Nc) d.T <<= 1; }
    uint64_t o = 0;
    d.pos_off = (uint32_t)o; o = align16(o + 8ull * d.N);
    d.prev_off = (uint32_t)o; o = align16(o + 8ull * d.N);
    d.crad_off = (uint32_t)o; o = align16(o + 4ull * d.Nc);
    d.acc0_off = (uint32_t)o; o = align16(o + 8ull * d.Nc);
    d.pin_off = (uint32_t)o; o = align16(o + 4ull * ((d.Np + 31) / 32));
    d.loff_off = (uint32_t)o; o = align16(o + 4ull * (d.depth + 1));
    d.links_off = (uint32_t)o; o = align16(o + 8ull * d.L);
    d.hdr_off = (uint32_t)o; o = align16(o + 16ull * d.ncol);
    d.geo_off = (uint32_t)o; o = align16(o + 16ull * d.ncol);
    d.pl_off = (uint32_t)o; o = align16(o + 16ull * d.nplanes);
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
    put(d.pos_off, pos.data(), 8ull * d.N);
    put(d.prev_off, prev.data(), 8ull * d.N);
    put(d.crad_off, s.c_r.data(), 4ull * d.Nc);
    put(d.acc0_off, s.c_acc.data(), 8ull * d.Nc);
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
    img.acc_pending = s.any_initial_acc();
    d.rmax = 0.0f;
    for (float r : s.c_r) d.rmax = std::max(d.rmax, r);
    img.skin0 = std::max(0.5f * d.rmax, 0.05f);
    return BD_OK;
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
        lds_bytes_ = std::min<size_t>(16ull * desc_.N, scene_max_lds());
        BD_HIP(hipStreamSynchronize(stream_));
        N_ = p.N; Np_ = p.Np;
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

    int download(HostScene &s, std::string &err) override {
        std::vector<float2> pos(N_), prev(N_);
        SceneState stt;
        if (N_) {
            BD_HIP(hipMemcpyAsync(pos.data(), chunk_ + desc_.pos_off, 8ull * N_, hipMemcpyDeviceToHost, stream_));
            BD_HIP(hipMemcpyAsync(prev.data(), chunk_ + desc_.prev_off, 8ull * N_, hipMemcpyDeviceToHost, stream_));
        }
        BD_HIP(hipMemcpyAsync(&stt, state_ptr(), sizeof(stt), hipMemcpyDeviceToHost, stream_));
        BD_HIP(hipStreamSynchronize(stream_));
        for (uint32_t i = 0; i < Np_; i++) {
            s.p_pos[i] = V2{pos[i].x, pos[i].y};
            s.p_prev[i] = V2{prev[i].x, prev[i].y};
        }
        for (uint32_t i = Np_; i < N_; i++) {
            s.c_pos[i - Np_] = V2{pos[i].x, pos[i].y};
            s.c_prev[i - Np_] = V2{prev[i].x, prev[i].y};
        }
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
    uint8_t *dev_ = nullptr;  // [0,256) desc, [256,260) order, [260,264) queue, [384,416) state
    uint8_t host_blob_[264];
    SceneDesc desc_{};
    bool acc_pending_ = false;
    int variant_ = 0;
    size_t lds_bytes_ = 0;
    uint32_t N_ = 0, Np_ = 0;
};

bool scene_engine_fits(const HostScene &s, const Plan &p) {
    return s.polys.empty() && p.N <= 16384 && p.N > 0;
}

Engine *make_scene_engine(hipStream_t s) { return new SceneEngine(s); }

}  // namespace bd
