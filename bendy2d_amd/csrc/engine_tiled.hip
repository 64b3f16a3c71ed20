// engine_tiled.hip -- TILED engine: temporal blocking of the link sweeps of
// large particle scenes in LDS (plan in tiled_plan.h).
//
// Per substep and link sweep, one launch per pass: each workgroup loads its
// tile's region (interior + halo) into LDS, applies the pass's dependency
// layers with a barrier between layers, and writes its interior to the other
// position buffer.  The last pass of the last sweep of a substep also runs
// the per-body phase (static colliders, bounds, Verlet) on the interior, so
// a whole substep of the 1024x1024 cloth (8 sweeps x 8 layers) takes 8
// launches instead of 65.  Tile t runs on blockIdx b with b % 8 = XCD, and
// tiles are grouped so that each XCD owns a contiguous Morton range: halo
// reads and the next pass's reads of the same tile stay in that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/bendy2d.h"
#include "bodies.h"
#include "common.h"
#include "engine.h"
#include "kernels.h"
#include "tiled_plan.h"

namespace bd {

__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t ntiles) {
    const uint32_t x = b & 7u, i = b >> 3, q = ntiles >> 3, r = ntiles & 7u;
    return x * q + (x < r ? x : r) + i;
}

template <int WG>
__global__ void __launch_bounds__(WG) k_tiled_pass(const float2 *__restrict__ src, float2 *__restrict__ dst,
                                                   float2 *__restrict__ prev, const TileInfo *__restrict__ tiles,
                                                   uint32_t ntiles, const uint32_t *__restrict__ halo,
                                                   const uint32_t *__restrict__ loff, const LinkS *__restrict__ links,
                                                   uint32_t nlayers, uint32_t fuse,
                                                   const uint32_t *__restrict__ pinmask,
                                                   const int4 *__restrict__ colhdr, const float4 *__restrict__ colgeo,
                                                   const float4 *__restrict__ planes, BodyParams bp) {
    extern __shared__ float2 s_reg[];
    const uint32_t t = xcd_tile(blockIdx.x, ntiles);
    const TileInfo ti = tiles[t];
    const int tid = threadIdx.x;
    for (uint32_t r = tid; r < ti.int_count; r += WG) s_reg[region_slot(r)] = src[ti.int_start + r];
    for (uint32_t h = tid; h < ti.halo_count; h += WG)
        s_reg[region_slot(ti.int_count + h)] = src[halo[ti.halo_off + h]];
    __syncthreads();
    const uint32_t *lo = loff + ti.loff_off;
    const LinkS *lk = links + ti.link_off;
    for (uint32_t L = 0; L < nlayers; L++) {
        const uint32_t e = lo[L + 1];
        for (uint32_t l = lo[L] + tid; l < e; l += WG) {
            LinkS rec = lk[l];
            const uint32_t a = rec.a & 0x7fffu, b = rec.b & 0x7fffu;
            float2 A = s_reg[a], B = s_reg[b];
            solve_particle_link(A.x, A.y, B.x, B.y, rec.t, (rec.a & kScenePinBit) != 0, (rec.b & kScenePinBit) != 0);
            s_reg[a] = A;
            s_reg[b] = B;
        }
        __syncthreads();
    }
    if (fuse) {
        for (uint32_t r = tid; r < ti.int_count; r += WG) {
            const uint32_t u = ti.int_start + r;
            float2 p = s_reg[region_slot(r)];
            if (!(pinmask && ((pinmask[u >> 5] >> (u & 31)) & 1u))) {
                float2 q = prev[u];
                particle_body(p, q, bp, colhdr, colgeo, planes);
                prev[u] = q;
            }
            dst[u] = p;
        }
    } else {
        for (uint32_t r = tid; r < ti.int_count; r += WG) dst[ti.int_start + r] = s_reg[region_slot(r)];
    }
}

// Per-body phase alone (a scene whose particles have no links).
__global__ void k_tiled_bodies(float2 *__restrict__ pos, float2 *__restrict__ prev, uint32_t N,
                               const uint32_t *__restrict__ pinmask, const int4 *__restrict__ colhdr,
                               const float4 *__restrict__ colgeo, const float4 *__restrict__ planes, BodyParams bp) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= N) return;
    if (pinmask && ((pinmask[u >> 5] >> (u & 31)) & 1u)) return;
    float2 p = pos[u], q = prev[u];
    particle_body(p, q, bp, colhdr, colgeo, planes);
    pos[u] = p;
    prev[u] = q;
}

template <int WG>
static void launch_pass(uint32_t grid, size_t lds, hipStream_t s, const float2 *src, float2 *dst, float2 *prev,
                        const TileInfo *tiles, uint32_t ntiles, const uint32_t *halo, const uint32_t *loff,
                        const LinkS *links, uint32_t nlayers, uint32_t fuse, const uint32_t *pin, const int4 *h,
                        const float4 *g, const float4 *pl, const BodyParams &bp) {
    hipLaunchKernelGGL((k_tiled_pass<WG>), dim3(grid), dim3(WG), lds, s, src, dst, prev, tiles, ntiles, halo, loff,
                       links, nlayers, fuse, pin, h, g, pl, bp);
}

struct DevPass {
    uint32_t nlayers = 0, ntiles = 0, max_region = 0;
    DevBuf<TileInfo> tiles;
    DevBuf<uint32_t> halo, loff;
    DevBuf<LinkS> links;
};

class TiledEngine : public Engine {
public:
    explicit TiledEngine(hipStream_t s, uint32_t tile_points, uint32_t wg) : stream_(s), tile_points_(tile_points), wg_(wg) {}
    int kind() const override { return BD_ENGINE_TILED; }

    int upload(const HostScene &s, const Plan &p, std::string &err) override {
        const uint32_t region_cap = kMaxRegionBytes / 8;
        TilePlan tp;
        int rc = build_tile_plan(s, p, tile_points_, region_cap, tp, err);
        if (rc) return rc;
        N_ = p.N;
        perm_ = tp.perm;
        passes_.clear();
        passes_.resize(tp.passes.size());
        for (size_t i = 0; i < tp.passes.size(); i++) {
            const TilePass &h = tp.passes[i];
            DevPass &d = passes_[i];
            d.nlayers = h.L1 - h.L0;
            d.ntiles = (uint32_t)h.tiles.size();
            d.max_region = h.max_region;
            BD_HIP(d.tiles.alloc(std::max<size_t>(h.tiles.size(), 1)));
            BD_HIP(d.halo.alloc(std::max<size_t>(h.halo.size(), 1)));
            BD_HIP(d.loff.alloc(std::max<size_t>(h.loff.size(), 1)));
            BD_HIP(d.links.alloc(std::max<size_t>(h.links.size(), 1)));
            BD_HIP(hipMemcpyAsync(d.tiles.p, h.tiles.data(), h.tiles.size() * sizeof(TileInfo), hipMemcpyHostToDevice, stream_));
            if (!h.halo.empty())
                BD_HIP(hipMemcpyAsync(d.halo.p, h.halo.data(), h.halo.size() * 4, hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(d.loff.p, h.loff.data(), h.loff.size() * 4, hipMemcpyHostToDevice, stream_));
            if (!h.links.empty())
                BD_HIP(hipMemcpyAsync(d.links.p, h.links.data(), h.links.size() * sizeof(LinkS), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipStreamSynchronize(stream_));  // host plan vectors are freed at return
        }
        std::vector<float2> pos(N_), prev(N_);
        for (uint32_t i = 0; i < N_; i++) {
            const V2 &a = s.p_pos[perm_[i]], &b = s.p_prev[perm_[i]];
            pos[i] = make_float2(a.x, a.y);
            prev[i] = make_float2(b.x, b.y);
        }
        BD_HIP(pos_[0].alloc(std::max<uint32_t>(N_, 1)));
        BD_HIP(pos_[1].alloc(std::max<uint32_t>(N_, 1)));
        BD_HIP(prev_.alloc(std::max<uint32_t>(N_, 1)));
        if (N_) {
            BD_HIP(hipMemcpyAsync(pos_[0].p, pos.data(), N_ * sizeof(float2), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(prev_.p, prev.data(), N_ * sizeof(float2), hipMemcpyHostToDevice, stream_));
        }
        cur_ = 0;
        has_pins_ = s.any_pinned();
        if (has_pins_) {
            std::vector<uint32_t> mask((N_ + 31) / 32, 0u);
            for (uint32_t i = 0; i < N_; i++)
                if (s.p_pin[perm_[i]]) mask[i >> 5] |= 1u << (i & 31);
            BD_HIP(pinmask_.alloc(mask.size()));
            BD_HIP(hipMemcpyAsync(pinmask_.p, mask.data(), mask.size() * 4, hipMemcpyHostToDevice, stream_));
        }
        ncol_ = (uint32_t)s.cols.size();
        nplanes_ = (uint32_t)(s.planes.size() / 3);
        if (ncol_) {
            std::vector<int4> hdr(ncol_);
            std::vector<float4> geo(ncol_), pl(std::max<uint32_t>(nplanes_, 1));
            for (uint32_t k = 0; k < ncol_; k++) {
                const Collider &c = s.cols[k];
                hdr[k] = make_int4(c.kind, (int)c.plane_off, (int)c.nv, 0);
                geo[k] = make_float4(c.cx, c.cy, c.r, 0.0f);
            }
            for (uint32_t k = 0; k < nplanes_; k++)
                pl[k] = make_float4(s.planes[3 * k], s.planes[3 * k + 1], s.planes[3 * k + 2], 0.0f);
            BD_HIP(colhdr_.alloc(ncol_));
            BD_HIP(colgeo_.alloc(ncol_));
            BD_HIP(planes_.alloc(pl.size()));
            BD_HIP(hipMemcpyAsync(colhdr_.p, hdr.data(), ncol_ * sizeof(int4), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(colgeo_.p, geo.data(), ncol_ * sizeof(float4), hipMemcpyHostToDevice, stream_));
            BD_HIP(hipMemcpyAsync(planes_.p, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice, stream_));
        }
        static bool attr_done = false;
        if (!attr_done) {
            BD_HIP(hipFuncSetAttribute((const void *)k_tiled_pass<256>, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxRegionBytes));
            BD_HIP(hipFuncSetAttribute((const void *)k_tiled_pass<512>, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxRegionBytes));
            BD_HIP(hipFuncSetAttribute((const void *)k_tiled_pass<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxRegionBytes));
            attr_done = true;
        }
        BD_HIP(hipStreamSynchronize(stream_));
        return BD_OK;
    }

    int update(const StepParams &sp, Stats &st, std::string &err) override {
        BodyParams bp{};
        bp.dt = sp.delta; bp.gex = sp.gex; bp.gey = sp.gey; bp.gx = sp.gx; bp.gy = sp.gy;
        bp.lox = sp.lox; bp.loy = sp.loy; bp.hix = sp.hix; bp.hiy = sp.hiy;
        bp.N = N_; bp.Np = N_; bp.Nc = 0; bp.first_acc = 0; bp.ncol = ncol_; bp.nplanes = nplanes_;
        const uint32_t *pin = has_pins_ ? pinmask_.p : nullptr;
        for (uint32_t step = 0; step < sp.sub_steps; step++) {
            if (passes_.empty()) {
                if (N_) {
                    hipLaunchKernelGGL(k_tiled_bodies, dim3((N_ + 255) / 256), dim3(256), 0, stream_, pos_[cur_].p,
                                       prev_.p, N_, pin, colhdr_.p, colgeo_.p, planes_.p, bp);
                    st.launches++;
                }
            } else {
                for (uint32_t it = 0; it < sp.link_iters; it++) {
                    for (size_t k = 0; k < passes_.size(); k++) {
                        const DevPass &d = passes_[k];
                        const uint32_t fuse = (it + 1 == sp.link_iters && k + 1 == passes_.size()) ? 1u : 0u;
                        const size_t lds = (size_t)d.max_region * sizeof(float2);
                        const float2 *src = pos_[cur_].p;
                        float2 *dst = pos_[cur_ ^ 1].p;
                        if (wg_ == 1024)
                            launch_pass<1024>(d.ntiles, lds, stream_, src, dst, prev_.p, d.tiles.p, d.ntiles, d.halo.p,
                                              d.loff.p, d.links.p, d.nlayers, fuse, pin, colhdr_.p, colgeo_.p, planes_.p, bp);
                        else if (wg_ == 512)
                            launch_pass<512>(d.ntiles, lds, stream_, src, dst, prev_.p, d.tiles.p, d.ntiles, d.halo.p,
                                             d.loff.p, d.links.p, d.nlayers, fuse, pin, colhdr_.p, colgeo_.p, planes_.p, bp);
                        else
                            launch_pass<256>(d.ntiles, lds, stream_, src, dst, prev_.p, d.tiles.p, d.ntiles, d.halo.p,
                                             d.loff.p, d.links.p, d.nlayers, fuse, pin, colhdr_.p, colgeo_.p, planes_.p, bp);
                        cur_ ^= 1;
                        st.launches++;
                    }
                }
            }
            st.substeps++;
        }
        st.tile_passes = (uint32_t)passes_.size();
        BD_HIP(hipGetLastError());
        return BD_OK;
    }

    size_t state_count() const override { return 2ull * N_; }
    hipStream_t stream() const override { return stream_; }

    int copy_state(void *dst, hipMemcpyKind kind, hipStream_t st, std::string &err) override {
        float2 *d = (float2 *)dst;
        if (N_) {
            BD_HIP(hipMemcpyAsync(d, pos_[cur_].p, N_ * sizeof(float2), kind, st));
            BD_HIP(hipMemcpyAsync(d + N_, prev_.p, N_ * sizeof(float2), kind, st));
        }
        return BD_OK;
    }

    int unpack(const float2 *src, HostScene &s, std::string &err) override {
        if (!host_shape_matches(s, N_, 0, 0, err)) return BD_ERR_STATE;
        const float2 *pos = src, *prev = src + N_;
        for (uint32_t i = 0; i < N_; i++) {
            s.p_pos[perm_[i]] = V2{pos[i].x, pos[i].y};
            s.p_prev[perm_[i]] = V2{prev[i].x, prev[i].y};
        }
        return BD_OK;
    }

    int sync(std::string &err) override {
        BD_HIP(hipStreamSynchronize(stream_));
        return BD_OK;
    }

    static constexpr int kMaxRegionBytes = 128 * 1024;

private:
    hipStream_t stream_;
    uint32_t tile_points_, wg_;
    uint32_t N_ = 0, ncol_ = 0, nplanes_ = 0;
    int cur_ = 0;
    bool has_pins_ = false;
    std::vector<uint32_t> perm_;
    std::vector<DevPass> passes_;
    DevBuf<float2> pos_[2], prev_;
    DevBuf<uint32_t> pinmask_;
    DevBuf<int4> colhdr_;
    DevBuf<float4> colgeo_, planes_;
};

bool tiled_engine_fits(const HostScene &s, const Plan &p) {
    return p.Nc == 0 && p.Ng == 0 && s.polys.empty() && p.N > 16384 && p.depth >= 1;
}

Engine *make_tiled_engine(hipStream_t s, uint32_t tile_points, uint32_t wg) {
    return new TiledEngine(s, tile_points ? tile_points : 2048, wg ? wg : 1024);
}

}  // namespace bd
