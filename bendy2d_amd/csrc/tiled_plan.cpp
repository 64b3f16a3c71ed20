// tiled_plan.cpp -- host planner of the TILED engine (see tiled_plan.h).
#include "tiled_plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/bendy2d.h"

namespace bd {

static uint32_t spread16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Morton order of the particles' current positions over their bounding box.
static void morton_order(const std::vector<V2> &pos, std::vector<uint32_t> &perm) {
    const size_t n = pos.size();
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (const V2 &p : pos) {
        if (std::isfinite(p.x)) { x0 = std::min(x0, (double)p.x); x1 = std::max(x1, (double)p.x); }
        if (std::isfinite(p.y)) { y0 = std::min(y0, (double)p.y); y1 = std::max(y1, (double)p.y); }
    }
    if (!(x1 >= x0)) { x0 = 0; x1 = 1; }
    if (!(y1 >= y0)) { y0 = 0; y1 = 1; }
    double span = std::max(x1 - x0, y1 - y0);
    if (!(span > 0)) span = 1;
    std::vector<uint64_t> key(n);
    for (size_t i = 0; i < n; i++) {
        double fx = std::isfinite(pos[i].x) ? (pos[i].x - x0) / span : 0.0;
        double fy = std::isfinite(pos[i].y) ? (pos[i].y - y0) / span : 0.0;
        uint32_t qx = (uint32_t)std::min(65535.0, std::max(0.0, fx * 65535.0));
        uint32_t qy = (uint32_t)std::min(65535.0, std::max(0.0, fy * 65535.0));
        key[i] = ((uint64_t)(spread16(qx) | (spread16(qy) << 1)) << 32) | i;
    }
    std::sort(key.begin(), key.end());
    perm.resize(n);
    for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)(key[i] & 0xffffffffu);
}

namespace {

struct Incidence {
    std::vector<uint32_t> off;    // per internal point
    std::vector<uint32_t> link;   // link positions (layer-CSR order), sorted by layer per point
};

// Builds one pass [L0, L1) for every tile; false if a region exceeds the cap.
bool build_pass(const Plan &p, const TilePlan &tp, const Incidence &inc, const std::vector<uint32_t> &link_layer,
                const std::vector<uint32_t> &ia, const std::vector<uint32_t> &ib, uint32_t L0, uint32_t L1,
                uint32_t region_cap, TilePass &out, std::vector<int32_t> &loc, std::vector<uint32_t> &lstamp,
                uint32_t &stamp_base) {
    const uint32_t N = p.N, T = tp.tile_points;
    const uint32_t nl = L1 - L0;
    out = TilePass();
    out.L0 = L0; out.L1 = L1;
    std::vector<uint32_t> R;
    std::vector<uint32_t> E;
    std::vector<std::vector<uint32_t>> per_layer(nl);
    for (uint32_t t = 0; t < tp.ntiles; t++) {
        const uint32_t s0 = t * T, s1 = std::min(N, s0 + T);
        const uint32_t stamp = ++stamp_base;
        R.clear();
        for (uint32_t q = s0; q < s1; q++) { loc[q] = (int32_t)(q - s0); R.push_back(q); }
        for (auto &v : per_layer) v.clear();
        bool ok = true;
        for (uint32_t d = L1; d-- > L0;) {
            const size_t rn = R.size();
            for (size_t i = 0; i < rn; i++) {
                const uint32_t q = R[i];
                // the (at most one) link of layer d at point q
                const uint32_t *b = inc.link.data() + inc.off[q], *e = inc.link.data() + inc.off[q + 1];
                const uint32_t *it = std::lower_bound(b, e, d, [&](uint32_t k, uint32_t dd) { return link_layer[k] < dd; });
                if (it == e || link_layer[*it] != d) continue;
                const uint32_t k = *it;
                if (lstamp[k] == stamp) continue;
                lstamp[k] = stamp;
                per_layer[d - L0].push_back(k);
                const uint32_t o = ia[k] == q ? ib[k] : ia[k];
                if (loc[o] < 0) { loc[o] = (int32_t)R.size(); R.push_back(o); }
            }
            if (((R.size() + 15) & ~size_t(15)) > region_cap) { ok = false; break; }  // slots fill whole blocks of 16
        }
        if (ok) {
            TileInfo ti{};
            ti.int_start = s0; ti.int_count = s1 - s0;
            ti.halo_off = (uint32_t)out.halo.size();
            ti.halo_count = (uint32_t)(R.size() - (s1 - s0));
            ti.loff_off = (uint32_t)out.loff.size();
            ti.link_off = (uint32_t)out.links.size();
            for (size_t i = s1 - s0; i < R.size(); i++) out.halo.push_back(R[i]);
            uint32_t cnt = 0;
            for (uint32_t d = 0; d < nl; d++) {
                out.loff.push_back(cnt);
                std::vector<uint32_t> &v = per_layer[d];
                std::sort(v.begin(), v.end());
                // The links of one layer share no point, so their order is
                // free: deal them out by the bank of their first end (slot
                // mod 16, one 8-byte slot spans two banks), so that
                // neighbouring lanes load from different banks.
                std::vector<uint32_t> by_bank[16];
                for (uint32_t k : v) by_bank[region_slot((uint32_t)loc[ia[k]]) & 15u].push_back(k);
                size_t at[16] = {};
                for (size_t left = v.size(); left;) {
                    for (uint32_t q = 0; q < 16; q++) {
                        if (at[q] == by_bank[q].size()) continue;
                        const uint32_t k = by_bank[q][at[q]++];
                        left--;
                        const Link12 &l = p.links[k];
                        LinkS r;
                        r.a = (uint16_t)(region_slot((uint32_t)loc[ia[k]]) | ((l.a & kPinBit) ? kScenePinBit : 0u));
                        r.b = (uint16_t)(region_slot((uint32_t)loc[ib[k]]) | ((l.b & kPinBit) ? kScenePinBit : 0u));
                        r.t = l.t;
                        out.links.push_back(r);
                    }
                }
                cnt += (uint32_t)v.size();
            }
            out.loff.push_back(cnt);
            out.link_count += cnt;
            out.max_region = std::max<uint32_t>(out.max_region, (uint32_t)((R.size() + 15) & ~size_t(15)));
            out.tiles.push_back(ti);
        }
        for (uint32_t q : R) loc[q] = -1;
        if (!ok) return false;
    }
    return true;
}

}  // namespace

int build_tile_plan(const HostScene &s, const Plan &p, uint32_t tile_points, uint32_t region_cap, TilePlan &out,
                    std::string &err) {
    if (p.Nc || p.Ng) { err = "the TILED engine holds particles and particle links only"; return BD_ERR_CAPACITY; }
    if (region_cap > 32767) region_cap = 32767;
    if (tile_points == 0 || tile_points > region_cap) { err = "tile size exceeds the region capacity"; return BD_ERR_ARG; }
    const uint32_t N = p.N;
    out = TilePlan();
    morton_order(s.p_pos, out.perm);
    out.inv.assign(N, 0);
    for (uint32_t i = 0; i < N; i++) out.inv[out.perm[i]] = i;
    out.tile_points = tile_points;
    out.ntiles = (N + tile_points - 1) / tile_points;
    const uint64_t L = p.links.size();
    std::vector<uint32_t> ia(L), ib(L), link_layer(L);
    for (uint32_t d = 0; d < p.depth; d++)
        for (uint32_t k = p.layer_off[d]; k < p.layer_off[d + 1]; k++) link_layer[k] = d;
    Incidence inc;
    inc.off.assign(N + 1, 0);
    for (uint64_t k = 0; k < L; k++) {
        ia[k] = out.inv[p.links[k].a & kIndexMask];
        ib[k] = out.inv[p.links[k].b & kIndexMask];
        inc.off[ia[k] + 1]++;
        inc.off[ib[k] + 1]++;
    }
    for (uint32_t i = 0; i < N; i++) inc.off[i + 1] += inc.off[i];
    inc.link.resize(2 * L);
    {
        std::vector<uint32_t> fill(inc.off.begin(), inc.off.end() - 1);
        for (uint64_t k = 0; k < L; k++) {  // k ascends in layer order, so lists are sorted by layer
            inc.link[fill[ia[k]]++] = (uint32_t)k;
            inc.link[fill[ib[k]]++] = (uint32_t)k;
        }
    }
    std::vector<int32_t> loc(N, -1);
    std::vector<uint32_t> lstamp(L, 0);
    uint32_t stamp_base = 0;
    const uint32_t D = p.depth;
    if (D == 0) return BD_OK;
    // largest run of layers per pass that fits, halving from the whole sweep
    for (uint32_t per = D;; per = (per + 1) / 2) {
        std::vector<TilePass> passes;
        bool ok = true;
        for (uint32_t L0 = 0; L0 < D && ok; L0 += per) {
            TilePass tp;
            ok = build_pass(p, out, inc, link_layer, ia, ib, L0, std::min(D, L0 + per), region_cap, tp, loc, lstamp,
                            stamp_base);
            if (ok) passes.push_back(std::move(tp));
        }
        if (ok) {
            out.passes = std::move(passes);
            out.max_region = 0;
            for (const TilePass &tp : out.passes) out.max_region = std::max(out.max_region, tp.max_region);
            return BD_OK;
        }
        if (per <= 2) break;
    }
    err = "tile regions exceed the LDS capacity even for 2 layers per pass";
    return BD_ERR_CAPACITY;
}

uint64_t tile_plan_selfcheck(const HostScene &s, const Plan &p, const TilePlan &tp) {
    const uint32_t N = p.N;
    std::vector<V2> seq(N), cur(N), nxt(N);
    for (uint32_t i = 0; i < N; i++) seq[i] = cur[i] = s.p_pos[tp.perm[i]];
    // sequential sweep in layer-CSR order (the reference's order, by layering)
    for (const Link12 &l : p.links) {
        V2 &A = seq[tp.inv[l.a & kIndexMask]], &B = seq[tp.inv[l.b & kIndexMask]];
        solve_particle_link(A.x, A.y, B.x, B.y, l.t, (l.a & kPinBit) != 0, (l.b & kPinBit) != 0);
    }
    std::vector<V2> reg;
    for (const TilePass &ps : tp.passes) {
        const uint32_t nl = ps.L1 - ps.L0;
        for (const TileInfo &ti : ps.tiles) {
            reg.assign((ti.int_count + ti.halo_count + 15) & ~15u, V2{0.0f, 0.0f});
            for (uint32_t r = 0; r < ti.int_count; r++) reg[region_slot(r)] = cur[ti.int_start + r];
            for (uint32_t h = 0; h < ti.halo_count; h++)
                reg[region_slot(ti.int_count + h)] = cur[ps.halo[ti.halo_off + h]];
            for (uint32_t d = 0; d < nl; d++)
                for (uint32_t k = ps.loff[ti.loff_off + d]; k < ps.loff[ti.loff_off + d + 1]; k++) {
                    const LinkS &l = ps.links[ti.link_off + k];
                    V2 &A = reg[l.a & 0x7fffu], &B = reg[l.b & 0x7fffu];
                    solve_particle_link(A.x, A.y, B.x, B.y, l.t, (l.a & kScenePinBit) != 0, (l.b & kScenePinBit) != 0);
                }
            for (uint32_t r = 0; r < ti.int_count; r++) nxt[ti.int_start + r] = reg[region_slot(r)];
        }
        std::swap(cur, nxt);
    }
    uint64_t bad = 0;
    for (uint32_t i = 0; i < N; i++)
        bad += (std::memcmp(&cur[i].x, &seq[i].x, 4) != 0) + (std::memcmp(&cur[i].y, &seq[i].y, 4) != 0);
    return bad;
}

}  // namespace bd
