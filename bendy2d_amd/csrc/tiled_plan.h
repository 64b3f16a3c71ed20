// tiled_plan.h -- temporal blocking of the link sweeps for large particle
// scenes (the TILED engine).
//
// Particles are renumbered in Morton order of their initial positions and cut
// into tiles of consecutive points (a tile's interior).  A pass applies a run
// of consecutive dependency layers [L0, L1) of one sweep to every tile in LDS.
// For a tile with interior I, going backwards over the pass's layers,
//     R_{L1-1} = I,   E_d = links of layer d touching R_d,
//     R_{d-1} = R_d + both ends of E_d,
// and the tile loads R_{L0-1} (interior then halo), applies E_d for d = L0..L1-1
// and writes back I.  Every point of R_d is exact after layer d (its layer-d
// link, if any, is in E_d with both ends exact before it), so the interior
// leaves the pass exactly as the sequential sweep would leave it.  Halo points
// are recomputed by every tile that needs them; only interiors are written,
// into a second position buffer (tiles read the pass's input buffer), so
// tiles never race.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pack.h"
#include "scene.h"

namespace bd {

// LDS slot of a tile region's point m: the low four bits of m are XORed
// with the next four, so that the points of one layer's links, which often
// share low Morton bits (all left ends of horizontal links have an even x),
// spread over the 32 four-byte banks.  It permutes each aligned block of 16,
// so a region of n points needs ceil(n / 16) * 16 slots.
BD_HD uint32_t region_slot(uint32_t m) { return m ^ ((m >> 4) & 15u); }

struct TileInfo {
    uint32_t int_start, int_count;   // interior: internal points [int_start, int_start + int_count)
    uint32_t halo_off, halo_count;   // halo: internal indices halo[halo_off ...]
    uint32_t loff_off, link_off;     // layer offsets loff[loff_off ... + nlayers], links from link_off
    uint32_t pad0, pad1;
};

struct TilePass {
    uint32_t L0 = 0, L1 = 0;         // layers [L0, L1) of one sweep
    std::vector<TileInfo> tiles;
    std::vector<uint32_t> halo;
    std::vector<uint32_t> loff;      // per tile: L1 - L0 + 1 offsets (relative to the tile's link_off)
    std::vector<LinkS> links;        // region slots (region_slot of interior-then-halo indices)
    uint32_t max_region = 0;
    uint64_t link_count = 0;         // links applied per pass (with halo redundancy)
};

struct TilePlan {
    std::vector<uint32_t> perm;      // internal index -> particle index
    std::vector<uint32_t> inv;       // particle index -> internal index
    uint32_t tile_points = 0, ntiles = 0;
    std::vector<TilePass> passes;    // passes of one sweep, in order
    uint32_t max_region = 0;
};

// Builds the plan for a scene of particles and particle links only.
// region_cap bounds the points of any tile region (LDS capacity).  Passes
// cover as many layers as fit (the whole sweep when possible).  Returns BD_OK
// or BD_ERR_CAPACITY (err says why) when no pass of 2+ layers fits.
int build_tile_plan(const HostScene &s, const Plan &p, uint32_t tile_points, uint32_t region_cap,
                    TilePlan &out, std::string &err);

// Runs one sweep of the plan on the host with the engine's link arithmetic
// and compares it with the sequential sweep over the same positions.
// Returns the number of differing f32 values (0: the schedule is exact).
uint64_t tile_plan_selfcheck(const HostScene &s, const Plan &p, const TilePlan &tp);

}  // namespace bd
