// pack.h -- device layout of scenes run by the SCENE engine (one workgroup
// per scene) and the batched form.
//
// Every scene lives in one contiguous chunk of device memory (so scenes can be
// admitted, evicted and repacked independently):
//
//   pos[N] float2 | prev[N] float2 | crad[Nc] f32 | acc0[Nc+Ng] float2 |
//   pinmask[ceil(Np/32)] u32 | layer_off[depth+1] u32 | links[L] LinkS |
//   collider hdr[ncol] int4 | geo[ncol] float4 | planes[nplanes] float4 |
//   polygons[P] uint4 | centres[P] float2 | static mask[ceil(Ng/32)] u32 |
//   circle-pair workspace (if Nc >= 2)
//
// Scene-local point order: particles [0, Np), circles [Np, Np+Nc), soft
// polygon points [Np+Nc, N), polygon by polygon.  A polygon record is
// (first point, point count, is_static, 0).
#pragma once

#include <cstdint>

#include "common.h"

namespace bd {

// Scene-local link record (8 bytes).  a, b: point indices < 32768; bit 15 of
// a / b marks a pinned end.  A link whose a >= Np is a circle link.
struct LinkS { uint16_t a, b; float t; };

constexpr uint32_t kScenePinBit = 0x8000u;
constexpr uint32_t kSceneMaxPoints = 32767;
constexpr uint32_t kCandCap = 32;  // candidate list capacity per scene: kCandCap * Nc entries

// Scene descriptor: read-only in the kernel.
struct SceneDesc {
    uint64_t base;               // device address of the scene chunk
    uint32_t N, Np, Nc, depth;
    uint32_t L, ncol, nplanes, T;  // T: hash buckets (power of two) for circle pairs
    uint32_t pos_off, prev_off, crad_off, acc0_off;
    uint32_t pin_off, loff_off, links_off, hdr_off;
    uint32_t geo_off, pl_off, ws_off, ws_bytes;
    uint32_t steps;              // updates to run in this launch
    uint32_t sub_steps, link_iters;
    uint32_t flags;              // kFlag*
    float delta, gex, gey, gx, gy;
    float lox, loy, hix, hiy;
    float rmax;
    uint32_t state_slot;         // index of this scene's SceneState
    uint32_t budget;             // device wall-clock ticks after which the scene stops at the
                                 // end of an update (0: run all `steps`); SceneState::last_steps
                                 // tells how many ran
    float runi;                  // every circle's radius, when kFlagUniformR is set
    uint32_t pad0;
    uint64_t ws_base;            // device address of the circle-pair workspace when it is not
                                 // in the chunk (0: base + ws_off)
    uint32_t Ng, P;              // soft polygon points and polygons
    uint32_t poly_off, ctr_off;  // polygon records (uint4) and their centres (float2)
    uint32_t gst_off, pad1;      // static-polygon point mask (one bit per polygon point)
};
static_assert(sizeof(SceneDesc) <= 256, "a scene descriptor fits its 256-byte slot");

constexpr uint32_t kFlagPins = 1u;        // some particle is pinned
constexpr uint32_t kFlagAcc = 2u;         // initial circle acc pending (first substep)
constexpr uint32_t kFlagLds = 4u;         // positions held in LDS
constexpr uint32_t kFlagUniformR = 8u;    // all circles share one radius (runi)
constexpr uint32_t kFlagSequential = 16u; // every circle-pair sweep by the sequential fallback

// Per-scene state that persists between launches (written by the kernel).
struct SceneState {
    float skin;          // base candidate margin of the circle-pair schedule
    uint8_t list_valid;  // candidate list built and still usable
    uint8_t ws_ready;    // per-circle skin state initialised
    uint16_t sub_phase;  // substeps of the current update already run (a launch's time budget
                         // stopped the scene inside an update)
    uint32_t err;        // nonzero: the circle-pair schedule failed
    uint32_t rounds;     // schedule rounds executed (statistics)
    uint32_t redos;      // sweeps redone with larger skins
    uint32_t rebuilds;   // candidate lists built
    uint32_t pairs;      // candidate pairs in the last list
    uint32_t fallbacks;  // sweeps run by the sequential fallback (statistics)
    uint64_t t_links, t_build, t_sweep, t_bodies;  // phase times, device wall-clock ticks (statistics)
    uint32_t since_build;  // sweeps since the candidate list was built
    // the last candidate list (statistics): outliers, of them searched against
    // every circle, longest list, outlier skin threshold, largest skin
    uint32_t n_out, n_far, max_list;
    float s_cell, smax;
    uint64_t t_count, t_lists, t_sort;  // list build split: counts, lists, sorting (ticks)
    uint32_t kept;  // sweeps kept although circles left their skins (no missed pair touched)
    uint32_t unchecked;  // sweeps redone because too many circles left their skins, or too far, to check
    uint32_t last_steps;  // updates the scene's last launch completed
    float smin;           // smallest skin of the last candidate list
};
static_assert(sizeof(SceneState) <= 128, "SceneState must fit the single-scene device blob");

// Byte offsets of the circle-pair workspace inside ws (relative to ws_off).
struct CCLayout {
    uint32_t ref, save, path, skin, need, cnt, off, cand, traj, head, stamp, wl, key, rec, start, fill, items, bytes;
};

BD_HD uint32_t align16(uint64_t v) { return (uint32_t)((v + 15) & ~uint64_t(15)); }

BD_HD CCLayout cc_layout(uint32_t Nc, uint32_t T) {
    CCLayout c{};
    uint64_t o = 0;
    c.ref = (uint32_t)o; o = align16(o + 8ull * Nc);
    c.save = (uint32_t)o; o = align16(o + 8ull * Nc);
    c.path = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.skin = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.need = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.cnt = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.off = (uint32_t)o; o = align16(o + 4ull * (Nc + 1));
    c.cand = (uint32_t)o; o = align16(o + 2ull * kCandCap * Nc);
    c.traj = (uint32_t)o; o = align16(o + 8ull * kCandCap * Nc);
    c.head = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.stamp = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.wl = (uint32_t)o; o = align16(o + 4ull * 2 * (Nc / 2 + 1));
    c.key = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.rec = (uint32_t)o; o = align16(o + 16ull * Nc);
    c.start = (uint32_t)o; o = align16(o + 4ull * (T + 1));
    c.fill = (uint32_t)o; o = align16(o + 4ull * T);
    c.items = (uint32_t)o; o = align16(o + 4ull * Nc);
    c.bytes = (uint32_t)o;
    return c;
}

}  // namespace bd
