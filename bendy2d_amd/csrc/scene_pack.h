// scene_pack.h -- host-side helpers of the SCENE engine shared with the
// batched form and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "engine.h"
#include "pack.h"
#include "scene.h"

namespace bd {

// A scene's device chunk image: the first init_bytes are copied from the host,
// the rest (circle-pair workspace) is scratch the kernel initialises.
struct SceneImage {
    std::vector<uint8_t> bytes;
    SceneDesc desc{};
    uint64_t init_bytes = 0, total_bytes = 0;
    bool acc_pending = false;
    float skin0 = 0.05f;
};

int build_scene_image(const HostScene &s, const Plan &p, SceneImage &img, std::string &err);

// The bodies of one scene from a host copy of the first init_bytes of its
// chunk: particles, circles and polygon points by index, the acc of circles
// and polygon points when the chunk was built, each polygon's last centre
// (h.polys gets one record per polygon, holding only that centre), and the
// gravity and substeps the scene runs with (for the acc that static polygon
// points accumulate).
void scene_chunk_to_host(const SceneDesc &d, const uint8_t *chunk, HostScene &h);
void set_scene_params(SceneDesc &d, const StepParams &sp, uint32_t steps, bool acc_pending);

// Kernel variants by workgroup size.  A scene of N points keeps pos and prev
// in LDS when 16*N bytes fit the launch's dynamic LDS, pos alone when 8*N
// fit, else works on its chunk in global memory.
int scene_variant_count();
int scene_variant_wg(int v);
int variant_for_wg(int wg);
size_t scene_max_lds();
// steps: updates each scene runs in this launch (at most its desc.steps).
int launch_scene_variant(int v, uint32_t blocks, size_t lds_bytes, hipStream_t stream, const SceneDesc *d,
                         SceneState *st, const uint32_t *order, uint32_t n, uint32_t *queue, uint32_t steps,
                         std::string &err);
int scene_kernels_init(std::string &err);
// Resident workgroups per CU for variant v at that LDS size.
int scene_blocks_per_cu(int v, size_t lds_bytes);

}  // namespace bd
