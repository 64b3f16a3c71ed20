// scene_image.hip -- one scene in caller-owned device memory (include/bendy2d.h,
// bd_scene_image_*).
//
// An image is the SCENE engine's chunk of one scene behind a 512-byte header:
//
//   [0, 256)    SceneDesc: launch parameters of the scene; base, ws_base,
//               steps, delta and the initial-acc flag are set by each update
//   [256, 272)  words: order[0] = 0 (the one scene), work-queue counter,
//               updates-ran flag, format version
//   [384, 512)  SceneState: the circle-pair schedule's state between updates
//   [512, ...)  the chunk's initialised part: pos, prev, radii, initial acc,
//               pin mask, link layers, links, static colliders, polygons
//
// An update is two launches on the caller's stream: k_image_begin (one lane)
// writes the per-call fields of the header on the device, then the SCENE
// kernel runs the scene for `steps` updates.  Nothing is allocated, copied
// from the host or waited for, so the pair can be captured in a HIP graph.
// The workspace is a second caller buffer; its contents do not carry over
// between calls (every call rebuilds the candidate lists from the positions,
// which does not change results), so one workspace can serve many images.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/bendy2d.h"
#include "engine.h"
#include "pack.h"
#include "scene_pack.h"

namespace bd {
int solver_export(const bd_solver *s, HostScene &out, std::string &err);
}

using namespace bd;

namespace {

constexpr uint64_t kImgWords = 256, kImgState = 384, kImgChunk = 512;
constexpr uint32_t kImgVersion = 0x42440002u;  // "BD", layout 2
constexpr uint32_t kErrWorkspace = 16u;        // SceneState::err: the workspace was too small

static_assert(sizeof(SceneState) <= kImgChunk - kImgState, "scene state fits the header");

struct ImageWords {
    uint32_t order0, queue, ran, version;  // ran: updates run so far
};

int fail(int code, const std::string &msg) {
    set_last_error(msg);
    return code;
}

// Launch shape of a single scene on an otherwise idle device, as the SCENE
// engine picks it: about two bodies per lane; positions, previous positions
// and (when they fit) link records in LDS.
void launch_shape(const SceneDesc &d, uint32_t &wg, uint32_t &lds) {
    uint32_t want = 64;
    while (want < 1024 && want * 2 < d.N) want <<= 1;
    wg = want;
    const size_t with_links = 16ull * d.N + 8ull * d.L;
    lds = (uint32_t)(with_links <= scene_max_lds() ? with_links : std::min<size_t>(16ull * d.N, scene_max_lds()));
}

int image_of(const bd_solver *s, SceneImage &img, std::string &err) {
    HostScene h;
    int rc = solver_export(s, h, err);
    if (rc) return rc;
    Plan p;
    rc = build_plan(h, p, err);
    if (rc) return rc;
    if (!scene_engine_fits(h, p)) {
        err = "scene images hold scenes of 1 to 16384 bodies";
        return BD_ERR_CAPACITY;
    }
    rc = build_scene_image(h, p, img, err);
    if (rc) return rc;
    // the step parameters that do not depend on dt
    set_scene_params(img.desc, make_params(h, 0.0f, img.acc_pending), 0, img.acc_pending);
    return BD_OK;
}

}  // namespace

namespace bd {

// Per-call header fields, written on the device so that the update needs no
// host copy.  delta = dt * (1 / sub_steps) exactly as src/solver.rs:107-108
// computes it (correctly rounded division, no contraction).
__global__ void k_image_begin(uint8_t *img, uint8_t *ws, uint64_t ws_cap, float dt, uint32_t steps) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    SceneDesc *d = (SceneDesc *)img;
    ImageWords *w = (ImageWords *)(img + kImgWords);
    SceneState *st = (SceneState *)(img + kImgState);
    d->base = (uint64_t)(uintptr_t)(img + kImgChunk);
    d->ws_base = (uint64_t)(uintptr_t)ws;
    const float mult = 1.0f / (float)d->sub_steps;
    d->delta = dt * mult;
    // the initial circle acc belongs to the first update that runs
    if (w->ran) d->flags &= ~kFlagAcc;
    w->order0 = 0u;
    w->queue = 0u;
    // the workspace holds nothing from an earlier call
    st->list_valid = 0u;
    st->ws_ready = 0u;
    if (w->version != kImgVersion || (d->Nc >= 2 && d->ws_bytes > ws_cap)) {
        d->steps = 0u;
        st->err = kErrWorkspace;
        return;
    }
    d->steps = steps;
    w->ran += steps;
}

// Shared by the workspace and the allocation-free update: validates the call
// on the host, then launches k_image_begin and the scene kernel.
int image_update(const bd_scene_image_info *info, void *image, void *ws, uint64_t ws_cap, float dt, uint32_t steps,
                 void *stream) {
    if (!info || !image) return fail(BD_ERR_ARG, "null image or info");
    if ((uintptr_t)image & 255u) return fail(BD_ERR_ARG, "a scene image must start on a 256-byte boundary");
    if (info->n_circles >= 2 && (!ws || ws_cap < info->workspace_bytes))
        return fail(BD_ERR_ARG, "workspace smaller than the image's workspace_bytes");
    if ((uintptr_t)ws & 15u) return fail(BD_ERR_ARG, "the workspace must start on a 16-byte boundary");
    const int v = variant_for_wg((int)info->workgroup);
    if (v < 0) return fail(BD_ERR_ARG, "workgroup must be 64, 128, 256, 512 or 1024");
    if (steps == 0) return BD_OK;
    std::string err;
    int rc = scene_kernels_init(err);
    if (rc) return fail(rc, err);
    hipStream_t st = (hipStream_t)stream;
    uint8_t *img = (uint8_t *)image;
    hipLaunchKernelGGL(k_image_begin, dim3(1), dim3(64), 0, st, img, (uint8_t *)ws, ws_cap, dt, steps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(BD_ERR_DEVICE, std::string("image kernel launch: ") + hipGetErrorString(e));
    rc = launch_scene_variant(v, 1, info->lds_bytes, st, (const SceneDesc *)img, (SceneState *)(img + kImgState),
                              (const uint32_t *)(img + kImgWords), 1u, (uint32_t *)(img + kImgWords + 4), steps, err);
    return rc ? fail(rc, err) : BD_OK;
}

// Workspace of the allocation-free update: one array per device, holding the
// circle-pair workspace of any scene an image can hold (16,384 circles need
// 6.4 MiB).  It lives until the process ends.  Host code never uses the
// array's name as an address: scratch_address() asks the runtime for the
// device's copy once per device.
constexpr uint64_t kScratchBytes = 8ull << 20;
__device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// The current device's address of `scratch` (resolved on the first call
// with that device current).
int scratch_address(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(BD_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(BD_ERR_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BD_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return BD_OK;
}
}  // namespace

}  // namespace bd

extern "C" {

int32_t bd_scene_image_update_noalloc(const bd_scene_image_info *info, void *image, float dt, uint32_t steps,
                                      void *hip_stream) {
    if (!info || !image) return BD_ERR_ARG;
    if (info->workspace_bytes > kScratchBytes)
        return fail(BD_ERR_CAPACITY, "the allocation-free update holds at most 8 MiB of workspace "
                                     "(scenes of up to 16384 circles); use bd_scene_image_update with a workspace");
    void *ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    return image_update(info, image, ws, kScratchBytes, dt, steps, hip_stream);
}

int32_t bd_scene_image_describe(const bd_solver *s, bd_scene_image_info *info) {
    if (!s || !info) return BD_ERR_ARG;
    SceneImage img;
    std::string err;
    int rc = image_of(s, img, err);
    if (rc) return fail(rc, err);
    // set the kernels' LDS attribute now, outside any graph capture, when a
    // device is present (describing needs none)
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {
        rc = scene_kernels_init(err);
        if (rc) return fail(rc, err);
        void *ws = nullptr;
        rc = scratch_address(&ws);  // the allocation-free update's workspace on this device
        if (rc) return rc;
    }
    std::memset(info, 0, sizeof(*info));
    info->image_bytes = kImgChunk + img.init_bytes;
    info->workspace_bytes = img.desc.Nc >= 2 ? img.desc.ws_bytes : 0u;
    info->n_bodies = img.desc.N;
    info->n_circles = img.desc.Nc;
    launch_shape(img.desc, info->workgroup, info->lds_bytes);
    return BD_OK;
}

int32_t bd_scene_image_write(const bd_solver *s, void *host_image, uint64_t bytes) {
    if (!s || !host_image) return BD_ERR_ARG;
    SceneImage img;
    std::string err;
    int rc = image_of(s, img, err);
    if (rc) return fail(rc, err);
    if (bytes < kImgChunk + img.init_bytes) return fail(BD_ERR_ARG, "buffer smaller than the image");
    uint8_t *b = (uint8_t *)host_image;
    std::memset(b, 0, kImgChunk);
    SceneDesc d = img.desc;
    d.base = 0;
    d.ws_base = 0;
    d.state_slot = 0;
    d.budget = 0;
    std::memcpy(b, &d, sizeof(d));
    ImageWords w{0u, 0u, 0u, kImgVersion};
    std::memcpy(b + kImgWords, &w, sizeof(w));
    SceneState st{};
    st.skin = img.skin0;
    std::memcpy(b + kImgState, &st, sizeof(st));
    std::memcpy(b + kImgChunk, img.bytes.data(), img.init_bytes);
    return BD_OK;
}

int32_t bd_scene_image_update(const bd_scene_image_info *info, void *image, void *workspace,
                              uint64_t workspace_bytes, float dt, uint32_t steps, void *hip_stream) {
    return image_update(info, image, workspace, workspace_bytes, dt, steps, hip_stream);
}

int32_t bd_scene_image_read(bd_solver *s, const void *host_image, uint64_t bytes) {
    if (!s || !host_image || bytes < kImgChunk) return BD_ERR_ARG;
    const uint8_t *b = (const uint8_t *)host_image;
    SceneDesc d;
    ImageWords w;
    SceneState st;
    std::memcpy(&d, b, sizeof(d));
    std::memcpy(&w, b + kImgWords, sizeof(w));
    std::memcpy(&st, b + kImgState, sizeof(st));
    if (w.version != kImgVersion) return fail(BD_ERR_ARG, "not a scene image");
    if (bytes < kImgChunk + (uint64_t)d.ws_off) return fail(BD_ERR_ARG, "image truncated");
    HostScene h;
    scene_chunk_to_host(d, b + kImgChunk, h);
    std::string err;
    int rc = solver_import_state(s, h, w.ran, err);
    if (rc) return fail(rc, err);
    if (st.err == kErrWorkspace) return fail(BD_ERR_CAPACITY, "an update was given a workspace smaller than the image needs");
    if (st.err) return fail(BD_ERR_COLLISION, "circle pair schedule failed for this scene");
    return BD_OK;
}

}  // extern "C"
