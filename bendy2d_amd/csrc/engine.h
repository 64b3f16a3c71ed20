// engine.h -- interface between the C ABI layer and the device engines.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bendy2d.h"
#include "scene.h"

namespace bd {

// Per-solver counters, mirrored into bd_stats.
struct Stats {
    uint64_t updates = 0, substeps = 0, launches = 0;
    uint64_t cc_rounds = 0, cc_redos = 0, cc_pairs = 0;
    float cc_margin = 0.0f;
    uint32_t tile_passes = 0;
};

// Per-update scalar parameters, computed on the host exactly as the
// reference does (src/solver.rs:107-108, src/particle.rs:52-54).
struct StepParams {
    float delta;            // dt * (1.0f / sub_steps)
    uint32_t sub_steps, link_iters;
    float gx, gy;           // raw gravity
    float gex, gey;         // 0.0f + gravity: the acc every body holds after apply_gravity
    float lox, loy, hix, hiy;  // bounds.pos and bounds.pos + bounds.size
    bool first_acc;         // initial circle / polygon-point acc still pending
};

StepParams make_params(const HostScene &s, float dt, bool acc_pending);

class Engine {
public:
    virtual ~Engine() {}
    // Copy the host scene into device buffers following the plan.
    virtual int upload(const HostScene &s, const Plan &p, std::string &err) = 0;
    // Run one update (all substeps).
    virtual int update(const StepParams &sp, Stats &st, std::string &err) = 0;
    // Copy positions back into the host scene (and static polygon centres).
    virtual int download(HostScene &s, std::string &err) {
        std::vector<float2> buf(state_count());
        int rc = copy_state(buf.data(), hipMemcpyDeviceToHost, stream(), err);
        if (rc) return rc;
        if (hipStreamSynchronize(stream()) != hipSuccess) {
            err = "hipStreamSynchronize failed";
            return BD_ERR_DEVICE;
        }
        return unpack(buf.data(), s, err);
    }
    // What a download reads, as state_count() float2: pos[N], prev[N], then
    // the engine's extra words (polygon centres, the circle-pair state);
    // copy_state queues its copy into dst (device or host memory, by kind) on
    // stream, unpack fills the host scene from a host copy of it.
    virtual size_t state_count() const = 0;
    virtual int copy_state(void *dst, hipMemcpyKind kind, hipStream_t stream, std::string &err) = 0;
    virtual int unpack(const float2 *src, HostScene &s, std::string &err) = 0;
    virtual hipStream_t stream() const = 0;
    virtual int sync(std::string &err) = 0;
    virtual int kind() const = 0;
};

Engine *make_global_engine(hipStream_t stream);
Engine *make_scene_engine(hipStream_t stream);
Engine *make_tiled_engine(hipStream_t stream, uint32_t tile_points, uint32_t wg);

// A download writes the device's bodies into the host scene by index; the
// host scene must still have the shape the engine was uploaded with.
inline bool host_shape_matches(const HostScene &s, uint64_t Np, uint64_t Nc, uint64_t Ng, std::string &err) {
    if (s.p_pos.size() == Np && s.c_pos.size() == Nc && s.g_pos.size() == Ng) return true;
    err = "host scene changed shape since the device copy was made";
    return false;
}

// True when the SCENE engine (one workgroup, LDS-resident) can run this plan.
bool scene_engine_fits(const HostScene &s, const Plan &p);
bool tiled_engine_fits(const HostScene &s, const Plan &p);

// Thread-local detailed error message used by bd_last_error().
void set_last_error(const std::string &msg);

}  // namespace bd
struct bd_solver;
namespace bd {
int solver_export(const bd_solver *s, HostScene &out, std::string &err);
int solver_import_state(bd_solver *s, const HostScene &src, uint64_t updates, std::string &err);
const HostScene &solver_host(const bd_solver *s);

#define BD_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                   \
            return BD_ERR_DEVICE;                                                      \
        }                                                                              \
    } while (0)

// Device buffer with RAII release.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    hipError_t alloc(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e == hipSuccess) n = count; else p = nullptr;
        return e;
    }
};

}  // namespace bd
