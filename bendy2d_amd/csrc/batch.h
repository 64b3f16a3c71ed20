// batch.h -- launch machinery shared by the batched form and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bendy2d.h"
#include "engine.h"
#include "pack.h"

namespace bd {

// Device-side tables of a set of resident scenes and the persistent launch
// over them: descriptors, the per-class work order, queue counters, two
// streams (one per size class) and an optional captured graph.
struct BatchCore {
    bd_batch_config cfg{};
    std::vector<SceneDesc> descs;     // indexed by scene slot
    std::vector<uint32_t> order;      // tiny, small, then large class, each largest first
    std::vector<uint64_t> order_ticks;   // per scene slot: its device ticks at the last ordering
    std::vector<uint32_t> order_rounds;  // per scene slot: its pair rounds at the last ordering
    uint32_t n_tiny = 0, n_small = 0, n_large = 0;
    bool descs_dirty = true, order_dirty = true;
    DevBuf<uint8_t> d_desc, d_order, d_queue;
    size_t d_desc_cap = 0;
    SceneState *states_ptr = nullptr;
    hipStream_t stream = nullptr, side = nullptr, side2 = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_steps = 0;
    int num_cu = 256;
    float last_ms = 0.0f;
    uint32_t last_launches = 0;
    bool timing_pending = false;

    ~BatchCore();
    void release();
    int init_device(std::string &err);
    void invalidate_graph();
    void build_order(const std::vector<uint32_t> &active);
    // the same classes, each ordered by cost[slot] (largest first, ties by
    // bodies plus links); the launch graph stays valid
    void order_by(const std::vector<uint64_t> &cost);
    int upload_tables(std::string &err);
    int enqueue_launch(uint32_t steps, std::string &err);
    // `launches` launches of `steps` updates each.  The device time between
    // the first call with open_timing and the last with close_timing becomes
    // last_ms (one update may take several run calls).
    int run(uint32_t launches, uint32_t steps, std::string &err, bool open_timing = true, bool close_timing = true);
    int finish_timing(std::string &err);
    SceneState *states_dev() { return states_ptr; }
};

}  // namespace bd
