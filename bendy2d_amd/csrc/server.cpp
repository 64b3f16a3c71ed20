// server.cpp -- serving: continuous batching of scene-advance requests over
// scenes kept resident in device memory (include/bendy2d.h, bd_server_*).
//
// A scheduler thread repeats: take newly submitted requests; pick the scenes
// whose oldest request is pending (FIFO by submit time); make each resident
// (admitting it into the device arena, evicting least recently used idle
// scenes to host memory when the budget is full); pack them into one compact
// descriptor table, each with min(remaining, steps_per_launch) updates; run one
// persistent batched launch; retire requests whose steps are done.  Requests
// submitted during a launch join the next one.
//
// A scene whose updates take far longer than the launch budget (one update
// cannot be cut short: a pile whose circles move several diameters per
// substep, swept sequentially, can take a second) would hold every launch it
// joins.  Such heavy scenes run in launches of their own on a second stream,
// beside the others: the scheduler starts one when none is in flight and
// retires it when its stream is done, while ordinary launches go on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bendy2d.h"
#include "batch.h"
#include "engine.h"
#include "scene_pack.h"

using namespace bd;

namespace {

// First-fit allocator over one device allocation, 256-byte granules.
struct Arena {
    uint8_t *base = nullptr;
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> bytes
    bool alloc(uint64_t bytes, uint64_t &off) {
        bytes = (bytes + 255) & ~uint64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                uint64_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used += bytes;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = (bytes + 255) & ~uint64_t(255);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    int64_t id, scene;
    uint32_t steps, remaining;
    float dt;
    double t_submit, t_start = -1.0;
};

struct ServedScene {
    int64_t id;
    SceneDesc desc{};               // base / steps / delta / state_slot set per launch
    std::vector<uint8_t> host;      // chunk init bytes while not resident
    uint64_t init_bytes = 0, total_bytes = 0;
    float skin0 = 0.05f;
    StepParams params{};
    bool acc_pending = false;
    bool resident = false;
    uint64_t off = 0;
    uint32_t slot = 0;
    SceneState saved{};
    std::deque<Request> queue;
    uint64_t last_used = 0;
    uint64_t updates_run = 0;
    bool failed = false;            // its circle-pair schedule failed: requests finish with BD_ERR_COLLISION
    double ms_per_update = 0.0;     // device time per update, recent launches (phase timers)
    uint64_t ticks_seen = 0;        // the phase timers' total when last read
    bool heavy_inflight = false;    // in the heavy launch now running
};

}  // namespace

struct bd_server {
    bd_server_config cfg{};
    BatchCore core;
    BatchCore heavy;                     // launches of heavy scenes, on their own stream
    std::vector<ServedScene *> heavy_launch;
    bool heavy_busy = false;
    double heavy_t_start = 0.0;
    double heavy_ms = 100.0;             // a scene above this device time per update is heavy
    double tick_s = 1e-8;                // seconds per device wall-clock tick
    Arena arena;
    DevBuf<SceneState> slots;
    std::vector<uint32_t> free_slots;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;                  // incoming, done, scenes map structure
    std::condition_variable cv, idle_cv;
    std::deque<Request> incoming;
    std::deque<bd_request_done> done;
    std::unordered_map<int64_t, std::unique_ptr<ServedScene>> scenes;
    std::mutex launch_mu;           // held by the scheduler while a launch and its bookkeeping run
    std::thread worker;
    std::atomic<bool> stop{false};
    int64_t next_request = 1, next_scene = 1;
    uint64_t pending = 0;           // requests submitted and not finished
    std::string fatal;

    // statistics (scheduler thread; read under mu)
    uint64_t launches = 0, pu = 0, n_done = 0, evictions = 0, readmissions = 0, waiting = 0, heavy_launches = 0;
    double busy_s = 0.0;
    uint32_t budget_ticks = 0;           // per-scene launch budget in device wall-clock ticks
    std::vector<SceneState> readback;    // states of the last launch's slot range

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static int fail(int code, const std::string &msg) {
    set_last_error(msg);
    return code;
}

static constexpr uint32_t kSlotCap = 1u << 20;

// Device work of the scheduler thread.  Returns false on a fatal error.
static bool admit(bd_server *S, ServedScene &sc, const std::vector<ServedScene *> &keep, std::string &err) {
    if (sc.resident) return true;
    uint64_t off;
    while (!S->arena.alloc(sc.total_bytes, off)) {
        // evict the least recently used idle resident scene (the map can grow
        // under submit, so it is scanned under the server mutex; queues and
        // residency are only changed by this thread)
        ServedScene *victim = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (auto &kv : S->scenes) {
                ServedScene *v = kv.second.get();
                if (!v->resident || !v->queue.empty()) continue;
                if (std::find(keep.begin(), keep.end(), v) != keep.end()) continue;
                if (!victim || v->last_used < victim->last_used) victim = v;
            }
        }
        if (!victim) return false;  // waits for space
        victim->host.resize(victim->init_bytes);
        if (hipMemcpyAsync(victim->host.data(), S->arena.base + victim->off, victim->init_bytes, hipMemcpyDeviceToHost,
                           S->core.stream) != hipSuccess ||
            hipMemcpyAsync(&victim->saved, S->slots.p + victim->slot, sizeof(SceneState), hipMemcpyDeviceToHost,
                           S->core.stream) != hipSuccess ||
            hipStreamSynchronize(S->core.stream) != hipSuccess) {
            err = "eviction copy failed";
            return false;
        }
        S->arena.release(victim->off, victim->total_bytes);
        S->free_slots.push_back(victim->slot);
        victim->resident = false;
        S->evictions++;
    }
    if (S->free_slots.empty()) {
        S->arena.release(off, sc.total_bytes);
        return false;
    }
    sc.slot = S->free_slots.back();
    S->free_slots.pop_back();
    sc.off = off;
    // the circle-pair workspace is not kept across eviction (only the chunk's
    // initialised part and the state travel), so its lists and per-circle
    // skin history start over; results do not depend on them
    sc.saved.list_valid = 0u;
    sc.saved.ws_ready = 0u;
    if (hipMemcpyAsync(S->arena.base + off, sc.host.data(), sc.init_bytes, hipMemcpyHostToDevice, S->core.stream) != hipSuccess ||
        hipMemcpyAsync(S->slots.p + sc.slot, &sc.saved, sizeof(SceneState), hipMemcpyHostToDevice, S->core.stream) != hipSuccess ||
        hipStreamSynchronize(S->core.stream) != hipSuccess) {
        err = "admission copy failed";
        return false;
    }
    if (sc.updates_run) S->readmissions++;
    std::vector<uint8_t>().swap(sc.host);
    sc.resident = true;
    return true;
}

// Bookkeeping of a finished launch: every launched scene reports whether its
// schedule failed, the updates it ran (under a time budget) and its phase
// timers; requests whose steps are done are retired.  Returns false (with
// err) when the state readback fails.
static bool retire(bd_server *S, BatchCore &c, const std::vector<ServedScene *> &launch, uint32_t spl,
                   std::string &err) {
    uint32_t lo_slot = UINT32_MAX, hi_slot = 0;
    for (ServedScene *sc : launch) { lo_slot = std::min(lo_slot, sc->slot); hi_slot = std::max(hi_slot, sc->slot); }
    std::vector<SceneState> &rb = S->readback;
    rb.resize(hi_slot - lo_slot + 1);
    if (hipMemcpyAsync(rb.data(), S->slots.p + lo_slot, rb.size() * sizeof(SceneState), hipMemcpyDeviceToHost,
                       c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        err = "scene state readback failed";
        return false;
    }
    const double t_done = S->now();
    std::vector<bd_request_done> fin;
    uint64_t pu = 0;
    for (ServedScene *sc : launch) {
        Request &rq = sc->queue.front();
        const SceneState &ss = rb[sc->slot - lo_slot];
        // whole updates run (under a time budget a scene may stop between
        // substeps, even before finishing one update: it resumes there)
        uint32_t ran = std::min(rq.remaining, spl);
        if (S->budget_ticks && sc->desc.N) ran = std::min(ran, ss.last_steps);
        rq.remaining -= ran;
        sc->updates_run += ran;
        sc->acc_pending = false;
        sc->last_used = S->launches;
        sc->heavy_inflight = false;
        pu += (uint64_t)sc->desc.N * sc->params.sub_steps * ran;
        // device time per update from the scene's phase timers (a launch that
        // ended inside an update adds to the next whole one)
        const uint64_t ticks = ss.t_links + ss.t_build + ss.t_sweep + ss.t_bodies;
        if (ran) {
            if (ticks >= sc->ticks_seen) {
                const double ms = (double)(ticks - sc->ticks_seen) * S->tick_s * 1e3 / ran;
                sc->ms_per_update = sc->ms_per_update > 0.0 ? 0.5 * (sc->ms_per_update + ms) : ms;
            }
            sc->ticks_seen = ticks;
        }
        if (ss.err) {
            // the request ends here with the error; the scene's later
            // requests are failed before the next launch
            sc->failed = true;
            rq.remaining = 0;
        }
        if (rq.remaining == 0) {
            bd_request_done d{};
            d.request_id = rq.id; d.scene_id = sc->id; d.steps = rq.steps;
            d.status = sc->failed ? BD_ERR_COLLISION : BD_OK;
            d.t_submit = rq.t_submit; d.t_start = rq.t_start; d.t_done = t_done;
            d.particle_updates = sc->failed ? 0 : (uint64_t)sc->desc.N * sc->params.sub_steps * rq.steps;
            fin.push_back(d);
            sc->queue.pop_front();
        }
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->launches++;
        S->pu += pu;
        S->busy_s += c.last_ms * 1e-3;
        for (auto &d : fin) S->done.push_back(d);
        S->n_done += fin.size();
        S->pending -= fin.size();
    }
    if (!fin.empty()) S->idle_cv.notify_all();
    return true;
}

// Descriptors of the launch's scenes (each at most spl updates of its oldest
// request) into core c.
static void describe_launch(bd_server *S, BatchCore &c, const std::vector<ServedScene *> &launch, uint32_t spl,
                            double t_start) {
    c.descs.resize(launch.size());
    for (size_t i = 0; i < launch.size(); i++) {
        ServedScene *sc = launch[i];
        Request &rq = sc->queue.front();
        if (rq.t_start < 0) rq.t_start = t_start;
        StepParams sp = sc->params;
        float mult = 1.0f / (float)sp.sub_steps;  // src/solver.rs:107-108
        sp.delta = rq.dt * mult;
        SceneDesc d = sc->desc;
        d.base = (uint64_t)(uintptr_t)(S->arena.base + sc->off);
        d.state_slot = sc->slot;
        d.budget = S->budget_ticks;
        set_scene_params(d, sp, std::min(rq.remaining, spl), sc->acc_pending);
        c.descs[i] = d;
    }
    std::vector<uint32_t> all(launch.size());
    std::iota(all.begin(), all.end(), 0u);
    c.build_order(all);
    c.descs_dirty = true;
}

static void scheduler(bd_server *S) {
    std::string err;
    const uint32_t spl = std::max<uint32_t>(S->cfg.steps_per_launch, 1);
    for (;;) {
        std::vector<ServedScene *> run;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait_for(lk, std::chrono::milliseconds(2),
                           [&] { return S->stop.load() || !S->incoming.empty() || S->pending > 0; });
            if (S->stop.load()) return;
            while (!S->incoming.empty()) {
                Request r = S->incoming.front();
                S->incoming.pop_front();
                S->scenes[r.scene]->queue.push_back(r);
            }
            // scenes with a pending request, oldest request first
            for (auto &kv : S->scenes)
                if (!kv.second->queue.empty() && !kv.second->heavy_inflight) run.push_back(kv.second.get());
        }
        // a finished heavy launch is retired first
        if (S->heavy_busy && hipStreamQuery(S->heavy.stream) != hipErrorNotReady) {
            std::lock_guard<std::mutex> lg(S->launch_mu);
            int rc = S->heavy.finish_timing(err);
            if (rc != BD_OK || !retire(S, S->heavy, S->heavy_launch, spl, err)) { S->fatal = err; return; }
            S->heavy_busy = false;
            S->heavy_launch.clear();
            continue;  // its scenes may have further requests
        }
        if (run.empty()) {
            if (S->heavy_busy) std::this_thread::sleep_for(std::chrono::microseconds(200));
            continue;
        }
        // a scene whose schedule failed runs no more: its requests finish with
        // the error (the kernel would skip its circle pairs, so results would
        // no longer be the reference's)
        {
            std::vector<bd_request_done> fin;
            const double t_now = S->now();
            for (ServedScene *sc : run) {
                if (!sc->failed) continue;
                while (!sc->queue.empty()) {
                    const Request &rq = sc->queue.front();
                    bd_request_done d{};
                    d.request_id = rq.id; d.scene_id = sc->id; d.steps = rq.steps; d.status = BD_ERR_COLLISION;
                    d.t_submit = rq.t_submit; d.t_start = rq.t_start < 0 ? t_now : rq.t_start; d.t_done = t_now;
                    fin.push_back(d);
                    sc->queue.pop_front();
                }
            }
            if (!fin.empty()) {
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    for (auto &d : fin) S->done.push_back(d);
                    S->n_done += fin.size();
                    S->pending -= fin.size();
                }
                S->idle_cv.notify_all();
                run.erase(std::remove_if(run.begin(), run.end(), [](const ServedScene *sc) { return sc->failed; }),
                          run.end());
                if (run.empty()) continue;
            }
        }
        std::sort(run.begin(), run.end(), [](const ServedScene *a, const ServedScene *b) {
            return a->queue.front().t_submit < b->queue.front().t_submit ||
                   (a->queue.front().t_submit == b->queue.front().t_submit && a->id < b->id);
        });
        if (S->cfg.max_active && run.size() > S->cfg.max_active) run.resize(S->cfg.max_active);
        std::lock_guard<std::mutex> lg(S->launch_mu);
        std::vector<ServedScene *> launch, heavy;
        uint64_t waiting = 0;
        for (ServedScene *sc : run) {
            // heavy scenes wait for the heavy stream to be free
            const bool is_heavy = sc->ms_per_update > S->heavy_ms;
            if (is_heavy && S->heavy_busy) continue;
            if (admit(S, *sc, run, err)) (is_heavy ? heavy : launch).push_back(sc);
            else if (!err.empty()) { S->fatal = err; return; }
            else waiting++;
        }
        S->waiting = waiting;
        const double t_start = S->now();
        if (!heavy.empty()) {
            // started, not waited for: retired by a later pass of this loop
            describe_launch(S, S->heavy, heavy, spl, t_start);
            int rc = S->heavy.run(1, spl, err);
            if (rc != BD_OK) { S->fatal = err; return; }
            for (ServedScene *sc : heavy) sc->heavy_inflight = true;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                S->heavy_launches++;
            }
            S->heavy_launch = heavy;
            S->heavy_busy = true;
        }
        if (launch.empty()) {
            // nothing to run now (heavy scenes waiting for their stream, or
            // scenes waiting for memory): do not spin
            if (heavy.empty()) std::this_thread::sleep_for(std::chrono::microseconds(200));
            continue;
        }
        BatchCore &c = S->core;
        describe_launch(S, c, launch, spl, t_start);
        int rc = c.run(1, spl, err);
        if (rc == BD_OK) rc = c.finish_timing(err);
        if (rc != BD_OK || !retire(S, c, launch, spl, err)) { S->fatal = err; return; }
    }
}

extern "C" {

bd_server *bd_server_new(const bd_server_config *cfg) {
    if (!cfg || cfg->memory_budget < 4096) { set_last_error("memory budget below 4 KiB"); return nullptr; }
    std::unique_ptr<bd_server> S(new bd_server());
    S->cfg = *cfg;
    S->core.cfg = cfg->launch;
    if (!S->core.cfg.wg_small) S->core.cfg = bd_batch_config{128, 1024, 2048, 32 * 1024, 144 * 1024, 1, 0, 400, 64, 8 * 1024};
    S->core.cfg.use_graph = 0;
    S->core.cfg.steps_per_launch = std::max<uint32_t>(cfg->steps_per_launch, 1);
    if (S->core.cfg.lds_tiny > scene_max_lds()) S->core.cfg.lds_tiny = (uint32_t)scene_max_lds();
    if (variant_for_wg((int)S->core.cfg.wg_small) < 0 || variant_for_wg((int)S->core.cfg.wg_large) < 0 ||
        (S->core.cfg.split_tiny && variant_for_wg((int)S->core.cfg.wg_tiny) < 0)) {
        set_last_error("workgroup sizes must be 64, 128, 256, 512 or 1024");
        return nullptr;
    }
    std::string err;
    if (S->core.init_device(err)) { set_last_error(err); return nullptr; }
    S->heavy.cfg = S->core.cfg;
    if (S->heavy.init_device(err)) { set_last_error(err); return nullptr; }
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) {
        set_last_error("device wall clock rate unavailable");
        return nullptr;
    }
    S->tick_s = 1.0 / (1000.0 * (double)khz);
    if (cfg->launch_budget_us) {
        const uint64_t ticks = (uint64_t)cfg->launch_budget_us * (uint64_t)khz / 1000u;
        S->budget_ticks = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(ticks, 1), UINT32_MAX);
    }
    // heavy: ten launch budgets of device time per update (at least 100 ms)
    // unless the caller says
    S->heavy_ms = cfg->heavy_update_us ? cfg->heavy_update_us * 1e-3
                                       : std::max(100.0, 10.0 * cfg->launch_budget_us * 1e-3);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { set_last_error("hipMemGetInfo failed"); return nullptr; }
    uint64_t budget = std::min<uint64_t>(cfg->memory_budget, (uint64_t)(free_b * 0.9));
    budget &= ~uint64_t(255);
    if (hipMalloc((void **)&S->arena.base, budget) != hipSuccess) { set_last_error("arena allocation failed"); return nullptr; }
    S->arena.size = budget;
    S->arena.free_[0] = budget;
    if (S->slots.alloc(kSlotCap) != hipSuccess) { set_last_error("state slots allocation failed"); return nullptr; }
    S->core.states_ptr = S->slots.p;
    S->heavy.states_ptr = S->slots.p;
    S->free_slots.reserve(kSlotCap);
    for (uint32_t i = kSlotCap; i-- > 0;) S->free_slots.push_back(i);
    S->t0 = std::chrono::steady_clock::now();
    bd_server *raw = S.release();
    raw->worker = std::thread(scheduler, raw);
    return raw;
}

void bd_server_free(bd_server *S) {
    if (!S) return;
    S->stop.store(true);
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    if (S->core.stream) (void)hipStreamSynchronize(S->core.stream);
    if (S->heavy.stream) (void)hipStreamSynchronize(S->heavy.stream);
    if (S->arena.base) (void)hipFree(S->arena.base);
    delete S;
}

double bd_server_now(const bd_server *S) { return S ? S->now() : 0.0; }

int64_t bd_server_submit_create(bd_server *S, const bd_solver *scene, uint32_t steps, float dt, int64_t *scene_id) {
    if (!S || !scene) return BD_ERR_ARG;
    std::string err;
    HostScene h;
    int rc = solver_export(scene, h, err);
    if (rc) return fail(rc, err);
    Plan p;
    rc = build_plan(h, p, err);
    if (rc) return fail(rc, err);
    if (p.N > 16384) return fail(BD_ERR_CAPACITY, "served scenes hold at most 16384 bodies");
    SceneImage img;
    rc = build_scene_image(h, p, img, err);
    if (rc) return fail(rc, err);
    std::unique_ptr<ServedScene> sc(new ServedScene());
    sc->desc = img.desc;
    sc->host = std::move(img.bytes);
    sc->init_bytes = img.init_bytes;
    sc->total_bytes = img.total_bytes;
    sc->skin0 = img.skin0;
    sc->saved.skin = img.skin0;
    sc->params = make_params(h, dt, img.acc_pending);
    sc->acc_pending = img.acc_pending;
    // gravity and substeps in the stored descriptor too: reading the scene
    // back derives the acc of static polygon points from them
    set_scene_params(sc->desc, sc->params, 0, sc->acc_pending);
    if (sc->total_bytes > S->arena.size) return fail(BD_ERR_BUDGET, "scene larger than the memory budget");
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->fatal.empty()) return fail(BD_ERR_DEVICE, S->fatal);
    int64_t sid = S->next_scene++, rid = S->next_request++;
    sc->id = sid;
    S->scenes[sid] = std::move(sc);
    S->incoming.push_back(Request{rid, sid, steps, steps, dt, S->now()});
    S->pending++;
    if (scene_id) *scene_id = sid;
    S->cv.notify_one();
    return rid;
}

int64_t bd_server_submit_continue(bd_server *S, int64_t scene_id, uint32_t steps, float dt) {
    if (!S) return BD_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->fatal.empty()) return fail(BD_ERR_DEVICE, S->fatal);
    if (!S->scenes.count(scene_id)) return fail(BD_ERR_INDEX, "unknown scene id");
    int64_t rid = S->next_request++;
    S->incoming.push_back(Request{rid, scene_id, steps, steps, dt, S->now()});
    S->pending++;
    S->cv.notify_one();
    return rid;
}

int32_t bd_server_poll(bd_server *S, bd_request_done *out, int32_t max) {
    if (!S || (!out && max > 0)) return BD_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    int32_t n = 0;
    while (n < max && !S->done.empty()) {
        out[n++] = S->done.front();
        S->done.pop_front();
    }
    return n;
}

int32_t bd_server_wait_idle(bd_server *S, double timeout_s) {
    if (!S) return BD_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    bool ok = S->idle_cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                  [&] { return S->pending == 0 || !S->fatal.empty(); });
    if (!S->fatal.empty()) return fail(BD_ERR_DEVICE, S->fatal);
    return ok ? BD_OK : fail(BD_ERR_STATE, "requests still pending at timeout");
}

int32_t bd_server_read_scene(bd_server *S, int64_t scene_id, bd_solver *out) {
    if (!S || !out) return BD_ERR_ARG;
    ServedScene *sc;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->scenes.find(scene_id);
        if (it == S->scenes.end()) return fail(BD_ERR_INDEX, "unknown scene id");
        sc = it->second.get();
    }
    std::unique_lock<std::mutex> lg(S->launch_mu);  // no ordinary launch in flight
    // nor a heavy launch holding this scene: the scheduler retires it
    while (sc->heavy_inflight && S->fatal.empty()) {
        lg.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(500));
        lg.lock();
    }
    const SceneDesc &d = sc->desc;
    std::vector<uint8_t> buf(sc->init_bytes);
    SceneState st = sc->saved;
    if (sc->resident) {
        if (hipMemcpyAsync(buf.data(), S->arena.base + sc->off, sc->init_bytes, hipMemcpyDeviceToHost, S->core.stream) != hipSuccess ||
            hipMemcpyAsync(&st, S->slots.p + sc->slot, sizeof(st), hipMemcpyDeviceToHost, S->core.stream) != hipSuccess ||
            hipStreamSynchronize(S->core.stream) != hipSuccess)
            return fail(BD_ERR_DEVICE, "scene readback failed");
    } else {
        buf = sc->host;
    }
    HostScene h;
    scene_chunk_to_host(d, buf.data(), h);
    std::string err;
    int rc = solver_import_state(out, h, sc->updates_run, err);
    if (rc) return fail(rc, err);
    if (st.err) return fail(BD_ERR_COLLISION, "circle pair schedule failed for this scene");
    return BD_OK;
}

int32_t bd_server_get_stats(bd_server *S, bd_server_stats *out) {
    if (!S || !out) return BD_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(out, 0, sizeof(*out));
    for (auto &kv : S->scenes) out->resident_scenes += kv.second->resident ? 1 : 0;
    out->known_scenes = S->scenes.size();
    out->bytes_used = S->arena.used;
    out->bytes_budget = S->arena.size;
    out->launches = S->launches;
    out->particle_updates = S->pu;
    out->requests_done = S->n_done;
    out->requests_pending = S->pending;
    out->evictions = S->evictions;
    out->readmissions = S->readmissions;
    out->waiting_admission = S->waiting;
    out->busy_s = S->busy_s;
    out->heavy_launches = S->heavy_launches;
    return BD_OK;
}

}  // extern "C"
