/*
 * bendy2d.h -- C ABI of the MI355X-native bendy2d solver step.
 *
 * The crate Murrent/bendy2d 0.0.1 exposes a Rust API (src/lib.rs:1-6):
 * Solver, Particle, ParticleLink/CircleLink/Link, Circle, Polygon, Bounds.
 * This header is the boundary a Rust (or any FFI) facade binds to replace
 * Solver::update (src/solver.rs:106-116) with hand-written HIP kernels for
 * gfx950.  Plain C types only: pointers, sizes, POD records.
 *
 * Semantics: results of bd_update are bit-identical to the reference's f32
 * arithmetic for every phase (gravity, links, circle/polygon collisions,
 * bounds, Verlet integration), including the sequential Gauss-Seidel order of
 * link and circle-pair solves.  Documented deviations:
 *   - links are validated before anything moves; an invalid link makes
 *     bd_update return BD_ERR_LINK (the reference panics mid-sweep,
 *     src/link.rs:19-21,37-39);
 *   - extensions (bd_ext_*) exist that the reference does not have; with
 *     their defaults the solver behaves exactly like the reference.
 *
 * Threading: one bd_solver is one HIP stream and is not thread-safe (the
 * reference's update takes &mut self).  Distinct handles may be used from
 * distinct threads.  All functions use the HIP device that is current on the
 * calling thread when the handle is created.
 */
#ifndef BENDY2D_H
#define BENDY2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Plain records mirroring the crate's Copy types                      */
/* ------------------------------------------------------------------ */

/* nalgebra::Vector2<f32> */
typedef struct bd_vec2 { float x, y; } bd_vec2;

/* Particle {pos, prev_pos, acc}: src/particle.rs:4-9 */
typedef struct bd_particle { bd_vec2 pos, prev_pos, acc; } bd_particle;

/* Link {particle_a, particle_b, target_distance}: src/link.rs:4-10.
 * usize indices are carried as uint64; particle_a < particle_b is required
 * (src/link.rs:6). */
typedef struct bd_link { uint64_t particle_a, particle_b; float target_distance; } bd_link;

/* Circle {point, radius}: src/circle.rs:4-8 */
typedef struct bd_circle { bd_particle point; float radius; } bd_circle;

/* Bounds {pos, size}: src/solver.rs:13-17 */
typedef struct bd_bounds { bd_vec2 pos, size; } bd_bounds;

/* Polygon {particles, particle_links, is_static, center, scale}:
 * src/polygon.rs:7-14.  Input descriptor over caller-owned arrays. */
typedef struct bd_polygon_in {
    const bd_particle *particles; uint64_t n_particles;
    const bd_link *particle_links; uint64_t n_links;
    int32_t is_static;
    bd_vec2 center;
    float scale;
} bd_polygon_in;

/* Scalar fields of a stored Polygon (src/polygon.rs:7-14). */
typedef struct bd_polygon_info {
    uint64_t n_particles, n_links;
    int32_t is_static;
    bd_vec2 center;
    float scale;
} bd_polygon_info;

/* ------------------------------------------------------------------ */
/* Error codes (the reference has none: it panics or yields NaN)       */
/* ------------------------------------------------------------------ */
#define BD_OK 0
#define BD_ERR_ARG (-1)        /* null handle, bad size or value */
#define BD_ERR_LINK (-2)       /* a link with particle_a >= particle_b or out of range */
#define BD_ERR_INDEX (-3)      /* index out of range */
#define BD_ERR_DEVICE (-4)     /* a HIP call failed */
#define BD_ERR_NOMEM (-5)      /* host or device allocation failed */
#define BD_ERR_CAPACITY (-6)   /* a size limit of the engine was exceeded */
#define BD_ERR_COLLISION (-7)  /* the circle-pair schedule did not converge */
#define BD_ERR_NODEVICE (-8)   /* no HIP device is visible */
#define BD_ERR_BUDGET (-9)     /* serving: admission refused by the memory budget */
#define BD_ERR_STATE (-10)     /* call not valid in the current state */

/* Static message for a code; bd_last_error() gives the last detailed message
 * produced on the calling thread. */
const char *bd_error_string(int32_t code);
const char *bd_last_error(void);

/* ------------------------------------------------------------------ */
/* Solver lifecycle and public fields                                  */
/* ------------------------------------------------------------------ */
typedef struct bd_solver bd_solver;

/* Solver::new, with its defaults: gravity (0, 98.2), bounds (0,0)+(100,100),
 * bounds_active true, sub_steps 1 (src/solver.rs:34-50).  NULL on failure. */
bd_solver *bd_solver_new(void);
void bd_solver_free(bd_solver *s);
/* #[derive(Clone)] on Solver (src/solver.rs:19): deep copy of the state. */
bd_solver *bd_solver_clone(const bd_solver *s);

/* Snapshot: the whole solver (bodies, links, colliders, public fields and
 * extension settings) as a versioned byte image with a checksum; the
 * reference's only analogue is #[derive(Clone)] on Solver (src/solver.rs:19).
 * bd_solver_save returns the snapshot's size in bytes (buf NULL: size only)
 * or a negative BD_ERR_*; bd_solver_load returns a new solver, or NULL
 * (bd_last_error) when the bytes are not a valid snapshot.  A run continued
 * from a loaded snapshot equals the uninterrupted run bit for bit. */
int64_t bd_solver_save(const bd_solver *s, void *buf, uint64_t cap);
bd_solver *bd_solver_load(const void *buf, uint64_t bytes);

/* pub gravity / pub bounds / pub bounds_active (src/solver.rs:21-23).
 * bounds_active is stored and returned but, as in the reference, never read
 * by bd_update (src/solver.rs:155-165). */
void bd_set_gravity(bd_solver *s, bd_vec2 g);
bd_vec2 bd_get_gravity(const bd_solver *s);
void bd_set_bounds(bd_solver *s, bd_bounds b);
bd_bounds bd_get_bounds(const bd_solver *s);
void bd_set_bounds_active(bd_solver *s, int32_t active);
int32_t bd_get_bounds_active(const bd_solver *s);

/* ------------------------------------------------------------------ */
/* Adders: return the new index (>= 0) or a negative BD_ERR_*          */
/* ------------------------------------------------------------------ */
/* Solver::add_particle(pos) -> Particle::new(pos): src/solver.rs:52-54 */
int64_t bd_add_particle(bd_solver *s, bd_vec2 pos);
/* Solver::add_circle(circle): src/solver.rs:55-57 */
int64_t bd_add_circle(bd_solver *s, const bd_circle *c);
/* Solver::add_polygon(polygon): src/solver.rs:58-60 */
int64_t bd_add_polygon(bd_solver *s, const bd_polygon_in *p);
/* Solver::add_particle_link(link): src/solver.rs:62-64 (unchecked until update) */
int64_t bd_add_particle_link(bd_solver *s, const bd_link *l);
/* Solver::add_circle_link(link): src/solver.rs:65-67 */
int64_t bd_add_circle_link(bd_solver *s, const bd_link *l);

/* ------------------------------------------------------------------ */
/* Getters (copy out; the reference lends &Vec / Option<&T>)           */
/* ------------------------------------------------------------------ */
/* get_particle_len / get_circles_len / get_polygons_len: src/solver.rs:69-77 */
uint64_t bd_get_particle_len(const bd_solver *s);
uint64_t bd_get_circles_len(const bd_solver *s);
uint64_t bd_get_polygons_len(const bd_solver *s);
/* get_particle_links / get_circle_links: src/solver.rs:79-84.  Copy up to
 * cap records, return the total count. */
uint64_t bd_get_particle_links(const bd_solver *s, bd_link *out, uint64_t cap);
uint64_t bd_get_circle_links(const bd_solver *s, bd_link *out, uint64_t cap);
/* get_particles / get_circles: src/solver.rs:86-91 */
uint64_t bd_get_particles(const bd_solver *s, bd_particle *out, uint64_t cap);
uint64_t bd_get_circles(const bd_solver *s, bd_circle *out, uint64_t cap);
/* get_particle / get_circle: src/solver.rs:96-101.  1 = Some (out filled),
 * 0 = None. */
int32_t bd_get_particle(const bd_solver *s, uint64_t index, bd_particle *out);
int32_t bd_get_circle(const bd_solver *s, uint64_t index, bd_circle *out);
/* get_polygon / get_polygons: src/solver.rs:92-94,102-104.  info gets the
 * scalar fields; the arrays are copied by the two calls below. */
int32_t bd_get_polygon(const bd_solver *s, uint64_t index, bd_polygon_info *info);
uint64_t bd_get_polygon_particles(const bd_solver *s, uint64_t index, bd_particle *out, uint64_t cap);
uint64_t bd_get_polygon_links(const bd_solver *s, uint64_t index, bd_link *out, uint64_t cap);

/* ------------------------------------------------------------------ */
/* The hot path                                                        */
/* ------------------------------------------------------------------ */
/* Solver::update(dt): src/solver.rs:106-116.  Runs sub_steps substeps of
 * gravity -> links -> dynamic collisions -> bounds -> integrate on the GPU.
 * Returns BD_OK or a negative BD_ERR_*; on BD_ERR_LINK nothing has moved.
 * On a failure during the step (BD_ERR_COLLISION, BD_ERR_DEVICE) the
 * substeps that completed stay applied and getters show them.  Adders and
 * other structural changes return the error instead of changing anything
 * while the device state cannot be read back.
 * The call is asynchronous with respect to the host except where the circle
 * pair schedule needs a readback; getters synchronise. */
int32_t bd_update(bd_solver *s, float dt);

/* ------------------------------------------------------------------ */
/* Polygon constructors (host code, bit-identical to the crate)        */
/* ------------------------------------------------------------------ */
/* Polygon::circle(radius, pos, point_count, is_static): src/polygon.rs:17-82.
 * out_particles holds point_count records, out_links 2*point_count. */
int32_t bd_polygon_circle(float radius, bd_vec2 pos, uint64_t point_count, int32_t is_static,
                          bd_particle *out_particles, bd_link *out_links, bd_vec2 *out_center);
/* Polygon::new(points, is_static): src/polygon.rs:84-123.  out_links holds
 * n records. */
int32_t bd_polygon_new(const bd_vec2 *points, uint64_t n, int32_t is_static,
                       bd_particle *out_particles, bd_link *out_links, bd_vec2 *out_center);

/* ------------------------------------------------------------------ */
/* Extensions (not in the reference; defaults reproduce it exactly)    */
/* ------------------------------------------------------------------ */
/* EXT1: number of substeps per update (reference: fixed 1, src/solver.rs:29,47).
 * delta = dt * (1.0f / n), as src/solver.rs:107-108 computes it. */
int32_t bd_ext_set_sub_steps(bd_solver *s, uint32_t n);
uint32_t bd_ext_get_sub_steps(const bd_solver *s);
/* EXT2: link sweeps per substep (reference: one sweep, src/solver.rs:143-153). */
int32_t bd_ext_set_link_iterations(bd_solver *s, uint32_t k);
uint32_t bd_ext_get_link_iterations(const bd_solver *s);
/* EXT3: pin a solver particle.  A pinned particle never moves; a link to it
 * moves only its free end, by the whole correction. */
int32_t bd_ext_set_pinned(bd_solver *s, uint64_t index, int32_t pinned);
/* EXT6: a particle with an explicit previous position (initial velocity). */
int64_t bd_ext_add_particle_with_prev(bd_solver *s, bd_vec2 pos, bd_vec2 prev_pos);
/* EXT4: static colliders (infinite mass), applied to every body after the
 * reference's dynamic collisions and before bounds. */
int64_t bd_ext_add_static_circle(bd_solver *s, bd_vec2 center, float radius);
int64_t bd_ext_add_static_polygon(bd_solver *s, const bd_vec2 *vertices, uint64_t n);
uint64_t bd_ext_get_static_collider_len(const bd_solver *s);
/* Read back collider i: kind 0 = circle (center, radius), kind 1 = polygon
 * (its n_vertices vertices, as given, copied to vertices if cap allows).
 * Returns the kind, or a negative BD_ERR_*. */
int32_t bd_ext_get_static_collider(const bd_solver *s, uint64_t i, bd_vec2 *center, float *radius,
                                   bd_vec2 *vertices, uint64_t cap, uint64_t *n_vertices);
/* Copy pin flags of particles [first, first+n) into out (1 = pinned). */
int32_t bd_ext_get_pinned_range(const bd_solver *s, uint64_t first, uint8_t *out, uint64_t n);

/* Bulk adders for scene set-up at scale (same semantics as repeated adds).
 * prev may be NULL (prev = pos).  Return the index of the first new item. */
int64_t bd_ext_add_particles(bd_solver *s, const bd_vec2 *pos, const bd_vec2 *prev, uint64_t n);
int64_t bd_ext_add_particle_links(bd_solver *s, const uint32_t *a, const uint32_t *b,
                                  const float *target, uint64_t n);
int64_t bd_ext_add_circles(bd_solver *s, const bd_circle *c, uint64_t n);
int64_t bd_ext_add_circle_links(bd_solver *s, const uint32_t *a, const uint32_t *b,
                                const float *target, uint64_t n);
/* Bulk pin: pinned[i] != 0 pins particle first+i. */
int32_t bd_ext_set_pinned_range(bd_solver *s, uint64_t first, const uint8_t *pinned, uint64_t n);

/* Engine selection.  AUTO picks SCENE (one workgroup, state in LDS, all
 * substeps in one launch) when the scene fits it, else GLOBAL (one launch
 * per link layer over the whole device), else TILED where the link graph
 * allows it. */
#define BD_ENGINE_AUTO 0
#define BD_ENGINE_GLOBAL 1
#define BD_ENGINE_SCENE 2
#define BD_ENGINE_TILED 3
int32_t bd_ext_set_engine(bd_solver *s, int32_t engine);
int32_t bd_ext_get_engine_used(const bd_solver *s);
/* TILED engine shape: interior points per tile (64..16384) and workgroup
 * size (256, 512 or 1024).  Defaults 2048 and 1024. */
int32_t bd_ext_set_tiling(bd_solver *s, uint32_t tile_points, uint32_t workgroup);
/* Host check of the TILED schedule for this scene: builds the tile plan
 * (region_cap points per tile region at most), runs one link sweep of it on
 * the host with the engine's arithmetic and compares it with the sequential
 * sweep.  Returns the number of differing values (0 = exact) or a negative
 * BD_ERR_*; passes receives the passes per sweep. */
int64_t bd_ext_tiled_selfcheck(const bd_solver *s, uint32_t tile_points, uint32_t region_cap, uint32_t *passes);

/* Readback of every update (no counterpart in the reference, whose getters
 * lend live state, src/solver.rs:86-104).  Off (default): a getter after an
 * update copies the state back then, synchronously.  On: each bd_update also
 * queues a copy of the new state into pinned host memory, snapshotted on the
 * device and copied on a second stream so that it overlaps the next update's
 * kernels; a getter then only waits for that copy.  Values are the same
 * either way. */
int32_t bd_ext_set_readback(bd_solver *s, int32_t on);

/* Wait for all queued device work of this solver. */
int32_t bd_ext_sync(bd_solver *s);

/* Counters of the last plan and of the work issued so far. */
typedef struct bd_stats {
    uint64_t n_points;          /* particles + circle points + polygon points */
    uint64_t n_links;           /* all links, all groups */
    uint64_t link_layers;       /* dependency layers of one link sweep */
    uint64_t updates;           /* bd_update calls that ran */
    uint64_t substeps;          /* substeps executed */
    uint64_t kernel_launches;   /* kernel launches issued */
    uint64_t cc_rounds;         /* circle-pair schedule rounds executed */
    uint64_t cc_redos;          /* circle-pair sweeps redone with a larger margin */
    uint64_t cc_pairs;          /* candidate circle pairs in the last sweep */
    float cc_margin;            /* current candidate margin */
    int32_t engine;             /* engine used by the last update */
    uint32_t tile_passes;       /* TILED engine: passes per link sweep */
} bd_stats;
int32_t bd_ext_get_stats(const bd_solver *s, bd_stats *out);

/* Link-layer planner (exposed for tests and tools).  For links (a[i], b[i])
 * over n_points points, in sequence order, layer[i] = 1 + max(layer of the
 * last earlier link touching a[i] or b[i]) (0-based output).  order receives
 * the link indices sorted by (layer, sequence index).  Returns the number of
 * layers, or a negative BD_ERR_*. */
int64_t bd_plan_layers(const uint32_t *a, const uint32_t *b, uint64_t n_links, uint64_t n_points,
                       uint32_t *layer, uint64_t *order);

/* ------------------------------------------------------------------ */
/* Batched form: many independent scenes in one persistent launch      */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (it has one Solver per world).  A batch
 * holds copies of independent scenes (particles, circles, soft polygons and
 * static colliders; at most 16384 bodies each) resident in device memory; one persistent kernel launch per size
 * class advances all of them, one workgroup per scene taken from a work
 * queue, longest first (by bodies plus links, then by each scene's measured
 * device time).  Results per scene are bit-identical to bd_update on the
 * same scene. */
typedef struct bd_batch bd_batch;

/* Launch configuration.  Scenes with at most split_points bodies run on
 * workgroups of wg_small threads, larger ones on wg_large; lds_small /
 * lds_large bytes of LDS per workgroup decide residency (pos and prev in LDS
 * when 16 B per body fit, pos alone when 8 B fit, else global memory);
 * steps_per_launch updates run per kernel launch; use_graph replays the launch
 * sequence as a captured hipGraph.  A third class takes the smallest scenes:
 * those of at most split_tiny bodies run on workgroups of wg_tiny threads with
 * lds_tiny bytes of LDS (split_tiny = 0: no third class). */
typedef struct bd_batch_config {
    uint32_t wg_small, wg_large;
    uint32_t split_points;
    uint32_t lds_small, lds_large;
    uint32_t steps_per_launch;
    int32_t use_graph;
    uint32_t split_tiny, wg_tiny, lds_tiny;
} bd_batch_config;

bd_batch *bd_batch_new(void);
void bd_batch_free(bd_batch *b);
/* Copy a scene (its state now) into the batch; returns its id (>= 0). */
int64_t bd_batch_add_scene(bd_batch *b, const bd_solver *scene);
uint64_t bd_batch_len(const bd_batch *b);
uint64_t bd_batch_total_points(const bd_batch *b);
/* Advance every scene by `steps` updates of dt (each update runs the scene's
 * own sub_steps and link iterations).  Each call first reads the scenes'
 * device timers back (one stream synchronisation) to rank the queues. */
int32_t bd_batch_update(bd_batch *b, float dt, uint32_t steps);
int32_t bd_batch_sync(bd_batch *b);
/* Copy scene id's positions back into `out`, a solver holding the same scene
 * (e.g. the one passed to bd_batch_add_scene, or its clone). */
int32_t bd_batch_read_scene(bd_batch *b, int64_t id, bd_solver *out);
int32_t bd_batch_set_config(bd_batch *b, const bd_batch_config *cfg);
int32_t bd_batch_get_config(const bd_batch *b, bd_batch_config *cfg);
/* Time candidate configurations on this batch (each for `steps` updates of
 * dt on a scratch copy of the state, which is restored afterwards, after the
 * same updates once untimed) and keep the fastest: workgroup sizes, the
 * small/large split, LDS budgets, updates per launch.  The smallest class
 * keeps its configured shape.  Returns the number of configurations tried. */
int32_t bd_batch_autotune(bd_batch *b, float dt, uint32_t steps);
/* Device time of the last bd_batch_update in milliseconds (HIP events on the
 * batch's stream around the launches), and the launches it took. */
int32_t bd_batch_last_timing(const bd_batch *b, float *ms, uint32_t *launches);
/* Circle-pair schedule counters summed over scenes; err_scenes counts scenes
 * whose schedule failed (their results are not valid). */
int32_t bd_batch_cc_stats(bd_batch *b, uint64_t *rounds, uint64_t *redos, uint64_t *rebuilds,
                          uint64_t *err_scenes);
/* Sweeps run by the sequential fallback (a sweep whose candidate lists
 * failed 8 attempts, or did not fit the workspace at the smallest skin,
 * runs the reference's (i, j) loop row by row instead), summed over scenes,
 * and the scenes that ran at least one. */
int32_t bd_batch_cc_fallbacks(bd_batch *b, uint64_t *fallbacks, uint64_t *scenes);
/* Circle-pair schedule state of one scene: base skin, error kind (0 none,
 * 3 stalled schedule; the kinds 1 and 2 of earlier versions, lists beyond
 * the workspace and no valid sweep in the attempts allowed, now run the
 * sequential fallback instead), rounds, redos, rebuilds and candidate
 * pairs. */
int32_t bd_batch_scene_status(bd_batch *b, int64_t id, float *skin, uint32_t *err, uint32_t *rounds,
                              uint32_t *redos, uint32_t *rebuilds, uint32_t *pairs);
/* Device time a scene spent per phase so far, in seconds: link layers,
 * candidate-list builds, pair sweeps (rounds and checks), bodies. */
int32_t bd_batch_scene_times(bd_batch *b, int64_t id, double *links, double *build, double *sweep, double *bodies);
/* The scene's last candidate-list build (statistics): outlier circles (skin
 * above the cell threshold), of them those searched against every circle,
 * the longest list, the threshold and the largest skin; the device time so
 * far of list builds split into counting, list filling and sorting, in
 * seconds; the sweeps kept although circles moved beyond their skins
 * (every pair missing from the lists was checked not to touch), and those
 * redone because too many circles, or some too far, left their skins to
 * check, and the sweeps run by the sequential fallback. */
int32_t bd_batch_scene_lists(bd_batch *b, int64_t id, uint32_t *n_out, uint32_t *n_far, uint32_t *max_list,
                             float *s_cell, float *smax, double *t_count, double *t_lists, double *t_sort,
                             uint32_t *kept, uint32_t *unchecked, uint32_t *fallbacks);

/* ------------------------------------------------------------------ */
/* Serving: continuous batching of scene-advance requests              */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference.  A server keeps scenes resident in device
 * memory within a byte budget and runs a scheduler thread: every launch
 * advances each scene that has a pending request by up to steps_per_launch
 * updates, in one persistent batched launch over all of them; requests that
 * arrive meanwhile join the next launch.  A scene's requests run in arrival
 * order.  When the budget is full, idle scenes (no pending request) are
 * evicted to host memory, least recently used first, and brought back when a
 * request names them; a new scene that cannot fit waits.  Results are
 * bit-identical to bd_update on the same scene, alone or in any mix.  A scene
 * whose circle-pair schedule fails (an engine capacity limit) stops running:
 * the request in flight and every later request for it finish with status
 * BD_ERR_COLLISION and particle_updates 0. */
typedef struct bd_server bd_server;

typedef struct bd_server_config {
    uint64_t memory_budget;     /* bytes of device memory for resident scenes */
    uint32_t steps_per_launch;  /* updates per launch for each active scene */
    uint32_t max_active;        /* scenes per launch at most (0 = no limit) */
    bd_batch_config launch;     /* launch shape (wg sizes, split, LDS); steps_per_launch and use_graph ignored */
    uint32_t launch_budget_us;  /* a scene stops after the substep during which it passed this much device
                                   time in a launch, so slow scenes do not hold up the rest (0: no limit);
                                   it resumes there in its next launch */
    uint32_t heavy_update_us;   /* a scene whose updates take more device time than this runs in launches
                                   of its own on a second stream (0: ten launch budgets, at least 100 ms) */
} bd_server_config;

/* A finished request.  Times are seconds on the server's monotonic clock
 * (bd_server_now). */
typedef struct bd_request_done {
    int64_t request_id, scene_id;
    uint32_t steps;
    int32_t status;             /* BD_OK or a negative BD_ERR_* */
    double t_submit, t_start, t_done;
    uint64_t particle_updates;  /* bodies x substeps x steps */
} bd_request_done;

typedef struct bd_server_stats {
    uint64_t resident_scenes, known_scenes, bytes_used, bytes_budget;
    uint64_t launches, particle_updates, requests_done, requests_pending;
    uint64_t evictions, readmissions, waiting_admission;
    double busy_s;              /* device time of all launches */
    uint64_t heavy_launches;    /* launches of heavy scenes on the second stream (of `launches`) */
} bd_server_stats;

bd_server *bd_server_new(const bd_server_config *cfg);
void bd_server_free(bd_server *srv);
double bd_server_now(const bd_server *srv);
/* New scene (a copy of `scene`) advanced by `steps` updates of dt.  Returns
 * the request id; *scene_id receives the id later requests use. */
int64_t bd_server_submit_create(bd_server *srv, const bd_solver *scene, uint32_t steps, float dt, int64_t *scene_id);
/* Advance a known scene by `steps` more updates of dt. */
int64_t bd_server_submit_continue(bd_server *srv, int64_t scene_id, uint32_t steps, float dt);
/* Move up to max finished requests into out; returns how many. */
int32_t bd_server_poll(bd_server *srv, bd_request_done *out, int32_t max);
/* Block until no request is pending (or timeout_s passes: BD_ERR_STATE). */
int32_t bd_server_wait_idle(bd_server *srv, double timeout_s);
/* Copy a scene's current state into `out` (a solver holding the same scene). */
int32_t bd_server_read_scene(bd_server *srv, int64_t scene_id, bd_solver *out);
int32_t bd_server_get_stats(bd_server *srv, bd_server_stats *out);

/* ------------------------------------------------------------------ */
/* Scene images: one scene stepped in caller-owned device memory        */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference.  For callers that manage device memory
 * and streams themselves (code under HIP graph capture, serving loops): a
 * scene image is one scene (1 to 16384 bodies of any kind) in the
 * SCENE engine's device layout behind a 512-byte header.  The caller copies
 * the host image to device memory it owns (256-byte aligned) and advances it
 * with bd_scene_image_update on its own stream: two kernel launches, no
 * allocation, no host copy, no synchronisation, so the call can be captured
 * in a HIP graph and replayed.  The circle-pair workspace is a second caller
 * buffer (16-byte aligned) of at least workspace_bytes; its contents do not
 * carry over between calls, so one workspace can serve any number of images
 * as long as calls that share it do not overlap.  Results are bit-identical
 * to bd_update on the same scene.  Copy the image back and pass it to
 * bd_scene_image_read to see positions through a solver holding the scene. */
typedef struct bd_scene_image_info {
    uint64_t image_bytes;      /* device bytes of the image */
    uint64_t workspace_bytes;  /* device bytes of workspace an update needs (0: none) */
    uint32_t n_bodies, n_circles;
    uint32_t workgroup;        /* launch shape for the scene: threads per workgroup */
    uint32_t lds_bytes;        /* and dynamic LDS bytes */
} bd_scene_image_info;

int32_t bd_scene_image_describe(const bd_solver *s, bd_scene_image_info *info);
/* Write the scene's image (its state now) into host memory of image_bytes. */
int32_t bd_scene_image_write(const bd_solver *s, void *host_image, uint64_t bytes);
/* Advance the image `steps` updates of dt on hip_stream (a hipStream_t; NULL:
 * the null stream).  Returns once the kernels are queued.  A workspace
 * smaller than the image needs is refused (BD_ERR_ARG), and, should the
 * image header disagree with info, the device leaves the scene untouched and
 * bd_scene_image_read reports BD_ERR_CAPACITY. */
int32_t bd_scene_image_update(const bd_scene_image_info *info, void *image, void *workspace,
                              uint64_t workspace_bytes, float dt, uint32_t steps, void *hip_stream);
/* Copy positions from a host copy of an image into `s`, a solver holding the
 * same scene (e.g. the one described).  BD_ERR_COLLISION if the scene's
 * circle-pair schedule failed. */
int32_t bd_scene_image_read(bd_solver *s, const void *host_image, uint64_t bytes);
/* bd_scene_image_update without a workspace argument.  No counterpart in the
 * reference: for callers that cannot allocate device memory at all (code
 * under graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * workspace is the library's static per-device array of 8 MiB, enough for
 * every scene an image holds (16,384 circles need 6.4 MiB); larger
 * workspace_bytes are refused with BD_ERR_CAPACITY.  It allocates nothing.
 * Every call on one device shares that device's array, so calls that overlap
 * on one device (concurrent streams or threads) race; calls in sequence on
 * one stream do not, since no call relies on what an earlier one left there. */
int32_t bd_scene_image_update_noalloc(const bd_scene_image_info *info, void *image, float dt, uint32_t steps,
                                      void *hip_stream);

/* ------------------------------------------------------------------ */
/* Scene generators (synthetic inputs of the benchmark workloads)       */
/* ------------------------------------------------------------------ */
/* Link emission order of a grid: COLOUR emits 8 conflict-free classes
 * (H even columns, H odd, V even rows, V odd, D even rows, D odd, AD even,
 * AD odd; row-major inside a class), so a sweep has 8 dependency layers;
 * PARTICLE emits H, V, D, AD per particle in row-major order (deep layers). */
#define BD_ORDER_COLOUR 0
#define BD_ORDER_PARTICLE 1

/* A grid of nx * ny bodies, body (r, c) at (x0 + c*spacing, y0 - r*spacing)
 * plus U(-jitter, jitter) per axis (seeded), index first + r*nx + c, with
 * structural (H, V: rest spacing) and shear (D, AD: rest sqrtf(2)*spacing)
 * links.  as_circles = 0: particles + particle links (pin_top pins row 0);
 * as_circles = 1: circles of the given radius + circle links.  Returns the
 * index of the first body. */
typedef struct bd_grid_desc {
    uint32_t nx, ny;
    float spacing, x0, y0, jitter;
    uint64_t seed;
    int32_t order;
    int32_t pin_top;
    int32_t as_circles;
    float radius;
} bd_grid_desc;
int64_t bd_scene_add_grid(bd_solver *s, const bd_grid_desc *g);

/* Free circles of one radius, uniform in [lo, hi]^2, with initial velocity
 * N(0, vsigma) per axis carried as prev = pos - v * dt_sub.  Returns the
 * index of the first circle. */
int64_t bd_scene_add_gas(bd_solver *s, uint64_t n, float radius, float lo, float hi, float vsigma,
                         float dt_sub, uint64_t seed);

/* The collision-heavy workload: n circles of radius r in a [0, box]^2 box
 * walled by four static polygons (and bounds), n_static static circles with
 * radius U(4, 16), gravity (0, -9.81), sub_steps 4.  (n = 262144, r = 0.5,
 * box = 1024, n_static = 64 is the full-size scene.) */
int32_t bd_scene_gas_box(bd_solver *s, uint64_t n, float radius, float box, uint32_t n_static, uint64_t seed);

/* One softbody scene of the batched workload, fully determined by seed:
 * side = round(4 * 32^u), u ~ U(0,1) (16 .. 16384 circles of radius 0.35,
 * spacing 1, structural + shear circle links in colour order), 1..8 static
 * colliders alternating circles (radius U(2,10)) and convex polygons (3..8
 * vertices on a jittered circle), bounds (0,0)+(W,W) with W = 3*side + 40,
 * gravity (0,-9.81), sub_steps 8, one link sweep.  max_side > 0 caps the
 * side.  Returns the number of circles. */
int64_t bd_scene_softbody(bd_solver *s, uint64_t seed, uint32_t max_side);

/* ------------------------------------------------------------------ */
/* Device information                                                  */
/* ------------------------------------------------------------------ */
/* Returns the number of visible HIP devices (0 if none) without creating a
 * context on them. */
int32_t bd_device_count(void);
/* Library build identification string. */
const char *bd_version(void);
/* Measurement helper, no counterpart in the reference: one plain device copy
 * of `bytes` (a multiple of 16) between two fresh buffers by a streaming
 * kernel, timed with HIP events (*ms).  Profiler runs read the HBM byte
 * counters of this kernel to scale those of the solver kernels, and bytes x 2
 * / time is the achievable copy bandwidth. */
int32_t bd_ext_copy_calibration(uint64_t bytes, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* BENDY2D_H */
