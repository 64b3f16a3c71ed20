#!/usr/bin/env python3
"""Benchmark: particle-updates/s of the batched form vs the plain per-scene
path on one MI355X, plus the 1024x1024 cloth (plain path, HBM roofline) and,
when built, the serving loop's p50/p99 request latency.

One JSON line on stdout.  `value` is the batched form's aggregate
particle-updates/s (PU/s) on the 4096-scene softbody workload, inputs resident
in HBM; a step is one update of every scene (8 substeps each).  A PU is one
body advanced through one substep.

Algorithmic bytes per substep of a scene (HBM roofline, not MFMA: nothing is a
dense contraction): 32 B per body (pos + prev read and written), 44 B per link
solve (12 B record + 2 x 8 B reads + 2 x 8 B writes), plus 32 B per body when
circle-pair collision is on.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
DT = 1.0 / 60.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_softbody_scenes(n, seed0=1000, max_side=0):
    from bendy2d_amd import Solver, workloads
    scenes, pts, links = [], 0, 0
    for i in range(n):
        s = Solver()
        workloads.softbody(s, seed0 + i, max_side=max_side)
        st = s.stats()
        pts += st["n_points"]
        links += st["n_links"]
        scenes.append(s)
    return scenes, pts, links


def bench_batched(scenes, pts, links, steps, warmup, tune=True, launch=None):
    """Times `steps` updates of the whole batch (rounded up to whole launches of
    steps_per_launch updates) after `warmup` untimed ones."""
    from bendy2d_amd import Batch
    b = Batch()
    t0 = time.time()
    for s in scenes:
        b.add_scene(s)
    b.update(DT, 0)  # upload
    b.sync()
    log(f"batch: {len(scenes)} scenes, {pts} bodies, {links} links, upload {time.time() - t0:.1f}s")
    if launch:
        b.config = launch
    elif tune:
        t0 = time.time()
        tried = b.autotune(DT, 2)
        log(f"autotune: {tried} configs in {time.time() - t0:.1f}s -> {b.config}")
    spl = max(1, int(b.config["steps_per_launch"]))
    steps = -(-steps // spl) * spl
    b.update(DT, warmup)
    b.sync()
    t0 = time.perf_counter()
    b.update(DT, steps)
    b.sync()
    wall = time.perf_counter() - t0
    dev_ms, launches = b.last_timing()
    replays = steps // spl
    cc = b.cc_stats()
    substeps = 8
    pu = pts * substeps * steps
    # bytes per substep: 64 B per body (integrate + collision) + 44 B per link
    alg_bytes = (64.0 * pts + 44.0 * links) * substeps * steps
    return {
        "pu_per_s": pu / wall, "wall_s": wall, "device_ms": dev_ms, "kernel_launches": launches,
        "replays": replays, "steps": steps, "alg_bytes": alg_bytes, "pu": pu, "config": b.config, "cc": cc,
        "batch": b,
    }


def bench_plain_per_scene(scenes, pts, updates=1):
    """The drop-in bd_update called on each scene in turn (AUTO engine)."""
    for s in scenes:          # warm: plan + upload + first launch
        s.update(DT)
    for s in scenes:
        s.sync()
    t0 = time.perf_counter()
    for _ in range(updates):
        for s in scenes:
            s.update(DT)
    for s in scenes:
        s.sync()
    wall = time.perf_counter() - t0
    return {"pu_per_s": pts * 8 * updates / wall, "wall_s": wall}


def bench_cloth(n, steps, warmup):
    from bendy2d_amd import Solver, workloads
    s = Solver()
    dt, _ = workloads.cloth_large(s, n=n)
    s.update(dt)
    for _ in range(warmup):
        s.update(dt)
    s.sync()
    st = s.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.update(dt)
    s.sync()
    wall = time.perf_counter() - t0
    N, L = st["n_points"], st["n_links"]
    sub, iters = 4, 8
    pu = N * sub * steps
    alg = (32.0 * N + 44.0 * L * iters) * sub * steps
    # measured HBM traffic: one k_tiled_pass launch per link sweep, each
    # moving the profiled bytes (tiles with their halos, link records)
    per_pass = profiled_json("k_tiled_pass_cloth1024.md", "traffic_bytes_per_pass") if n == 1024 else None
    traffic = per_pass * iters * sub * steps if per_pass else None
    return {"pu_per_s": pu / wall, "ms_per_step": 1e3 * wall / steps,
            "achieved_GBps": traffic / wall / 1e9 if traffic else None,
            "frac": traffic / wall / HBM_PEAK if traffic else None,
            "traffic_per_pass": per_pass,
            "traffic_source": "profiles/k_tiled_pass_cloth1024.md: rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                              "k_tiled_pass launch, scaled by a 1 GiB copy kernel",
            "alg_GBps": alg / wall / 1e9, "algorithmic_frac": alg / wall / HBM_PEAK, "engine": s.engine_used,
            "bodies": N, "links": L, "bytes_per_pu": alg / pu}


def bench_gas(steps, warmup):
    """262,144 free circles (r 0.5) in a 1024 box with 64 static circles and 4
    wall polygons, circle-pair collision on, 4 substeps: the plain path."""
    from bendy2d_amd import Solver, workloads
    s = Solver()
    dt, _ = workloads.gas_box(s)
    s.update(dt)
    for _ in range(warmup):
        s.update(dt)
    s.sync()
    st0 = s.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.update(dt)
    s.sync()
    wall = time.perf_counter() - t0
    st = s.stats()
    N, sub = st["n_points"], 4
    pu = N * sub * steps
    alg = 64.0 * N * sub * steps  # 32 B integrate + 32 B binning and resolve per body and substep
    return {"pu_per_s": pu / wall, "ms_per_step": 1e3 * wall / steps, "alg_GBps": alg / wall / 1e9,
            "frac": alg / wall / HBM_PEAK, "engine": s.engine_used, "bodies": N, "updates": steps,
            "after_updates": 1 + warmup, "kernel_launches_per_update": (st["kernel_launches"] - st0["kernel_launches"]) / steps,
            "cc_rounds_per_update": (st["cc_rounds"] - st0["cc_rounds"]) / steps,
            "cc_redos": st["cc_redos"] - st0["cc_redos"], "cc_pairs": st["cc_pairs"], "cc_margin": st["cc_margin"]}


def bench_cloth_small(steps=1000):
    """The 32x32 pinned cloth with a static circle, 8 substeps and 8 link
    sweeps, its whole 1,000-update run from rest on the plain path."""
    from bendy2d_amd import Solver, workloads
    s = Solver()
    dt, _ = workloads.cloth_pinned(s)
    c = s.clone()
    c.update(dt)  # plan and upload outside the timed run
    c.sync()
    c.close()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.update(dt)
    s.sync()
    wall = time.perf_counter() - t0
    N = s.stats()["n_points"]
    return {"pu_per_s": N * 8 * steps / wall, "ms_per_step": 1e3 * wall / steps, "run_s": wall,
            "engine": s.engine_used, "bodies": N, "updates": steps}


def bench_readback():
    """Frame time of update + reading every particle back (the reference's
    callers read positions each frame, src/solver.rs:86-104), with the
    readback path on (the copy queued behind each update on a second stream)
    and off (a synchronous copy when read): the 32x32 pinned cloth (100
    frames) and the 1024x1024 cloth (10 frames)."""
    from bendy2d_amd import Solver, workloads
    out = {}
    for name, build, frames in (("cloth_32", lambda s: workloads.cloth_pinned(s), 100),
                                ("cloth_1024", lambda s: workloads.cloth_large(s), 10)):
        row = {}
        for mode in ("off", "on"):
            s = Solver()
            dt, _ = build(s)
            s.set_readback(mode == "on")
            s.update(dt)
            s.get_particles()
            t0 = time.perf_counter()
            for _ in range(frames):
                s.update(dt)
                s.get_particles()
            row[f"ms_per_frame_readback_{mode}"] = (time.perf_counter() - t0) * 1e3 / frames
            s.close()
        s = Solver()
        dt, _ = build(s)
        s.update(dt)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(frames):
            s.update(dt)
        s.sync()
        row["ms_per_frame_no_read"] = (time.perf_counter() - t0) * 1e3 / frames
        s.close()
        out[name] = row
    return out


def bench_serving(seconds, rate, launch_cfg, steps_per_launch=64, budget=64 << 30, seed=4, drain_limit=30.0,
                  launch_budget_us=20000):
    """Poisson stream of requests: half create a new softbody scene (generated
    as in the batched workload, seeded by request number), half continue a
    uniformly chosen known scene; each advances its scene K ~ U{10..600}
    updates.  Arrivals stop after `seconds`; the server then drains for at
    most `drain_limit` seconds.  Latency percentiles cover every submitted
    request: one still pending at the end counts with its age then, a lower
    bound on its latency (reported as censored)."""
    import numpy as np
    from bendy2d_amd import Server, Solver, workloads
    rng = np.random.default_rng(seed)
    n = int(rate * seconds * 1.2) + 16
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals = arrivals[arrivals < seconds]
    kinds = rng.random(len(arrivals)) < 0.5       # True: create
    ks = rng.integers(10, 601, len(arrivals))
    picks = rng.random(len(arrivals))
    srv = Server(memory_budget=budget, steps_per_launch=steps_per_launch, launch_budget_us=launch_budget_us,
                 launch={k: launch_cfg[k] for k in ("wg_small", "wg_large", "split_points", "lds_small", "lds_large",
                                                    "split_tiny", "wg_tiny", "lds_tiny")})
    known, lag, submitted, done = [], [], {}, []
    t0 = time.perf_counter()
    last_log = 0.0
    for i, ta in enumerate(arrivals):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        lag.append(time.perf_counter() - t0 - ta)
        if kinds[i] or not known:
            s = Solver()
            workloads.softbody(s, 500000 + i)
            rid, sid = srv.create(s, int(ks[i]), DT)
            known.append(sid)
            s.close()
        else:
            rid = srv.advance(known[int(picks[i] * len(known))], int(ks[i]), DT)
        submitted[rid] = srv.now()
        if ta - last_log >= 5.0:
            done += srv.poll()
            st = srv.stats()
            log(f"serving t={ta:.0f}s submitted={i + 1} done={len(done)} pending={st['requests_pending']} "
                f"resident={st['resident_scenes']} GB={st['bytes_used'] / 2**30:.1f} busy={st['busy_s']:.1f}s")
            last_log = ta
    t_arrive_end = time.perf_counter() - t0
    while True:
        idle = srv.wait_idle(5.0)
        done += srv.poll()
        el = time.perf_counter() - t0
        st = srv.stats()
        log(f"serving drain t={el:.0f}s done={len(done)} pending={st['requests_pending']} busy={st['busy_s']:.1f}s")
        if idle or el - t_arrive_end >= drain_limit:
            break
    t_end = time.perf_counter() - t0
    t_end_srv = srv.now()
    st = srv.stats()
    srv.close()
    if not submitted:
        return None
    finished = {d["request_id"]: d for d in done}
    lat = np.array([finished[r]["t_done"] - finished[r]["t_submit"] if r in finished else t_end_srv - ts
                    for r, ts in submitted.items()])
    censored = len(submitted) - len(finished)
    pu = sum(d["particle_updates"] for d in done)

    def pct(q):
        v = float(np.percentile(lat, q))
        # a percentile at or above the smallest censored age is only a lower bound
        lower = censored > 0 and v >= min(t_end_srv - submitted[r] for r in submitted if r not in finished)
        return v * 1e3, lower
    p50, p50_lb = pct(50)
    p99, p99_lb = pct(99)
    return {
        "requests": len(submitted), "completed": len(finished), "censored": censored,
        "failed": sum(1 for d in done if d["status"] != 0),
        "rate_req_per_s": rate, "arrival_seconds": seconds, "drain_seconds": t_end - t_arrive_end,
        "p50_ms": p50, "p99_ms": p99, "p50_is_lower_bound": p50_lb, "p99_is_lower_bound": p99_lb,
        "pu_per_s": pu / t_end, "pu_per_s_device_busy": st["particle_updates"] / max(st["busy_s"], 1e-9),
        "device_busy_frac": st["busy_s"] / t_end, "launches": st["launches"],
        "resident_scenes": st["resident_scenes"], "bytes_used": st["bytes_used"], "evictions": st["evictions"],
        "driver_lag_ms_p99": float(np.percentile(np.array(lag), 99) * 1e3), "steps_per_launch": steps_per_launch,
        "launch_budget_us": launch_budget_us,
    }


def copy_calibration(nbytes=1 << 30):
    """One plain device copy of nbytes by the library's streaming copy kernel
    (k_copy_f4), for scaling the HBM byte counters of a profiler run (its
    FETCH_SIZE and WRITE_SIZE should each read nbytes) and for the achievable
    copy bandwidth (read + write bytes over the kernel's time)."""
    import ctypes
    from bendy2d_amd import lib
    ms = ctypes.c_float()
    rc = lib().bd_ext_copy_calibration(nbytes, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"bd_ext_copy_calibration failed: {rc}")
    return {"bytes": nbytes, "ms": ms.value, "copy_GBps": 2.0 * nbytes / (ms.value * 1e-3) / 1e9}


def _oracle_imports():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from helpers import to_oracle
    return to_oracle


def cpu_leg(seconds_scenes=12.0):
    """The CPU baselines, run in a child process pinned to one core before it
    loads anything (it never touches the GPU): the f32 oracle (gcc -O2
    -ffp-contract=off, one thread), a C restatement of the reference's solver,
    timed with process CPU time on bounded samples of each workload."""
    to_oracle = _oracle_imports()
    from bendy2d_amd import Solver, workloads
    out = {}
    label = "f32 oracle (C, gcc -O2 -ffp-contract=off), 1 thread"

    # 4,096-scene workload: the reference's all-pairs circle loop, and the
    # candidate-list form of it (same results bit for bit), one update each
    for mode, secs in (("all_pairs", seconds_scenes), ("candidate_list", 0.5 * seconds_scenes)):
        pu, busy, n = 0, 0.0, 0
        while busy < secs and n < 4096:
            s = Solver()
            workloads.softbody(s, 1000 + n)
            o = to_oracle(s, 32, candidate_mode=(mode == "candidate_list"))
            npts = s.stats()["n_points"]
            t0 = time.process_time()
            o.update(DT)
            busy += time.process_time() - t0
            pu += npts * 8
            n += 1
            s.close()
        loop = ("the reference's all-pairs circle loop (src/solver.rs:168-177)" if mode == "all_pairs"
                else "candidate lists in place of the all-pairs loop (bit-identical results)")
        out[f"softbody4096_{mode}"] = {
            "value": pu / busy, "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "sample": f"{label}, {loop}: scenes 0..{n - 1} of the 4096-scene workload, update 1 (8 substeps) "
                      f"each, {busy:.1f} s CPU"}

    # 32x32 pinned cloth: the whole 1,000-update run
    s = Solver()
    dt, steps = workloads.cloth_pinned(s)
    o = to_oracle(s, 32)
    t0 = time.process_time()
    for _ in range(steps):
        o.update(dt)
    t = time.process_time() - t0
    out["cloth_32"] = {"value": 1024 * 8 * steps / t, "unit": "particle-updates/s", "cores": 1, "kind": "port",
                       "sample": f"{label}: the whole {steps}-update run, {t:.2f} s CPU"}
    s.close()

    # 1024x1024 cloth: the first updates from rest until 20 s, extrapolated
    s = Solver()
    dt, _ = workloads.cloth_large(s)
    o = to_oracle(s, 32)
    n, t = 0, 0.0
    while t < 20.0 and n < 25:
        t0 = time.process_time()
        o.update(dt)
        t += time.process_time() - t0
        n += 1
    N = s.stats()["n_points"]
    out["cloth_1024"] = {"value": N * 4 * n / t, "unit": "particle-updates/s", "cores": 1, "kind": "port",
                         "sample": f"{label}: updates 1..{n} ({4 * n} substeps) from rest, {t:.1f} s CPU; "
                                   "the 500-update run extrapolated at this rate"}
    s.close()

    # 262,144-circle gas box: update 1 with candidate lists, and a slice of the
    # reference's all-pairs loop timed and extrapolated to its C(C-1)/2 tests
    s = Solver()
    dt, _ = workloads.gas_box(s)
    N = s.stats()["n_points"]
    o = to_oracle(s, 32, candidate_mode=True)
    t0 = time.process_time()
    o.update(dt)
    t = time.process_time() - t0
    out["gas_262144_candidate_list"] = {
        "value": N * 4 / t, "unit": "particle-updates/s", "cores": 1, "kind": "port",
        "sample": f"{label}, candidate lists in place of the all-pairs loop (bit-identical results): update 1 "
                  f"(4 substeps), {t:.2f} s CPU"}
    o2 = to_oracle(s, 32)
    rows, tests, t = 0, 0, 0.0
    while t < 5.0 and rows < N:
        t0 = time.process_time()
        tests += o2.all_pairs_rows(rows, rows + 256)
        t += time.process_time() - t0
        rows += 256
    rows = min(rows, N)
    per_substep = (N * (N - 1) / 2) / (tests / t)
    out["gas_262144_all_pairs"] = {
        "value": N / per_substep, "unit": "particle-updates/s", "cores": 1, "kind": "port",
        "sample": f"{label}, the reference's all-pairs circle loop (src/solver.rs:168-177): rows 0..{rows - 1} of "
                  f"the first substep's loop ({tests} pair tests) in {t:.2f} s CPU, extrapolated to the "
                  f"{N * (N - 1) // 2} tests of one substep ({per_substep:.0f} s per substep)"}
    s.close()
    return out


def host_cpu_info(core):
    info = {"core": core, "host_cpus": os.cpu_count(), "cpu_model": "unknown", "mhz": None, "max_mhz": None}
    try:
        with open("/proc/cpuinfo") as f:
            cur = None
            for line in f:
                if line.startswith("processor"):
                    cur = int(line.split(":")[1])
                elif line.startswith("model name") and info["cpu_model"] == "unknown":
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                elif line.startswith("cpu MHz") and cur == core:
                    info["mhz"] = float(line.split(":")[1])
    except OSError:
        pass
    for key, name in (("mhz_scaling", "scaling_cur_freq"), ("max_mhz", "cpuinfo_max_freq")):
        try:
            with open(f"/sys/devices/system/cpu/cpu{core}/cpufreq/{name}") as f:
                info[key] = int(f.read()) / 1000.0
        except (OSError, ValueError):
            pass
    return info


def cpu_baselines():
    """Runs cpu_leg() in a child process pinned to one core (the last of this
    process's affinity set) before it imports anything; returns its results
    with the host's CPU model, core and clock."""
    import subprocess
    core = sorted(os.sched_getaffinity(0))[-1]
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", str(core)],
                       capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        log(f"cpu leg failed ({r.returncode}): {r.stderr[-2000:]}")
        return None
    data = json.loads(r.stdout.strip().splitlines()[-1])
    host = host_cpu_info(core)
    for v in data.values():
        v["pinned_core"] = core
        v["cpu_model"] = host["cpu_model"]
        v["mhz"] = host.get("mhz_scaling") or host.get("mhz")
        v["max_mhz"] = host.get("max_mhz")
        v["host_cpus"] = host["host_cpus"]
    return data


def profiled_json(name, key):
    """A figure from the JSON block at the end of a committed rocprofv3
    summary under profiles/ (FETCH_SIZE / WRITE_SIZE scaled by the copy
    kernel); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        text = open(path).read()
        block = text.rsplit("```json", 1)[1].split("```", 1)[0]
        return float(json.loads(block)[key])
    except (OSError, IndexError, KeyError, ValueError):
        return None


def profiled_traffic_per_update():
    """HBM bytes per update of the 4,096-scene batch."""
    return profiled_json("k_scenes_batched_softbody4096.md", "traffic_bytes_per_update")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed updates of the batch (rounded up to whole launches)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scenes", type=int, default=4096)
    ap.add_argument("--cloth", type=int, default=1024)
    ap.add_argument("--gas-steps", type=int, default=30,
                    help="timed updates of the 262,144-circle gas box, plain path (0 = skip it and the 32x32 cloth)")
    ap.add_argument("--plain-sample", type=int, default=0, help="scenes for the plain path (0 = all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="arrival window of the serving run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-drain", type=float, default=30.0, help="longest wait for the serving backlog")
    ap.add_argument("--serve-rate-sustained", type=float, default=100.0,
                    help="request rate of a second serving run that the device keeps up with (0 = skip)")
    ap.add_argument("--serve-seconds-sustained", type=float, default=30.0)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--dense-at", type=int, default=300,
                    help="also time 10 updates of the batch from this update on (0 = skip)")
    ap.add_argument("--launch", type=str, default="", help="batch launch config as JSON (skips the autotuner)")
    ap.add_argument("--no-window100", dest="window100", action="store_false",
                    help="skip timing updates 1-100 of a second copy of the batch")
    ap.add_argument("--only", choices=("all", "batched", "cloth", "copy", "gas", "cloth32"), default="all",
                    help="run one part alone (for profiler runs); only 'all' prints the JSON line")
    ap.add_argument("--cpu-leg", type=int, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg is not None:
        # the CPU baselines' child process: pinned to one core before it loads
        # anything, it never touches the GPU
        os.sched_setaffinity(0, {args.cpu_leg})
        print(json.dumps(cpu_leg()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box of this run has one MI355X; --gpus {args.gpus} is not run. "
              "Scenes are independent, so N GPUs would each take 1/N of them (no exchange).", file=sys.stderr)
        sys.exit(2)
    from bendy2d_amd import lib
    if lib().bd_device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    launch = json.loads(args.launch) if args.launch else None

    if args.only == "copy":
        log(f"copy calibration: {copy_calibration()}")
        return
    if args.only == "cloth":
        log(f"cloth {args.cloth}^2: {bench_cloth(args.cloth, args.steps, args.warmup)}")
        return
    if args.only == "gas":
        log(f"gas 262144: {bench_gas(args.steps, args.warmup)}")
        return
    if args.only == "cloth32":
        log(f"cloth 32^2: {bench_cloth_small()}")
        return

    t0 = time.time()
    scenes, pts, links = build_softbody_scenes(args.scenes)
    log(f"built {args.scenes} scenes in {time.time() - t0:.1f}s")
    bat = bench_batched(scenes, pts, links, args.steps, args.warmup, tune=not args.no_tune, launch=launch)
    log(f"batched: {bat['pu_per_s']:.3e} PU/s, {bat['wall_s'] * 1e3 / bat['steps']:.3f} ms/step, "
        f"device {bat['device_ms']:.1f} ms over {bat['replays']} launches, cc {bat['cc']}")
    copy = copy_calibration()
    log(f"copy calibration: {copy}")
    if args.only == "batched":
        return
    b = bat["batch"]
    done = args.warmup + bat["steps"]

    # the plain path (bd_update per scene, the drop-in) on a sample of the
    # scenes in the state the batch has reached (every k-th scene)
    k_sample = 1 if not args.plain_sample else max(1, len(scenes) // args.plain_sample)
    sample_ids = list(range(0, len(scenes), k_sample))

    def plain_at_batch_state():
        from bendy2d_amd import Solver
        b.sync()
        sample = []
        for i in sample_ids:
            c = scenes[i].clone()
            b.read_scene(i, c)
            sample.append(c)
        spts = sum(c.stats()["n_points"] for c in sample)
        r = bench_plain_per_scene(sample, spts)
        for c in sample:
            c.close()
        return r, spts

    plain, _ = plain_at_batch_state()
    plain_update = done + 2
    log(f"plain per-scene (update {plain_update}): {plain['pu_per_s']:.3e} PU/s over {len(sample_ids)} scenes")

    # the error window: updates 1..100 of the batch, timed whole on a copy of
    # the scenes from rest (the batch itself continues from where it is)
    window100 = None
    if args.window100:
        from bendy2d_amd import Batch
        w = Batch()
        for s in scenes:
            w.add_scene(s)
        w.config = bat["config"]
        w.update(DT, 0)
        w.sync()
        t0 = time.perf_counter()
        w.update(DT, 100)
        w.sync()
        el = time.perf_counter() - t0
        window100 = {"updates": "1-100", "pu_per_s": pts * 8 * 100 / el, "ms_per_update": el * 1e3 / 100,
                     "cc": w.cc_stats()}
        w.close()
        log(f"window 1-100: {window100}")

    # the same batch once the large softbodies have collapsed into piles in
    # dense contact, with the plain path on the same states
    dense = None
    if args.dense_at > done:
        b.update(DT, args.dense_at - done)
        b.sync()
        t0 = time.perf_counter()
        b.update(DT, 10)
        b.sync()
        el = time.perf_counter() - t0
        done = args.dense_at + 10
        dense = {"updates": f"{args.dense_at + 1}-{args.dense_at + 10}", "pu_per_s": pts * 8 * 10 / el,
                 "ms_per_update": el * 1e3 / 10, "cc": b.cc_stats()}
        dplain, _ = plain_at_batch_state()
        dense["plain_per_scene"] = {"value": dplain["pu_per_s"], "scenes": len(sample_ids),
                                    "update": args.dense_at + 12}
        dense["batched_vs_plain"] = dense["pu_per_s"] / dplain["pu_per_s"]
        log(f"dense regime: {dense}")

    cloth = bench_cloth(args.cloth, 10, 2) if args.cloth else None
    if cloth:
        log(f"cloth {args.cloth}^2: {cloth}")
    gas = bench_gas(args.gas_steps, 1) if args.gas_steps else None
    if gas:
        log(f"gas 262144: {gas}")
    cloth32 = bench_cloth_small() if args.gas_steps else None
    if cloth32:
        log(f"cloth 32^2: {cloth32}")
    readback = bench_readback() if args.gas_steps else None
    if readback:
        log(f"readback: {readback}")

    serving = None
    if args.serve_seconds > 0 or args.serve_rate_sustained > 0:
        # free the batched workload's device memory first
        b.close()
        for s in scenes:
            s.close()
    if args.serve_seconds > 0:
        serving = bench_serving(args.serve_seconds, args.serve_rate, bat["config"], drain_limit=args.serve_drain)
        log(f"serving: {serving}")
    # the same serving loop at an offered load the device sustains: the
    # specified 2,000 requests/s offer more particle-updates than one GPU
    # runs, so its queue grows and its percentiles are lower bounds
    serving_sustained = None
    if args.serve_seconds_sustained > 0 and args.serve_rate_sustained > 0:
        serving_sustained = bench_serving(args.serve_seconds_sustained, args.serve_rate_sustained, bat["config"],
                                          seed=5, drain_limit=45.0)
        log(f"serving (sustained load): {serving_sustained}")

    cpus = None if args.no_cpu else cpu_baselines()
    cpu = None
    if cpus:
        log(f"cpu: {json.dumps(cpus)}")
        cpu = dict(cpus["softbody4096_all_pairs"])
        cpu["candidate_list"] = cpus["softbody4096_candidate_list"]
        if cloth:
            cloth["cpu_baseline"] = cpus["cloth_1024"]
        if gas:
            gas["cpu_baseline"] = dict(cpus["gas_262144_all_pairs"])
            gas["cpu_baseline"]["candidate_list"] = cpus["gas_262144_candidate_list"]
        if cloth32:
            cloth32["cpu_baseline"] = cpus["cloth_32"]

    # dominant kernel: one launch (graph replay) of k_scenes, its size classes
    # running concurrently on their streams, timed by HIP events on the
    # batch's stream around all timed launches.  `achieved` is the HBM traffic
    # rocprofv3 measured per launch over the same window (FETCH_SIZE +
    # WRITE_SIZE, scaled by a copy kernel); the algorithmic model's bytes,
    # which charge HBM for state that stays in LDS, are reported beside it.
    per_launch_s = bat["device_ms"] / 1e3 / max(bat["replays"], 1)
    alg_per_launch = bat["alg_bytes"] / max(bat["replays"], 1)
    spl = int(bat["config"]["steps_per_launch"])
    tpu = profiled_traffic_per_update() if args.scenes == 4096 else None
    traffic = tpu * spl if tpu else None
    achieved = traffic / per_launch_s if traffic else None
    out = {
        "metric": "particle-updates/s batched vs plain (1 GPU) + p50/p99 request latency",
        "value": bat["pu_per_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": bat["steps"],
        "warmup": args.warmup,
        "ms_per_step": bat["wall_s"] * 1e3 / bat["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.scenes} independent softbody scenes (grid of circles r=0.35, side log-uniform "
                        "4..128, structural+shear circle links, 1-8 static circle/convex-polygon colliders, "
                        "circle-pair collision on), 8 substeps and 1 link sweep per update, batched persistent "
                        f"launch replayed as a hipGraph; timed: updates {args.warmup + 1}-{args.warmup + bat['steps']} "
                        "from rest (the softbodies falling, the first contacts from about update 90); updates 1-100 "
                        "and the dense-contact regime (updates 301-310) are in window_1_100 and dense_regime",
            "scenes": args.scenes, "bodies": pts, "links": links, "substeps": 8,
            "launch": {k: v for k, v in bat["config"].items()},
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved / 1e9 if achieved else None, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK if achieved else None, "traffic": traffic,
            "traffic_source": "profiles/k_scenes_batched_softbody4096.md: rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                              "update over the same window, scaled by a 1 GiB copy kernel, times this run's "
                              "updates per launch",
            "algorithmic_bytes_per_launch": alg_per_launch,
            "algorithmic_GBps": alg_per_launch / per_launch_s / 1e9,
            "algorithmic_frac": alg_per_launch / per_launch_s / HBM_PEAK,
            "measured_copy_GBps": copy["copy_GBps"],
            "kernel": f"k_scenes, one launch = its size classes concurrently, {spl} updates, "
                      f"{per_launch_s * 1e3:.3f} ms per launch",
        },
        "cpu_baseline": cpu,
        "plain_per_scene": {"value": plain["pu_per_s"], "scenes": len(sample_ids), "update": plain_update,
                            "path": "bd_update per scene (drop-in), AUTO engine, on the batch's state after the "
                                    "timed window"} if plain else None,
        "batched_vs_plain": bat["pu_per_s"] / plain["pu_per_s"] if plain else None,
        "window_1_100": window100,
        "dense_regime": dense,
        "cloth_1024": cloth,
        "gas_262144": gas,
        "cloth_32": cloth32,
        "readback": readback,
        "serving": serving,
        "serving_sustained_load": serving_sustained,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
