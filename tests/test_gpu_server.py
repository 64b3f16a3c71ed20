"""Serving on the GPU: each request's result is bit-identical to the f32
oracle running that scene alone, whatever else shares the launches, also when
the memory budget forces scenes out to host memory and back."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff, polygon_scene, state, state_acc, to_oracle
from bendy2d_amd import Server, Solver, workloads

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def scene(seed, max_side=14):
    s = Solver()
    workloads.softbody(s, seed, max_side=max_side)
    return s


def oracle_state(s, steps_dts):
    o = to_oracle(s, 32)
    for steps, dt in steps_dts:
        for _ in range(steps):
            assert o.update(dt) == 0
    return state(o)


@pytest.mark.parametrize("budget_us", [0, 1])
def test_mix_matches_oracle(budget_us):
    # budget_us=1: every scene stops after one update per launch, so requests
    # finish over several launches in uneven steps
    rng = np.random.default_rng(3)
    srv = Server(memory_budget=1 << 30, steps_per_launch=4, launch_budget_us=budget_us)
    scenes, ids, history = [], [], []
    for i in range(12):
        s = scene(200 + i)
        k = int(rng.integers(1, 12))
        _, sid = srv.create(s, k, DT)
        scenes.append(s); ids.append(sid); history.append([(k, DT)])
    # continues, some with another dt, interleaved with creates
    for j in range(20):
        i = int(rng.integers(0, len(ids)))
        k = int(rng.integers(1, 9))
        dt = DT if j % 3 else 1.0 / 120.0
        srv.advance(ids[i], k, dt)
        history[i].append((k, dt))
    assert srv.wait_idle(120)
    done = srv.poll()
    assert len(done) == 32 and all(d["status"] == 0 for d in done)
    assert all(d["t_done"] >= d["t_start"] >= d["t_submit"] for d in done)
    for i, s in enumerate(scenes):
        want = oracle_state(s, history[i])
        srv.read_scene(ids[i], s)
        got = state(s)
        assert bits_equal(got, want), f"scene {i}: {first_diff(got, want)}"
    st = srv.stats()
    assert st["requests_done"] == 32 and st["requests_pending"] == 0
    if budget_us:
        # one update per scene and launch: at least as many launches as the
        # longest scene history has updates
        assert st["launches"] >= max(sum(k for k, _ in h) for h in history)
    srv.close()


def test_polygon_scenes_served():
    # soft polygon scenes among plain ones, over several requests each, with a
    # one-update time budget so that they finish in uneven launches
    srv = Server(memory_budget=1 << 30, steps_per_launch=3, launch_budget_us=1)
    scenes = [polygon_scene(600 + i, n_polys=2 + i) for i in range(4)] + [scene(610), scene(611)]
    ids, history = [], []
    for i, s in enumerate(scenes):
        _, sid = srv.create(s, 4 + i, DT)
        ids.append(sid); history.append([(4 + i, DT)])
    for i in (0, 2, 4, 1):
        srv.advance(ids[i], 5, 1.0 / 120.0)
        history[i].append((5, 1.0 / 120.0))
    assert srv.wait_idle(120)
    done = srv.poll()
    assert len(done) == 10 and all(d["status"] == 0 for d in done)
    for i, s in enumerate(scenes):
        o = to_oracle(s, 32)
        for steps, dt in history[i]:
            for _ in range(steps):
                assert o.update(dt) == 0
        srv.read_scene(ids[i], s)
        got, want = state_acc(s), state_acc(o)
        assert bits_equal(got, want), f"scene {i}: {first_diff(got, want)}"
    srv.close()


def test_eviction_and_readmission():
    # a budget of about three scene chunks: scenes are evicted to host memory
    # between their requests and brought back, without changing results
    s0 = scene(300, max_side=8)
    one = Server(memory_budget=1 << 30)
    one.create(s0.clone(), 1, DT)
    assert one.wait_idle(60)
    chunk = one.stats()["bytes_used"]
    one.close()
    srv = Server(memory_budget=3 * chunk + 1024, steps_per_launch=2)
    scenes = [scene(300, max_side=8) for i in range(6)]  # same size: exactly three fit
    ids = []
    for s in scenes:
        ids.append(srv.create(s, 3, DT)[1])
        assert srv.wait_idle(60)
    for sid in ids:
        srv.advance(sid, 2, DT)
        assert srv.wait_idle(60)
    st = srv.stats()
    assert st["evictions"] > 0 and st["readmissions"] > 0
    for i, s in enumerate(scenes):
        want = oracle_state(s, [(3, DT), (2, DT)])
        srv.read_scene(ids[i], s)
        assert bits_equal(state(s), want), f"scene {i}"
    srv.close()


def test_alone_equals_in_mix():
    target = scene(777, max_side=24)
    alone = Server(memory_budget=1 << 30, steps_per_launch=3)
    sid = alone.create(target.clone(), 10, DT)[1]
    assert alone.wait_idle(60)
    a = target.clone()
    alone.read_scene(sid, a)
    alone.close()
    mix = Server(memory_budget=1 << 30, steps_per_launch=3)
    others = [scene(900 + i, max_side=32) for i in range(30)]
    for o in others[:15]:
        mix.create(o, 7, DT)
    sid2 = mix.create(target.clone(), 10, DT)[1]
    for o in others[15:]:
        mix.create(o, 13, DT)
    assert mix.wait_idle(60)
    b = target.clone()
    mix.read_scene(sid2, b)
    mix.close()
    assert bits_equal(state(a), state(b))


def test_heavy_scene_runs_beside_the_others(monkeypatch):
    # a 12,000-circle gas whose every sweep runs sequentially (tens of ms per
    # update, above a heavy threshold of 5 ms) among small scenes: after its
    # first launch it runs in launches of its own on the second stream (the
    # small scenes' requests then finish first when the two streams sit on
    # different hardware queues); every result equals the oracle
    heavy = Solver()
    workloads.gas_box(heavy, n=12000, radius=0.5, box=160.0, n_static=2, seed=21)
    heavy.set_sub_steps(1)
    srv = Server(memory_budget=1 << 30, steps_per_launch=4, launch_budget_us=1000, heavy_update_us=5000)
    monkeypatch.setenv("BD_SCENE_SEQUENTIAL_SWEEP", "1")
    r_heavy, sid_heavy = srv.create(heavy, 6, DT)
    monkeypatch.delenv("BD_SCENE_SEQUENTIAL_SWEEP")
    light = [scene(700 + i, max_side=10) for i in range(6)]
    ids = [srv.create(s, 24, DT) for s in light]
    assert srv.wait_idle(300)
    done = {d["request_id"]: d for d in srv.poll()}
    assert all(d["status"] == 0 for d in done.values()) and len(done) == 7
    assert srv.stats()["heavy_launches"] >= 4
    for s, (_, sid), k in [(heavy, (r_heavy, sid_heavy), 6)] + [(s, i, 24) for s, i in zip(light, ids)]:
        want = oracle_state(s, [(k, DT)])
        srv.read_scene(sid, s)
        got = state(s)
        assert bits_equal(got, want), f"{first_diff(got, want)}"
    srv.close()


def test_overfull_lists_run_sequentially():
    # 100 circles of radius 1 stacked within 0.05 of one point: every pair is a
    # candidate (99 per circle), more than the workspace's 32 per circle at
    # any skin, so each sweep of this scene runs the sequential fallback.  Its
    # requests must still equal the oracle, as must a scene in the same
    # launches.
    bad = Solver()
    bad.gravity = (0.0, 0.0)
    bad.bounds = ((-1000.0, -1000.0), (2000.0, 2000.0))
    rng = np.random.default_rng(7)
    for _ in range(100):
        bad.add_circle(rng.uniform(-0.05, 0.05, 2).astype(np.float32), 1.0)
    good = scene(31)
    srv = Server(memory_budget=1 << 30, steps_per_launch=2)
    r_bad, sid_bad = srv.create(bad, 3, DT)
    r_good, sid_good = srv.create(good, 5, DT)
    assert srv.wait_idle(60)
    r_bad2 = srv.advance(sid_bad, 2, DT)
    assert srv.wait_idle(60)
    done = {d["request_id"]: d for d in srv.poll()}
    assert all(done[r]["status"] == 0 for r in (r_bad, r_bad2, r_good))
    for s, sid, steps in ((bad, sid_bad, 5), (good, sid_good, 5)):
        want = oracle_state(s, [(steps, DT)])
        srv.read_scene(sid, s)
        got = state(s)
        assert bits_equal(got, want), f"{first_diff(got, want)}"
    srv.close()
