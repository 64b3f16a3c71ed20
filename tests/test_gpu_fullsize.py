"""Parity at the configurations' full sizes on the GPU: the 32x32 pinned
cloth's whole 1,000-update run, the 1024x1024 cloth (TILED engine), the
262,144-circle gas box (GLOBAL engine) over its 100-update error window, the
4,096-scene batch and a served scene of the largest size.  Each is
bit-identical to the f32 oracle: live where the oracle finishes in seconds,
through the committed state checksums of tests/golden/ where it does not."""
import json
import os

import numpy as np
import pytest

from helpers import ROOT, bits_equal, first_diff, state, to_oracle
from bendy2d_amd import BD_ENGINE_GLOBAL, BD_ENGINE_SCENE, BD_ENGINE_TILED, Batch, Server, Solver, workloads

import sys
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from statehash import block_hashes, state_hash  # noqa: E402

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.mark.parametrize("engine", [BD_ENGINE_SCENE, BD_ENGINE_GLOBAL])
def test_cloth_32_whole_run(engine):
    # 32x32 cloth, top row pinned, static circle R=8 20 below the centre,
    # g=(0,-9.81), dt=1/60, 8 substeps, 8 link sweeps, 1,000 updates from rest
    s = Solver()
    s.set_engine(engine)
    dt, steps = workloads.cloth_pinned(s)
    assert steps == 1000 and s.stats()["n_points"] == 1024
    o = to_oracle(s, 32)
    for k in range(steps):
        s.update(dt)
        assert o.update(dt) == 0
        if (k + 1) % 250 == 0:
            a, b = state(s), state(o)
            assert bits_equal(a, b), f"update {k + 1}: first diff {first_diff(a, b)}"
    assert s.engine_used == engine
    # the cloth reached the collider: some particle sits on the circle's rim
    p = s.get_particles()
    d = np.hypot(p[:, 0] - 15.5, p[:, 1] + 35.5)
    assert d.min() < 8.05


def test_cloth_1024_full_size():
    s = Solver()
    dt, _ = workloads.cloth_large(s, n=1024)
    o = to_oracle(s, 32)
    for _ in range(2):
        s.update(dt)
        assert o.update(dt) == 0
    assert s.engine_used == BD_ENGINE_TILED
    a, b = state(s), state(o)
    assert bits_equal(a, b), f"first diff {first_diff(a, b)}"


def test_gas_262144_full_size():
    s = Solver()
    dt, _ = workloads.gas_box(s)  # 262,144 circles r=0.5, 1024 box, 64 static circles, 4 walls
    o = to_oracle(s, 32, candidate_mode=True)
    for k in range(3):
        s.update(dt)
        assert o.update(dt) == 0
    assert s.engine_used == BD_ENGINE_GLOBAL
    a, b = state(s), state(o)
    assert bits_equal(a, b), f"first diff {first_diff(a, b)}"


def test_gas_262144_100_update_window():
    # the whole error window of the collision-heavy configuration at full size,
    # checked against the oracle's state checksums after updates 1..100
    # (tests/golden/make_gas_box_hashes.py wrote them)
    with open(os.path.join(ROOT, "tests", "golden", "gas_box_262144_hashes.json")) as f:
        gold = json.load(f)
    s = Solver()
    dt, _ = workloads.gas_box(s)
    last = max(int(k) for k in gold["checkpoints"])
    for k in range(1, last + 1):
        s.update(dt)
        want = gold["checkpoints"].get(str(k))
        if want is None:
            continue
        x = state(s)
        if state_hash(x) != want["sha256"]:
            bad = [i for i, (g, w) in enumerate(zip(block_hashes(x, gold["block_bodies"]), want["blocks"])) if g != w]
            raise AssertionError(f"update {k}: state differs from the oracle in body blocks {bad[:10]} "
                                 f"(of {len(want['blocks'])}, {gold['block_bodies']} bodies each)")
    assert s.engine_used == BD_ENGINE_GLOBAL
    st = s.stats()
    assert st["cc_rounds"] > 0


def test_served_scene_of_full_size():
    # a served request on a scene of the serving workload's largest size
    # (side 128: 16,384 circles) advancing 300 updates in steps of a launch
    # budget, beside smaller scenes, equals the oracle running it alone
    seed = next(x for x in range(1000, 10000) if workloads.softbody_side(x) == 128)
    big = Solver()
    workloads.softbody(big, seed)
    assert big.stats()["n_points"] == 16384
    o = to_oracle(big, 32, candidate_mode=True)
    srv = Server(memory_budget=2 << 30, steps_per_launch=32, launch_budget_us=20000)
    _, sid = srv.create(big, 300, DT)
    others = []
    for i in range(6):
        s = Solver()
        workloads.softbody(s, 300 + i, max_side=40)
        others.append(srv.create(s, 120, DT))
    for _ in range(300):   # the oracle runs while the server works
        assert o.update(DT) == 0
    assert srv.wait_idle(100)
    done = srv.poll()
    assert len(done) == 7 and all(d["status"] == 0 for d in done)
    srv.read_scene(sid, big)
    a, b = state(big), state(o)
    assert bits_equal(a, b), f"first diff {first_diff(a, b)}"
    srv.close()


def test_batch_4096_full_size_100_updates():
    scenes, sizes = [], []
    for i in range(4096):
        s = Solver()
        workloads.softbody(s, 1000 + i)
        scenes.append(s)
        sizes.append(s.stats()["n_points"])
    order = np.argsort(sizes)
    sample = sorted({0, 1, 2, int(order[len(order) // 2]), int(order[-1])})
    oracles = {i: to_oracle(scenes[i], 32, candidate_mode=True) for i in sample}
    b = Batch()
    for s in scenes:
        b.add_scene(s)
    b.update(DT, 100)
    b.sync()
    assert b.cc_stats()["err_scenes"] == 0
    for i in sample:
        for _ in range(100):
            assert oracles[i].update(DT) == 0
        b.read_scene(i, scenes[i])
        a, w = state(scenes[i]), state(oracles[i])
        assert bits_equal(a, w), f"scene {i} ({sizes[i]} bodies): first diff {first_diff(a, w)}"
    b.close()
