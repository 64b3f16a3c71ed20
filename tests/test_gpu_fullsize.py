"""Parity at the configurations' full sizes on the GPU: the 1024x1024 cloth
(TILED engine), the 262,144-circle gas box (GLOBAL engine) and the
4,096-scene batch.  Each is bit-identical to the f32 oracle over as many
updates as the oracle finishes in seconds; the batch also runs its whole
100-update window with every scene checked for a failed schedule and a
sample of scenes (the largest among them) checked bit for bit."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff, state, to_oracle
from bendy2d_amd import BD_ENGINE_GLOBAL, BD_ENGINE_TILED, Batch, Solver, workloads

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


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
