"""The sequential fallback of the circle-pair sweep on the GPU.

A sweep whose candidate lists keep failing runs the reference's (i, j) loop
row by row from a grid of the circles' current positions instead
(src/solver.rs:168-177).  BD_SCENE_SEQUENTIAL_SWEEP=1 makes every sweep of
a scene take that path, so these tests hold it to the oracle bit for bit on
scenes that exercise it: gases, dense piles where circles are pushed far
within one sweep, coincident circles (NaN), radii that differ, circles
linked to each other, soft polygons beside them."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff, polygon_scene, state, state_acc, to_oracle
from bendy2d_amd import BD_ENGINE_SCENE, Batch, Solver, workloads

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.fixture
def sequential(monkeypatch):
    monkeypatch.setenv("BD_SCENE_SEQUENTIAL_SWEEP", "1")


def run(s, steps, dt=DT, check_every=None):
    s.set_engine(BD_ENGINE_SCENE)
    o = to_oracle(s, 32)
    for k in range(steps):
        s.update(dt)
        assert o.update(dt) == 0
        if check_every and (k + 1) % check_every == 0:
            a, b = state_acc(s), state_acc(o)
            assert bits_equal(a, b), f"update {k + 1}: {first_diff(a, b)}"
    a, b = state_acc(s), state_acc(o)
    assert bits_equal(a, b), f"final: {first_diff(a, b)}"
    assert s.engine_used == BD_ENGINE_SCENE


def test_gas(sequential):
    s = Solver()
    workloads.gas_box(s, n=600, radius=0.5, box=64.0, n_static=3, seed=11)
    run(s, 20, check_every=10)


def test_dense_nan(sequential):
    # bounds clamp circles onto one corner point: coincident circles give NaN
    s = Solver()
    workloads.gas_box(s, n=600, radius=0.5, box=40.0, n_static=3, seed=11)
    run(s, 12, check_every=4)


def test_pile_with_mixed_radii_and_links(sequential):
    # circles of radius 0.2 to 0.9 dropped in a narrow box: the pile pushes
    # circles several radii within one sweep, so rows gather again and
    # circles change cells; a chain of circle links runs through it
    rng = np.random.default_rng(7)
    s = Solver()
    for k in range(400):
        s.add_circle((float(rng.uniform(1.0, 11.0)), float(rng.uniform(1.0, 60.0))), float(rng.uniform(0.2, 0.9)))
    for k in range(0, 60, 3):
        s.add_circle_link(k, k + 1, 1.2)
    s.gravity = (0.0, -40.0)
    s.bounds = ((0.0, 0.0), (12.0, 80.0))
    s.set_sub_steps(4)
    run(s, 60, check_every=20)


def test_more_candidates_than_ranked(sequential):
    # 300 circles of radius 1 within 0.05 of one point: a row gathers 299
    # candidates, more than the 256 the wavefront ranks, so one lane finds
    # them in order by walking the buckets
    rng = np.random.default_rng(17)
    s = Solver()
    for _ in range(300):
        s.add_circle(rng.uniform(-0.05, 0.05, 2).astype(np.float32), 1.0)
    s.gravity = (0.0, 0.0)
    s.bounds = ((-1000.0, -1000.0), (2000.0, 2000.0))
    run(s, 3)


def test_polygons_beside_circles(sequential):
    run(polygon_scene(41, n_polys=4), 30, check_every=15)


def test_batch_all_sequential(sequential):
    # every scene of a batch on the sequential path, against the oracle
    scenes = []
    for seed in (900, 901, 902):
        s = Solver()
        workloads.softbody(s, seed, max_side=24)
        scenes.append(s)
    s = Solver()
    workloads.gas_box(s, n=300, radius=0.5, box=30.0, n_static=2, seed=5)
    scenes.append(s)
    want = []
    for s in scenes:
        o = to_oracle(s, 32)
        for _ in range(10):
            assert o.update(DT) == 0
        want.append(state(o))
    # the fallback's grid in LDS (default launch) and in the workspace (no LDS)
    for cfg in (None, dict(wg_small=64, wg_large=1024, split_points=64, lds_small=0, lds_large=0,
                           steps_per_launch=4, use_graph=0, split_tiny=0)):
        cl = [x.clone() for x in scenes]
        b = Batch()
        for s in cl:
            b.add_scene(s)
        if cfg:
            b.config = cfg
        b.update(DT, 10)
        b.sync()
        st = b.cc_stats()
        assert st["err_scenes"] == 0 and st["fallback_scenes"] == len(cl)
        for i, s in enumerate(cl):
            b.read_scene(i, s)
            assert bits_equal(state(s), want[i]), f"cfg {cfg} scene {i}: {first_diff(state(s), want[i])}"
