"""HIP path vs the f32 oracle on the same seeded scenes: bit-exact.

Every case builds a scene through the C ABI, mirrors it into the oracle,
runs both for a number of updates and requires identical f32 bits for every
position, previous position and polygon centre.
"""
import numpy as np
import pytest

from helpers import bits_equal, first_diff, polygon_scene, state, state_acc, to_oracle
from bendy2d_amd import BD_ENGINE_GLOBAL, BD_ENGINE_SCENE, BD_ENGINE_TILED, BD_ORDER_PARTICLE, Solver, workloads

ENGINES = [BD_ENGINE_GLOBAL, BD_ENGINE_SCENE]

pytestmark = pytest.mark.gpu


def run_pair(s, steps, dt, candidate_mode=False, check_every=None, engine=None):
    o = to_oracle(s, 32, candidate_mode=candidate_mode)
    for k in range(steps):
        s.update(dt)
        if engine is not None and k == 0:
            assert s.engine_used == engine
        assert o.update(dt) == 0
        if check_every and (k + 1) % check_every == 0:
            a, b = state(s), state(o)
            assert bits_equal(a, b), f"step {k + 1}: first diff {first_diff(a, b)}"
    a, b = state(s), state(o)
    assert bits_equal(a, b), f"final: first diff {first_diff(a, b)}"
    return s, o


@pytest.mark.parametrize("engine", ENGINES)
def test_cloth_pinned_small(engine):
    s = Solver()
    s.set_engine(engine)
    dt, _ = workloads.cloth_pinned(s, n=12, sub_steps=2, link_iters=3)
    run_pair(s, 40, dt, check_every=10)


@pytest.mark.parametrize("engine", ENGINES)
def test_cloth_per_particle_order(engine):
    s = Solver()
    s.set_engine(engine)
    dt, _ = workloads.cloth_pinned(s, n=10, order=BD_ORDER_PARTICLE, sub_steps=2, link_iters=2)
    run_pair(s, 20, dt)


@pytest.mark.parametrize("engine", ENGINES)
def test_reference_defaults(engine):
    # Solver::new defaults: gravity (0, 98.2), bounds (0,0)+(100,100), 1 substep
    s = Solver()
    s.set_engine(engine)
    rng = np.random.default_rng(5)
    for i in range(50):
        s.add_particle(rng.uniform(5, 95, 2))
    for i in range(49):
        s.add_particle_link(i, i + 1, float(rng.uniform(0.5, 3.0)))
    s.add_particle_link(3, 40, 10.0)
    s.add_particle_link(3, 40, 10.0)  # duplicate links are legal (applied twice)
    run_pair(s, 100, 1.0 / 60.0, check_every=25)


@pytest.mark.parametrize("engine", ENGINES)
def test_gas_small(engine):
    s = Solver()
    s.set_engine(engine)
    dt, _ = workloads.gas_box(s, n=600, radius=0.5, box=64.0, n_static=3, seed=11)
    run_pair(s, 30, dt, check_every=10)


@pytest.mark.parametrize("engine", ENGINES)
def test_gas_dense_nan(engine):
    # dense enough that bounds clamp circles onto one corner point: coincident
    # circles normalise a zero vector and yield NaN, as in the reference
    s = Solver()
    s.set_engine(engine)
    dt, _ = workloads.gas_box(s, n=600, radius=0.5, box=40.0, n_static=3, seed=11)
    run_pair(s, 12, dt, check_every=4)


def test_global_gas_hot_colliders():
    # random overlaps turn into speed, and circles expelled from the static
    # colliders jam into rings: candidate lists span several margin levels,
    # components go to every replay tier (lane, LDS wavefront, workgroup), and
    # circles leave their margins, so the per-pair check at each skipped pair's
    # turn runs
    s = Solver()
    s.set_engine(BD_ENGINE_GLOBAL)
    dt, _ = workloads.gas_box(s, n=12000, radius=0.5, box=220.0, n_static=10, seed=5)
    s, _ = run_pair(s, 12, dt, candidate_mode=True, check_every=4, engine=BD_ENGINE_GLOBAL)
    assert s.stats()["cc_rounds"] > 0


def test_global_gas_pile():
    # a dense gas falling into a pile on the floor: large components
    s = Solver()
    s.set_engine(BD_ENGINE_GLOBAL)
    dt, _ = workloads.gas_box(s, n=4000, radius=0.5, box=80.0, n_static=0, seed=6)
    run_pair(s, 60, dt, candidate_mode=True, check_every=20, engine=BD_ENGINE_GLOBAL)


def test_global_overlap_chain():
    # a row of circles each overlapping the next by half a radius: every pair
    # of the row is in contact, so one sweep pushes every circle far beyond
    # the base margin (overruns, the per-pair check, margins on several
    # levels), and one component spans the whole row
    s = Solver()
    s.set_engine(BD_ENGINE_GLOBAL)
    s.gravity = (0.0, 0.0)
    s.bounds = ((-5000.0, -5000.0), (10000.0, 10000.0))
    for k in range(300):
        p = np.array([0.5 * k, 0.01 * (k % 7)], np.float32)
        s.add_circle(p, 0.5)
    for k in range(40):  # a few quiet circles around, on the base level
        p = np.array([-30.0 + 1.5 * k, 20.0], np.float32)
        s.add_circle(p, 0.5)
    run_pair(s, 6, 1.0 / 60.0, check_every=2, engine=BD_ENGINE_GLOBAL)


def test_global_missed_pair_forces_a_redo():
    # The row's sweep pushes its front circle about one radius forward (each
    # pair (k, k+1) finds k already pushed, so the push grows to r).  A sentinel
    # 1.3 beyond the front is outside both reaches at the base margin
    # (2 * (0.5 + 0.0625) = 1.125), so the pair (front, sentinel) is not a
    # candidate, yet the front reaches it before that pair's turn: the sweep's
    # check must find the pair, undo the sweep and redo it with wider margins,
    # and the result must still equal the reference's loop bit for bit.
    s = Solver()
    s.set_engine(BD_ENGINE_GLOBAL)
    s.gravity = (0.0, 0.0)
    s.bounds = ((-5000.0, -5000.0), (10000.0, 10000.0))
    for k in range(300):
        s.add_circle(np.array([0.5 * k, 0.0], np.float32), 0.5)
    s.add_circle(np.array([0.5 * 299 + 1.3, 0.0], np.float32), 0.5)
    s, o = run_pair(s, 4, 1.0 / 60.0, check_every=1, engine=BD_ENGINE_GLOBAL)
    assert s.stats()["cc_redos"] > 0
    assert o.get_circles()[300, 0] > 0.5 * 299 + 1.3  # the sentinel was pushed


@pytest.mark.parametrize("engine", ENGINES)
def test_circles_links_acc(engine):
    s = Solver()
    s.set_engine(engine)
    rng = np.random.default_rng(3)
    for i in range(40):
        p = rng.uniform(10, 90, 2)
        v = rng.normal(0, 0.3, 2)
        s.add_circle(p, float(rng.uniform(0.5, 3.0)), prev_pos=p - v, acc=rng.normal(0, 50, 2))
    for i in range(0, 38, 2):
        s.add_circle_link(i, i + 1, 4.0)
    s.add_circle_link(0, 39, 20.0)
    run_pair(s, 60, 1.0 / 60.0, check_every=20)


@pytest.mark.parametrize("engine", ENGINES)
def test_polygons(engine):
    s = Solver()
    s.set_engine(engine)
    s.add_polygon_circle(8.0, (30.0, 30.0), 10, False)
    s.add_polygon_circle(6.0, (42.0, 38.0), 7, False)
    s.add_polygon_new([(10.0, 80.0), (60.0, 80.0), (60.0, 90.0), (10.0, 90.0)], True)
    s.add_polygon_circle(5.0, (35.0, 60.0), 3, False)  # n = 3: duplicate links
    run_pair(s, 120, 1.0 / 60.0, check_every=20)


@pytest.mark.parametrize("engine", ENGINES)
def test_mixed_scene(engine):
    s = Solver()
    s.set_engine(engine)
    workloads.add_grid(s, 8, 6, x0=20.0, y0=90.0, pin_top=True)
    workloads.add_grid(s, 5, 5, x0=60.0, y0=70.0, as_circles=True, radius=0.45)
    s.add_polygon_circle(6.0, (50.0, 40.0), 8, False)
    s.add_polygon_circle(5.0, (52.0, 55.0), 6, False)
    s.add_static_circle((30.0, 20.0), 6.0)
    s.add_static_polygon([(60.0, 5.0), (90.0, 5.0), (75.0, 25.0)])
    s.set_sub_steps(3)
    s.set_link_iterations(2)
    run_pair(s, 60, 1.0 / 60.0, check_every=20)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("seed", [11, 12])
def test_polygon_pile(engine, seed):
    # soft polygons landing on each other and on static polygons, among
    # circles and a cloth; some polygon points start with an acc
    s = polygon_scene(seed, n_polys=7)
    s.set_engine(engine)
    o = to_oracle(s, 32)
    for k in range(90):
        s.update(1.0 / 60.0)
        assert o.update(1.0 / 60.0) == 0
        if (k + 1) % 30 == 0:
            a, b = state_acc(s), state_acc(o)
            assert bits_equal(a, b), f"step {k + 1}: first diff {first_diff(a, b)}"
    assert s.engine_used == engine


@pytest.mark.parametrize("link_iters", [1, 3])
def test_polygon_wider_than_the_workgroup(link_iters):
    # a 300-point ring (the SCENE engine picks 256 lanes for 330 bodies, so
    # lanes take several of its points per edge) squeezed by two small ones
    s = Solver()
    s.set_engine(BD_ENGINE_SCENE)
    s.add_polygon_circle(30.0, (50.0, 50.0), 300, False)
    s.add_polygon_circle(4.0, (50.0, 78.0), 12, False)
    s.add_polygon_circle(4.0, (22.0, 50.0), 18, False)
    s.set_sub_steps(2)
    s.set_link_iterations(link_iters)
    run_pair(s, 60, 1.0 / 60.0, check_every=20, engine=BD_ENGINE_SCENE)


@pytest.mark.parametrize("tile,wg", [(4096, 512), (1024, 256), (2048, 1024)])
def test_tiled_cloth(tile, wg):
    # 160x160 jittered pinned cloth, 4 substeps, 8 sweeps, plus a static circle
    s = Solver()
    s.set_engine(BD_ENGINE_TILED)
    s.set_tiling(tile, wg)
    dt, _ = workloads.cloth_large(s, n=160)
    s.add_static_circle((80.0, -200.0), 30.0)
    run_pair(s, 30, dt, check_every=10, engine=BD_ENGINE_TILED)


def test_tiled_deep_layers():
    # per-particle link order: hundreds of layers, several passes per sweep
    s = Solver()
    s.set_engine(BD_ENGINE_TILED)
    s.set_tiling(1024, 256)
    dt, _ = workloads.cloth_large(s, n=140, order=BD_ORDER_PARTICLE, sub_steps=2, link_iters=2)
    run_pair(s, 10, dt, engine=BD_ENGINE_TILED)


def test_auto_engine_choice():
    small = Solver()
    workloads.cloth_pinned(small, n=32)
    small.update(1 / 60)
    assert small.engine_used == BD_ENGINE_SCENE
    big = Solver()
    workloads.cloth_large(big, n=200)
    big.update(1 / 60)
    assert big.engine_used == BD_ENGINE_TILED
    poly = Solver()
    poly.add_polygon_circle(5.0, (50.0, 50.0), 8)
    poly.update(1 / 60)
    assert poly.engine_used == BD_ENGINE_SCENE
