"""Batched form on the GPU: every scene of a batch bit-identical to the f32
oracle, for every launch configuration (workgroup sizes, LDS residency
modes, graph replay, updates per launch)."""
import numpy as np
import pytest

from helpers import bits_equal, first_diff, polygon_scene, state, state_acc, to_oracle
from bendy2d_amd import Batch, Solver, workloads

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def make_scenes(n, max_side=20, seed0=100):
    scenes = []
    for i in range(n):
        s = Solver()
        workloads.softbody(s, seed0 + i, max_side=max_side)
        scenes.append(s)
    # a pinned particle cloth with a static circle, and a circle gas, too
    s = Solver()
    workloads.cloth_pinned(s, n=12, sub_steps=3, link_iters=2)
    scenes.append(s)
    s = Solver()
    workloads.gas_box(s, n=300, radius=0.5, box=40.0, n_static=2, seed=5)
    scenes.append(s)
    return scenes


def oracle_run(scenes, steps):
    outs = []
    for s in scenes:
        o = to_oracle(s, 32)
        for _ in range(steps):
            assert o.update(DT) == 0
        outs.append(state(o))
    return outs


# the tiny class (scenes of at most split_tiny bodies on their own workgroups)
# on, off, and covering the whole small class
CONFIGS = [
    dict(wg_small=64, wg_large=256, split_points=256, lds_small=16384, lds_large=147456, steps_per_launch=1, use_graph=1,
         split_tiny=0),
    dict(wg_small=128, wg_large=1024, split_points=2048, lds_small=32768, lds_large=65536, steps_per_launch=3, use_graph=1,
         split_tiny=100, wg_tiny=64, lds_tiny=8192),
    dict(wg_small=256, wg_large=512, split_points=100000, lds_small=4096, lds_large=4096, steps_per_launch=2, use_graph=0,
         split_tiny=400, wg_tiny=128, lds_tiny=0),
    dict(wg_small=64, wg_large=1024, split_points=64, lds_small=0, lds_large=0, steps_per_launch=4, use_graph=0,
         split_tiny=0),
    dict(wg_small=256, wg_large=1024, split_points=1024, lds_small=32768, lds_large=147456, steps_per_launch=2,
         use_graph=1, split_tiny=4096, wg_tiny=64, lds_tiny=16384),
]


@pytest.mark.parametrize("cfg", range(len(CONFIGS)))
def test_batch_matches_oracle(cfg):
    steps = 12
    scenes = make_scenes(14)
    want = oracle_run(scenes, steps)
    b = Batch()
    for s in scenes:
        b.add_scene(s)
    b.config = CONFIGS[cfg]
    b.update(DT, 5)
    b.update(DT, steps - 5)
    b.sync()
    st = b.cc_stats()
    assert st["err_scenes"] == 0
    for i, s in enumerate(scenes):
        b.read_scene(i, s)
        got = state(s)
        assert bits_equal(got, want[i]), f"scene {i}: {first_diff(got, want[i])}"


def test_batch_polygon_scenes():
    # scenes with soft polygons (and circles, cloth, static colliders) beside
    # plain ones, in every size class; acc of static polygon points included
    scenes = [polygon_scene(500 + i, n_polys=3 + i % 5, circles=i % 2 == 0, cloth=i % 3 != 2) for i in range(6)]
    scenes += make_scenes(3, max_side=12)
    want = []
    for s in scenes:
        o = to_oracle(s, 32)
        for _ in range(20):
            assert o.update(DT) == 0
        want.append(state_acc(o))
    for cfg in (CONFIGS[1], CONFIGS[3]):
        cl = [x.clone() for x in scenes]
        b = Batch()
        for x in cl:
            b.add_scene(x)
        b.config = cfg
        b.update(DT, 7)
        b.update(DT, 13)
        b.sync()
        for i, x in enumerate(cl):
            b.read_scene(i, x)
            got = state_acc(x)
            assert bits_equal(got, want[i]), f"scene {i}: {first_diff(got, want[i])}"


def test_batch_autotune_keeps_state():
    scenes = make_scenes(10, max_side=16)
    want = oracle_run(scenes, 6)
    b = Batch()
    for s in scenes:
        b.add_scene(s)
    tried = b.autotune(DT, 2)
    assert tried > 5
    b.update(DT, 6)
    b.sync()
    for i, s in enumerate(scenes):
        b.read_scene(i, s)
        assert bits_equal(state(s), want[i]), f"scene {i}"


def test_batch_large_scene_global_residency():
    # a 128x128 softbody (16384 circles): pos alone fits LDS at 128 KiB, and
    # with a small LDS budget it runs from global memory
    s = Solver()
    workloads.softbody(s, 7, max_side=0)
    big = Solver()
    dt, _, n = workloads.softbody(big, 3)  # whatever size seed 3 draws
    scenes = [s, big]
    c = Solver()
    workloads.add_grid(c, 128, 128, x0=200.0, y0=400.0, as_circles=True, radius=0.35)
    c.gravity = (0.0, -9.81)
    c.bounds = ((0.0, 0.0), (424.0, 424.0))
    c.set_sub_steps(2)
    scenes.append(c)
    want = oracle_run(scenes, 3)
    for lds in (147456, 65536):
        cl = [x.clone() for x in scenes]
        b = Batch()
        for x in cl:
            b.add_scene(x)
        b.config = dict(wg_small=256, wg_large=1024, split_points=1024, lds_small=lds, lds_large=lds,
                        steps_per_launch=1, use_graph=1)
        b.update(DT, 3)
        b.sync()
        for i, x in enumerate(cl):
            b.read_scene(i, x)
            assert bits_equal(state(x), want[i]), f"lds {lds} scene {i}: {first_diff(state(x), want[i])}"


def circle_column(wide, tall, seed):
    """wide x tall free circles (r 0.35 at spacing 0.8) stacked in a box one
    column wide, 8 substeps: the stack settles into lasting dense contact,
    where circles keep pushing each other around within every sweep."""
    s = Solver()
    rng = np.random.default_rng(seed)
    for r in range(tall):
        for c in range(wide):
            s.add_circle((1.0 + 0.8 * c + rng.uniform(-0.05, 0.05), 1.0 + 0.8 * r), 0.35)
    s.gravity = (0.0, -9.81)
    s.bounds = ((0.0, 0.0), (0.8 * wide + 1.2, 0.8 * tall + 10.0))
    s.set_sub_steps(8)
    s.set_link_iterations(1)
    return s


def test_batch_dense_contact_exact():
    # in dense contact, circles leave their skins in many sweeps; such a sweep
    # is kept only when every pair missing from the lists provably did not
    # touch at its turn, so the result must stay bit-identical to the
    # oracle's all-pairs loop
    scenes = [circle_column(12, 120, 1), circle_column(10, 140, 2)]
    b = Batch()
    for s in scenes:
        b.add_scene(s)
    oracles = [to_oracle(s, 32) for s in scenes]
    done = 0
    n_out = 0
    for chunk in (60, 60, 60, 60):
        b.update(DT, chunk)
        b.sync()
        for o in oracles:
            for _ in range(chunk):
                assert o.update(DT) == 0
        done += chunk
        n_out = max([n_out] + [b.scene_lists(i)["n_out"] for i in range(len(scenes))])
        for i, s in enumerate(scenes):
            b.read_scene(i, s)
            a, w = state(s), state(oracles[i])
            assert bits_equal(a, w), f"scene {i} step {done}: first diff {first_diff(a, w)}"
    st = [b.scene_status(i) for i in range(len(scenes))]
    assert all(x["err"] == 0 for x in st)
    assert sum(x["rounds"] for x in st) > 0
    kept = sum(b.scene_lists(i)["kept"] for i in range(len(scenes)))
    print("sweeps kept past skins:", kept, "redos:", sum(x["redos"] for x in st), "outliers seen:", n_out)
    # the lists' outlier paths (circles whose skins outgrew the cells) ran
    assert n_out > 0


_placement_oracle = {}


@pytest.mark.parametrize("lds_small, placement", [
    (49152, "LDS beside positions and previous positions"),
    (24576, "previous positions parked, state in their place"),
    (16384, "positions parked (positions alone fit), state in their place"),
    (8192, "positions in global memory, state in the workspace"),
])
def test_round_state_placements(lds_small, placement):
    # the pair rounds keep each circle's next partner, stamp and list cursor
    # in LDS when the scene's LDS budget leaves 8 bytes a circle, else in
    # place of the previous positions or positions (parked in global memory
    # for the sweep), else in the workspace; for a stack of about 1,400
    # circles in lasting contact these budgets select each case in turn, and
    # every one must match the oracle bit for bit
    scenes = [circle_column(12, 120, 3), circle_column(10, 140, 4)]
    steps = 120
    if "want" not in _placement_oracle:
        _placement_oracle["want"] = oracle_run(scenes, steps)
    want = _placement_oracle["want"]
    b = Batch()
    for s in scenes:
        b.add_scene(s)
    b.config = {"wg_small": 256, "wg_large": 1024, "split_points": 2048, "lds_small": lds_small,
                "lds_large": 147456, "steps_per_launch": 2, "use_graph": 0, "split_tiny": 0, "wg_tiny": 64,
                "lds_tiny": 8192}
    b.update(DT, steps)
    b.sync()
    for i, s in enumerate(scenes):
        b.read_scene(i, s)
        a = state(s)
        assert bits_equal(a, want[i]), f"{placement}: scene {i} first diff {first_diff(a, want[i])}"
    st = [b.scene_status(i) for i in range(len(scenes))]
    assert all(x["err"] == 0 for x in st), st
    assert sum(x["rounds"] for x in st) > 0
