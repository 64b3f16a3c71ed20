"""TILED engine schedule (CPU): for every tile, the backward dependency
region and per-layer link sets must make one sweep of tiles (interiors
written, halos recomputed) bit-identical to the sequential sweep."""
import numpy as np
import pytest

from bendy2d_amd import BD_ORDER_COLOUR, BD_ORDER_PARTICLE, Solver, workloads


@pytest.mark.parametrize("n,order,tile,cap", [
    (64, BD_ORDER_COLOUR, 256, 16384),
    (96, BD_ORDER_COLOUR, 1024, 2048),
    (128, BD_ORDER_PARTICLE, 512, 4096),   # deep layers: several passes per sweep
    (200, BD_ORDER_COLOUR, 4096, 16384),
])
def test_tiled_schedule_is_exact(n, order, tile, cap):
    s = Solver()
    workloads.cloth_large(s, n=n, order=order)
    bad, passes = s.tiled_selfcheck(tile, cap)
    assert bad == 0
    if order == BD_ORDER_COLOUR:
        assert passes == 1
    else:
        assert passes >= 1


def test_tiled_schedule_random_graph():
    # irregular link graph with pins: random particles, random a<b links
    rng = np.random.default_rng(7)
    s = Solver()
    n = 3000
    pos = rng.uniform(0, 60, (n, 2)).astype(np.float32)
    s.add_particles(pos)
    # links between nearby particles (k nearest by index in x-sorted order)
    order = np.argsort(pos[:, 0])
    a, b = [], []
    for i in range(n - 3):
        for d in (1, 2, 3):
            x, y = sorted((int(order[i]), int(order[i + d])))
            a.append(x); b.append(y)
    a, b = np.array(a), np.array(b)
    perm = rng.permutation(len(a))
    s.add_particle_links(a[perm], b[perm], rng.uniform(0.5, 2.0, len(a)).astype(np.float32))
    s.set_pinned_range(0, (rng.random(n) < 0.05).astype(np.uint8))
    bad, passes = s.tiled_selfcheck(256, 2048)
    assert bad == 0 and passes >= 1
