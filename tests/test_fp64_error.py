"""Error of the f32 step against the FP64 twin of the oracle (CPU).  The GPU
path is bit-identical to the f32 oracle (the GPU tests), so these figures are
also the GPU's error against FP64.  Smooth motion (the pinned cloth) stays
close; collision scenes are chaotic, so f32 and f64 runs drift apart
macroscopically within 100 updates (reported, bounded loosely)."""
import numpy as np

from helpers import state, to_oracle
from bendy2d_amd import Solver, workloads


def f32_f64_gap(s, steps, dt, candidate_mode=False):
    o32 = to_oracle(s, 32, candidate_mode=candidate_mode)
    o64 = to_oracle(s, 64, candidate_mode=candidate_mode)
    for _ in range(steps):
        assert o32.update(dt) == 0
        assert o64.update(dt) == 0
    a = state(o32).reshape(-1, 4)[:, :2]
    b = state(o64).reshape(-1, 4)[:, :2]
    d = np.abs(a - b)
    return float(np.nanmax(d)), int((~np.isfinite(d)).any(1).sum())


def test_pinned_cloth_1000_updates_close_to_fp64():
    # the 32x32 pinned cloth over a static circle, 8 substeps x 8 sweeps, 1000 updates
    s = Solver()
    dt, steps = workloads.cloth_pinned(s)
    gap, bad = f32_f64_gap(s, steps, dt)
    print(f"32x32 cloth, {steps} updates: max |f32 - f64| = {gap:.3e} (extent 32)")
    assert bad == 0 and gap < 1e-4 * 32.0


def test_softbody_collisions_drift_from_fp64():
    # a 12x12 circle softbody with colliders, 100 updates of 8 substeps: collision
    # chains amplify the last-bit differences; the gap is reported, not small
    s = Solver()
    dt, _, n = workloads.softbody(s, 1001)
    gap, bad = f32_f64_gap(s, 100, dt, candidate_mode=True)
    extent = s.bounds[1][0]
    print(f"softbody ({n} circles), 100 updates: max |f32 - f64| = {gap:.3e} = {gap / extent:.2e} x extent")
    assert bad == 0 and gap < extent
