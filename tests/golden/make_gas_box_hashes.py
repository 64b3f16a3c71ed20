"""Writes gas_box_262144_hashes.json: the f32 oracle's state checksums for the
262,144-circle gas box (bendy2d_amd.workloads.gas_box, seed 2) after updates
1..100, so the GPU test can check the whole 100-update window at full size
without rerunning the oracle for minutes on the GPU host.

The oracle runs in candidate-list mode, which equals the reference's
all-pairs circle loop bit for bit (tests/test_oracle_kat.py checks that).
Run from the repository root: python tests/golden/make_gas_box_hashes.py
(about 8 minutes on one core)."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [HERE, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), ROOT]

from helpers import state, to_oracle  # noqa: E402
from statehash import block_hashes, state_hash  # noqa: E402
from bendy2d_amd import Solver, workloads  # noqa: E402

CHECKPOINTS = (1, 2, 3, 5, 10, 25, 50, 75, 100)
BLOCK = 4096


def main():
    s = Solver()
    dt, _ = workloads.gas_box(s)
    o = to_oracle(s, 32, candidate_mode=True)
    out = {"workload": "gas_box(n=262144, radius=0.5, box=1024, n_static=64, seed=2), 4 substeps, dt=1/60",
           "block_bodies": BLOCK, "checkpoints": {}}
    t0 = time.time()
    for k in range(1, CHECKPOINTS[-1] + 1):
        assert o.update(dt) == 0
        if k in CHECKPOINTS:
            x = state(o)
            out["checkpoints"][str(k)] = {"sha256": state_hash(x), "blocks": block_hashes(x, BLOCK)}
            print(f"update {k}: {time.time() - t0:.0f}s", flush=True)
    out["oracle_redos"] = int(o.cc_redos())
    with open(os.path.join(HERE, "gas_box_262144_hashes.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
