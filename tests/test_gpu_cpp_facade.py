"""The C++ host interface (include/bendy2d.hpp) on the GPU: a scene built from
the crate's types with Solver::new's defaults, advanced 120 updates, is
bit-identical to the f32 oracle (tests/cpp/facade_parity.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_parity")


def test_cpp_facade_matches_oracle():
    assert os.path.exists(BIN), "tests/cpp/facade_parity is built by __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade parity OK" in r.stdout
