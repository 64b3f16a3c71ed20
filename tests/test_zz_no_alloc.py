"""The allocation-free scene-image update against the workspace update it
varies: the same scenes, the same calls, bit-identical images (and equal to
the oracle), on two calls in a row, the second with a smaller scene; and the
size limit refused before any device work."""
import ctypes

import numpy as np
import pytest

from helpers import bits_equal, first_diff, state, to_oracle
from hipmem import DevBuf, Stream
from bendy2d_amd import SceneImage, Solver, _lib, lib, workloads

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def positions(img, dev, solver):
    host = np.zeros_like(img.host)
    dev.download(host)
    SceneImage.read(solver, host)
    return state(solver)


def run_both(seed, max_side, steps, stream):
    s = Solver()
    workloads.softbody(s, seed, max_side=max_side)
    o = to_oracle(s, 32, candidate_mode=True)
    img = SceneImage(s)
    a, b = DevBuf(img.info.image_bytes), DevBuf(img.info.image_bytes)
    ws = DevBuf(img.workspace_bytes)
    a.upload(img.host)
    b.upload(img.host)
    img.update(a.ptr, ws.ptr, ws.nbytes, DT, steps, stream.h)
    rc = lib().bd_scene_image_update_noalloc(ctypes.byref(img.info), b.ptr, ctypes.c_float(DT), steps, stream.h)
    assert rc == 0, lib().bd_last_error()
    stream.sync()
    for _ in range(steps):
        assert o.update(DT) == 0
    got_ws, got_na = positions(img, a, s.clone()), positions(img, b, s.clone())
    assert bits_equal(got_na, got_ws), f"noalloc vs workspace: {first_diff(got_na, got_ws)}"
    assert bits_equal(got_na, state(o)), f"noalloc vs oracle: {first_diff(got_na, state(o))}"
    for d in (a, b, ws):
        d.free()
    return s.stats()["n_points"]


def test_noalloc_matches_workspace_update():
    stream = Stream()
    n1 = run_both(5001, 40, 12, stream)   # 1,600 circles
    n2 = run_both(5002, 10, 7, stream)    # a smaller scene next, same scratch
    assert n2 < n1
    stream.destroy()


def test_noalloc_refuses_oversized_workspace():
    s = Solver()
    workloads.softbody(s, 5003, max_side=8)
    img = SceneImage(s)
    info = _lib.SceneImageInfo.from_buffer_copy(img.info)
    info.workspace_bytes = (8 << 20) + 16
    dev = DevBuf(img.info.image_bytes)
    dev.upload(img.host)
    rc = lib().bd_scene_image_update_noalloc(ctypes.byref(info), dev.ptr, ctypes.c_float(DT), 1, None)
    assert rc == _lib.BD_ERR_CAPACITY
    assert b"8 MiB" in lib().bd_last_error()
    # nothing ran: the image is as written
    host = np.zeros_like(img.host)
    dev.download(host)
    assert np.array_equal(host[512:], img.host[512:])
    dev.free()
