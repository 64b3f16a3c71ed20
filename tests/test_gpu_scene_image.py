"""Scene images on the GPU: one scene in device memory the caller owns,
stepped on the caller's stream (also inside a captured HIP graph), must equal
the f32 oracle bit for bit, whatever workspace it is given between calls."""
import ctypes

import numpy as np
import pytest

from helpers import bits_equal, first_diff, polygon_scene, state, state_acc, to_oracle
from hipmem import DevBuf, Graph, Stream
from bendy2d_amd import SceneImage, Solver, _lib, lib, workloads

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def softbody(seed, max_side):
    s = Solver()
    workloads.softbody(s, seed, max_side=max_side)
    return s


def check_image(img, dev, solver, oracle):
    host = np.zeros_like(img.host)
    dev.download(host)
    SceneImage.read(solver, host)
    a, b = state(solver), state(oracle)
    assert bits_equal(a, b), f"first diff {first_diff(a, b)}"


def test_image_updates_match_oracle():
    s = softbody(4001, 24)
    o = to_oracle(s, 32, candidate_mode=True)
    img = SceneImage(s)
    assert img.info.n_circles == s.stats()["n_points"] and img.workspace_bytes > 0
    dev, ws, stream = DevBuf(img.info.image_bytes), DevBuf(img.workspace_bytes), Stream()
    ws.fill(0xA5)  # the workspace's contents must not matter
    dev.upload(img.host)
    for steps, dt in ((3, DT), (1, 1.0 / 120.0), (20, DT)):
        img.update(dev.ptr, ws.ptr, ws.nbytes, dt, steps, stream.h)
        stream.sync()
        for _ in range(steps):
            assert o.update(dt) == 0
        check_image(img, dev, s, o)
    for b in (dev, ws):
        b.free()
    stream.destroy()


def test_polygon_scene_image():
    # soft polygons, circles, a cloth and static colliders in one image; the
    # acc that static polygon points gather over the image's updates comes back
    s = polygon_scene(4010, n_polys=5)
    o = to_oracle(s, 32)
    img = SceneImage(s)
    dev, ws, stream = DevBuf(img.info.image_bytes), DevBuf(img.workspace_bytes), Stream()
    dev.upload(img.host)
    for steps in (1, 4, 15):
        img.update(dev.ptr, ws.ptr, ws.nbytes, DT, steps, stream.h)
        stream.sync()
        for _ in range(steps):
            assert o.update(DT) == 0
        host = np.zeros_like(img.host)
        dev.download(host)
        SceneImage.read(s, host)
        a, b = state_acc(s), state_acc(o)
        assert bits_equal(a, b), f"after {steps}: first diff {first_diff(a, b)}"
    for b in (dev, ws):
        b.free()
    stream.destroy()


def test_images_share_one_workspace():
    # two scenes stepped alternately with one workspace: nothing carries over
    scenes = [softbody(4002, 16), softbody(4003, 20)]
    oracles = [to_oracle(s, 32, candidate_mode=True) for s in scenes]
    imgs = [SceneImage(s) for s in scenes]
    devs = [DevBuf(i.info.image_bytes) for i in imgs]
    ws = DevBuf(max(i.workspace_bytes for i in imgs))
    stream = Stream()
    for i, d in zip(imgs, devs):
        d.upload(i.host)
    for rnd in range(4):
        for i, d, o in zip(imgs, devs, oracles):
            i.update(d.ptr, ws.ptr, ws.nbytes, DT, 5, stream.h)
            for _ in range(5):
                assert o.update(DT) == 0
    stream.sync()
    for i, d, s, o in zip(imgs, devs, scenes, oracles):
        check_image(i, d, s, o)
    for b in devs + [ws]:
        b.free()
    stream.destroy()


def test_image_in_a_captured_graph():
    # the update captured once and replayed: every replay advances the scene by
    # the captured steps, checked against the oracle after each replay
    s = softbody(4004, 18)
    o = to_oracle(s, 32, candidate_mode=True)
    img = SceneImage(s)
    dev, ws, stream = DevBuf(img.info.image_bytes), DevBuf(img.workspace_bytes), Stream()
    dev.upload(img.host)
    g = Graph(stream)
    with g:
        img.update(dev.ptr, ws.ptr, ws.nbytes, DT, 4, stream.h)
    for replay in range(3):
        g.launch()
        stream.sync()
        for _ in range(4):
            assert o.update(DT) == 0
        check_image(img, dev, s, o)
    g.destroy()
    for b in (dev, ws):
        b.free()
    stream.destroy()


def test_image_refuses_a_small_workspace():
    s = softbody(4005, 12)
    img = SceneImage(s)
    dev, ws = DevBuf(img.info.image_bytes), DevBuf(16)
    dev.upload(img.host)
    rc = lib().bd_scene_image_update(ctypes.byref(img.info), dev.ptr, ws.ptr, 16, ctypes.c_float(DT), 1, None)
    assert rc == _lib.BD_ERR_ARG
    rc = lib().bd_scene_image_update(ctypes.byref(img.info), ctypes.c_void_p(dev.ptr.value + 16), ws.ptr,
                                     img.workspace_bytes, ctypes.c_float(DT), 1, None)
    assert rc == _lib.BD_ERR_ARG  # not 256-byte aligned
    for b in (dev, ws):
        b.free()
