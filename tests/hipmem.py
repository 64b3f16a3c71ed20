"""Device memory, streams and graphs for tests that act as a caller owning
them: ctypes over the HIP runtime that libbendy2d.so itself links, so
pointers and streams are valid for the library."""
import ctypes
import os

from bendy2d_amd import lib

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib()  # load libbendy2d.so first, then the runtime it resolved
        path = None
        with open(f"/proc/{os.getpid()}/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        H = ctypes.CDLL(path or "libamdhip64.so")
        v, pv = ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)
        for name, args in (("hipMalloc", [pv, ctypes.c_size_t]), ("hipFree", [v]),
                           ("hipMemcpy", [v, v, ctypes.c_size_t, ctypes.c_int]),
                           ("hipMemset", [v, ctypes.c_int, ctypes.c_size_t]),
                           ("hipDeviceSynchronize", []), ("hipStreamCreate", [pv]), ("hipStreamDestroy", [v]),
                           ("hipStreamSynchronize", [v]), ("hipStreamBeginCapture", [v, ctypes.c_int]),
                           ("hipStreamEndCapture", [v, pv]), ("hipGraphInstantiate", [pv, v, v, v, ctypes.c_size_t]),
                           ("hipGraphLaunch", [v, v]), ("hipGraphExecDestroy", [v]), ("hipGraphDestroy", [v])):
            f = getattr(H, name)
            f.restype = ctypes.c_int
            f.argtypes = args
        _hip = H
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


class DevBuf:
    """hipMalloc'd bytes (256-byte aligned, as hipMalloc returns)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    def upload(self, host):
        check(hip().hipMemcpy(self.ptr, host.ctypes.data_as(ctypes.c_void_p), host.nbytes, 1), "hipMemcpy H2D")

    def download(self, host):
        check(hip().hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), self.ptr, host.nbytes, 2), "hipMemcpy D2H")

    def fill(self, byte):
        check(hip().hipMemset(self.ptr, byte, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


class Stream:
    def __init__(self):
        self.h = ctypes.c_void_p()
        check(hip().hipStreamCreate(ctypes.byref(self.h)), "hipStreamCreate")

    def sync(self):
        check(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        if self.h:
            hip().hipStreamDestroy(self.h)
            self.h = ctypes.c_void_p()


class Graph:
    """Captures the work queued on `stream` inside a with-block."""

    def __init__(self, stream):
        self.stream, self.graph, self.exec = stream, ctypes.c_void_p(), ctypes.c_void_p()

    def __enter__(self):
        check(hip().hipStreamBeginCapture(self.stream.h, 0), "hipStreamBeginCapture")
        return self

    def __exit__(self, *exc):
        check(hip().hipStreamEndCapture(self.stream.h, ctypes.byref(self.graph)), "hipStreamEndCapture")
        if exc[0] is None:
            check(hip().hipGraphInstantiate(ctypes.byref(self.exec), self.graph, None, None, 0),
                  "hipGraphInstantiate")
        return False

    def launch(self):
        check(hip().hipGraphLaunch(self.exec, self.stream.h), "hipGraphLaunch")

    def destroy(self):
        if self.exec:
            hip().hipGraphExecDestroy(self.exec)
        if self.graph:
            hip().hipGraphDestroy(self.graph)
