"""The reference's device module surface (crates/riga-engine/src/gpu.rs:490-570)
on gfx950: availability, device name and architecture, the worst-of and
two-factor local-vol propagation calls and their resident forms, so callers of
``riga_engine::gpu`` (examples/gpu_bench.rs, examples/gpu_lv_bench.rs) find the
same operations.  Inputs are float32 like the reference's."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .api import arch as _arch, device_name as _device_name, gpu_backend
from .types import RigaError

_F32 = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _p(a: np.ndarray):
    return a.ctypes.data_as(_F32)


def ensure_available() -> None:
    """gpu.rs:500-503."""
    gpu_backend()


def device_name() -> str:
    return _device_name()


def arch() -> str:
    return _arch()


def worst_of(a, b) -> np.ndarray:
    """gpu.rs:515-517: elementwise ``a < b ? a : b``."""
    a, b = _f32(a), _f32(b)
    if a.shape != b.shape:
        raise RigaError("shape_mismatch", "worst-of inputs must have equal length")
    gpu_backend()
    out = np.zeros_like(a)
    err = N.Error()
    N.check(N.lib().riga_worst_of(N.Device.get().handle, _p(a), _p(b), a.size, _p(out), C.byref(err)), err)
    return out


class ResidentWorstOf:
    """gpu.rs:362-405: inputs uploaded once; ``reduce`` runs the kernel."""

    def __init__(self, a, b):
        a, b = _f32(a), _f32(b)
        if a.shape != b.shape:
            raise RigaError("shape_mismatch", "worst-of inputs must have equal length")
        gpu_backend()
        self.n = a.size
        h = C.c_void_p()
        err = N.Error()
        N.check(N.lib().riga_resident_worst_of_create(N.Device.get().handle, _p(a), _p(b), self.n, C.byref(h),
                                                      C.byref(err)), err)
        self.h = h

    def reduce(self) -> None:
        err = N.Error()
        N.check(N.lib().riga_resident_worst_of_reduce(self.h, C.byref(err)), err)

    def download(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.float32)
        err = N.Error()
        N.check(N.lib().riga_resident_worst_of_download(self.h, _p(out), C.byref(err)), err)
        return out

    def synchronize(self) -> None:
        err = N.Error()
        N.check(N.lib().riga_resident_worst_of_synchronize(self.h, C.byref(err)), err)

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            N.lib().riga_resident_worst_of_destroy(self.h)
            self.h = None

    __del__ = close


def resident_worst_of(a, b) -> ResidentWorstOf:
    """gpu.rs:520-522."""
    return ResidentWorstOf(a, b)


def _lv_inputs(normals, rows, ls_axis, paths, steps):
    n, r, ls = _f32(normals).reshape(-1), _f32(rows).reshape(-1), _f32(ls_axis).reshape(-1)
    gx = ls.size // 2
    if ls.size != 2 * gx or gx < 2 or n.size < paths * steps * 2 or r.size < 2 * steps * gx:
        raise RigaError("shape_mismatch", "normals [paths][steps][2], rows [2][steps][gx], ls_axis [2][gx]")
    return n, r, ls, gx


def lv_terminal_spots(normals, rows, ls_axis, s0, references, call_barrier: float, paths: int, steps: int,
                      dt: float, rate: float, dividends):
    """gpu.rs:538-566: terminal spots [paths][2] (flattened) and first KO steps."""
    n, r, ls, gx = _lv_inputs(normals, rows, ls_axis, paths, steps)
    gpu_backend()
    out = np.zeros(2 * paths, dtype=np.float32)
    ko = np.zeros(paths, dtype=np.int32)
    err = N.Error()
    N.check(N.lib().riga_lv_terminal_spots(
        N.Device.get().handle, _p(n), _p(r), _p(ls), gx, _p(_f32(s0)), _p(_f32(references)), call_barrier,
        paths, steps, dt, rate, _p(_f32(dividends)), _p(out), ko.ctypes.data_as(_I32), C.byref(err)), err)
    return out, ko


class ResidentLv:
    """gpu.rs:407-489: shocks, rows and axes uploaded once; ``run`` per scenario."""

    def __init__(self, normals, rows, ls_axis, paths: int, steps: int, dt: float):
        n, r, ls, gx = _lv_inputs(normals, rows, ls_axis, paths, steps)
        gpu_backend()
        self.paths = paths
        h = C.c_void_p()
        err = N.Error()
        N.check(N.lib().riga_resident_lv_create(N.Device.get().handle, _p(n), _p(r), _p(ls), gx, paths, steps, dt,
                                                C.byref(h), C.byref(err)), err)
        self.h = h

    def run(self, s0, references, call_barrier: float, rate: float, dividends) -> None:
        err = N.Error()
        N.check(N.lib().riga_resident_lv_run(self.h, _p(_f32(s0)), _p(_f32(references)), call_barrier, rate,
                                             _p(_f32(dividends)), C.byref(err)), err)

    def download(self):
        out = np.zeros(2 * self.paths, dtype=np.float32)
        ko = np.zeros(self.paths, dtype=np.int32)
        err = N.Error()
        N.check(N.lib().riga_resident_lv_download(self.h, _p(out), ko.ctypes.data_as(_I32), C.byref(err)), err)
        return out, ko

    def synchronize(self) -> None:
        err = N.Error()
        N.check(N.lib().riga_resident_lv_synchronize(self.h, C.byref(err)), err)

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            N.lib().riga_resident_lv_destroy(self.h)
            self.h = None

    __del__ = close


def resident_lv(normals, rows, ls_axis, paths: int, steps: int, dt: float) -> ResidentLv:
    """gpu.rs:525-535."""
    return ResidentLv(normals, rows, ls_axis, paths, steps, dt)
