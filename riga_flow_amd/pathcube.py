"""PathCube on demand (crates/riga-engine/src/pathcube.rs): the correlated
Brownian block ``[path, time, factor]`` drawn from ``ChaCha8Rng::seed_from_u64``
(``factors`` StandardNormals per cell, y = L z) and the GBM spot cube evolved
from it, both float32, generated on the device only when asked for; the JSON
archive in serde's field layout; and the terminal worst-of queries.

The device lane per path starts at the path's own word of the stream (the
Monte Carlo lane's stream parse), so block and cube equal the reference's
sequential generation.
"""
from __future__ import annotations

import ctypes as C
import json
import math
from dataclasses import dataclass
from decimal import Decimal
from typing import Sequence

import numpy as np

from . import _native as N
from .api import gpu_backend
from .types import RigaError


def _ryu32(x: float) -> str:
    """serde_json's text for an f32 (ryu's shortest round-trip digits and its f32 layout)."""
    if x != x or x in (math.inf, -math.inf):
        return "null"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    t = Decimal(np.format_float_scientific(np.float32(abs(x)), unique=True)).as_tuple()
    digits = "".join(map(str, t.digits)).lstrip("0")
    k = t.exponent
    stripped = digits.rstrip("0")
    k += len(digits) - len(stripped)
    digits = stripped
    length = len(digits)
    kk = length + k
    if 0 <= k and kk <= 13:
        return sign + digits + "0" * k + ".0"
    if 0 < kk <= 13:
        return sign + digits[:kk] + "." + digits[kk:]
    if -6 < kk <= 0:
        return sign + "0." + "0" * (-kk) + digits
    if length == 1:
        return sign + digits + "e" + str(kk - 1)
    return sign + digits[0] + "." + digits[1:] + "e" + str(kk - 1)


class _Cube:
    def __init__(self, paths: int, times: int, factors: int, data):
        self.paths, self.times, self.factors = int(paths), int(times), int(factors)
        self.data = np.ascontiguousarray(np.asarray(data, dtype=np.float32).reshape(paths, times, factors))

    def index(self, path: int, time: int, factor: int) -> int:
        return (path * self.times + time) * self.factors + factor

    def get(self, path: int, time: int, factor: int) -> float:
        return float(self.data[path, time, factor])

    def __eq__(self, other) -> bool:
        return (type(self) is type(other) and (self.paths, self.times, self.factors) ==
                (other.paths, other.times, other.factors) and np.array_equal(self.data, other.data))


class BrownianBlock(_Cube):
    """pathcube.rs:19-44: correlated standard normals, row-major [path][time][factor]."""

    def cell(self, path: int, time: int) -> np.ndarray:
        return self.data[path, time]


class PathCube(_Cube):
    """pathcube.rs:78-137: GBM spots, row-major [path][time][factor]."""

    def write_archive(self, path) -> None:
        """serde_json layout {"paths","times","factors","data"} (pathcube.rs:89-93)."""
        body = ",".join(_ryu32(float(v)) for v in self.data.reshape(-1))
        with open(path, "w") as fh:
            fh.write('{"paths":%d,"times":%d,"factors":%d,"data":[%s]}'
                     % (self.paths, self.times, self.factors, body))

    @classmethod
    def read_archive(cls, path) -> "PathCube":
        with open(path) as fh:
            doc = json.load(fh)
        return cls(doc["paths"], doc["times"], doc["factors"], np.asarray(doc["data"], dtype=np.float32))

    def terminal_worst(self, initial_spots: Sequence[float]) -> np.ndarray:
        """pathcube.rs:101-109: min over factors of S_T / initial, in f32."""
        init = np.asarray(initial_spots, dtype=np.float32)
        ratio = self.data[:, self.times - 1, :] / init
        out = np.full(self.paths, np.inf, dtype=np.float32)
        for f in range(self.factors):
            out = np.minimum(out, ratio[:, f])
        return out

    def worst_percentile_threshold(self, initial_spots: Sequence[float], percentile: float) -> float:
        """pathcube.rs:111-119."""
        worst = np.sort(self.terminal_worst(initial_spots), kind="stable")
        index = min(max(math.ceil(len(worst) * percentile) - 1, 0), len(worst) - 1)
        return float(worst[index])

    def terminal(self):
        """pathcube.rs:127-136: [factor][path] terminal spots."""
        return [self.data[:, self.times - 1, f].copy() for f in range(self.factors)]


def _f32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _run(seed, paths, times, factors, correlation, s0, vols, rate, q, dt, in_block, want_block, want_cube):
    gpu_backend()
    corr = None if correlation is None else N.darr(np.asarray(correlation, dtype=np.float64).reshape(-1))
    arrs = [None if v is None else N.darr(v) for v in (s0, vols, q)]
    blk = np.zeros((paths, times, factors), dtype=np.float32) if want_block else None
    cube = np.zeros((paths, times, factors), dtype=np.float32) if want_cube else None
    inb = None if in_block is None else np.ascontiguousarray(in_block, dtype=np.float32)
    err = N.Error()
    N.check(N.lib().riga_path_cube(
        N.Device.get().handle, seed, paths, times, factors,
        None if corr is None else N.dptr(corr),
        *[None if a is None else N.dptr(a) for a in arrs[:2]], rate,
        None if arrs[2] is None else N.dptr(arrs[2]), dt,
        None if inb is None else _f32p(inb), None if blk is None else _f32p(blk),
        None if cube is None else _f32p(cube), C.byref(err)), err)
    return blk, cube


def generate_correlated_block(seed: int, paths: int, times: int, correlation) -> BrownianBlock:
    """pathcube.rs:47-76 on the device."""
    F = len(correlation)
    blk, _ = _run(seed, paths, times, F, correlation, None, None, 0.0, None, 0.0, None, True, False)
    return BrownianBlock(paths, times, F, blk)


def gbm_evolve(s0, vols, rate: float, dividend_yields, dt: float, block: BrownianBlock) -> PathCube:
    """pathcube.rs:146-187 on the device."""
    F = len(s0)
    if len(vols) != F or len(dividend_yields) != F:
        raise RigaError("shape_mismatch", "s0, vols and dividend_yields must have equal length")
    if block.factors != F:
        raise RigaError("shape_mismatch", "Brownian block factor count must match the spot count")
    _, cube = _run(0, block.paths, block.times, F, None, s0, vols, rate, dividend_yields, dt, block.data,
                   False, True)
    return PathCube(block.paths, block.times, F, cube)


def path_cube(seed: int, paths: int, times: int, correlation, s0, vols, rate: float, dividend_yields,
              dt: float, with_block: bool = False):
    """Block and cube in one device pass (the block is never stored unless asked for)."""
    F = len(correlation)
    if len(s0) != F or len(vols) != F or len(dividend_yields) != F:
        raise RigaError("shape_mismatch", "s0, vols and dividend_yields must have equal length")
    blk, cube = _run(seed, paths, times, F, correlation, s0, vols, rate, dividend_yields, dt, None,
                     with_block, True)
    out = PathCube(paths, times, F, cube)
    return (out, BrownianBlock(paths, times, F, blk)) if with_block else out
