"""Explainability on the device: deterministic replay of individual paths with
their coupon strips, single-path reconstruction and the worst-percentile
PathCube query (crates/riga-engine/src/explain.rs).

The pricing launches never write per-path data; these calls re-run the same
simulation with records (and, for coupon strips, each path's daily worst-of)
written to device memory for the paths asked about only.  Every path keeps its
identity: on the QMC lanes path p is Sobol/Halton point p + 1, on the ChaCha8
lane it starts at its own word of the request's stream, so a single path
replays alone exactly as it does inside the full run.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N
from .api import PRECISIONS, _check_as_of, _Packed, gpu_backend
from .types import PriceRequest, RigaError


@dataclass
class PathIdentity:
    """explain.rs:22-30."""

    path_id: int
    rng_seed: int
    sobol_index: int = 0


@dataclass
class CouponPeriodRecord:
    """explain.rs:35-55 (nominal amounts; pv = paid * discount_factor)."""

    period: int
    years: float
    qualified_steps: int
    period_steps: int
    accrual: float
    memory_paid: float
    paid: float
    memory_accrued: float
    discount_factor: float
    pv: float
    knocked_out: bool


@dataclass
class PathRecord:
    """explain.rs:57-75."""

    identity: PathIdentity
    worst_final: float
    knocked_out: bool
    ko_time: float
    ki_hit: bool
    put_payoff: float
    coupon_pv: float
    coupon_by_period: List[float] = field(default_factory=list)
    coupon_ledger: List[CouponPeriodRecord] = field(default_factory=list)
    funding: float = 0.0
    pv: float = 0.0


@dataclass
class PathCubeQuery:
    """explain.rs:77-88."""

    percentile: float
    threshold_worst_final: float
    selected: int
    total: int
    sample: List[PathRecord]


def _periods(request: PriceRequest) -> int:
    p = request.trade.coupon_periods
    return 9 if p is None else max(1, p)


def _records(request: PriceRequest, ids: Sequence[int], recs, rows) -> List[PathRecord]:
    """explain.rs:263-313: the record fields from the device's record and strip."""
    dt = request.trade.maturity_years / request.trade.steps
    qmc = request.config.variance_reduction.kind != "none"
    P = _periods(request)
    out = []
    for i, pid in enumerate(ids):
        r = recs[i]
        ko = r.ko_step >= 0
        ledger = []
        if rows is not None:
            for q in range(P):
                w = rows[i * P + q]
                ledger.append(CouponPeriodRecord(int(w.period), w.years, int(w.qualified_steps),
                                                 int(w.period_steps), w.accrual, w.memory_paid, w.paid,
                                                 w.memory_accrued, w.discount_factor, w.pv, bool(w.knocked_out)))
        out.append(PathRecord(
            identity=PathIdentity(int(pid), request.config.seed, int(pid) + 1 if qmc else 0),
            worst_final=r.final_worst, knocked_out=ko,
            ko_time=(r.ko_step + 1) * dt if ko else math.nan,
            ki_hit=bool(r.hit_ki), put_payoff=r.put,
            coupon_pv=r.coupon_pv + r.memory_pv,  # coupon_nominal, note.rs:252
            coupon_by_period=[c.pv for c in ledger], coupon_ledger=ledger,
            funding=r.funding, pv=r.pv))
    return out


def _replay(request: PriceRequest, ids: Optional[Sequence[int]], precision: str, ledger: bool,
            max_dimension: int):
    gpu_backend()
    _check_as_of(request)
    p = _Packed(request, max_dimension)
    n = request.config.paths if ids is None else len(ids)
    if n == 0:
        raise RigaError("invalid_argument", "no paths to replay")
    idarr = None if ids is None else np.ascontiguousarray(np.asarray(ids, dtype=np.uint64))
    recs = (N.PathRecord * n)()
    rows = (N.CouponRow * (n * _periods(request)))() if ledger else None
    err = N.Error()
    N.check(N.lib().riga_replay_ledger(
        N.Device.get().handle, C.byref(p.r), PRECISIONS[precision],
        None if idarr is None else idarr.ctypes.data_as(C.POINTER(C.c_uint64)), n, recs, rows,
        C.byref(err)), err)
    return recs, rows


def run_paths(request: PriceRequest, precision: str = "fp64", max_dimension: int = 512,
              ledger: bool = True) -> List[PathRecord]:
    """explain.rs:90-316: every path's record, with its coupon strip."""
    recs, rows = _replay(request, None, precision, ledger, max_dimension)
    return _records(request, range(request.config.paths), recs, rows)


def reconstruct_path(request: PriceRequest, path_id: int, precision: str = "fp64",
                     max_dimension: int = 512) -> PathRecord:
    """explain.rs:319-334: one path from its identity alone."""
    if path_id >= request.config.paths or path_id < 0:
        raise RigaError("path_out_of_range", f"path_id {path_id} is outside the configured path count "
                                             f"{request.config.paths}")
    recs, rows = _replay(request, [path_id], precision, True, max_dimension)
    return _records(request, [path_id], recs, rows)[0]


def query_worst_percentile(request: PriceRequest, percentile: float, max_sample: int,
                           precision: str = "fp64", max_dimension: int = 512) -> PathCubeQuery:
    """explain.rs:336-372: the worst ``percentile`` fraction of paths by final worst-of."""
    if not (0.0 <= percentile <= 1.0):
        raise RigaError("invalid_percentile", "percentile must be within [0, 1]")
    gpu_backend()
    _check_as_of(request)
    p = _Packed(request, max_dimension)
    cap = max(1, max_sample)
    ids = np.zeros(cap, dtype=np.uint64)
    recs = (N.PathRecord * cap)()
    rows = (N.CouponRow * (cap * _periods(request)))()
    q = N.WorstQuery()
    err = N.Error()
    N.check(N.lib().riga_query_worst_percentile(
        N.Device.get().handle, C.byref(p.r), PRECISIONS[precision], percentile, max_sample, C.byref(q),
        ids.ctypes.data_as(C.POINTER(C.c_uint64)), recs, rows, C.byref(err)), err)
    k = int(q.n_sample)
    sample = _records(request, [int(i) for i in ids[:k]], recs, rows) if k else []
    return PathCubeQuery(q.percentile, q.threshold_worst_final, int(q.selected), int(q.total), sample)
