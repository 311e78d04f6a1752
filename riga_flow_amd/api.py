"""The pricing API of the gfx950 backend, mirroring riga-engine's surface.

    gpu_enabled()                 crates/riga-engine/src/lib.rs:205-207
    gpu_backend()                 crates/riga-engine/src/lib.rs:193-202 (fails loudly)
    device_name(), arch()         crates/riga-engine/src/gpu.rs:504-512
    validate(request)             crates/riga-engine/src/lib.rs:209-321 (F factors)
    price_put_leg(request)        crates/riga-engine/src/lib.rs:183-186
    price_batch(requests)         many requests in one launch (CRN items)
    replay_paths(request)         crates/riga-engine/src/explain.rs:90-316
    build_local_vol(request)      crates/riga-engine/src/lib.rs:326-343

``PriceResult`` is assembled exactly as note.rs:262-372 does (stats, ledger,
evidence strings, trace events, determinism envelope with the market
snapshot id of riga-core/src/lib.rs:47-55).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import math
from decimal import Decimal
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .types import (DeterminismEnvelope, LedgerEntry, LegAttribution, LocalVolGrid, PriceRequest,
                    PriceResult, RigaError, TransitionStats)

PRECISIONS = {"fp64": N.FP64, "fp32": N.FP32}


def gpu_enabled() -> bool:
    """This build always advertises the HIP backend."""
    return True


def gpu_backend() -> None:
    """Initialise the process-global device or raise ``gpu_unavailable``."""
    N.Device.get()


def device_name() -> str:
    return N.Device.get().name()


def arch() -> str:
    return N.Device.get().arch()


# ------------------------------------------------------------------ packing
class _Packed:
    """A ``riga_request`` plus the numpy buffers it points into."""

    def __init__(self, request: PriceRequest, max_dimension: int = 512):
        t, m, c = request.trade, request.market, request.config
        F = len(m.spots)
        keep = []

        def arr(v):
            a = N.darr(v)
            keep.append(a)
            return N.dptr(a)

        r = N.Request()
        r.factors = F
        r.spots = arr(m.spots)
        r.vols = arr(m.vols)
        r.dividend_yields = arr(m.dividend_yields)
        r.correlation = arr(np.asarray(m.correlation, dtype=np.float64).reshape(-1))
        r.risk_free_rate = m.risk_free_rate
        r.initial_spots = arr(t.initial_spots)
        r.maturity_years = t.maturity_years
        r.steps = t.steps
        r.strike = t.strike
        r.knock_in_barrier = t.knock_in_barrier
        r.has_knock_out = t.knock_out_barrier is not None
        r.knock_out_barrier = t.knock_out_barrier if t.knock_out_barrier is not None else 0.0
        r.has_coupon_barrier = t.coupon_barrier is not None
        r.coupon_barrier = t.coupon_barrier if t.coupon_barrier is not None else 0.0
        r.has_coupon_rate = t.coupon_rate is not None
        r.coupon_rate = t.coupon_rate if t.coupon_rate is not None else 0.0
        r.memory_coupon = bool(t.memory_coupon)
        r.coupon_periods = -1 if t.coupon_periods is None else t.coupon_periods
        r.fixed_first_coupon = bool(t.fixed_first_coupon)
        r.call_start_step = -1 if t.call_start_step is None else t.call_start_step
        r.paths = c.paths
        r.path_block_size = c.path_block_size
        r.seed = c.seed
        r.trace_level = N.TRACE[c.trace_level]
        r.ladder = {"none": N.VR_NONE, "halton": N.VR_HALTON, "sobol": N.VR_SOBOL}[
            c.variance_reduction.kind]
        r.ladder_dimension = c.variance_reduction.dimension
        r.max_dimension = max_dimension
        lv = request.local_vol
        if lv is not None:
            if lv.resident is not None:
                if len(lv.resident) != F:
                    raise RigaError("invalid_local_vol",
                                    "local_vol.resident must be aligned with market.spots")
                hs = (C.c_void_p * F)(*[s.handle.value for s in lv.resident])
                keep.append(hs)
                r.surfaces = C.cast(hs, C.POINTER(C.c_void_p))
            elif lv.calibrated is not None:
                grids = (N.LvGrid * F)()
                for f, g in enumerate(lv.calibrated):
                    v = np.asarray(g.v, dtype=np.float64)
                    grids[f] = N.LvGrid(len(g.t_axis), len(g.log_s), arr(g.t_axis), arr(g.log_s),
                                        arr(v.reshape(-1)))
                keep.append(grids)
                r.calibrated = C.cast(grids, C.POINTER(N.LvGrid))
            else:
                if len(lv.grids) != F:
                    raise RigaError("invalid_local_vol",
                                    "local_vol.grids must be aligned with market.spots")
                grids = (N.ImpliedGrid * F)()
                for f, g in enumerate(lv.grids):
                    v = np.asarray(g.vols, dtype=np.float64)
                    grids[f] = N.ImpliedGrid(len(g.maturities_years), len(g.strikes),
                                             arr(g.maturities_years), arr(g.strikes),
                                             arr(v.reshape(-1)))
                keep.append(grids)
                r.implied = C.cast(grids, C.POINTER(N.ImpliedGrid))
            r.grid_x = lv.grid_x
        self.r = r
        self.keep = keep


def _check_as_of(request: PriceRequest) -> None:
    """lib.rs:213-224: as_of is required and must match the market snapshot."""
    if not request.as_of:
        raise RigaError("missing_as_of", "as_of is required")
    if request.as_of != request.market.as_of:
        raise RigaError("as_of_mismatch", "request as_of must match the market snapshot as_of")


def validate(request: PriceRequest, max_dimension: int = 512) -> None:
    _check_as_of(request)
    p = _Packed(request, max_dimension)
    err = N.Error()
    N.check(N.lib().riga_validate(C.byref(p.r), C.byref(err)), err)


# ------------------------------------------------------------------ JSON / snapshot id
def _ryu(x: float) -> str:
    """Shortest round-trip float text in the layout serde_json (ryu) prints."""
    if x != x or x in (math.inf, -math.inf):
        return "null"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    t = Decimal(repr(abs(x))).as_tuple()  # repr is the shortest round-trip digit string
    digits = "".join(map(str, t.digits)).lstrip("0")
    k = t.exponent
    stripped = digits.rstrip("0")
    k += len(digits) - len(stripped)
    digits = stripped
    length = len(digits)
    kk = length + k  # 10^(kk-1) <= |x| < 10^kk
    if 0 <= k and kk <= 16:
        return sign + digits + "0" * k + ".0"
    if 0 < kk <= 16:
        return sign + digits[:kk] + "." + digits[kk:]
    if -5 < kk <= 0:
        return sign + "0." + "0" * (-kk) + digits
    if length == 1:
        return sign + digits + "e" + str(kk - 1)
    return sign + digits[0] + "." + digits[1:] + "e" + str(kk - 1)


def _json(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return _ryu(v)
    if isinstance(v, str):
        out = ['"']
        for ch in v:
            if ch == '"':
                out.append('\\"')
            elif ch == "\\":
                out.append("\\\\")
            elif ch == "\n":
                out.append("\\n")
            elif ch == "\r":
                out.append("\\r")
            elif ch == "\t":
                out.append("\\t")
            elif ord(ch) < 0x20:
                out.append("\\u%04x" % ord(ch))
            else:
                out.append(ch)
        out.append('"')
        return "".join(out)
    if isinstance(v, (list, tuple)):
        return "[" + ",".join(_json(x) for x in v) + "]"
    if isinstance(v, dict):
        return "{" + ",".join(_json(k) + ":" + _json(x) for k, x in v.items()) + "}"
    raise TypeError(type(v))


def market_snapshot_id(market) -> str:
    """``MarketSnapshot::snapshot_id`` (riga-core/src/lib.rs:47-55)."""
    doc = {
        "as_of": market.as_of,
        "spots": [float(x) for x in market.spots],
        "vols": [float(x) for x in market.vols],
        "risk_free_rate": float(market.risk_free_rate),
        "dividend_yields": [float(x) for x in market.dividend_yields],
        "correlation": [[float(x) for x in row] for row in market.correlation],
    }
    return "sha256:" + hashlib.sha256(_json(doc).encode()).hexdigest()


# ------------------------------------------------------------------ result assembly
def _assemble(request: PriceRequest, r: N.Result) -> PriceResult:
    """note.rs:262-372 from the device aggregates."""
    t = request.trade
    paths = int(r.paths)
    steps = t.steps
    dt = t.maturity_years / steps
    call_barrier = t.knock_out_barrier if t.knock_out_barrier is not None else math.inf
    coupon_barrier = t.coupon_barrier if t.coupon_barrier is not None else math.inf
    snapshot_id = market_snapshot_id(request.market)
    ko, ki, cp = int(r.ko_count), int(r.ki_count), int(r.coupon_path_count)
    ko_prob = ko / paths
    cp_prob = cp / paths
    stats = [
        TransitionStats("daily_global_ko", paths, ko, ko_prob, r.ko_mean, r.ko_variance),
        TransitionStats("final_fixing_eki", paths, ki, r.ki_probability, r.ki_mean, r.ki_variance),
        TransitionStats("memory_coupon", paths, cp, cp_prob, r.coupon_mean, r.coupon_variance),
    ]
    evidence = [f"trade:{t.trade_id}", f"market:{snapshot_id}", f"seed:{request.config.seed}"]
    ledger = [
        LedgerEntry("daily_global_ko",
                    f"all names >= {call_barrier:.6f} during callable observations",
                    paths, ko, r.leg_funding * ko_prob, 0.0, 0.0, list(evidence)),
        LedgerEntry("final_fixing_eki",
                    f"terminal worst <= {t.knock_in_barrier:.6f} and no prior autocall; "
                    f"put strike {t.strike:.6f}",
                    paths - ko, int(r.put_eligible_count), -r.leg_put, 0.0, 0.0, list(evidence)),
        LedgerEntry("memory_coupon",
                    f"daily worst-of >= coupon barrier {coupon_barrier:.6f}; memory carry "
                    f"enabled={'true' if t.memory_coupon else 'false'}",
                    paths, cp, (r.leg_coupon + r.leg_memory_coupon) / 10.0, 0.0, 0.0,
                    list(evidence)),
    ]
    trace = []
    if request.config.trace_level != "off":
        for i in range(r.n_trace):
            step = int(r.trace_step[i])
            time = (step + 1) * dt
            trace.append(f"daily_global_ko:path={int(r.trace_path[i])}:t={time:.6f}")
    legs = LegAttribution(0.0, r.leg_coupon, r.leg_memory_coupon, r.leg_put, r.leg_funding, 0.0,
                          r.leg_total)
    res = PriceResult(r.pv, r.mc_error, r.ki_probability, legs, stats, ledger, trace,
                      DeterminismEnvelope(request.config.seed, request.as_of, snapshot_id))
    res.device_ms = r.device_ms
    return res


def result_to_json(res: PriceResult) -> dict:
    """The serde field layout of ``PriceResult`` (riga-core/src/lib.rs:243-255)."""
    out = {
        "pv": res.pv, "mc_error": res.mc_error, "ki_probability": res.ki_probability,
        "legs": dict(vars(res.legs)),
        "stats": [dict(vars(s)) for s in res.stats],
        "ledger": [dict(vars(l)) for l in res.ledger],
    }
    if res.trace_events:
        out["trace_events"] = list(res.trace_events)
    out["determinism"] = dict(vars(res.determinism))
    return out


# ------------------------------------------------------------------ pricing
def _owen_mode(owen_table) -> int:
    """riga_launch_config.owen_table: None lets the library decide per seed (a table it
    holds already, or one worth building for the plan's points), True forces the
    coordinate table, an int k > 1 forces one covering at most the first k dimensions,
    False hashes all 32 Owen levels (the same integers)."""
    if owen_table is None:
        return 0
    if owen_table is True or owen_table is False:
        return 1 if owen_table else -1
    return max(2, int(owen_table))


def price_batch(requests: Sequence[PriceRequest], precision: str = "fp64",
                paths_per_thread: int = 0, max_dimension: int = 512,
                raw: bool = False, owen_table: Optional[bool] = None, normal_table: int = 0,
                waves_per_block: int = 0, fp32_full_ladder: bool = False, stack_split: int = 0,
                persistent: bool = False):
    """Price many requests in one launch; CRN scenarios share their Sobol points."""
    gpu_backend()
    for q in requests:
        _check_as_of(q)
    packed = [_Packed(q, max_dimension) for q in requests]
    n = len(packed)
    arr = (N.Request * n)(*[p.r for p in packed])
    res = (N.Result * n)()
    cfg = N.LaunchConfig(paths_per_thread, 0, _owen_mode(owen_table), normal_table, 0,
                         waves_per_block, 0, 1 if fp32_full_ladder else 0, stack_split,
                         1 if persistent else 0)
    err = N.Error()
    dev = N.Device.get()
    N.check(N.lib().riga_price_batch(dev.handle, arr, n, PRECISIONS[precision], C.byref(cfg), res,
                                     C.byref(err)), err)
    if raw:
        return list(res)
    return [_assemble(q, res[i]) for i, q in enumerate(requests)]


def autotune(requests: Sequence[PriceRequest], precision: str = "fp64", use_graph: bool = True,
             max_dimension: int = 512, owen_table: Optional[bool] = None, normal_table: int = 0,
             reps: int = 2) -> dict:
    """Time the batch at every launch shape (paths per lane x waves per workgroup) and
    keep the fastest for batches of its shape: plans and batches created afterwards
    with the default (0) shape use it.  Returns the trials and the choice."""
    gpu_backend()
    packed = [_Packed(q, max_dimension) for q in requests]
    n = len(packed)
    arr = (N.Request * n)(*[p.r for p in packed])
    cfg = N.LaunchConfig(0, 1 if use_graph else 0, _owen_mode(owen_table), normal_table)
    rep = N.TuneReport()
    err = N.Error()
    N.check(N.lib().riga_autotune(N.Device.get().handle, arr, n, PRECISIONS[precision],
                                  C.byref(cfg), reps, C.byref(rep), C.byref(err)), err)
    trials = [(t.paths_per_thread, t.waves_per_block, t.device_ms)
              for t in rep.trials[:rep.n_trials]]
    b = trials[rep.best]
    return {"paths_per_thread": b[0], "waves_per_block": b[1], "device_ms": b[2],
            "trials": trials}


class Plan:
    """A batch resident in device memory, launched any number of times."""

    def __init__(self, requests: Sequence[PriceRequest], precision: str = "fp64",
                 paths_per_thread: int = 0, use_graph: bool = False, max_dimension: int = 512,
                 owen_table: Optional[bool] = None, normal_table: int = 0, waves_per_block: int = 0,
                 serial_launches: bool = False, fp32_full_ladder: bool = False,
                 stack_split: int = 0, persistent: bool = False):
        gpu_backend()
        for q in requests:
            _check_as_of(q)
        self.requests = list(requests)
        packed = [_Packed(q, max_dimension) for q in requests]
        n = len(packed)
        arr = (N.Request * n)(*[p.r for p in packed])
        cfg = N.LaunchConfig(paths_per_thread, 1 if use_graph else 0, _owen_mode(owen_table),
                             normal_table, 0, waves_per_block, 1 if serial_launches else 0,
                             1 if fp32_full_ladder else 0, stack_split, 1 if persistent else 0)
        err = N.Error()
        h = C.c_void_p()
        N.check(N.lib().riga_plan_create(N.Device.get().handle, arr, n, PRECISIONS[precision],
                                         C.byref(cfg), C.byref(h), C.byref(err)), err)
        self.handle = h
        self._res = (N.Result * n)()
        info = N.PlanInfo()
        N.lib().riga_plan_query(h, C.byref(info))
        self.info = info

    @property
    def path_steps(self) -> float:
        return self.info.path_steps

    def launch(self) -> None:
        err = N.Error()
        N.check(N.lib().riga_plan_launch(self.handle, C.byref(err)), err)

    def wait(self, raw: bool = False):
        err = N.Error()
        N.check(N.lib().riga_plan_wait(self.handle, self._res, C.byref(err)), err)
        if raw:
            return list(self._res)
        return [_assemble(q, self._res[i]) for i, q in enumerate(self.requests)]

    def run(self, raw: bool = False):
        self.launch()
        return self.wait(raw)

    def close(self) -> None:
        if getattr(self, "handle", None):
            N.lib().riga_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def price_put_leg(request: PriceRequest, precision: str = "fp64",
                  max_dimension: int = 512) -> PriceResult:
    """``price_put_leg`` (lib.rs:183-186) on the device."""
    return price_batch([request], precision, max_dimension=max_dimension)[0]


def price_trade(request_json, precision: str = "fp64", max_dimension: int = 512) -> dict:
    """riga-mcp's ``price_trade`` (main.rs:1821-1831): a ``PriceRequest`` as JSON (a
    dict or text) in, the ``PriceResult`` JSON out; a request that does not decode is
    ``invalid_price_request``, engine errors keep their codes."""
    import json as _json
    from .types import request_from_json
    if isinstance(request_json, (str, bytes)):
        try:
            request_json = _json.loads(request_json)
        except ValueError as e:
            raise RigaError("invalid_price_request", f"request must match riga_engine::PriceRequest: {e}")
    req = request_from_json(request_json)
    return result_to_json(price_put_leg(req, precision, max_dimension=max_dimension))


def replay_paths(request: PriceRequest, precision: str = "fp64",
                 max_dimension: int = 512) -> Tuple[np.ndarray, PriceResult]:
    """Per-path records (explain.rs:90-316); the only call that writes paths to HBM."""
    gpu_backend()
    _check_as_of(request)
    p = _Packed(request, max_dimension)
    recs = (N.PathRecord * request.config.paths)()
    res = N.Result()
    err = N.Error()
    N.check(N.lib().riga_replay_paths(N.Device.get().handle, C.byref(p.r), PRECISIONS[precision],
                                      recs, C.byref(res), C.byref(err)), err)
    arr = np.frombuffer(recs, dtype=N.RECORD_DTYPE, count=request.config.paths).copy()
    return arr, _assemble(request, res)


class Surface:
    """A calibrated local-vol grid copied to device memory once (riga_surface_create),
    so requests name it instead of sending the grid (the reference keeps its LV
    rows resident the same way, crates/riga-engine/src/gpu.rs:280-327).  Plans and
    tickets that use it hold a reference through their requests."""

    def __init__(self, grid: LocalVolGrid):
        dev = N.Device.get()
        v = np.ascontiguousarray(np.asarray(grid.v, dtype=np.float64).reshape(-1))
        t = N.darr(grid.t_axis)
        x = N.darr(grid.log_s)
        g = N.LvGrid(len(grid.t_axis), len(grid.log_s), N.dptr(t), N.dptr(x), N.dptr(v))
        h = C.c_void_p()
        err = N.Error()
        N.check(N.lib().riga_surface_create(dev.handle, C.byref(g), C.byref(h), C.byref(err)), err)
        self.handle = h
        self.grid = grid

    def close(self) -> None:
        if getattr(self, "handle", None) and self.handle.value:
            N.lib().riga_surface_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def resident(request: PriceRequest, cache: Optional[dict] = None) -> PriceRequest:
    """The request with its calibrated grids made device-resident (one Surface per
    distinct grid object, shared through ``cache``)."""
    import dataclasses
    lv = request.local_vol
    if lv is None or lv.calibrated is None:
        return request
    cache = {} if cache is None else cache
    surfs = []
    for g in lv.calibrated:
        if id(g) not in cache:
            cache[id(g)] = (g, Surface(g))
        surfs.append(cache[id(g)][1])
    return dataclasses.replace(request, local_vol=dataclasses.replace(lv, resident=surfs))


def lv_calibrate(maturities, strikes, vols, spot: float, rate: float,
                 grid_x: int = 96) -> LocalVolGrid:
    """``DupireLocalVol::calibrate`` (riga-market/src/lib.rs:210-299) on the host."""
    m, k = N.darr(maturities), N.darr(strikes)
    v = N.darr(vols).reshape(-1)
    g = N.ImpliedGrid(len(m), len(k), N.dptr(m), N.dptr(k), N.dptr(v))
    t = np.zeros(len(m))
    x = np.zeros(grid_x)
    out = np.zeros(len(m) * grid_x)
    err = N.Error()
    N.check(N.lib().riga_lv_calibrate(C.byref(g), spot, rate, grid_x, N.dptr(t), N.dptr(x),
                                      N.dptr(out), C.byref(err)), err)
    return LocalVolGrid(t.tolist(), x.tolist(), out.reshape(len(m), grid_x).tolist())


def lv_sigma(grid: LocalVolGrid, t: float, s: float) -> float:
    ta, xa = N.darr(grid.t_axis), N.darr(grid.log_s)
    va = N.darr(grid.v).reshape(-1)
    g = N.LvGrid(len(ta), len(xa), N.dptr(ta), N.dptr(xa), N.dptr(va))
    return N.lib().riga_lv_sigma(C.byref(g), t, s)


def lv_row(grid: LocalVolGrid, t: float) -> List[float]:
    """``DupireLocalVol::row`` (riga-market/src/lib.rs:346-364)."""
    ta = grid.t_axis
    if len(ta) == 1:
        return list(grid.v[0])
    tc = min(max(t, ta[0]), ta[-1])
    i = 0
    while i + 1 < len(ta) - 1 and tc > ta[i + 1]:
        i += 1
    span = ta[i + 1] - ta[i]
    w = (tc - ta[i]) / span if span > 0.0 else 0.0
    return [(1.0 - w) * a + w * b for a, b in zip(grid.v[i], grid.v[i + 1])]


def lv_resample(grid: LocalVolGrid, times: Sequence[float]) -> LocalVolGrid:
    """Rows of ``grid`` at ``times`` with ``row(t)`` semantics (a 64 x n_x grid for the books)."""
    return LocalVolGrid([float(t) for t in times], list(grid.log_s), [lv_row(grid, t) for t in times])


def build_local_vol(request: PriceRequest) -> Optional[List[LocalVolGrid]]:
    """``build_local_vol`` (lib.rs:326-343): one calibrated surface per factor."""
    spec = request.local_vol
    if spec is None:
        return None
    if spec.calibrated is not None:
        return spec.calibrated
    return [lv_calibrate(g.maturities_years, g.strikes, g.vols, request.market.spots[f],
                         request.market.risk_free_rate, spec.grid_x)
            for f, g in enumerate(spec.grids)]


def debug_sobol(seed: int, ladder_dims: int, first_point: int, n_points: int, first_dim: int,
                n_dims: int):
    """Device Sobol integers (unscrambled, Owen-scrambled) and FP64 normals."""
    gpu_backend()
    cells = n_points * n_dims
    u = np.zeros(cells, dtype=np.uint32)
    s = np.zeros(cells, dtype=np.uint32)
    z = np.zeros(cells, dtype=np.float64)
    err = N.Error()
    N.check(N.lib().riga_debug_sobol(
        N.Device.get().handle, seed, ladder_dims, first_point, n_points, first_dim, n_dims,
        u.ctypes.data_as(C.POINTER(C.c_uint32)), s.ctypes.data_as(C.POINTER(C.c_uint32)),
        N.dptr(z), C.byref(err)), err)
    return (u.reshape(n_points, n_dims), s.reshape(n_points, n_dims), z.reshape(n_points, n_dims))


def debug_sobol_f32(seed: int, ladder_dims: int, first_point: int, n_points: int, n_nodes: int,
                    factors: int, owen_table: bool = True, full_ladder: bool = False):
    """The FP32 tier's ladder as its path kernels run it: the integers it turns into
    uniforms (24 Owen levels with low byte 0x80, all 32 levels within 2^-10 of 0 or 1)
    and its FP32 normals, [point][node * factors + f]."""
    gpu_backend()
    dims = n_nodes * factors
    s = np.zeros(n_points * dims, dtype=np.uint32)
    z = np.zeros(n_points * dims, dtype=np.float32)
    err = N.Error()
    N.check(N.lib().riga_debug_sobol_f32(
        N.Device.get().handle, seed, ladder_dims, first_point, n_points, n_nodes, factors,
        (1 if owen_table else 0) | (2 if full_ladder else 0), s.ctypes.data_as(C.POINTER(C.c_uint32)),
        z.ctypes.data_as(C.POINTER(C.c_float)), C.byref(err)), err)
    return s.reshape(n_points, dims), z.reshape(n_points, dims)


def debug_halton(seed: int, ladder_dims: int, first_point: int, n_points: int, first_dim: int,
                 n_dims: int):
    """The device's Halton ladder: permuted digits [point][dim][32] (least significant
    first, 0xFFFFFFFF after the last), radical inverses and FP64 normals [point][dim]."""
    gpu_backend()
    cells = n_points * n_dims
    dig = np.zeros(cells * 32, dtype=np.uint32)
    u = np.zeros(cells, dtype=np.float64)
    z = np.zeros(cells, dtype=np.float64)
    err = N.Error()
    N.check(N.lib().riga_debug_halton(
        N.Device.get().handle, seed, ladder_dims, first_point, n_points, first_dim, n_dims,
        dig.ctypes.data_as(C.POINTER(C.c_uint32)), N.dptr(u), N.dptr(z), C.byref(err)), err)
    return (dig.reshape(n_points, n_dims, 32), u.reshape(n_points, n_dims), z.reshape(n_points, n_dims))


def debug_exp_log(x) -> Tuple[np.ndarray, np.ndarray]:
    """The FP64 tier's exp and log of each x, as its path kernels evaluate them."""
    gpu_backend()
    xs = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    e = np.zeros_like(xs)
    lg = np.zeros_like(xs)
    err = N.Error()
    N.check(N.lib().riga_debug_exp_log(N.Device.get().handle, N.dptr(xs), xs.size, N.dptr(e), N.dptr(lg),
                                       C.byref(err)), err)
    return e, lg


def debug_mc_starts(seed: int, draws_per_path: int, n_paths: int) -> np.ndarray:
    """Device stream parse of the Monte Carlo lane: the u64 word of
    ``ChaCha8Rng::seed_from_u64(seed)`` where each path's first StandardNormal
    draw begins (path p starts at draw ``p * draws_per_path``)."""
    gpu_backend()
    out = np.zeros(n_paths, dtype=np.uint64)
    err = N.Error()
    N.check(N.lib().riga_debug_mc_starts(
        N.Device.get().handle, seed, draws_per_path, n_paths,
        out.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(err)), err)
    return out
