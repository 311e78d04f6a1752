"""Serving loop (continuous batching) and its open-loop load generator.

``Server`` wraps the device-owning scheduler of libriga_hip.so
(riga_server_*): ``submit`` copies a ticket of requests and returns at once;
tickets are admitted in arrival order while the device-memory budget allows,
batched into one plan, run on their own stream (overlapping plans already on
the device) and shed after the queue limit.  This replaces the reference's
blocking per-call pricing in its MCP handlers (riga-mcp/src/main.rs:1821-1831,
2005-2035) and its ``price_book_async`` tasks (main.rs:1005-1042, 2541-2575).

``poisson_load`` drives a server with Poisson arrivals of requests shaped like
the reference's tool calls: 1-64 trades of a 20-asset local-vol universe, each
with 1-10 CRN bump scenarios, 16k/65k/262k/1M paths, FP32 or FP64 tier.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N
from . import api, workloads as W
from .types import LocalVolGrid, LocalVolSpec, MarketSnapshot, PriceRequest, RigaError

DONE, SHED, FAILED = 2, 3, 4


@dataclass
class TicketResult:
    ticket: int
    state: int
    latency_s: float
    submit_s: float
    start_s: float
    end_s: float
    n_requests: int
    batch_tickets: int
    path_steps: float
    device_start_s: float = -1.0  # its plan's span on the device clock (s since the context began)
    device_end_s: float = -1.0
    results: list = field(default_factory=list)
    error: Optional[RigaError] = None


class Server:
    def __init__(self, hbm_budget_gib: float = 64.0, max_inflight: int = 2,
                 max_batch_requests: int = 4096, max_queue_seconds: float = 0.0,
                 paths_per_thread: int = 1, normal_table: int = 0, max_dimension: int = 7000,
                 max_batch_path_steps: float = 0.0):
        api.gpu_backend()
        cfg = N.ServerConfig(int(hbm_budget_gib * (1 << 30)), max_inflight, max_batch_requests,
                             max_queue_seconds, paths_per_thread, normal_table,
                             float(max_batch_path_steps))
        h = C.c_void_p()
        err = N.Error()
        N.check(N.lib().riga_server_create(N.Device.get().handle, C.byref(cfg), C.byref(h),
                                           C.byref(err)), err)
        self.handle = h
        self.max_dimension = max_dimension
        self._pending = {}

    def pack(self, requests: Sequence[PriceRequest]) -> "PackedTicket":
        """Serialise a ticket's requests into C structs (the client's side of the wire)."""
        return PackedTicket(list(requests), [api._Packed(q, self.max_dimension) for q in requests])

    def submit(self, requests, precision: str = "fp64") -> int:
        pt = requests if isinstance(requests, PackedTicket) else self.pack(requests)
        t = C.c_uint64()
        err = N.Error()
        N.check(N.lib().riga_server_submit(self.handle, pt.array, len(pt.requests),
                                           api.PRECISIONS[precision], C.byref(t), C.byref(err)), err)
        self._pending[t.value] = pt.requests
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0, assemble: bool = True) -> Optional[TicketResult]:
        reqs = self._pending[ticket]
        res = (N.Result * len(reqs))()
        info = N.TicketInfo()
        err = N.Error()
        rc = N.lib().riga_server_wait(self.handle, ticket, timeout, res, C.byref(info), C.byref(err))
        if rc == 1:
            return None
        del self._pending[ticket]
        out = TicketResult(ticket, info.state, info.latency_s, info.submit_s, info.start_s,
                           info.end_s, info.n_requests, info.batch_tickets, info.path_steps,
                           info.device_start_s, info.device_end_s)
        if rc == 0:
            out.results = [api._assemble(q, res[i]) for i, q in enumerate(reqs)] if assemble \
                else list(res)
        else:
            out.error = RigaError(err.code.decode() or "failed", err.message.decode())
        return out

    def stats(self) -> dict:
        s = N.ServerStats()
        N.lib().riga_server_stats_get(self.handle, C.byref(s))
        return {k: getattr(s, k) for k, _ in N.ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None):
            N.lib().riga_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PackedTicket:
    """A ticket's requests with their C structs; submit copies them, so one packed
    ticket can be submitted many times."""

    def __init__(self, requests: List[PriceRequest], packed: list):
        self.requests = requests
        self._packed = packed
        self.array = (N.Request * len(packed))(*[p.r for p in packed])


# ---------------------------------------------------------------- resident market
class Universe:
    """A 20-asset local-vol market: spots, vols, dividend yields, one
    correlation and one calibrated 64 x 128 Dupire grid per asset (5y horizon)."""

    def __init__(self, n_assets: int = 20, seed: int = 5, horizon: float = 5.0):
        rng = np.random.default_rng(seed)
        self.n = n_assets
        self.corr = W.universe_correlation(n_assets)
        self.vols = rng.uniform(0.15, 0.40, size=n_assets)
        self.q = rng.uniform(0.0, 0.03, size=n_assets)
        self.r = 0.03
        self.sigmas = [0.2 + float(x) for x in rng.uniform(-0.05, 0.05, size=n_assets)]
        self.horizon = horizon
        self.grids = [W.lv_grid_for(W.svi_implied(self.sigmas[a]), 100.0, self.r, horizon)
                      for a in range(n_assets)]
        # bumped surfaces are part of the resident market: calibrated once
        self.spot_grids = {}
        for a in range(n_assets):
            for rel in (0.01, -0.01):
                self.spot_grids[(a, rel)] = W.lv_grid_for(W.svi_implied(self.sigmas[a]),
                                                          100.0 + rel * 100.0, self.r, horizon)
        self.vega_grids = [W.lv_grid_for(W.svi_implied(self.sigmas[a], bump=0.01), 100.0, self.r,
                                         horizon) for a in range(n_assets)]
        self.surfaces = None

    def make_resident(self) -> None:
        """Copy every surface of the market (base and bumped) to device memory once;
        requests built afterwards name them instead of carrying the grids."""
        grids = list(self.grids) + list(self.spot_grids.values()) + list(self.vega_grids)
        self.surfaces = {id(g): api.Surface(g) for g in grids}

    def request(self, assets: List[int], years: float, paths: int, spot_bump=None,
                vega_asset: int = -1, trade_rng=None, trade_id: str = "t") -> PriceRequest:
        F = len(assets)
        steps = max(1, int(round(252 * years)))
        spots = [100.0] * F
        grids = [self.grids[a] for a in assets]
        if spot_bump is not None:
            f, rel = spot_bump
            spots[f] = 100.0 + rel * 100.0
            grids[f] = self.spot_grids[(assets[f], rel)]
        vols = [float(self.vols[a]) for a in assets]
        if vega_asset >= 0:
            vols[vega_asset] += 0.01
            grids[vega_asset] = self.vega_grids[assets[vega_asset]]
        rng = trade_rng or np.random.default_rng(0)
        market = MarketSnapshot(W.AS_OF, spots, vols, self.r, [float(self.q[a]) for a in assets],
                                [[float(self.corr[a][b]) for b in assets] for a in assets])
        trade = W.phoenix_trade(trade_id, F, years, steps, ko=float(rng.uniform(0.95, 1.05)),
                                coupon_barrier=float(rng.uniform(0.60, 0.80)),
                                eki=float(rng.uniform(0.50, 0.65)),
                                coupon=float(rng.uniform(0.005, 0.012)))
        res = [self.surfaces[id(g)] for g in grids] if self.surfaces else None
        return PriceRequest(trade, market, W.sobol_config(paths, steps, F, seed=1), W.AS_OF,
                            LocalVolSpec(calibrated=grids, resident=res))


def sample_request(u: Universe, rng, path_choices=(16_384, 65_536, 262_144, 1_048_576),
                   max_trades: int = 64, max_bumps: int = 10) -> List[PriceRequest]:
    """One serving request: 1-64 trades x 1-10 CRN bumps (base, spot +-1%, vega +1pt)."""
    n_trades = int(rng.integers(1, max_trades + 1))
    n_bumps = int(rng.integers(1, max_bumps + 1))
    paths = int(rng.choice(path_choices))
    out = []
    for t in range(n_trades):
        F = int(rng.integers(1, 6))
        assets = sorted(rng.choice(u.n, size=F, replace=False).tolist())
        years = float(rng.uniform(0.5, 5.0))
        tseed = int(rng.integers(1 << 31))  # one set of trade terms for all its bumps
        bumps = [None]
        for f in range(F):
            bumps += [("spot", f, 0.01), ("spot", f, -0.01), ("vega", f)]
        for b in bumps[:n_bumps]:
            trng_b = np.random.default_rng(tseed)
            if b is None:
                q = u.request(assets, years, paths, trade_rng=trng_b, trade_id=f"t{t}")
            elif b[0] == "spot":
                q = u.request(assets, years, paths, spot_bump=(b[1], b[2]), trade_rng=trng_b,
                              trade_id=f"t{t}")
            else:
                q = u.request(assets, years, paths, vega_asset=b[1], trade_rng=trng_b,
                              trade_id=f"t{t}")
            out.append(q)
    return out


def poisson_load(server: Server, requests: List[List[PriceRequest]], tiers: List[str], qps: float,
                 duration_s: float, seed: int = 0) -> dict:
    """Open-loop Poisson arrivals at ``qps`` for ``duration_s``; requests are taken in
    turn, serialised before the clock starts (as a client would send them)."""
    requests = [r if isinstance(r, PackedTicket) else server.pack(r) for r in requests]
    rng = np.random.default_rng(seed)
    tickets = []
    t0 = time.perf_counter()
    t_next = t0
    i = 0
    while True:
        t_next += float(rng.exponential(1.0 / qps))
        if t_next - t0 > duration_s:
            break
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        k = i % len(requests)
        tickets.append(server.submit(requests[k], tiers[k]))
        i += 1
    submitted = time.perf_counter() - t0
    done = [server.wait(t, assemble=False) for t in tickets]
    wall = time.perf_counter() - t0
    ok = [d for d in done if d.state == DONE]
    lat = sorted(d.latency_s for d in ok)
    ps = sum(d.path_steps for d in ok)
    tier_of = [tiers[i % len(requests)] for i in range(len(tickets))]

    def censored(ds):
        # shed and failed tickets never finish: they count as later than every
        # completed one, so a percentile that lands on them is reported as None
        xs = sorted(d.latency_s if d.state == DONE else math.inf for d in ds)
        out = {}
        for q in (50, 99):
            if not xs:
                out[f"p{q}_s"] = None
                continue
            k = min(len(xs) - 1, int(math.ceil(q / 100.0 * len(xs))) - 1)
            out[f"p{q}_s"] = xs[k] if xs[k] < math.inf else None
        return out

    per_tier = {}
    for tier in sorted(set(tier_of)):
        ds = [d for d, t in zip(done, tier_of) if t == tier]
        per_tier[tier] = {"submitted": len(ds), "completed": sum(d.state == DONE for d in ds), **censored(ds)}
    return {
        "qps": qps, "submitted": len(tickets), "completed": len(ok),
        "shed": sum(d.state == SHED for d in done), "failed": sum(d.state == FAILED for d in done),
        "p50_s": float(np.percentile(lat, 50)) if lat else None,
        "p99_s": float(np.percentile(lat, 99)) if lat else None,
        "censored": censored(done), "per_tier": per_tier,
        "throughput_path_steps_per_s": ps / wall, "wall_s": wall, "arrival_window_s": submitted,
        "mean_batch_tickets": float(np.mean([d.batch_tickets for d in ok])) if ok else None,
    }
