"""Synthetic workloads for the Phoenix hot path (all seeds fixed).

Each builder returns reference-shaped ``PriceRequest`` objects:

* ``flat_book``       3 assets, flat vols 0.20/0.25/0.30, rho 0.5, 1y x 252 daily steps,
                      65,536 Sobol paths (seed 1), KO 100% from day 21, 0.8% monthly
                      memory coupon at 70%, EKI 60%.
* ``lv_book``         the same terms on Dupire local vol from SVI smiles
                      (a=0.02, b=0.1, rho=-0.5, m=0, sigma=0.2+-U(0,0.05)) on
                      12 maturities x 41 strikes (50-150%), calibrated with 128
                      log-spot knots and resampled to 64 tenors; 3y x 756 steps,
                      262,144 paths.
* ``crn_scenarios``   the local-vol trade under base, spot +-1% per asset and vega
                      +1 vol point per asset (10 scenarios sharing 1,048,576 paths).
* ``uneven_book``     2,048 trades from a 20-asset universe (1-5 underlyings,
                      0.5-5y daily, KO 95-105%, coupon barrier 60-80%, EKI 50-65%,
                      coupon 0.5-1.2% a month), 65,536 paths each.

Build decisions (the workload text leaves these open): strike 1.0 (= notional);
"KO from day 21" is ``call_start_step = 21`` (note.rs:166); monthly coupons are
the reference's day-in accrual (note.rs:156-163) with 12 periods a year; the
QMC dimension is steps*F with the 512 cap of qmc.rs:9 lifted so no draw repeats;
SVI parameters are annualised implied variance, w = sigma_imp^2 T.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .types import (ImpliedVolGrid, LocalVolGrid, LocalVolSpec, MarketSnapshot, PhoenixTrade,
                    PriceRequest, SimulationConfig, VarianceReduction)

AS_OF = "2026-01-02"
SVI_MATURITIES = [1 / 12, 2 / 12, 3 / 12, 6 / 12, 9 / 12, 1.0, 1.5, 2.0, 2.5, 3.0, 4.0, 5.0]
SVI_STRIKES = [50.0 + 2.5 * i for i in range(41)]  # 50 .. 150


def phoenix_trade(trade_id: str, F: int, years: float, steps: int, *, ko: float = 1.0,
                  coupon_barrier: float = 0.70, eki: float = 0.60, coupon: float = 0.008,
                  call_start: int = 21, periods: int = None, strike: float = 1.0,
                  s0: float = 100.0) -> PhoenixTrade:
    return PhoenixTrade(
        trade_id=trade_id, initial_spots=[s0] * F, maturity_years=years, steps=steps,
        strike=strike, knock_in_barrier=eki, knock_out_barrier=ko, coupon_barrier=coupon_barrier,
        coupon_rate=coupon, memory_coupon=True,
        coupon_periods=periods if periods is not None else max(1, int(round(12 * years))),
        fixed_first_coupon=False, call_start_step=call_start)


def flat_market(F: int = 3, vols=(0.20, 0.25, 0.30), rho: float = 0.5, r: float = 0.03,
                q: float = 0.01, s0: float = 100.0) -> MarketSnapshot:
    corr = [[1.0 if i == j else rho for j in range(F)] for i in range(F)]
    vols = [float(vols[i % len(vols)]) for i in range(F)]  # cycled beyond the given list
    return MarketSnapshot(AS_OF, [s0] * F, vols, r, [q] * F, corr)


def sobol_config(paths: int, steps: int, F: int, seed: int = 1, trace: str = "aggregated",
                 kind: str = "sobol") -> SimulationConfig:
    return SimulationConfig(paths=paths, path_block_size=4096, seed=seed, trace_level=trace,
                            variance_reduction=VarianceReduction(kind, steps * F))


def flat_book(paths: int = 65_536, steps: int = 252, years: float = 1.0) -> PriceRequest:
    """Plain FP64 worst-of Phoenix on flat vols (the CPU-runnable case)."""
    F = 3
    return PriceRequest(
        trade=phoenix_trade("flat-3", F, years, steps),
        market=flat_market(F),
        config=sobol_config(paths, steps, F),
        as_of=AS_OF,
    )


# ---------------------------------------------------------------- local vol
def svi_sigmas(F: int, seed: int = 20260102) -> List[float]:
    rng = np.random.default_rng(seed)
    return [0.2 + float(rng.uniform(-0.05, 0.05)) for _ in range(F)]


def svi_implied(sigma: float, spot: float = 100.0, a: float = 0.02, b: float = 0.1,
                rho: float = -0.5, m: float = 0.0, bump: float = 0.0) -> ImpliedVolGrid:
    """Implied vols sigma_imp(k) = sqrt(a + b(rho(k-m) + sqrt((k-m)^2 + sigma^2))), k = ln(K/S0)."""
    vols = []
    for _ in SVI_MATURITIES:
        row = []
        for K in SVI_STRIKES:
            k = math.log(K / spot)
            var = a + b * (rho * (k - m) + math.sqrt((k - m) ** 2 + sigma * sigma))
            row.append(math.sqrt(var) + bump)
        vols.append(row)
    return ImpliedVolGrid(list(SVI_MATURITIES), list(SVI_STRIKES), vols)


def lv_grid_for(implied: ImpliedVolGrid, spot: float, rate: float, horizon: float,
                grid_x: int = 128, n_times: int = 64) -> LocalVolGrid:
    """Calibrate with ``grid_x`` knots, then resample to ``n_times`` tenors on [0, horizon]."""
    from . import api
    g = api.lv_calibrate(implied.maturities_years, implied.strikes, implied.vols, spot, rate, grid_x)
    times = [horizon * i / (n_times - 1) for i in range(n_times)]
    return api.lv_resample(g, times)


def lv_book(paths: int = 262_144, steps: int = 756, years: float = 3.0, F: int = 3,
            r: float = 0.03, q: float = 0.01) -> PriceRequest:
    """Dupire local vol, 3 underlyings, 3y daily."""
    market = flat_market(F, r=r, q=q)
    sig = svi_sigmas(F)
    grids = [lv_grid_for(svi_implied(sig[f]), market.spots[f], r, years) for f in range(F)]
    return PriceRequest(
        trade=phoenix_trade("lv-3", F, years, steps),
        market=market,
        config=sobol_config(paths, steps, F),
        as_of=AS_OF,
        local_vol=LocalVolSpec(calibrated=grids),
    )


def crn_scenarios(paths: int = 1_048_576, steps: int = 756, years: float = 3.0,
                  F: int = 3) -> Tuple[List[str], List[PriceRequest]]:
    """Base, spot +-1% per asset, vega +1 vol point per asset (riga-risk/src/lib.rs:117-186,
    260-279): every scenario keeps the seed; spot bumps recalibrate the local vol with the
    bumped spot (build_local_vol uses market.spots); vega shifts every implied node."""
    base = lv_book(paths, steps, years, F)
    r = base.market.risk_free_rate
    sig = svi_sigmas(F)
    names, out = ["base"], [base]
    for f in range(F):
        for sgn, tag in ((1.0, "+"), (-1.0, "-")):
            req = base.clone()
            req.market.spots[f] = base.market.spots[f] + sgn * base.market.spots[f] * 0.01
            grids = list(base.local_vol.calibrated)
            grids[f] = lv_grid_for(svi_implied(sig[f]), req.market.spots[f], r, years)
            req.local_vol = LocalVolSpec(calibrated=grids)
            names.append(f"spot[{f}]{tag}")
            out.append(req)
    for f in range(F):
        req = base.clone()
        req.market.vols[f] = base.market.vols[f] + 0.01
        grids = list(base.local_vol.calibrated)
        grids[f] = lv_grid_for(svi_implied(sig[f], bump=0.01), req.market.spots[f], r, years)
        req.local_vol = LocalVolSpec(calibrated=grids)
        names.append(f"vol[{f}]+")
        out.append(req)
    return names, out


# ---------------------------------------------------------------- uneven book
def universe_correlation(n: int = 20, k: int = 4, seed: int = 7) -> np.ndarray:
    """normalise(B B^T + diag(d)) with B in R^{n x k}."""
    rng = np.random.default_rng(seed)
    B = rng.normal(0.0, 0.5, size=(n, k))
    d = rng.uniform(0.2, 0.6, size=n)
    C = B @ B.T + np.diag(d)
    s = np.sqrt(np.diag(C))
    C = C / np.outer(s, s)
    np.fill_diagonal(C, 1.0)
    return (C + C.T) / 2.0


def uneven_book(n_trades: int = 2048, paths: int = 65_536, seed: int = 11,
                n_assets: int = 20) -> List[PriceRequest]:
    rng = np.random.default_rng(seed)
    corr = universe_correlation(n_assets)
    vols_u = rng.uniform(0.15, 0.40, size=n_assets)
    q_u = rng.uniform(0.0, 0.03, size=n_assets)
    r = 0.03
    out = []
    for i in range(n_trades):
        F = int(rng.integers(1, 6))
        names = sorted(rng.choice(n_assets, size=F, replace=False).tolist())
        years = float(rng.uniform(0.5, 5.0))
        steps = max(1, int(round(252 * years)))
        sub = [[float(corr[a][b]) for b in names] for a in names]
        market = MarketSnapshot(AS_OF, [100.0] * F, [float(vols_u[a]) for a in names], r,
                                [float(q_u[a]) for a in names], sub)
        trade = phoenix_trade(f"book-{i}", F, years, steps,
                              ko=float(rng.uniform(0.95, 1.05)),
                              coupon_barrier=float(rng.uniform(0.60, 0.80)),
                              eki=float(rng.uniform(0.50, 0.65)),
                              coupon=float(rng.uniform(0.005, 0.012)),
                              call_start=21)
        out.append(PriceRequest(trade, market, sobol_config(paths, steps, F, seed=1), AS_OF))
    return out


def path_steps(req: PriceRequest) -> int:
    return req.config.paths * req.trade.steps


def max_dimension_for(req: PriceRequest) -> int:
    """Ladder cap lifted to the request's own steps * F (no wrapped dimensions)."""
    return max(512, req.trade.steps * req.factors)
