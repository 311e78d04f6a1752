"""Loader of the captured term-sheet fixture (``load_termsheet1_fixture``).

Restates crates/riga-engine/src/fixture.rs:45-460: ACT/365 year fraction,
weekday step count, linear curve interpolation, the coupon schedule after
as_of, the call start step, and the captured implied smiles as one
``ImpliedVolGrid`` per asset (maturities as ACT/365 years, sorted).
"""
from __future__ import annotations

import json
from dataclasses import dataclass
from typing import List, Optional, Tuple

from .types import (ImpliedVolGrid, LocalVolSpec, MarketSnapshot, PhoenixTrade, PriceRequest,
                    RigaError, SimulationConfig, VarianceReduction)


@dataclass
class FixtureExpected:
    """``FixtureExpected`` (riga-engine/src/lib.rs:59-65)."""

    put_pv: float
    relative_tolerance: float
    ki_probability: float
    ki_probability_tolerance: float


@dataclass
class LoadedFixture:
    name: str
    request: PriceRequest
    expected: FixtureExpected
    market_is_pinned: bool
    source_model: str
    engine_model: str


def tenor_years(tenor: str) -> Optional[float]:
    """fixture.rs:54-62: '3M' -> 0.25, '2Y' -> 2.0."""
    try:
        if tenor.endswith("M"):
            return float(tenor[:-1]) / 12.0
        if tenor.endswith("Y"):
            return float(tenor[:-1])
    except ValueError:
        return None
    return None


def civil_day_number(date: str) -> int:
    """Days from 0000-03-01 of an ISO date (fixture.rs:64-94, days_from_civil)."""
    parts = date.split("-")
    if len(parts) != 3:
        raise RigaError("fixture_invalid_date", f"invalid ISO date {date}")
    try:
        year, month, day = (int(p) for p in parts)
    except ValueError:
        raise RigaError("fixture_invalid_date", f"invalid ISO date {date}")
    if not (1 <= month <= 12 and 1 <= day <= 31):
        raise RigaError("fixture_invalid_date", f"invalid ISO date {date}")
    year -= 1 if month <= 2 else 0
    era = year // 400
    yoe = year - era * 400
    mp = month + (-3 if month > 2 else 9)
    doy = (153 * mp + 2) // 5 + day - 1
    doe = yoe * 365 + yoe // 4 - yoe // 100 + doy
    return era * 146_097 + doe


def weekday_steps(start: str, end: str) -> int:
    """Weekdays in (start, end] (fixture.rs:96-108)."""
    a, b = civil_day_number(start), civil_day_number(end)
    if b < a:
        raise RigaError("fixture_invalid_date", "end date must not precede start date")
    return sum(1 for d in range(a + 1, b + 1) if (d + 3) % 7 < 5)


def remaining_business_steps(as_of: str, maturity: str) -> Tuple[float, int]:
    a, b = civil_day_number(as_of), civil_day_number(maturity)
    if b <= a:
        raise RigaError("fixture_invalid_date", "maturity must be after as_of")
    return (b - a) / 365.0, weekday_steps(as_of, maturity)


def interpolate_curve(points: List[Tuple[float, float]], time: float) -> float:
    """fixture.rs:125-148 (linear, flat outside)."""
    if not points:
        raise RigaError("fixture_missing_field", "market curve has no points")
    if not all(points[i][0] < points[i + 1][0] for i in range(len(points) - 1)):
        raise RigaError("fixture_invalid_field", "market curve tenors must be strictly increasing")
    if time <= points[0][0]:
        return points[0][1]
    for (t0, v0), (t1, v1) in zip(points, points[1:]):
        if time <= t1:
            w = (time - t0) / (t1 - t0)
            return v0 + w * (v1 - v0)
    return points[-1][1]


def load_termsheet1_fixture(path: str) -> LoadedFixture:
    with open(path, "rb") as fh:
        root = json.loads(fh.read())
    if root.get("fixture") != "termsheet1":
        raise RigaError("fixture_identity", "expected fixture id termsheet1")
    trade = root["trade_economics"]
    assets = trade["reference_assets"]
    if len(assets) != 2:
        raise RigaError("fixture_shape", "first build requires two reference assets")
    names = [a.get("symbol", "") for a in assets]
    md = root["market_data"]

    def points(asset):
        rows = md["vol_surfaces"][asset]
        out = []
        for row in rows:
            years = tenor_years(row["tenor"])
            if years is None:
                raise RigaError("fixture_invalid_field", f"invalid vol tenor {row['tenor']}")
            out.append((years, float(row["atm_implied_vol"])))
        return out

    vol0, vol1 = points(names[0]), points(names[1])
    as_of = root.get("as_of", "")
    t, steps = remaining_business_steps(as_of, trade["final_fixing_date"])
    if steps == 0:
        raise RigaError("fixture_invalid_date", "maturity interval has no weekday observations")
    v0 = interpolate_curve(vol0, t)
    v1 = interpolate_curve(vol1, t)
    curve = [(tenor_years(p["tenor"]), float(p["zero_rate"])) for p in md["discount_curve"]["points"]]
    rate = interpolate_curve(curve, t)
    future = [row for row in trade["schedule"] if row.get("observation", "") > as_of]
    if not future:
        raise RigaError("fixture_invalid_schedule", "no future coupon observations remain after as_of")
    fixed_first = future[0].get("coupon_type") == "fixed"
    call_start = trade["callable"]["period_start"]
    call_start_step = 0 if call_start <= as_of else weekday_steps(as_of, call_start)
    spots = [md.get("as_of_spots", {}).get(n, md["initial_fixings"]["spots"].get(n)) for n in names]
    local_vol = None
    surfaces = md.get("local_vol_surfaces")
    if isinstance(surfaces, dict) and surfaces:
        grids = []
        for n in names:
            node = surfaces[n]
            rows = []
            for k, date in enumerate(node["maturities"]):
                years = (civil_day_number(date) - civil_day_number(as_of)) / 365.0
                rows.append((years, [float(v) for v in node["vols"][k]]))
            rows.sort(key=lambda r: r[0])
            grids.append(ImpliedVolGrid([r[0] for r in rows], [float(s) for s in node["strikes"]],
                                        [r[1] for r in rows]))
        local_vol = LocalVolSpec(grids=grids)
    run = root.get("run", {})
    b = trade["barriers"]
    request = PriceRequest(
        local_vol=local_vol,
        trade=PhoenixTrade(
            trade_id=trade.get("trade_id", "termsheet1"),
            initial_spots=[float(md["initial_fixings"]["spots"][n]) for n in names],
            maturity_years=t,
            steps=steps,
            strike=float(b["exercise"]["level_pct"]),
            knock_in_barrier=float(b["knock_in"]["level_pct"]),
            knock_out_barrier=b.get("call", {}).get("level_pct"),
            coupon_barrier=b.get("coupon", {}).get("level_pct"),
            coupon_rate=trade.get("coupon", {}).get("rate_per_period"),
            memory_coupon=bool(trade.get("coupon", {}).get("memory", False)),
            coupon_periods=len(future),
            fixed_first_coupon=fixed_first,
            call_start_step=call_start_step,
        ),
        market=MarketSnapshot(
            as_of=md.get("as_of", ""),
            spots=[float(s) for s in spots],
            vols=[v0, v1],
            risk_free_rate=rate,
            dividend_yields=[float(md["dividend_yields"]["annual_continuous"][n]) for n in names],
            correlation=[[float(x) for x in row] for row in md["correlation_matrix"]["values"]],
        ),
        config=SimulationConfig(
            paths=int(run.get("paths", 30_000)),
            path_block_size=int(run.get("path_block_size", 4_096)),
            seed=int(run.get("seed", 1_729)),
            trace_level="aggregated",
            variance_reduction=VarianceReduction("none"),
        ),
        as_of=as_of,
    )
    e = root["expected"]
    expected = FixtureExpected(float(e["put"]["value"]), float(e["put"]["tol_rel"]),
                               float(e["ki_prob"]["value"]), float(e["ki_prob"]["tol_abs"]))
    prov = root.get("provenance", {})
    pinned = (isinstance(root.get("to_capture"), list) and not root["to_capture"]
              and isinstance(prov, dict)
              and all(isinstance(v, str) and (v.startswith("pinned") or v.startswith("captured"))
                      for v in prov.values()))
    return LoadedFixture(
        name="termsheet1", request=request, expected=expected, market_is_pinned=pinned,
        source_model=root.get("model", {}).get("process", "unknown"),
        engine_model="dupire_local_vol_calibrated_cpu" if local_vol is not None
        else "gbm_flat_atm_proxy_from_initial_fixings")
