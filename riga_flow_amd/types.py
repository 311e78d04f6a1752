"""Wire types of the pricing path, mirroring the reference's serde types.

Field names, optionality and defaults follow
``crates/riga-core/src/lib.rs:5-255`` (``RigaError``, ``TraceLevel``,
``MarketSnapshot``, ``PhoenixTrade``, ``SimulationConfig``,
``VarianceReduction``, ``LegAttribution``, ``TransitionStats``,
``LedgerEntry``, ``PriceResult``) and
``crates/riga-engine/src/lib.rs:31-46`` (``LocalVolSpec``, ``PriceRequest``),
with ``crates/riga-market/src/lib.rs:84-91`` (``ImpliedVolGrid``).

The reference hard-wires two assets (``lib.rs:234-245``); these types carry
any number of factors ``F``.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import List, Optional


class RigaError(Exception):
    """``RigaError{code, message}`` (``crates/riga-core/src/lib.rs:5-19``)."""

    def __init__(self, code: str, message: str):
        super().__init__(f"{code}: {message}")
        self.code = code
        self.message = message

    def to_json(self) -> dict:
        return {"code": self.code, "message": self.message}


TRACE_LEVELS = ("off", "aggregated", "sampled", "full")  # riga-core/src/lib.rs:21-28


@dataclass
class VarianceReduction:
    """``VarianceReduction`` tagged by ``kind`` (riga-core/src/lib.rs:88-106)."""

    kind: str = "none"  # "none" | "halton" | "sobol"
    dimension: int = 0

    def to_json(self) -> dict:
        if self.kind == "none":
            return {"kind": "none"}
        return {"kind": self.kind, "dimension": self.dimension}


@dataclass
class MarketSnapshot:
    """``MarketSnapshot`` (riga-core/src/lib.rs:37-45); field order is the JSON order."""

    as_of: str
    spots: List[float]
    vols: List[float]
    risk_free_rate: float
    dividend_yields: List[float]
    correlation: List[List[float]]


@dataclass
class PhoenixTrade:
    """``PhoenixTrade`` (riga-core/src/lib.rs:57-75)."""

    trade_id: str
    initial_spots: List[float]
    maturity_years: float
    steps: int
    strike: float
    knock_in_barrier: float
    knock_out_barrier: Optional[float] = None
    coupon_barrier: Optional[float] = None
    coupon_rate: Optional[float] = None
    memory_coupon: bool = False
    coupon_periods: Optional[int] = None
    fixed_first_coupon: bool = False
    call_start_step: Optional[int] = None


@dataclass
class SimulationConfig:
    """``SimulationConfig`` with the reference defaults (riga-core/src/lib.rs:77-118)."""

    paths: int = 30_000
    path_block_size: int = 4_096
    seed: int = 1_729
    trace_level: str = "aggregated"
    variance_reduction: VarianceReduction = field(default_factory=VarianceReduction)


@dataclass
class ImpliedVolGrid:
    """``ImpliedVolGrid`` (riga-market/src/lib.rs:84-91): vols[n_maturity][n_strike]."""

    maturities_years: List[float]
    strikes: List[float]
    vols: List[List[float]]


@dataclass
class LocalVolGrid:
    """A calibrated Dupire grid (``DupireLocalVol`` fields, riga-market/src/lib.rs:119-123).

    Used when a workload supplies an already-calibrated and resampled grid
    (t_axis x log_s), e.g. the 64 x 128 grids of the 3-asset local-vol book.
    """

    t_axis: List[float]
    log_s: List[float]
    v: List[List[float]]


@dataclass
class LocalVolSpec:
    """``LocalVolSpec`` (riga-engine/src/lib.rs:31-34): one implied grid per factor.

    ``calibrated`` optionally carries pre-calibrated grids instead, one per factor.
    ``grid_x`` is the dense log-strike axis length used by calibration
    (``GRID_X = 96`` in the reference, riga-market/src/lib.rs:28).
    """

    grids: List[ImpliedVolGrid] = field(default_factory=list)
    calibrated: Optional[List[LocalVolGrid]] = None
    grid_x: int = 96
    # device-resident copies of ``calibrated`` (``api.Surface``, one per factor);
    # when present the device path names them instead of sending the grids
    resident: Optional[list] = None


@dataclass
class PriceRequest:
    """``PriceRequest`` (riga-engine/src/lib.rs:36-46)."""

    trade: PhoenixTrade
    market: MarketSnapshot
    config: SimulationConfig
    as_of: str
    local_vol: Optional[LocalVolSpec] = None

    @property
    def factors(self) -> int:
        return len(self.market.spots)

    def clone(self) -> "PriceRequest":
        return dataclasses.replace(
            self,
            trade=dataclasses.replace(self.trade, initial_spots=list(self.trade.initial_spots)),
            market=dataclasses.replace(
                self.market,
                spots=list(self.market.spots),
                vols=list(self.market.vols),
                dividend_yields=list(self.market.dividend_yields),
                correlation=[list(r) for r in self.market.correlation],
            ),
            config=dataclasses.replace(
                self.config,
                variance_reduction=dataclasses.replace(self.config.variance_reduction),
            ),
            local_vol=None
            if self.local_vol is None
            else LocalVolSpec(
                grids=[
                    ImpliedVolGrid(list(g.maturities_years), list(g.strikes), [list(r) for r in g.vols])
                    for g in self.local_vol.grids
                ],
                calibrated=None
                if self.local_vol.calibrated is None
                else [
                    LocalVolGrid(list(c.t_axis), list(c.log_s), [list(r) for r in c.v])
                    for c in self.local_vol.calibrated
                ],
                grid_x=self.local_vol.grid_x,
                resident=None if self.local_vol.resident is None else list(self.local_vol.resident),
            ),
        )


@dataclass
class LegAttribution:
    """``LegAttribution`` (riga-core/src/lib.rs:201-217)."""

    par_redemption: float
    coupon: float
    memory_coupon: float
    down_and_in_put: float
    funding: float
    discounting: float
    total: float


@dataclass
class TransitionStats:
    """``TransitionStats`` (riga-core/src/lib.rs:219-229)."""

    transition_id: str
    input_count: int
    output_count: int
    probability: float
    mean_payoff: float
    variance_payoff: float
    mean_delta: float = 0.0
    mean_vega: float = 0.0


@dataclass
class LedgerEntry:
    """``LedgerEntry`` (riga-core/src/lib.rs:231-241)."""

    transition_id: str
    condition: str
    population_before: int
    population_after: int
    pv_impact: float
    delta_impact: float
    vega_impact: float
    evidence_refs: List[str]


@dataclass
class DeterminismEnvelope:
    """``DeterminismEnvelope`` (riga-core/src/lib.rs:30-35)."""

    seed: int
    as_of: str
    market_snapshot_id: str


@dataclass
class PriceResult:
    """``PriceResult`` (riga-core/src/lib.rs:243-255)."""

    pv: float
    mc_error: float
    ki_probability: float
    legs: LegAttribution
    stats: List[TransitionStats]
    ledger: List[LedgerEntry]
    trace_events: List[str]
    determinism: DeterminismEnvelope


# ---------------------------------------------------------------- PriceRequest JSON
def _bad(msg: str) -> RigaError:
    return RigaError("invalid_price_request", f"request must match riga_engine::PriceRequest: {msg}")


def _num(doc, key, where, optional=False):
    if key not in doc or doc[key] is None:
        if optional:
            return None
        raise _bad(f"missing field `{key}` in {where}")
    v = doc[key]
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise _bad(f"invalid type for `{key}` in {where}: expected f64")
    return float(v)


def _uint(doc, key, where, optional=False):
    if key not in doc or doc[key] is None:
        if optional:
            return None
        raise _bad(f"missing field `{key}` in {where}")
    v = doc[key]
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise _bad(f"invalid value for `{key}` in {where}: expected usize")
    return v


def _vec(doc, key, where):
    if key not in doc:
        raise _bad(f"missing field `{key}` in {where}")
    v = doc[key]
    if not isinstance(v, list) or any(isinstance(x, bool) or not isinstance(x, (int, float)) for x in v):
        raise _bad(f"invalid type for `{key}` in {where}: expected a sequence of f64")
    return [float(x) for x in v]


def _str(doc, key, where):
    if key not in doc:
        raise _bad(f"missing field `{key}` in {where}")
    if not isinstance(doc[key], str):
        raise _bad(f"invalid type for `{key}` in {where}: expected a string")
    return doc[key]


def _bool(doc, key, where, default=None):
    if key not in doc:
        if default is not None:
            return default
        raise _bad(f"missing field `{key}` in {where}")
    if not isinstance(doc[key], bool):
        raise _bad(f"invalid type for `{key}` in {where}: expected a boolean")
    return doc[key]


def request_from_json(doc) -> "PriceRequest":
    """serde's decoding of ``PriceRequest`` (riga-engine/src/lib.rs:36-46 with the
    riga-core field rules): required fields, ``Option`` fields absent or null,
    ``#[serde(default)]`` fields defaulted, snake_case enums, the
    ``variance_reduction`` kind tag.  Errors are ``invalid_price_request``."""
    if not isinstance(doc, dict):
        raise _bad("expected a JSON object")
    for k in ("trade", "market", "config"):
        if not isinstance(doc.get(k), dict):
            raise _bad(f"missing field `{k}`" if k not in doc else f"invalid type for `{k}`: expected a struct")
    t, m, c = doc["trade"], doc["market"], doc["config"]
    trade = PhoenixTrade(
        trade_id=_str(t, "trade_id", "trade"), initial_spots=_vec(t, "initial_spots", "trade"),
        maturity_years=_num(t, "maturity_years", "trade"), steps=_uint(t, "steps", "trade"),
        strike=_num(t, "strike", "trade"), knock_in_barrier=_num(t, "knock_in_barrier", "trade"),
        knock_out_barrier=_num(t, "knock_out_barrier", "trade", optional=True),
        coupon_barrier=_num(t, "coupon_barrier", "trade", optional=True),
        coupon_rate=_num(t, "coupon_rate", "trade", optional=True),
        memory_coupon=_bool(t, "memory_coupon", "trade"),
        coupon_periods=_uint(t, "coupon_periods", "trade", optional=True),
        fixed_first_coupon=_bool(t, "fixed_first_coupon", "trade", default=False),
        call_start_step=_uint(t, "call_start_step", "trade", optional=True))
    corr = m.get("correlation")
    if not isinstance(corr, list) or not all(isinstance(r, list) for r in corr):
        raise _bad("missing field `correlation` in market" if corr is None
                   else "invalid type for `correlation` in market: expected a sequence of sequences")
    market = MarketSnapshot(
        as_of=_str(m, "as_of", "market"), spots=_vec(m, "spots", "market"), vols=_vec(m, "vols", "market"),
        risk_free_rate=_num(m, "risk_free_rate", "market"),
        dividend_yields=_vec(m, "dividend_yields", "market"),
        correlation=[_vec({"row": r}, "row", "market.correlation") for r in corr])
    level = _str(c, "trace_level", "config")
    if level not in TRACE_LEVELS:
        raise _bad(f"unknown variant `{level}` for trace_level, expected one of {', '.join(TRACE_LEVELS)}")
    vr_doc = c.get("variance_reduction", {"kind": "none"})
    if not isinstance(vr_doc, dict) or vr_doc.get("kind") not in ("none", "halton", "sobol"):
        raise _bad("variance_reduction must be {\"kind\": \"none\" | \"halton\" | \"sobol\"}")
    vr = VarianceReduction(vr_doc["kind"], 0 if vr_doc["kind"] == "none"
                           else _uint(vr_doc, "dimension", "variance_reduction"))
    config = SimulationConfig(paths=_uint(c, "paths", "config"), path_block_size=_uint(c, "path_block_size", "config"),
                              seed=_uint(c, "seed", "config"), trace_level=level, variance_reduction=vr)
    lv = None
    if doc.get("local_vol") is not None:
        lvd = doc["local_vol"]
        if not isinstance(lvd, dict) or not isinstance(lvd.get("grids"), list):
            raise _bad("local_vol must be {\"grids\": [...]}")
        grids = []
        for g in lvd["grids"]:
            if not isinstance(g, dict) or not isinstance(g.get("vols"), list):
                raise _bad("each local_vol grid needs maturities_years, strikes and vols")
            grids.append(ImpliedVolGrid(_vec(g, "maturities_years", "grid"), _vec(g, "strikes", "grid"),
                                        [_vec({"row": r}, "row", "grid.vols") for r in g["vols"]]))
        lv = LocalVolSpec(grids=grids)
    return PriceRequest(trade=trade, market=market, config=config, as_of=_str(doc, "as_of", "request"),
                        local_vol=lv)


def request_to_json(req: "PriceRequest") -> dict:
    """serde's encoding of ``PriceRequest`` (field order as declared; ``local_vol``
    omitted when absent)."""
    t, m, c = req.trade, req.market, req.config
    out = {
        "trade": {"trade_id": t.trade_id, "initial_spots": list(t.initial_spots), "maturity_years": t.maturity_years,
                  "steps": t.steps, "strike": t.strike, "knock_in_barrier": t.knock_in_barrier,
                  "knock_out_barrier": t.knock_out_barrier, "coupon_barrier": t.coupon_barrier,
                  "coupon_rate": t.coupon_rate, "memory_coupon": t.memory_coupon,
                  "coupon_periods": t.coupon_periods, "fixed_first_coupon": t.fixed_first_coupon,
                  "call_start_step": t.call_start_step},
        "market": {"as_of": m.as_of, "spots": list(m.spots), "vols": list(m.vols),
                   "risk_free_rate": m.risk_free_rate, "dividend_yields": list(m.dividend_yields),
                   "correlation": [list(r) for r in m.correlation]},
        "config": {"paths": c.paths, "path_block_size": c.path_block_size, "seed": c.seed,
                   "trace_level": c.trace_level, "variance_reduction": c.variance_reduction.to_json()},
        "as_of": req.as_of,
    }
    if req.local_vol is not None:
        out["local_vol"] = {"grids": [{"maturities_years": list(g.maturities_years), "strikes": list(g.strikes),
                                       "vols": [list(r) for r in g.vols]} for g in req.local_vol.grids]}
    return out
