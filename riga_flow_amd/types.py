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
