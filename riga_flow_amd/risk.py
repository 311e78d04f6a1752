"""Common-random-number finite-difference risk on the device.

Mirrors ``crates/riga-risk/src/lib.rs``: ``CentralDifferenceConfig`` (lib.rs:30-58),
``price_central_greeks`` (lib.rs:117-186), ``crn_delta_convergence`` (lib.rs:212-258),
``crn_variance_reduction_factor`` (lib.rs:296-336) and ``bump_volatility_request``
(lib.rs:260-279).  The reference prices every scenario with a separate blocking
``price_put_leg`` call; here every scenario of a report goes into one batch, so the
scenarios that share a seed and a schedule become CRN scenarios of one device item
and read one set of Sobol points.  Results equal pricing each scenario alone.

Local vol follows the reference: a request carrying implied grids is recalibrated
per scenario (spot bumps move the calibration's spot, vega shifts every implied
node).  Requests with only pre-calibrated grids cannot be vega-bumped and are
rejected with the reference's ``risk_surface_shape`` code.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

from . import api
from .types import ImpliedVolGrid, LocalVolSpec, PriceRequest, PriceResult, RigaError


@dataclass
class CentralDifferenceConfig:
    """``CentralDifferenceConfig::default`` (riga-risk/src/lib.rs:47-58)."""

    spot_relative_bump: float = 0.01
    volatility_absolute_bump: float = 0.01
    rate_absolute_bump: float = 0.0001
    correlation_absolute_bump: float = 0.01
    random_number_policy: str = "common_random_numbers"


@dataclass
class FiniteDifferenceEstimate:
    value: float
    mc_error_upper_bound: float


@dataclass
class AssetGreeks:
    asset_index: int
    spot_bump: float
    volatility_bump: float
    delta: FiniteDifferenceEstimate
    gamma: FiniteDifferenceEstimate
    vega: FiniteDifferenceEstimate
    spot_up: PriceResult
    spot_down: PriceResult
    vol_up: PriceResult
    vol_down: PriceResult


@dataclass
class RiskReport:
    base_price: PriceResult
    seed: int
    random_number_policy: str
    assets: List[AssetGreeks] = field(default_factory=list)


@dataclass
class ConvergenceEstimate:
    value: float
    standard_error: float
    replicates: int

    def band_95(self) -> float:
        return 1.96 * self.standard_error


def _validate(request: PriceRequest, config: CentralDifferenceConfig) -> None:
    """validate_configuration (riga-risk/src/lib.rs:338-385), then the engine's own checks."""
    import math

    def finite_pos(v):
        return isinstance(v, (int, float)) and math.isfinite(v) and v > 0.0

    if config.random_number_policy != "common_random_numbers":
        raise RigaError("no_crn_rejected", "central finite differences require CommonRandomNumbers; "
                        "independent streams are rejected")
    if not finite_pos(config.spot_relative_bump) or config.spot_relative_bump >= 1.0:
        raise RigaError("invalid_spot_bump", "spot_relative_bump must be finite, positive, and less than one")
    if not finite_pos(config.volatility_absolute_bump):
        raise RigaError("invalid_volatility_bump", "volatility_absolute_bump must be finite and positive")
    if not finite_pos(config.rate_absolute_bump):
        raise RigaError("invalid_rate_bump", "rate_absolute_bump must be finite and positive")
    if not finite_pos(config.correlation_absolute_bump):
        raise RigaError("invalid_correlation_bump", "correlation_absolute_bump must be finite and positive")
    if len(request.market.spots) != len(request.market.vols) or not request.market.spots:
        raise RigaError("risk_input_shape", "spots and vols must be non-empty and the same length")
    api.validate(request, max_dimension=max(512, request.trade.steps * len(request.market.spots)))


def _with_seed(request: PriceRequest, seed: int) -> PriceRequest:
    q = request.clone()
    q.config.seed = seed
    return q


def _spot_bumped(request: PriceRequest, asset: int, delta: float, seed: int) -> PriceRequest:
    q = _with_seed(request, seed)
    q.market.spots[asset] += delta
    return q


def bump_volatility_request(request: PriceRequest, asset_index: int, bump: float) -> PriceRequest:
    """riga-risk/src/lib.rs:260-279: vols[asset] += bump and, under local vol, every
    implied node of that asset's grid shifted by the same amount (recalibrated)."""
    q = request.clone()
    q.market.vols[asset_index] += bump
    if q.local_vol is not None:
        lv = q.local_vol
        if lv.calibrated is not None and not lv.grids:
            raise RigaError("risk_surface_shape",
                            "a vega bump needs the implied grids; this request carries only "
                            "calibrated local-vol grids")
        if asset_index >= len(lv.grids):
            raise RigaError("risk_surface_shape", f"local-vol grid missing for asset {asset_index}")
        g = lv.grids[asset_index]
        grids = list(lv.grids)
        grids[asset_index] = ImpliedVolGrid(list(g.maturities_years), list(g.strikes),
                                            [[v + bump for v in row] for row in g.vols])
        q.local_vol = LocalVolSpec(grids=grids, calibrated=None, grid_x=lv.grid_x)
    return q


def price_central_greeks(request: PriceRequest, config: Optional[CentralDifferenceConfig] = None,
                         precision: str = "fp64", max_dimension: int = 512) -> RiskReport:
    """Delta, gamma and vega per asset with CRN (riga-risk/src/lib.rs:117-186), all
    1 + 4F scenarios in one device batch."""
    config = config or CentralDifferenceConfig()
    _validate(request, config)
    seed = request.config.seed
    F = len(request.market.spots)
    scen = [_with_seed(request, seed)]
    bumps = []
    for f in range(F):
        h = request.market.spots[f] * config.spot_relative_bump
        bumps.append(h)
        scen.append(_spot_bumped(request, f, h, seed))
        scen.append(_spot_bumped(request, f, -h, seed))
        scen.append(_with_seed(bump_volatility_request(request, f, config.volatility_absolute_bump), seed))
        scen.append(_with_seed(bump_volatility_request(request, f, -config.volatility_absolute_bump), seed))
    res = api.price_batch(scen, precision, max_dimension=max_dimension)
    base = res[0]
    report = RiskReport(base, seed, "common_random_numbers")
    vb = config.volatility_absolute_bump
    for f in range(F):
        up, dn, vu, vd = res[1 + 4 * f: 5 + 4 * f]
        h = bumps[f]
        report.assets.append(AssetGreeks(
            f, h, vb,
            FiniteDifferenceEstimate((up.pv - dn.pv) / (2.0 * h), (up.mc_error + dn.mc_error) / (2.0 * h)),
            FiniteDifferenceEstimate((up.pv - 2.0 * base.pv + dn.pv) / (h * h),
                                     (up.mc_error + 2.0 * base.mc_error + dn.mc_error) / (h * h)),
            FiniteDifferenceEstimate((vu.pv - vd.pv) / (2.0 * vb), (vu.mc_error + vd.mc_error) / (2.0 * vb)),
            up, dn, vu, vd))
    return report


def _sample_variance(values: List[float]) -> float:
    # riga-risk/src/lib.rs:281-288
    if len(values) < 2:
        return 0.0
    n = float(len(values))
    mean = sum(values) / n
    return sum((v - mean) * (v - mean) for v in values) / (n - 1.0)


def crn_delta_convergence(request: PriceRequest, config: Optional[CentralDifferenceConfig] = None,
                          replicates: int = 8, precision: str = "fp64",
                          max_dimension: int = 512) -> ConvergenceEstimate:
    """riga-risk/src/lib.rs:212-258: one seed per replicate (1,000,000 + r), CRN within
    the up/down pair; every replicate in one batch."""
    config = config or CentralDifferenceConfig()
    _validate(request, config)
    if replicates == 0:
        raise RigaError("invalid_replicates", "replicates must be greater than zero")
    h = request.market.spots[0] * config.spot_relative_bump
    scen = []
    for r in range(replicates):
        seed = 1_000_000 + r
        scen += [_spot_bumped(request, 0, h, seed), _spot_bumped(request, 0, -h, seed)]
    res = api.price_batch(scen, precision, max_dimension=max_dimension)
    samples = [(res[2 * r].pv - res[2 * r + 1].pv) / (2.0 * h) for r in range(replicates)]
    n = float(replicates)
    mean = sum(samples) / n
    var = _sample_variance(samples) if replicates > 1 else 0.0
    return ConvergenceEstimate(mean, (var / n) ** 0.5, replicates)


def crn_variance_reduction_factor(request: PriceRequest, config: Optional[CentralDifferenceConfig] = None,
                                  replicates: int = 16, precision: str = "fp64",
                                  max_dimension: int = 512, details: bool = False):
    """riga-risk/src/lib.rs:296-336: the delta of asset 0 estimated per replicate with
    one seed for up and down (CRN, seed 1,000 + r) and with two independent seeds
    (50,000 + 3r and 50,000 + 3r + 1); the factor is var(independent) / var(CRN).
    All 4 x replicates valuations run in one batch."""
    config = config or CentralDifferenceConfig()
    _validate(request, config)
    if replicates < 2:
        raise RigaError("risk_replicates_required",
                        "at least two independent replicates are required to estimate a variance ratio")
    h = request.market.spots[0] * config.spot_relative_bump
    scen = []
    for r in range(replicates):
        scen += [_spot_bumped(request, 0, h, 1_000 + r), _spot_bumped(request, 0, -h, 1_000 + r),
                 _spot_bumped(request, 0, h, 50_000 + 3 * r), _spot_bumped(request, 0, -h, 50_000 + 3 * r + 1)]
    res = api.price_batch(scen, precision, max_dimension=max_dimension)
    crn = [(res[4 * r].pv - res[4 * r + 1].pv) / (2.0 * h) for r in range(replicates)]
    ind = [(res[4 * r + 2].pv - res[4 * r + 3].pv) / (2.0 * h) for r in range(replicates)]
    crn_var, ind_var = _sample_variance(crn), _sample_variance(ind)
    if not (crn_var > 0.0) or crn_var != crn_var:
        raise RigaError("zero_crn_variance",
                        "CRN replicate variance must be positive to report a variance-reduction ratio")
    factor = ind_var / crn_var
    if details:
        return factor, crn, ind
    return factor
