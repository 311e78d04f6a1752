"""Common-random-number finite-difference risk on the device.

Mirrors ``crates/riga-risk/src/lib.rs``: ``CentralDifferenceConfig`` (lib.rs:30-58),
``price_central_greeks`` (lib.rs:117-186), ``crn_delta_convergence`` (lib.rs:212-258),
``crn_variance_reduction_factor`` (lib.rs:296-336), ``bump_volatility_request``
(lib.rs:260-279), ``validate_configuration`` (lib.rs:338-456), ``price_risk_batch``
with transition risk, DV01 and correlation sensitivity (lib.rs:458-628),
``pnl_explain`` (lib.rs:630-680) and ``aggregate_portfolio`` (lib.rs:682-720); and
the MCP adapters' scenario sets: theta and cross-gamma (riga-mcp/src/main.rs:1355-1424),
trade-design bumps (main.rs:1426-1537) and the path-count convergence ladder
(main.rs:1216-1278).  The reference prices every scenario with a separate blocking
``price_put_leg`` call; here every scenario of a report goes into one batch, so the
scenarios that share a seed and a schedule become CRN scenarios of one device item
and read one set of Sobol points (or one ChaCha8 stream).  Results equal pricing each
scenario alone.

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
    """validate_configuration (riga-risk/src/lib.rs:338-456), then the engine's own
    checks.  The reference's risk lane takes exactly two assets; here any F >= 2
    (the correlation scenarios bump rho[0][1] as the reference does)."""
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
    m = request.market
    if len(m.spots) != len(m.vols) or not m.spots:
        raise RigaError("risk_input_shape", "spot and volatility arrays must be non-empty and have the same length")
    F = len(m.spots)
    if F < 2 or len(m.correlation) != F or any(len(row) != F for row in m.correlation):
        raise RigaError("risk_input_shape", "the risk lane requires at least two assets and an F x F "
                                            "correlation matrix")
    rho = m.correlation[0][1]
    cb = config.correlation_absolute_bump
    if not math.isfinite(rho) or rho - cb <= -1.0 or rho + cb >= 1.0:
        raise RigaError("invalid_correlation_bump",
                        "correlation up/down scenarios must remain strictly inside (-1, 1)")
    for i, (spot, vol) in enumerate(zip(m.spots, m.vols)):
        if not math.isfinite(spot) or spot <= 0.0:
            raise RigaError("invalid_spot_bump", f"spot[{i}] must be finite and positive")
        h = spot * config.spot_relative_bump
        if not math.isfinite(h) or h == 0.0 or spot - h <= 0.0:
            raise RigaError("invalid_spot_bump", f"spot bump for asset {i} is not representable and positive")
        if not math.isfinite(vol) or vol < config.volatility_absolute_bump:
            raise RigaError("invalid_volatility_bump",
                            f"volatility[{i}] must be finite and at least volatility_absolute_bump")
    if request.config.paths < 2:
        raise RigaError("risk_paths_required",
                        "at least two paths are required to report Monte Carlo risk uncertainty")
    lv = request.local_vol
    if lv is not None and lv.grids:
        ok = len(lv.grids) == F and all(
            g.vols and all(row and all(math.isfinite(v) and v >= config.volatility_absolute_bump for v in row)
                           for row in g.vols) for g in lv.grids)
        if not ok:
            raise RigaError("invalid_volatility_bump",
                            "local-vol surface must align with assets and every implied-vol node must remain "
                            "non-negative under the down bump")
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


def _central_scenarios(request: PriceRequest, config: CentralDifferenceConfig) -> List[PriceRequest]:
    """Base, then per asset spot up/down and vol up/down, every one on the base seed."""
    seed = request.config.seed
    scen = [_with_seed(request, seed)]
    for f in range(len(request.market.spots)):
        h = request.market.spots[f] * config.spot_relative_bump
        scen.append(_spot_bumped(request, f, h, seed))
        scen.append(_spot_bumped(request, f, -h, seed))
        scen.append(_with_seed(bump_volatility_request(request, f, config.volatility_absolute_bump), seed))
        scen.append(_with_seed(bump_volatility_request(request, f, -config.volatility_absolute_bump), seed))
    return scen


def price_central_greeks(request: PriceRequest, config: Optional[CentralDifferenceConfig] = None,
                         precision: str = "fp64", max_dimension: int = 512) -> RiskReport:
    """Delta, gamma and vega per asset with CRN (riga-risk/src/lib.rs:117-186), all
    1 + 4F scenarios in one device batch."""
    config = config or CentralDifferenceConfig()
    _validate(request, config)
    res = api.price_batch(_central_scenarios(request, config), precision, max_dimension=max_dimension)
    return _greeks_report(request, config, res)


def _greeks_report(request: PriceRequest, config: CentralDifferenceConfig, res: List[PriceResult]) -> RiskReport:
    seed = request.config.seed
    F = len(request.market.spots)
    bumps = [request.market.spots[f] * config.spot_relative_bump for f in range(F)]
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


# ---------------------------------------------------------------------------- batch risk
@dataclass
class BatchAssetGreeks:
    """riga-risk/src/lib.rs:468-477."""

    asset_index: int
    delta: float
    gamma: float
    vega: float
    delta_mc_error: float
    vega_mc_error: float


@dataclass
class TransitionRisk:
    """riga-risk/src/lib.rs:479-493."""

    id: str
    pv_impact: float
    delta_impact: float
    vega_impact: float


@dataclass
class BatchRiskReport:
    """riga-risk/src/lib.rs:495-517."""

    base_pv: float
    scenario_count: int
    seed: int
    assets: List[BatchAssetGreeks]
    transitions: List[TransitionRisk]
    rate_up_pv: float
    rate_down_pv: float
    correlation_up_pv: float
    correlation_down_pv: float
    rate_bump: float
    correlation_bump: float
    pv_mc_error: float

    def dv01(self) -> float:
        """lib.rs:557-560."""
        return (self.rate_up_pv - self.rate_down_pv) / (2.0 * self.rate_bump)

    def correlation_sensitivity(self) -> float:
        """lib.rs:562-565."""
        return (self.correlation_up_pv - self.correlation_down_pv) / (2.0 * self.correlation_bump)


def transition_risk(report: RiskReport) -> List[TransitionRisk]:
    """lib.rs:519-554: per ledger transition, its PV impact differenced across the
    CRN spot and vol bump pairs already priced."""
    def impact(ledger, tid):
        for e in ledger:
            if e.transition_id == tid:
                return e.pv_impact
        return 0.0

    out = []
    for entry in report.base_price.ledger:
        d = v = 0.0
        for a in report.assets:
            if a.spot_bump > 0.0:
                d += (impact(a.spot_up.ledger, entry.transition_id)
                      - impact(a.spot_down.ledger, entry.transition_id)) / (2.0 * a.spot_bump)
            if a.volatility_bump > 0.0:
                v += (impact(a.vol_up.ledger, entry.transition_id)
                      - impact(a.vol_down.ledger, entry.transition_id)) / (2.0 * a.volatility_bump)
        out.append(TransitionRisk(entry.transition_id, entry.pv_impact, d, v))
    return out


def _rate_corr_scenarios(request: PriceRequest, config: CentralDifferenceConfig) -> List[PriceRequest]:
    """lib.rs:597-614: rate +/- and rho[0][1] +/- on the base seed."""
    out = []
    for d in (config.rate_absolute_bump, -config.rate_absolute_bump):
        q = _with_seed(request, request.config.seed)
        q.market.risk_free_rate += d
        out.append(q)
    for d in (config.correlation_absolute_bump, -config.correlation_absolute_bump):
        q = _with_seed(request, request.config.seed)
        q.market.correlation = [list(r) for r in q.market.correlation]
        q.market.correlation[0][1] += d
        q.market.correlation[1][0] += d
        out.append(q)
    return out


def _batch_report(request, config, res) -> BatchRiskReport:
    F = len(request.market.spots)
    greeks = _greeks_report(request, config, res[:1 + 4 * F])
    ru, rd, cu, cd = res[1 + 4 * F:5 + 4 * F]
    assets = [BatchAssetGreeks(a.asset_index, a.delta.value, a.gamma.value, a.vega.value,
                               a.delta.mc_error_upper_bound, a.vega.mc_error_upper_bound) for a in greeks.assets]
    return BatchRiskReport(greeks.base_price.pv, 1 + 4 * F + 4, request.config.seed, assets,
                           transition_risk(greeks), ru.pv, rd.pv, cu.pv, cd.pv, config.rate_absolute_bump,
                           config.correlation_absolute_bump, greeks.base_price.mc_error)


def price_risk_batch(request: PriceRequest, config: Optional[CentralDifferenceConfig] = None,
                     precision: str = "fp64", max_dimension: int = 512) -> BatchRiskReport:
    """lib.rs:573-628: the 1 + 4F + 4 CRN valuations (base; spot and vol up/down per
    asset; rate and correlation up/down) in one device batch."""
    config = config or CentralDifferenceConfig()
    _validate(request, config)
    scen = _central_scenarios(request, config) + _rate_corr_scenarios(request, config)
    res = api.price_batch(scen, precision, max_dimension=max_dimension)
    return _batch_report(request, config, res)


@dataclass
class FactorMove:
    """lib.rs:630-635."""

    label: str
    contribution: float


def pnl_explain(report: BatchRiskReport, spot_moves, vol_moves, rate_move: float,
                corr_move: float) -> List[FactorMove]:
    """lib.rs:637-680: delta*ds + gamma*ds^2/2, vega*dvol, dv01*dr, corr_sens*drho."""
    out = []
    for a in report.assets:
        i = a.asset_index
        ds = spot_moves[i] if i < len(spot_moves) else 0.0
        dv = vol_moves[i] if i < len(vol_moves) else 0.0
        if ds != 0.0 or a.delta != 0.0 or a.gamma != 0.0:
            out.append(FactorMove(f"spot[{i}]", a.delta * ds + 0.5 * a.gamma * ds * ds))
        if dv != 0.0 or a.vega != 0.0:
            out.append(FactorMove(f"vol[{i}]", a.vega * dv))
    if rate_move != 0.0:
        out.append(FactorMove("rate", report.dv01() * rate_move))
    if corr_move != 0.0:
        out.append(FactorMove("correlation", report.correlation_sensitivity() * corr_move))
    return out


@dataclass
class PortfolioRisk:
    """lib.rs:682-692."""

    trades: int = 0
    pv: float = 0.0
    delta: List[float] = field(default_factory=list)
    gamma: List[float] = field(default_factory=list)
    vega: List[float] = field(default_factory=list)
    dv01: float = 0.0
    correlation_sensitivity: float = 0.0


def aggregate_portfolio(rows) -> PortfolioRisk:
    """lib.rs:694-720: notional-weighted sums of (notional, BatchRiskReport) rows."""
    out = PortfolioRisk(trades=len(rows))
    for notional, rep in rows:
        out.pv += notional * rep.base_pv
        out.dv01 += notional * rep.dv01()
        out.correlation_sensitivity += notional * rep.correlation_sensitivity()
        for a in rep.assets:
            i = a.asset_index
            while len(out.delta) <= i:
                out.delta.append(0.0)
                out.gamma.append(0.0)
                out.vega.append(0.0)
            out.delta[i] += notional * a.delta
            out.gamma[i] += notional * a.gamma
            out.vega[i] += notional * a.vega
    return out


# ---------------------------------------------------------------------------- adapter batches
def _ko_rate(res: PriceResult) -> float:
    """main.rs:1245-1256: population_after / population_before of daily_global_ko."""
    for e in res.ledger:
        if e.transition_id == "daily_global_ko":
            return e.population_after / e.population_before if e.population_before > 0 else 0.0
    return 0.0


def risk_detailed(request: PriceRequest, precision: str = "fp64", max_dimension: int = 512) -> dict:
    """riga-mcp/src/main.rs:1355-1424: the batch risk report plus theta (maturity one
    calendar day shorter, same steps) and the cross-gamma of every factor pair from a
    four-point stencil of 1% relative spot moves; every valuation in one batch."""
    config = CentralDifferenceConfig()
    _validate(request, config)
    F = len(request.market.spots)
    scen = _central_scenarios(request, config) + _rate_corr_scenarios(request, config)
    decayed = request.clone()
    decayed.trade.maturity_years = max(request.trade.maturity_years - 1.0 / 365.0, 1.0 / 365.0)
    scen.append(decayed)
    half = [s * 0.01 for s in request.market.spots]
    pairs = [(i, j) for i in range(F) for j in range(i + 1, F)]
    for i, j in pairs:
        for di, dj in ((1.0, 1.0), (1.0, -1.0), (-1.0, 1.0), (-1.0, -1.0)):
            q = request.clone()
            q.market.spots[i] += di * half[i]
            q.market.spots[j] += dj * half[j]
            scen.append(q)
    res = api.price_batch(scen, precision, max_dimension=max_dimension)
    n = 5 + 4 * F
    report = _batch_report(request, config, res[:n])
    theta = res[n].pv - report.base_pv
    cross = [[0.0] * F for _ in range(F)]
    for k, (i, j) in enumerate(pairs):
        pp, pm, mp, mm = (r.pv for r in res[n + 1 + 4 * k:n + 5 + 4 * k])
        cross[i][j] = cross[j][i] = (pp - pm - mp + mm) / (4.0 * half[i] * half[j])
    return {
        "base_pv": report.base_pv,
        "assets": [{"asset_index": a.asset_index, "delta": a.delta, "gamma": a.gamma, "vega": a.vega}
                   for a in report.assets],
        "dv01": report.dv01(), "rho": report.dv01(),
        "correlation_sensitivity": report.correlation_sensitivity(),
        "theta": theta, "fx_delta": 0.0, "cross_gamma": cross,
        "bucket_vegas": [{"bucket": f"#{a.asset_index}", "vega": a.vega} for a in report.assets],
        "scenario_count": len(scen),
    }


DESIGN_STEP = 0.01


def trade_design(request: PriceRequest, precision: str = "fp64", max_dimension: int = 512) -> dict:
    """riga-mcp/src/main.rs:1426-1537: the base and one valuation per trade field moved
    by 0.01 (strike, KI barrier, and the KO barrier, coupon barrier and coupon rate
    when present), on the base seed, in one batch."""
    def bumped(name):
        q = request.clone()
        t = q.trade
        if name == "strike":
            t.strike += DESIGN_STEP
        elif name == "knock_in_barrier":
            t.knock_in_barrier += DESIGN_STEP
        elif getattr(t, name) is not None:
            setattr(t, name, getattr(t, name) + DESIGN_STEP)
        return q

    names = ["strike", "knock_in_barrier", "knock_out_barrier", "coupon_barrier", "coupon_rate"]
    res = api.price_batch([request] + [bumped(n) for n in names], precision, max_dimension=max_dimension)

    def point(r):
        return {"pv": r.pv, "ki": r.ki_probability, "ko": _ko_rate(r), "coupon_pv": r.legs.coupon,
                "memory_pv": r.legs.memory_coupon, "put_pv": r.legs.down_and_in_put, "funding": r.legs.funding}

    base = point(res[0])
    detailed = request.config.trace_level != "off"
    params = []
    for n, r in zip(names, res[1:]):
        p = point(r)
        if detailed:
            v = {"step": DESIGN_STEP, "pv": p["pv"], "delta_pv": p["pv"] - base["pv"], "ki": p["ki"],
                 "delta_ki": p["ki"] - base["ki"], "ko": p["ko"], "delta_ko": p["ko"] - base["ko"],
                 "coupon_pv": p["coupon_pv"], "put_pv": p["put_pv"], "funding": p["funding"]}
        else:
            v = {"step": DESIGN_STEP, "delta_pv": p["pv"] - base["pv"]}
        v["parameter"] = n
        params.append(v)
    return {"base": base, "step": DESIGN_STEP, "detailed": detailed, "scenario_count": 1 + len(params),
            "parameters": params}


def convergence_ladder(max_paths: int) -> List[int]:
    """riga-mcp/src/main.rs:1216-1227: max/32 .. max/2 and max, deduplicated."""
    m = max(max_paths, 1)
    return sorted({max(m // d, 1) for d in (32, 16, 8, 4, 2)} | {m})


def mc_convergence(request: PriceRequest, ladder: Optional[List[int]] = None, precision: str = "fp64",
                   max_dimension: int = 512) -> dict:
    """riga-mcp/src/main.rs:1229-1278: PV, SE, 95% band, KI and KO rates per path
    count on the base seed (CRN-paired points), every point in one batch; with
    tracing off only the last point."""
    ladder = convergence_ladder(request.config.paths) if ladder is None else list(ladder)
    last = ladder[-1] if ladder else request.config.paths
    if request.config.trace_level == "off":
        ladder = [last]
    scen = []
    for n in ladder:
        q = request.clone()
        q.config.paths = n
        scen.append(q)
    res = api.price_batch(scen, precision, max_dimension=max_dimension)
    points = []
    for n, r in zip(ladder, res):
        se = r.mc_error
        points.append({"paths": n, "pv": r.pv, "se": se, "relative_se": se / abs(r.pv) if abs(r.pv) > 0 else 0.0,
                       "lower": r.pv - 1.96 * se, "upper": r.pv + 1.96 * se, "ki": r.ki_probability,
                       "ko": _ko_rate(r)})
    return {"points": points, "final": points[-1] if points else None,
            "trace_level": {"off": "Off", "aggregated": "Aggregated", "sampled": "Sampled",
                            "full": "Full"}.get(request.config.trace_level, request.config.trace_level)}
