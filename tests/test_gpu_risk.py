"""CRN finite-difference risk on the device (riga_flow_amd.risk) against the same
formulas applied to oracle prices of the same scenarios (riga-risk/src/lib.rs:117-186,
212-258, 296-336): FP64 PVs within 1e-10, so every Greek and the variance-reduction
factor follow to the stated tolerances; each scenario priced in the batch equals
pricing it alone."""
import math

import pytest

import oracle as O
from parity import rel
from riga_flow_amd import api, risk, workloads as W
from riga_flow_amd.types import LocalVolSpec, RigaError

pytestmark = pytest.mark.gpu


def implied_lv_request(paths=1024, steps=63, years=0.25, F=2):
    # the reference's own shape: two assets, implied grids calibrated per request
    req = W.lv_book(paths=paths, steps=steps, years=years, F=F)
    sig = W.svi_sigmas(F)
    req.local_vol = LocalVolSpec(grids=[W.svi_implied(sig[f]) for f in range(F)], grid_x=96)
    req.trade.call_start_step = 5
    return req


def test_central_greeks_equal_the_fd_of_oracle_prices():
    req = implied_lv_request()
    md = W.max_dimension_for(req)
    cfg = risk.CentralDifferenceConfig()
    rep = risk.price_central_greeks(req, cfg, "fp64", max_dimension=md)
    base = O.price(req, max_dim=md)
    assert rel(rep.base_price.pv, base.pv) <= 1e-10
    for a in rep.assets:
        f, h, vb = a.asset_index, a.spot_bump, a.volatility_bump
        up = O.price(risk._spot_bumped(req, f, h, req.config.seed), max_dim=md).pv
        dn = O.price(risk._spot_bumped(req, f, -h, req.config.seed), max_dim=md).pv
        vu = O.price(risk.bump_volatility_request(req, f, vb), max_dim=md).pv
        vd = O.price(risk.bump_volatility_request(req, f, -vb), max_dim=md).pv
        # PVs agree to 1e-10 relative (~1e-10 absolute), so a difference over 2h (h = 1)
        # or h^2 agrees to a few 1e-10 absolute
        assert abs(a.delta.value - (up - dn) / (2 * h)) <= 5e-10
        assert abs(a.gamma.value - (up - 2 * base.pv + dn) / (h * h)) <= 5e-10
        assert abs(a.vega.value - (vu - vd) / (2 * vb)) <= 5e-8
        assert a.delta.mc_error_upper_bound > 0
    # a scenario of the batch alone gives the same number (CRN grouping is exact)
    alone = api.price_put_leg(risk._spot_bumped(req, 1, -rep.assets[1].spot_bump, req.config.seed),
                              "fp64", max_dimension=md)
    assert alone.pv == rep.assets[1].spot_down.pv


def test_crn_variance_reduction_factor_matches_oracle_and_exceeds_one():
    req = implied_lv_request(paths=512, steps=42, years=1.0 / 6.0)
    md = W.max_dimension_for(req)
    R = 4
    factor, crn, ind = risk.crn_variance_reduction_factor(req, replicates=R, precision="fp64",
                                                          max_dimension=md, details=True)
    h = req.market.spots[0] * 0.01
    o_crn, o_ind = [], []
    for r in range(R):
        p = lambda seed, d: O.price(risk._spot_bumped(req, 0, d, seed), max_dim=md).pv
        o_crn.append((p(1_000 + r, h) - p(1_000 + r, -h)) / (2 * h))
        o_ind.append((p(50_000 + 3 * r, h) - p(50_000 + 3 * r + 1, -h)) / (2 * h))
    for a, b in zip(crn + ind, o_crn + o_ind):
        assert abs(a - b) <= 5e-10
    o_factor = risk._sample_variance(o_ind) / risk._sample_variance(o_crn)
    assert rel(factor, o_factor) <= 1e-4
    assert factor > 1.0


def test_delta_convergence_and_rejections():
    req = implied_lv_request(paths=512, steps=42, years=1.0 / 6.0)
    md = W.max_dimension_for(req)
    est = risk.crn_delta_convergence(req, replicates=3, precision="fp64", max_dimension=md)
    assert est.replicates == 3 and est.standard_error > 0 and math.isfinite(est.value)
    with pytest.raises(RigaError) as e:
        risk.crn_variance_reduction_factor(req, replicates=1)
    assert e.value.code == "risk_replicates_required"
    with pytest.raises(RigaError) as e:
        risk.price_central_greeks(req, risk.CentralDifferenceConfig(spot_relative_bump=1.5))
    assert e.value.code == "invalid_spot_bump"
    cal = W.lv_book(paths=512, steps=42, years=1.0 / 6.0, F=2)  # calibrated grids only
    with pytest.raises(RigaError) as e:
        risk.price_central_greeks(cal, max_dimension=md)
    assert e.value.code == "risk_surface_shape"
