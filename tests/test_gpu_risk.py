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


def termsheet_small():
    # riga-risk/src/lib.rs:724-732: the reference's risk tests run on this request
    # (the ChaCha8 Monte Carlo lane, implied grids recalibrated per scenario)
    import os
    from riga_flow_amd.fixture import load_termsheet1_fixture
    req = load_termsheet1_fixture(os.path.join(os.path.dirname(__file__), "golden", "termsheet1.json")).request
    req.config.paths = 512
    req.config.path_block_size = 128
    req.config.seed = 81_281
    return req


def _same(a, b):
    assert a.pv == b.pv and a.mc_error == b.mc_error and a.legs == b.legs
    assert [s.output_count for s in a.stats] == [s.output_count for s in b.stats]


def test_reference_central_greeks_tests_on_the_device():
    # riga-risk/src/lib.rs:734-811, 847-883
    req = termsheet_small()
    cfg = risk.CentralDifferenceConfig()
    first = risk.price_central_greeks(req, cfg)
    second = risk.price_central_greeks(req, cfg)
    assert first.base_price.pv == second.base_price.pv and first.seed == req.config.seed
    assert len(first.assets) == 2
    a = first.assets[0]
    h = req.market.spots[0] * cfg.spot_relative_bump
    up = req.clone()
    up.market.spots[0] += h
    _same(a.spot_up, api.price_put_leg(up, "fp64"))
    assert a.delta.value == (a.spot_up.pv - a.spot_down.pv) / (2.0 * h)
    assert a.gamma.value == (a.spot_up.pv - 2.0 * first.base_price.pv + a.spot_down.pv) / (h * h)
    assert a.delta.mc_error_upper_bound == (a.spot_up.mc_error + a.spot_down.mc_error) / (2.0 * h)
    assert a.gamma.mc_error_upper_bound == \
        (a.spot_up.mc_error + 2.0 * first.base_price.mc_error + a.spot_down.mc_error) / (h * h)
    assert a.vega.mc_error_upper_bound == (a.vol_up.mc_error + a.vol_down.mc_error) / (2.0 * 0.01)
    for asset in first.assets:
        for x in (asset.delta, asset.gamma, asset.vega):
            assert math.isfinite(x.mc_error_upper_bound)
    # sticky-strike vega: every implied node of the asset shifted, recalibrated
    vu, vd = req.clone(), req.clone()
    for q, d in ((vu, 0.01), (vd, -0.01)):
        q.market.vols[0] += d
        g = q.local_vol.grids[0]
        g.vols = [[v + d for v in row] for row in g.vols]
    _same(a.vol_up, api.price_put_leg(vu, "fp64"))
    _same(a.vol_down, api.price_put_leg(vd, "fp64"))
    assert abs(a.vega.value) > 1e-8


def test_reference_crn_tests_on_the_device():
    # riga-risk/src/lib.rs:835-845, 919-939
    req = termsheet_small()
    assert risk.crn_variance_reduction_factor(req, replicates=40) > 1.5
    coarse = risk.crn_delta_convergence(req, replicates=16)
    fine = risk.crn_delta_convergence(req, replicates=64)
    assert coarse.standard_error > 0 and fine.standard_error > 0
    assert fine.standard_error <= coarse.standard_error


def test_reference_batch_risk_tests_on_the_device():
    # riga-risk/src/lib.rs:941-978
    req = termsheet_small()
    cfg = risk.CentralDifferenceConfig()
    rep = risk.price_risk_batch(req, cfg)
    assert rep.scenario_count == 1 + 4 * 2 + 4 and len(rep.assets) == 2
    assert math.isfinite(rep.base_pv) and math.isfinite(rep.dv01()) and math.isfinite(rep.correlation_sensitivity())
    rate = [m for m in risk.pnl_explain(rep, [0.0, 0.0], [0.0, 0.0], 0.0001, 0.0) if m.label == "rate"][0]
    assert abs(rate.contribution - rep.dv01() * 0.0001) < 1e-15
    pf = risk.aggregate_portfolio([(10_000.0, rep), (5_000.0, rep)])
    assert pf.trades == 2 and abs(pf.pv - 15_000.0 * rep.base_pv) < 1e-9 and len(pf.delta) == 2
    h = req.market.spots[0] * cfg.spot_relative_bump
    bumped = req.clone()
    bumped.market.spots[0] += h
    actual = api.price_put_leg(bumped, "fp64").pv - rep.base_pv
    total = sum(m.contribution for m in risk.pnl_explain(rep, [h, 0.0], [0.0, 0.0], 0.0, 0.0))
    assert abs(total - actual) < 1e-3
    # the rate and correlation scenarios, each priced alone
    ru = req.clone()
    ru.market.risk_free_rate += 1e-4
    assert api.price_put_leg(ru, "fp64").pv == rep.rate_up_pv
    cd = req.clone()
    cd.market.correlation = [[1.0, req.market.correlation[0][1] - 0.01], [req.market.correlation[1][0] - 0.01, 1.0]]
    assert api.price_put_leg(cd, "fp64").pv == rep.correlation_down_pv
    # transition risk: the KO transition's delta from its own bump-pair impacts
    names = [t.id for t in rep.transitions]
    assert names == ["daily_global_ko", "final_fixing_eki", "memory_coupon"]


def test_adapter_batches_equal_their_scenarios_priced_alone():
    # riga-mcp/src/main.rs:1216-1278, 1355-1537 with every valuation in one batch
    req = termsheet_small()
    det = risk.risk_detailed(req)
    assert det["scenario_count"] == 13 + 1 + 4
    half = [s * 0.01 for s in req.market.spots]
    pv = {}
    for di, dj in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
        q = req.clone()
        q.market.spots[0] += di * half[0]
        q.market.spots[1] += dj * half[1]
        pv[(di, dj)] = api.price_put_leg(q, "fp64").pv
    cg = (pv[(1, 1)] - pv[(1, -1)] - pv[(-1, 1)] + pv[(-1, -1)]) / (4.0 * half[0] * half[1])
    assert det["cross_gamma"][0][1] == cg == det["cross_gamma"][1][0]
    dec = req.clone()
    dec.trade.maturity_years = max(req.trade.maturity_years - 1 / 365, 1 / 365)
    assert det["theta"] == api.price_put_leg(dec, "fp64").pv - det["base_pv"]
    td = risk.trade_design(req)
    assert td["scenario_count"] == 6 and [p["parameter"] for p in td["parameters"]] == \
        ["strike", "knock_in_barrier", "knock_out_barrier", "coupon_barrier", "coupon_rate"]
    q = req.clone()
    q.trade.strike += 0.01
    assert td["parameters"][0]["delta_pv"] == api.price_put_leg(q, "fp64").pv - td["base"]["pv"]
    conv = risk.mc_convergence(req)
    assert [p["paths"] for p in conv["points"]] == risk.convergence_ladder(512)
    q = req.clone()
    q.config.paths = 64
    r = api.price_put_leg(q, "fp64")
    p64 = [p for p in conv["points"] if p["paths"] == 64][0]
    assert p64["pv"] == r.pv and p64["se"] == r.mc_error and conv["final"]["paths"] == 512
