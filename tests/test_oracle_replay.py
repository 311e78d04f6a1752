"""The oracle's replay records and coupon strips (explain.rs:90-316) and its
PathCube restatement (pathcube.rs:47-187), pinned by the reference's own replay
tests (explain.rs:389-455, pathcube.rs:193-290) and by invariants that tie the
strips to the pricing legs (themselves pinned by the golden result).
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from riga_flow_amd import workloads as W
from riga_flow_amd.fixture import load_termsheet1_fixture
from riga_flow_amd.pathcube import PathCube, _ryu32

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def termsheet(paths=2000):
    req = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json")).request
    req.config.paths = paths
    req.config.path_block_size = 512
    return req


def test_records_aggregate_to_the_pricing_result():
    # explain.rs:389-403
    req = termsheet()
    res, rec, led = O.price(req, ledger=True)
    assert abs(rec["put"].mean() - res.leg_put) < 1e-12
    assert abs(rec["hit_ki"].mean() - res.ki_probability) < 1e-12
    # the strip's pv adds up to the path's coupon legs (coupon + memory)
    np.testing.assert_allclose(led["pv"].sum(axis=1), rec["coupon_pv"] + rec["memory_pv"], rtol=1e-14,
                               atol=1e-16)


def test_replay_continues_the_market_path_after_autocall():
    # explain.rs:405-425
    base = termsheet()
    base.trade.knock_out_barrier = None
    early = termsheet()
    early.trade.knock_out_barrier = 1e-6
    _, rb = O.price(base, records=True)
    _, re = O.price(early, records=True)
    assert (rb["put"] > 0).any()
    assert np.array_equal(rb["final_worst"], re["final_worst"])
    assert np.array_equal(rb["hit_ki"], re["hit_ki"])
    assert (re["ko_step"] >= 0).all() and (re["put"] == 0).all()


def test_strip_rows_follow_the_schedule():
    req = termsheet(500)
    res, rec, led = O.price(req, ledger=True)
    T, P = req.trade.steps, req.trade.coupon_periods
    dt = req.trade.maturity_years / T
    per = [max(T * (p + 1) // P - T * p // P, 1) for p in range(P)]
    assert (led["period"] == np.arange(P)).all()
    assert (led["period_steps"] == per).all()
    assert (led["qualified_steps"] <= led["period_steps"]).all()
    for i in range(len(rec)):
        ko = rec["ko_step"][i]
        rows = led[i]
        assert rows["knocked_out"].sum() == (1 if ko >= 0 else 0)
        if ko >= 0:
            q = int(min(ko * P // T, P - 1))
            assert rows["knocked_out"][q] == 1 and rows["years"][q] == (ko + 1) * dt
            assert (rows["years"][q + 1:] == 0).all() and (rows["pv"][q + 1:] == 0).all()
        paid = rows["paid"] > 0
        np.testing.assert_allclose(rows["paid"][paid], rows["accrual"][paid] + rows["memory_paid"][paid],
                                   rtol=1e-15)
        np.testing.assert_allclose(rows["pv"], rows["paid"] * rows["discount_factor"], rtol=1e-15)


def test_strips_when_every_day_qualifies_and_when_none_does():
    req = W.flat_book(paths=64, steps=120, years=120 / 252)
    req.trade.knock_out_barrier = None
    req.trade.coupon_periods = 6
    req.trade.coupon_barrier = 1e-9  # every observation qualifies
    _, rec, led = O.price(req, ledger=True)
    assert (led["qualified_steps"] == led["period_steps"]).all()
    np.testing.assert_allclose(led["accrual"], req.trade.coupon_rate, rtol=1e-14)
    assert (led["memory_paid"] == 0).all() and (led["memory_accrued"] == 0).all()
    req.trade.coupon_barrier = 1e9  # none does: memory builds every period
    req.trade.memory_coupon = True
    _, rec, led = O.price(req, ledger=True)
    assert (led["qualified_steps"] == 0).all() and (led["paid"] == 0).all()
    assert (led["memory_accrued"] == req.trade.coupon_rate).all()
    np.testing.assert_allclose(led["years"][0], (np.array([20, 40, 60, 80, 100, 120])) * (req.trade.maturity_years / 120),
                               rtol=1e-15)


def test_correlated_block_reproduces_the_target_correlation():
    # pathcube.rs:193-221
    rho = 0.4041
    blk = O.correlated_block(1729, 200_000, 1, [[1.0, rho], [rho, 1.0]])
    c = np.corrcoef(blk[:, 0, 0].astype(np.float64), blk[:, 0, 1].astype(np.float64))[0, 1]
    assert abs(c - rho) < 0.01


def test_block_is_the_stream_through_the_cholesky_factor():
    corr = [[1.0, 0.5, 0.2], [0.5, 1.0, 0.3], [0.2, 0.3, 1.0]]
    blk = O.correlated_block(7, 5, 4, corr)
    z, _ = O.std_normals(7, 5 * 4 * 3)
    L = O.cholesky(corr)
    for p in range(5):
        for t in range(4):
            k = (p * 4 + t) * 3
            np.testing.assert_array_equal(blk[p, t], O.chol_apply(L, z[k:k + 3]).astype(np.float32))
    again = O.correlated_block(7, 5, 4, corr)
    assert np.array_equal(blk, again)


def test_gbm_evolve_is_the_exponential_of_summed_log_steps():
    blk = O.correlated_block(3, 64, 50, [[1.0, 0.2], [0.2, 1.0]])
    s0, vols, r, q, dt = [100.0, 50.0], [0.2, 0.3], 0.03, [0.0, 0.01], 1 / 252
    cube = O.gbm_evolve(s0, vols, r, q, dt, blk)
    for f in range(2):
        drift = (r - q[f] - 0.5 * vols[f] ** 2) * dt
        logs = math.log(s0[f]) + np.cumsum(drift + vols[f] * math.sqrt(dt) * blk[:, :, f].astype(np.float64),
                                           axis=1)
        np.testing.assert_allclose(cube[:, :, f], np.exp(logs).astype(np.float32), rtol=3e-7)


def test_pathcube_archive_round_trips(tmp_path):
    # pathcube.rs:267-290
    blk = O.correlated_block(3, 8, 4, [[1.0, 0.2], [0.2, 1.0]])
    cube = PathCube(8, 4, 2, O.gbm_evolve([100.0, 50.0], [0.2, 0.3], 0.03, [0.0, 0.0], 0.01, blk))
    f = tmp_path / "cube.json"
    cube.write_archive(f)
    back = PathCube.read_archive(f)
    assert back == cube
    t5 = back.worst_percentile_threshold([100.0, 50.0], 0.05)
    assert math.isfinite(t5) and t5 > 0 and t5 <= back.worst_percentile_threshold([100.0, 50.0], 0.5)
    text = f.read_text()
    assert text.startswith('{"paths":8,"times":4,"factors":2,"data":[')


def test_f32_json_text():
    cases = {1.0: "1.0", 0.1: "0.1", 100.25: "100.25", 1e-7: "1e-7", 3.4028235e38: "3.4028235e38",
             123456.79: "123456.79", 0.00012: "0.00012"}
    for v, txt in cases.items():
        assert _ryu32(float(np.float32(v))) == txt, (v, _ryu32(float(np.float32(v))))
