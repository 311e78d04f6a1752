"""Replay, coupon strips, the worst-percentile query and the PathCube on the
device (explain.rs:90-372, pathcube.rs:47-187) against the oracle, plus the
reference's own replay and PathCube tests (explain.rs:389-455,
pathcube.rs:193-290) run on the device.

Bars: FP64 records and coupon strips match the oracle's (integers and KO/KI
flags exactly, amounts and worst-of levels to the FP64 tier's 1e-10 relative:
the path's exp and log are the device's); block floats bit-exact; cube floats
within one f32 ulp.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from riga_flow_amd import explain as X
from riga_flow_amd import pathcube as PC
from riga_flow_amd import workloads as W
from riga_flow_amd.fixture import load_termsheet1_fixture
from riga_flow_amd.types import RigaError

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def termsheet(paths=2000):
    req = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json")).request
    req.config.paths = paths
    req.config.path_block_size = 512
    return req


def _strip_array(records):
    return np.array([[(c.period, c.qualified_steps, c.period_steps, c.knocked_out) for c in r.coupon_ledger]
                     for r in records])


def _strip_amounts(records):
    return np.array([[(c.years, c.accrual, c.memory_paid, c.paid, c.memory_accrued, c.discount_factor, c.pv)
                      for c in r.coupon_ledger] for r in records])


@pytest.mark.parametrize("lane", ["none", "sobol"])
def test_run_paths_strips_equal_the_oracle(lane):
    req = termsheet(1500)
    if lane == "sobol":
        req.config.variance_reduction = W.VarianceReduction("sobol", req.trade.steps * 2)
    recs = X.run_paths(req)
    _, orec, oled = O.price(req, ledger=True)
    assert [r.identity.path_id for r in recs] == list(range(1500))
    assert all(r.identity.sobol_index == (i + 1 if lane == "sobol" else 0) for i, r in enumerate(recs))
    np.testing.assert_allclose([r.worst_final for r in recs], orec["final_worst"], rtol=1e-10, atol=0)
    assert np.array_equal([r.knocked_out for r in recs], orec["ko_step"] >= 0)
    ints = _strip_array(recs)
    assert np.array_equal(ints[..., 0], oled["period"])
    assert np.array_equal(ints[..., 1], oled["qualified_steps"])
    assert np.array_equal(ints[..., 2], oled["period_steps"])
    assert np.array_equal(ints[..., 3], oled["knocked_out"])
    amt = _strip_amounts(recs)
    for k, f in enumerate(("years", "accrual", "memory_paid", "paid", "memory_accrued", "discount_factor", "pv")):
        np.testing.assert_allclose(amt[..., k], oled[f], rtol=1e-10, atol=0)
    np.testing.assert_allclose([r.coupon_pv for r in recs], orec["coupon_pv"] + orec["memory_pv"], rtol=1e-10)


def test_reconstructed_paths_aggregate_to_the_pricing_result():
    # explain.rs:389-403 on the device
    from riga_flow_amd import api
    req = termsheet()
    recs = X.run_paths(req)
    res = api.price_put_leg(req, "fp64")
    n = len(recs)
    assert abs(sum(r.put_payoff for r in recs) / n - res.legs.down_and_in_put) < 1e-12
    assert abs(sum(r.ki_hit for r in recs) / n - res.ki_probability) < 1e-12


def test_replay_continues_terminal_market_path_after_autocall():
    # explain.rs:405-425 on the device
    base = termsheet()
    base.trade.knock_out_barrier = None
    early = termsheet()
    early.trade.knock_out_barrier = 1e-6
    a = X.run_paths(base, ledger=False)
    b = X.run_paths(early, ledger=False)
    assert len(a) == len(b) and any(r.put_payoff > 0 for r in a)
    for x, y in zip(a, b):
        assert x.identity == y.identity and x.worst_final == y.worst_final and x.ki_hit == y.ki_hit
        assert y.knocked_out and y.put_payoff == 0.0


def test_a_single_path_reconstructs_from_its_identity():
    # explain.rs:427-441 on the device, and the path equals its row of the full replay
    req = termsheet()
    rec = X.reconstruct_path(req, 17)
    again = X.reconstruct_path(req, 17)
    assert rec.identity.path_id == 17 and rec.identity.rng_seed == req.config.seed
    assert rec == again or (rec.pv == again.pv and rec.worst_final == again.worst_final)
    full = X.run_paths(req)[17]
    assert rec.worst_final == full.worst_final and rec.pv == full.pv
    assert _strip_amounts([rec]).tolist() == _strip_amounts([full]).tolist()
    with pytest.raises(RigaError) as e:
        X.reconstruct_path(req, 2000)
    assert e.value.code == "path_out_of_range"


def test_worst_percentile_query_selects_the_lowest_paths():
    # explain.rs:443-455 on the device, and the sample is the oracle's lowest paths
    req = termsheet()
    q = X.query_worst_percentile(req, 0.05, 16)
    assert q.total == 2000 and q.selected == 100 and len(q.sample) == 16
    assert all(r.worst_final <= q.threshold_worst_final for r in q.sample)
    _, orec = O.price(req, records=True)
    order = np.argsort(orec["final_worst"], kind="stable")
    assert [r.identity.path_id for r in q.sample] == [int(i) for i in order[:16]]
    assert q.threshold_worst_final == pytest.approx(orec["final_worst"][order[99]], rel=1e-10)
    assert all(len(r.coupon_ledger) == 5 for r in q.sample)
    with pytest.raises(RigaError) as e:
        X.query_worst_percentile(req, 1.5, 4)
    assert e.value.code == "invalid_percentile"


@pytest.mark.parametrize("seed,paths,times,corr", [
    (7, 256, 64, [[1.0, 0.5], [0.5, 1.0]]),
    (3, 8, 4, [[1.0, 0.2], [0.2, 1.0]]),
    (11, 1000, 37, [[1.0, 0.3, 0.1], [0.3, 1.0, 0.4], [0.1, 0.4, 1.0]]),
    (5, 3, 500, [[1.0]]),
])
def test_block_and_cube_equal_the_oracle(seed, paths, times, corr):
    F = len(corr)
    s0 = [100.0, 50.0, 75.0][:F]
    vols = [0.2, 0.3, 0.25][:F]
    q = [0.0, 0.01, 0.02][:F]
    cube, blk = PC.path_cube(seed, paths, times, corr, s0, vols, 0.03, q, 1 / 252, with_block=True)
    want_blk = O.correlated_block(seed, paths, times, corr)
    assert np.array_equal(blk.data, want_blk)
    want = O.gbm_evolve(s0, vols, 0.03, q, 1 / 252, want_blk)
    ulps = np.abs(cube.data.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1 and (ulps == 0).mean() > 0.99
    b2 = PC.generate_correlated_block(seed, paths, times, corr)
    assert b2 == blk
    assert PC.gbm_evolve(s0, vols, 0.03, q, 1 / 252, blk) == cube


def test_cholesky_z_reproduces_the_target_correlation():
    # pathcube.rs:193-221 on the device
    rho = 0.4041
    blk = PC.generate_correlated_block(1729, 200_000, 1, [[1.0, rho], [rho, 1.0]])
    c = np.corrcoef(blk.data[:, 0, 0].astype(np.float64), blk.data[:, 0, 1].astype(np.float64))[0, 1]
    assert abs(c - rho) < 0.01


def test_factor_count_mismatch_and_archive(tmp_path):
    # pathcube.rs:245-290 on the device
    blk = PC.generate_correlated_block(1, 4, 4, [[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    with pytest.raises(RigaError) as e:
        PC.gbm_evolve([100.0, 100.0], [0.2, 0.2], 0.01, [0.0, 0.0], 0.01, blk)
    assert e.value.code == "shape_mismatch"
    blk = PC.generate_correlated_block(3, 8, 4, [[1.0, 0.2], [0.2, 1.0]])
    cube = PC.gbm_evolve([100.0, 50.0], [0.2, 0.3], 0.03, [0.0, 0.0], 0.01, blk)
    f = tmp_path / "cube.json"
    cube.write_archive(f)
    back = PC.PathCube.read_archive(f)
    assert back == cube
    t = back.worst_percentile_threshold([100.0, 50.0], 0.05)
    assert math.isfinite(t) and 0 < t <= back.worst_percentile_threshold([100.0, 50.0], 0.5)
