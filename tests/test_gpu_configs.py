"""Each workload configuration at its own shape on the device, against the oracle.

* plain FP64 worst-of Phoenix, 3 assets flat vol, 65,536 Sobol paths x 252 steps
  (the full configuration): every path's KO step and KI flag equal the oracle's,
  per-path values within 1e-12, the priced aggregates within 1e-10;
* the 10 CRN scenarios (base, spot +-1% per asset, vega +1 vol point per asset) of
  the 3-asset local-vol 3y x 756-step trade on a 2,048-path subsample of its shared
  Sobol points: FP64 per scenario within 1e-10 with counts exact, and
  delta/gamma/vega from the device's prices against the same finite differences
  of the oracle's prices (riga-risk/src/lib.rs:117-186, 260-279); FP32 on the
  configuration's full 1,048,576 paths within 1e-4 of FP64 per scenario;
* the 2,048-trade uneven book of the 20-asset universe in one plan: FP64 and FP32
  results identical with the shared normal table on and off at 65,536 paths, FP32
  within 1e-4 of FP64 on 99% of the trades (all within 5e-4), and 8 sampled trades
  at their own maturities against the oracle at 1,024 paths (FP64);
* the serving workload over the resident 20-asset local-vol market, with
  requests at 16k, 65k, 262k and 1M paths: every ticket bit-identical alone and in
  the mix, in both tiers.

The oracle runs over contiguous path ranges (or one request each) in threads:
Sobol points are independent per path and the oracle's C calls release the GIL.
"""
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from parity import FP32_REL, assert_fp64_parity, rel
from riga_flow_amd import api, serving, workloads as W

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _pmap(fn, args):
    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        return list(ex.map(fn, args))


def test_plain_fp64_full_configuration_every_path_event_exact():
    req = W.flat_book(paths=65_536, steps=252)
    md = W.max_dimension_for(req)
    res = api.price_put_leg(req, "fp64", max_dimension=md)
    recs, rres = api.replay_paths(req, "fp64", max_dimension=md)
    n = req.config.paths
    cuts = np.linspace(0, n, THREADS + 1).astype(int)
    ranges = [(int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]
    # the full sequential oracle for the aggregates, path ranges for the records
    with ThreadPoolExecutor(max_workers=THREADS + 1) as ex:
        full = ex.submit(O.price, req, md)
        parts = list(ex.map(lambda r: O.price(req, max_dim=md, path_range=r, records=True)[1], ranges))
        ora = full.result()
    orecs = np.concatenate(parts)
    assert len(orecs) == n
    assert np.array_equal(recs["ko_step"], orecs["ko_step"])  # zero KO event deltas
    assert np.array_equal(recs["hit_ki"], orecs["hit_ki"])    # zero KI event deltas
    for f in ("final_worst", "funding", "put", "coupon_pv", "memory_pv", "pv"):
        np.testing.assert_allclose(recs[f], orecs[f], rtol=1e-12, atol=1e-15)
    assert_fp64_parity(res, ora)
    assert_fp64_parity(rres, ora)
    assert ora.ko_count > 10_000 and ora.ki_count > 100  # both events populated


def _fd_greeks(pv, spots):
    """Delta and gamma from spot +-1% (central), vega from +1 vol point (one-sided):
    riga-risk/src/lib.rs:117-186 with the workload's one-sided vega bump."""
    F = len(spots)
    base = pv[0]
    out = []
    for f in range(F):
        up, dn = pv[1 + 2 * f], pv[2 + 2 * f]
        h = 0.01 * spots[f]
        out.append(((up - dn) / (2 * h), (up - 2 * base + dn) / (h * h), (pv[1 + 2 * F + f] - base) / 0.01))
    return np.array(out)


def test_crn_ten_scenarios_on_the_756_step_trade_against_the_oracle():
    names, reqs = W.crn_scenarios(paths=2048, steps=756, years=3.0)
    assert len(reqs) == 10 and reqs[0].trade.steps == 756
    md = max(W.max_dimension_for(r) for r in reqs)
    f64 = api.price_batch(reqs, "fp64", max_dimension=md)
    ora = _pmap(lambda q: O.price(q, max_dim=md), reqs)
    for name, a, o in zip(names, f64, ora):
        assert_fp64_parity(a, o)
    spots = reqs[0].market.spots
    g_ora = _fd_greeks([o.pv for o in ora], spots)
    g64 = _fd_greeks([r.pv for r in f64], spots)
    # the finite differences of prices that agree within 1e-10 relative
    # (absolute 1e-10 / h^2 for the second difference)
    np.testing.assert_allclose(g64, g_ora, rtol=1e-7, atol=1e-9)


def test_crn_ten_scenarios_fp32_at_the_full_1m_paths_within_1e4():
    """FP32 on the configuration's own 1,048,576 shared paths against the FP64 tier
    on the same paths (bit-exact to the oracle above and at the plain
    configuration's full size): every scenario's PV within 1e-4.  At a few thousand
    paths a single KO or KI decision flipped by float rounding moves a PV by
    1/paths of a payoff, so the 1e-4 bar is a full-size one.  The Greeks' FP32
    errors carry the same flips through differences of nearby PVs; their bars:
    delta 1e-4, gamma 1e-4 (absolute), vega 1e-3 absolute."""
    names, reqs = W.crn_scenarios(paths=1_048_576, steps=756, years=3.0)
    md = max(W.max_dimension_for(r) for r in reqs)
    f32 = api.price_batch(reqs, "fp32", max_dimension=md)
    f64 = api.price_batch(reqs, "fp64", max_dimension=md)
    for name, a, b in zip(names, f32, f64):
        assert rel(a.pv, b.pv) <= FP32_REL, (name, a.pv, b.pv)
        assert rel(a.legs.funding, b.legs.funding) <= FP32_REL, name
    spots = reqs[0].market.spots
    g32 = _fd_greeks([r.pv for r in f32], spots)
    g64 = _fd_greeks([r.pv for r in f64], spots)
    assert np.all(np.abs(g32[:, 0] - g64[:, 0]) <= 1e-4), (g32[:, 0], g64[:, 0])
    assert np.all(np.abs(g32[:, 1] - g64[:, 1]) <= 1e-4), (g32[:, 1], g64[:, 1])
    assert np.all(np.abs(g32[:, 2] - g64[:, 2]) <= 1e-3), (g32[:, 2], g64[:, 2])


def test_uneven_book_of_2048_trades_normal_table_on_and_off_and_the_oracle():
    book = W.uneven_book()
    assert len(book) == 2048 and book[0].config.paths == 65_536
    md = max(W.max_dimension_for(q) for q in book)
    for tier in ("fp64", "fp32"):
        plan_on = api.Plan(book, tier, max_dimension=md, normal_table=1)
        assert plan_on.info.normal_tables == 1
        on = plan_on.run()
        plan_on.close()
        off = api.price_batch(book, tier, max_dimension=md, normal_table=-1)
        for i, (a, b) in enumerate(zip(on, off)):
            assert a.pv == b.pv and a.legs.down_and_in_put == b.legs.down_and_in_put, (tier, i)
            assert a.stats[0].output_count == b.stats[0].output_count, (tier, i)
        if tier == "fp64":
            f64_full = on
        else:
            f32_full = on
    # FP32 against FP64 on every trade's 65,536 paths: a KO/KI decision flipped by
    # float rounding moves a trade's PV by up to 1/65,536 of a payoff, so the bar
    # is on the book's distribution: 99% of trades within 1e-4, all within 5e-4
    err = np.array([rel(a.pv, b.pv) for a, b in zip(f32_full, f64_full)])
    assert np.quantile(err, 0.99) <= FP32_REL and err.max() <= 5 * FP32_REL, (np.quantile(err, 0.99), err.max())
    small = W.uneven_book(paths=1024)
    f64 = api.price_batch(small, "fp64", max_dimension=md)
    rng = np.random.default_rng(17)
    pick = sorted(rng.choice(len(small), size=8, replace=False).tolist())
    # the sample spans the book's factor counts and maturities
    assert len({small[i].factors for i in pick}) >= 3
    ora = _pmap(lambda i: O.price(small[i], max_dim=md), pick)
    for i, o in zip(pick, ora):
        assert_fp64_parity(f64[i], o)


def test_serving_over_the_resident_20_asset_market_alone_equals_mix():
    u = serving.Universe()  # 20 assets, 5y local-vol surfaces, bumped surfaces calibrated once
    u.make_resident()
    rng = np.random.default_rng(21)
    pool = []
    for paths in (16_384, 65_536, 262_144, 1_048_576):
        for _ in range(2 if paths < 1_000_000 else 1):
            pool.append(serving.sample_request(u, rng, path_choices=(paths,), max_trades=3, max_bumps=3))
    tiers = ["fp64" if i % 2 == 0 else "fp32" for i in range(len(pool))]
    assert {q.config.paths for reqs in pool for q in reqs} == {16_384, 65_536, 262_144, 1_048_576}
    srv = serving.Server(max_inflight=3, max_batch_requests=64)
    ids = [srv.submit(srv.pack(reqs), tier) for reqs, tier in zip(pool, tiers)]
    got = [srv.wait(t) for t in ids]
    stats = srv.stats()
    srv.close()
    assert all(g.state == serving.DONE for g in got)
    assert stats["completed_tickets"] == len(pool)
    for reqs, tier, g in zip(pool, tiers, got):
        alone = api.price_batch(reqs, tier, max_dimension=7000)
        assert len(g.results) == len(alone)
        for a, b in zip(g.results, alone):
            assert a.pv == b.pv and a.legs.down_and_in_put == b.legs.down_and_in_put
            assert a.stats[0].output_count == b.stats[0].output_count
            assert a.trace_events == b.trace_events
