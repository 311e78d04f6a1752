"""The continuous-batching server on the device: each ticket's result is the
same alone and in the mix (bit for bit, both tiers), plans overlap, stale
tickets are shed and the memory budget is enforced."""
import time

import pytest

from parity import rel
from riga_flow_amd import api, serving, workloads as W

pytestmark = pytest.mark.gpu

MD = 7000


def small_tickets():
    names, crn = W.crn_scenarios(paths=2048, steps=126, years=0.5)
    book = W.uneven_book(n_trades=4, paths=1024, seed=9)
    for q in book:
        q.trade.steps = min(q.trade.steps, 200)
        q.trade.maturity_years = q.trade.steps / 252.0
        q.trade.coupon_periods = max(1, round(12 * q.trade.maturity_years))
        q.config.variance_reduction = W.VarianceReduction("sobol", q.trade.steps * q.factors)
    flat = [W.flat_book(paths=4096, steps=252)]
    return [(crn, "fp64"), (book, "fp64"), (flat, "fp32"), (crn[:3], "fp32")]


def test_results_alone_equal_results_in_the_mix():
    tickets = small_tickets()
    srv = serving.Server(max_inflight=2, max_batch_requests=64)
    ids = [srv.submit(reqs, tier) for reqs, tier in tickets]
    got = [srv.wait(t) for t in ids]
    stats = srv.stats()
    srv.close()
    assert all(g.state == serving.DONE for g in got)
    assert stats["completed_tickets"] == len(tickets)
    for (reqs, tier), g in zip(tickets, got):
        alone = api.price_batch(reqs, tier, max_dimension=MD)
        assert len(g.results) == len(alone)
        for a, b in zip(g.results, alone):
            assert a.pv == b.pv and a.legs.down_and_in_put == b.legs.down_and_in_put
            assert a.stats[0].output_count == b.stats[0].output_count


def test_plans_overlap_on_the_device():
    big = [W.lv_book(paths=262_144, steps=756)]
    srv = serving.Server(max_inflight=3, max_batch_requests=1)
    ids = [srv.submit(big, "fp64") for _ in range(3)]
    done = [srv.wait(t) for t in ids]
    stats = srv.stats()
    srv.close()
    assert all(d.state == serving.DONE for d in done)
    assert stats["max_concurrent"] >= 2
    assert rel(done[0].results[0].pv, done[2].results[0].pv) == 0.0


def test_a_ticket_arriving_mid_run_joins_the_running_device_work():
    # A long FP64 ticket (ten CRN scenarios on 1,048,576 paths of the 3y trade,
    # about a second of device work) and, once it runs, a small FP32 ticket: the
    # small ticket's workgroups take slots as the long launch's workgroups end, so
    # on the device clock it starts after the long one and finishes before it,
    # and its result equals the one it gets alone.
    _, crn = W.crn_scenarios(paths=1_048_576, steps=756, years=3.0)
    small = [W.flat_book(paths=4096, steps=252)]
    srv = serving.Server(max_inflight=3, max_batch_requests=64)
    long_id = srv.submit(crn, "fp64")
    t0 = time.time()
    while srv.stats()["running_plans"] < 1 and time.time() - t0 < 60.0:
        time.sleep(0.005)
    time.sleep(0.05)
    small_id = srv.submit(small, "fp32")
    s = srv.wait(small_id)
    long_r = srv.wait(long_id)
    stats = srv.stats()
    srv.close()
    assert s.state == serving.DONE and long_r.state == serving.DONE
    print(f"long ticket device {long_r.device_start_s:.4f}-{long_r.device_end_s:.4f} s, "
          f"small ticket device {s.device_start_s:.4f}-{s.device_end_s:.4f} s", flush=True)
    assert long_r.device_start_s >= 0.0 and s.device_start_s >= 0.0
    assert long_r.device_start_s < s.device_start_s
    assert s.device_end_s < long_r.device_end_s
    assert stats["overlapped_plans"] >= 1
    alone = api.price_batch(small, "fp32", max_dimension=MD)
    assert s.results[0].pv == alone[0].pv


def test_stale_tickets_are_shed_and_budget_is_enforced():
    big = [W.lv_book(paths=262_144, steps=756)]
    srv = serving.Server(max_inflight=1, max_batch_requests=1, max_queue_seconds=0.002)
    first = srv.submit(big, "fp64")
    time.sleep(0.05)
    later = [srv.submit(big, "fp64") for _ in range(3)]
    res = [srv.wait(t) for t in [first] + later]
    srv.close()
    assert res[0].state == serving.DONE
    assert any(r.state == serving.SHED for r in res[1:])
    tiny = serving.Server(hbm_budget_gib=1e-4)
    r = tiny.wait(tiny.submit(big, "fp64"))
    tiny.close()
    assert r.state == serving.FAILED and r.error.code == "memory_budget"


def test_resident_surfaces_price_as_sent_grids():
    # a surface copied to device memory once prices exactly as the grid sent per call
    req = W.lv_book(paths=4096, steps=252, years=1.0)
    res_req = api.resident(req)
    for tier in ("fp64", "fp32"):
        a = api.price_put_leg(req, tier, max_dimension=MD)
        b = api.price_put_leg(res_req, tier, max_dimension=MD)
        assert a.pv == b.pv and a.legs.down_and_in_put == b.legs.down_and_in_put
        assert a.trace_events == b.trace_events


def test_resident_market_tickets_alone_equal_in_the_mix():
    import numpy as np
    u = serving.Universe(n_assets=6, horizon=2.0)
    u.make_resident()
    rng = np.random.default_rng(3)
    pool = [serving.sample_request(u, rng, path_choices=(1024, 2048), max_trades=3, max_bumps=4)
            for _ in range(4)]
    for reqs in pool:
        for q in reqs:  # keep the horizons within the market's 2y tenors and short
            q.trade.steps = min(q.trade.steps, 126)
            q.trade.maturity_years = q.trade.steps / 252.0
            q.config.variance_reduction = W.VarianceReduction("sobol", q.trade.steps * q.factors)
    tiers = ["fp64", "fp32", "fp64", "fp32"]
    srv = serving.Server(max_inflight=2, max_batch_requests=64)
    ids = [srv.submit(srv.pack(reqs), tier) for reqs, tier in zip(pool, tiers)]
    got = [srv.wait(t) for t in ids]
    srv.close()
    assert all(g.state == serving.DONE for g in got)
    for reqs, tier, g in zip(pool, tiers, got):
        alone = api.price_batch(reqs, tier, max_dimension=MD)
        for a, b in zip(g.results, alone):
            assert a.pv == b.pv and a.legs.down_and_in_put == b.legs.down_and_in_put


def test_monte_carlo_lane_tickets_alone_equal_in_the_mix():
    # the reference's default lane (ChaCha8) through the server, beside Sobol tickets:
    # each ticket's stream parse runs inside its own plan
    names, crn = W.crn_scenarios(paths=3000, steps=126, years=0.5)
    mc = [r.clone() for r in crn[:5]]
    for r in mc:
        r.config.variance_reduction = W.VarianceReduction("none")
    flat = W.flat_book(paths=5000, steps=63, years=0.25)
    flat.config.variance_reduction = W.VarianceReduction("none")
    tickets = [(mc, "fp64"), (crn[:2], "fp64"), ([flat], "fp32"), (mc[:2], "fp32")]
    srv = serving.Server(max_inflight=3, max_batch_requests=64)
    ids = [srv.submit(reqs, tier) for reqs, tier in tickets]
    got = [srv.wait(t) for t in ids]
    srv.close()
    assert all(g.state == serving.DONE for g in got)
    for (reqs, tier), g in zip(tickets, got):
        alone = api.price_batch(reqs, tier, max_dimension=MD)
        for a, b in zip(g.results, alone):
            assert a.pv == b.pv and a.stats[0].output_count == b.stats[0].output_count
