"""GPU parity at the edges of the input space: every factor count, ragged path
counts, a single step, trace off, the FP32 tier on the Halton lane and at the
reference's 512-dimension cap, and rejected requests.

Bars as in test_gpu_parity: FP64 within 1e-10 relative of the oracle with counts
exact; FP32 within 1e-4 relative.
"""
import math

import numpy as np
import pytest

import oracle as O
from parity import FP32_REL, assert_fp64_parity, rel
from riga_flow_amd import api, workloads as W
from riga_flow_amd.types import RigaError

pytestmark = pytest.mark.gpu


def _book(F, paths, steps, years, local_vol):
    if local_vol:
        req = W.lv_book(paths=paths, steps=steps, years=years, F=F)
    else:
        req = W.flat_book(paths=paths, steps=steps, years=years)
        if F != 3:
            m = W.flat_market(F, vols=(0.20, 0.25, 0.30, 0.22, 0.28)[:F])
            req.market = m
            req.trade = W.phoenix_trade("flat-f", F, years, steps)
            req.config = W.sobol_config(paths, steps, F)
    return req


@pytest.mark.parametrize("F", [1, 2, 4, 5])
@pytest.mark.parametrize("local_vol", [False, True])
def test_every_factor_count(F, local_vol):
    req = _book(F, paths=768, steps=126, years=0.5, local_vol=local_vol)
    md = W.max_dimension_for(req)
    ora = O.price(req, max_dim=md)
    assert_fp64_parity(api.price_put_leg(req, "fp64", max_dimension=md), ora)
    f32 = api.price_put_leg(req, "fp32", max_dimension=md)
    assert rel(f32.pv, ora.pv) <= FP32_REL


@pytest.mark.parametrize("paths", [1, 63, 65, 1000, 4097])
def test_ragged_path_counts(paths):
    # the last workgroup and the last wave are partly empty; counts stay exact
    req = _flat(paths, steps=63, years=0.25)
    md = W.max_dimension_for(req)
    ora = O.price(req, max_dim=md)
    res = api.price_put_leg(req, "fp64", max_dimension=md)
    assert_fp64_parity(res, ora)
    assert res.stats[0].input_count == paths


def _flat(paths, steps, years):
    req = W.flat_book(paths=paths, steps=steps, years=years)
    req.trade.call_start_step = min(req.trade.call_start_step, steps - 1)
    req.trade.coupon_periods = max(1, round(12 * years))
    return req


def test_single_step():
    # one step: the bridge is the root alone, no LDS levels
    req = _flat(2048, steps=1, years=1.0 / 252)
    md = W.max_dimension_for(req)
    ora = O.price(req, max_dim=md)
    assert_fp64_parity(api.price_put_leg(req, "fp64", max_dimension=md), ora)
    assert rel(api.price_put_leg(req, "fp32", max_dimension=md).pv, ora.pv) <= FP32_REL


def test_trace_off_has_no_events_and_same_price():
    req = _flat(1024, steps=126, years=0.5)
    md = W.max_dimension_for(req)
    on = api.price_put_leg(req, "fp64", max_dimension=md)
    req.config.trace_level = "off"
    off = api.price_put_leg(req, "fp64", max_dimension=md)
    assert off.trace_events == []
    assert on.pv == off.pv


def test_fp32_tier_on_the_halton_lane():
    req = _flat(2048, steps=126, years=0.5)
    req.config.variance_reduction = W.VarianceReduction("halton", min(512, 126 * 3))
    md = 512
    ora = O.price(req, max_dim=md)
    assert_fp64_parity(api.price_put_leg(req, "fp64", max_dimension=md), ora)
    assert rel(api.price_put_leg(req, "fp32", max_dimension=md).pv, ora.pv) <= FP32_REL


def test_fp32_tier_at_the_reference_dimension_cap():
    # 252 steps x 3 factors = 756 dimensions wrap at 512 as the reference does
    req = _flat(2048, steps=252, years=1.0)
    ora = O.price(req, max_dim=512)
    assert_fp64_parity(api.price_put_leg(req, "fp64", max_dimension=512), ora)
    assert rel(api.price_put_leg(req, "fp32", max_dimension=512).pv, ora.pv) <= FP32_REL


def test_rejected_requests_name_the_reference_codes():
    bad = _flat(1024, steps=63, years=0.25)
    bad.market.spots[0] = -1.0
    with pytest.raises(RigaError) as e:
        api.price_put_leg(bad, "fp64")
    assert e.value.code == "invalid_market"


def test_step_counts_beyond_the_packed_event_width_are_refused():
    # KO steps travel in 20 bits on the device: 2^20 steps and more are refused
    # before anything is laid out, in both tiers, rather than mispriced
    long = _flat(64, steps=1 << 20, years=4000.0)
    for tier in ("fp64", "fp32"):
        with pytest.raises(RigaError) as e:
            api.price_put_leg(long, tier)
        assert e.value.code == "invalid_schedule" and "1,048,575" in e.value.message


def test_concurrent_callers_get_their_own_results():
    # the C ABI is called from several host threads at once (the reference's MCP
    # handlers run on a multi-threaded runtime): each call's result equals the
    # same call made alone
    import threading
    reqs = []
    for i in range(6):
        r = _flat(2048 + 512 * i, steps=63, years=0.25)
        if i % 2:
            r.config.variance_reduction = W.VarianceReduction("none")
        reqs.append(r)
    alone = [api.price_put_leg(r, "fp64") for r in reqs]
    got = [None] * len(reqs)
    errors = []

    def worker(i):
        try:
            for _ in range(3):
                got[i] = api.price_put_leg(reqs[i], "fp64")
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(reqs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors and all(not t.is_alive() for t in threads)
    for a, b in zip(got, alone):
        assert a.pv == b.pv and a.stats[0].output_count == b.stats[0].output_count
