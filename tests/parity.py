"""Shared comparison helpers for the GPU-vs-oracle parity tests."""
import math

import numpy as np

import oracle as O

# FP64 tier: legs within 1e-10 relative (sums are double-double on the device,
# sequential in the oracle: differences are ~1e-15).  Integer populations exact.
FP64_REL = 1e-10
# FP32 tier: path arithmetic in float, stated bound on price error vs the FP64 oracle.
FP32_REL = 1e-4


def rel(a, b):
    if a == b:
        return 0.0
    return abs(a - b) / max(abs(b), 1e-300)


def assert_fp64_parity(res, ora, tol=FP64_REL):
    """GPU ``PriceResult`` vs ``OracleResult``: counts exact, legs/moments within tol."""
    assert res.legs.funding == res.legs.funding
    assert res.stats[0].output_count == ora.ko_count
    assert res.stats[1].output_count == ora.ki_count
    assert res.stats[2].output_count == ora.coupon_path_count
    assert res.ledger[1].population_after == ora.put_eligible_count
    pairs = [
        ("pv", res.pv, ora.pv), ("funding", res.legs.funding, ora.leg_funding),
        ("put", res.legs.down_and_in_put, ora.leg_put), ("coupon", res.legs.coupon, ora.leg_coupon),
        ("memory", res.legs.memory_coupon, ora.leg_memory_coupon),
        ("ki_prob", res.ki_probability, ora.ki_probability),
        ("ko_mean", res.stats[0].mean_payoff, ora.ko_mean),
        ("ki_mean", res.stats[1].mean_payoff, ora.ki_mean),
        ("cpn_mean", res.stats[2].mean_payoff, ora.cpn_mean),
    ]
    for name, a, b in pairs:
        if b == 0.0:
            assert abs(a) <= 1e-15, (name, a, b)
        else:
            assert rel(a, b) <= tol, (name, a, b, rel(a, b))
    for name, a, b in [("mc_error", res.mc_error, ora.mc_error),
                       ("ko_var", res.stats[0].variance_payoff, ora.ko_var),
                       ("ki_var", res.stats[1].variance_payoff, ora.ki_var),
                       ("cpn_var", res.stats[2].variance_payoff, ora.cpn_var)]:
        assert abs(a - b) <= 1e-9 * max(abs(b), 1e-6), (name, a, b)


def assert_json_close(actual, expected, path="result"):
    """The reference's golden comparison (riga-engine/src/lib.rs:429-456): same keys
    and lengths, numbers within 1e-12 absolute, everything else equal."""
    if isinstance(expected, dict):
        assert set(actual) == set(expected), path
        for k in expected:
            assert_json_close(actual[k], expected[k], f"{path}.{k}")
    elif isinstance(expected, list):
        assert len(actual) == len(expected), path
        for i, (a, e) in enumerate(zip(actual, expected)):
            assert_json_close(a, e, f"{path}[{i}]")
    elif isinstance(expected, (int, float)) and not isinstance(expected, bool):
        assert abs(actual - expected) <= 1e-12, (path, actual, expected)
    else:
        assert actual == expected, (path, actual, expected)
