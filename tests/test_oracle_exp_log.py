"""The FP64 exp and log the device's FP64 tier evaluates inside the path
(riga_flow_amd/csrc/exp_log.h), built on the host into the oracle, pinned by:

* correctly rounded values (Python's decimal module at 50 digits, whose exp and
  ln are correctly rounded): at most 0.52 ulp (exp in the normal range, log) on
  random inputs, 1 ulp for subnormal exp results;
* glibc's exp and log -- the reference's f64::exp / f64::ln -- on the inputs the
  path meets (Euler increments, spots, inverse-normal tails): identical for at
  least 99.5%, never more than 1 ulp apart;
* special values and identities (exp(0) = 1, log(1) = 0, log(2^k) = k ln 2,
  the overflow and underflow edges, NaN, infinities, subnormals).
"""
import math
import random
from decimal import Decimal, getcontext

import pytest

import oracle as O

getcontext().prec = 50


def _ulps(v, exact):
    return abs(float(Decimal(v) - exact)) / math.ulp(float(exact))


def _ulp_gap(a, b):
    return 0.0 if a == b else abs(a - b) / math.ulp(b)


def test_exp_against_correctly_rounded_values():
    rng = random.Random(11)
    worst = max(_ulps(O.shared_exp(x), Decimal(x).exp())
                for x in [rng.uniform(-3.0, 3.0) for _ in range(6000)] +
                [rng.uniform(-708.0, 709.7) for _ in range(6000)])
    assert worst <= 0.52, worst
    # subnormal results: one extra rounding
    sub = max(_ulps(O.shared_exp(x), Decimal(x).exp()) for x in [rng.uniform(-744.0, -708.5) for _ in range(2000)])
    assert sub <= 1.0, sub


def test_log_against_correctly_rounded_values():
    rng = random.Random(12)
    xs = [math.exp(rng.uniform(-8.0, 10.0)) for _ in range(5000)]
    xs += [2.0 ** rng.uniform(-1074.0, 1023.9) for _ in range(5000)]
    xs += [1.0 + rng.uniform(-0.3, 0.3) * 10.0 ** -rng.uniform(0.0, 15.0) for _ in range(3000)]
    worst = max(_ulps(O.shared_log(x), Decimal(x).ln()) for x in xs if x != 1.0)
    assert worst <= 0.52, worst


def test_agreement_with_glibc_on_the_path_inputs():
    rng = random.Random(13)
    cases = {
        "exp increments": ([rng.gauss(0.0, 0.05) for _ in range(100_000)], O.shared_exp, math.exp),
        "exp wide": ([rng.uniform(-40.0, 40.0) for _ in range(50_000)], O.shared_exp, math.exp),
        "log spots": ([100.0 * math.exp(rng.gauss(0.0, 0.4)) for _ in range(100_000)], O.shared_log, math.log),
        "log tails": ([rng.uniform(1e-12, 0.02425) for _ in range(50_000)], O.shared_log, math.log),
    }
    for name, (xs, ours, glibc) in cases.items():
        gaps = [_ulp_gap(ours(x), glibc(x)) for x in xs]
        same = sum(g == 0.0 for g in gaps) / len(gaps)
        assert same >= 0.995 and max(gaps) <= 1.0, (name, same, max(gaps))


def test_special_values_and_identities():
    inf, nan = float("inf"), float("nan")
    assert O.shared_exp(0.0) == 1.0 and O.shared_exp(-0.0) == 1.0
    assert O.shared_exp(inf) == inf and O.shared_exp(-inf) == 0.0 and math.isnan(O.shared_exp(nan))
    assert O.shared_exp(709.78) == math.exp(709.78) and O.shared_exp(709.79) == inf
    assert O.shared_exp(-745.0) == 5e-324 and O.shared_exp(-745.2) == 0.0
    assert O.shared_log(1.0) == 0.0 and O.shared_log(0.0) == -inf and O.shared_log(inf) == inf
    assert math.isnan(O.shared_log(-1.0)) and math.isnan(O.shared_log(nan))
    for k in (-1074, -1022, -60, -1, 1, 7, 300, 1023):
        assert _ulps(O.shared_log(2.0 ** k), Decimal(2.0 ** k).ln()) <= 0.5, k
    assert O.shared_log(5e-324) == math.log(5e-324)
    assert O.shared_log(1.7976931348623157e308) == math.log(1.7976931348623157e308)
    rng = random.Random(14)
    for _ in range(2000):
        x = rng.uniform(-30.0, 30.0)
        assert abs(O.shared_log(O.shared_exp(x)) - x) <= 4 * math.ulp(max(abs(x), 1e-300)) + 1e-15


def test_oracle_path_uses_glibc_unless_asked():
    """The reference's exp/log by default; the device's on request -- both price
    the same request within the path's last-digit sensitivity."""
    from riga_flow_amd import workloads as W
    req = W.flat_book(paths=256, steps=64)
    md = W.max_dimension_for(req)
    a = O.price(req, max_dim=md)
    b = O.price(req, max_dim=md, device_math=True)
    assert O.lib().ro_get_device_math() == 1
    c = O.price(req, max_dim=md, device_math=False)
    assert a.pv == c.pv
    assert a.ko_count == b.ko_count and abs(a.pv - b.pv) <= 1e-12 * abs(a.pv)
