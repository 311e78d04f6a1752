"""riga_price_put_leg_noalloc (the allocation-free pricing call) against
riga_price_put_leg, both through the C ABI: every result field identical, on two
calls in a row -- the second with a smaller input -- in both tiers, on the Sobol
and the Monte Carlo lanes; requests beyond the scratch array are refused."""
import ctypes as C
import os

import pytest

from riga_flow_amd import _native as N
from riga_flow_amd import api, workloads as W
from riga_flow_amd.fixture import load_termsheet1_fixture
from riga_flow_amd.types import RigaError

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = [f for f, _ in N.Result._fields_ if f != "device_ms"]


def _call(fn, req, md, tier):
    p = api._Packed(req, md)
    res = N.Result()
    err = N.Error()
    N.check(fn(N.Device.get().handle, C.byref(p.r), api.PRECISIONS[tier], C.byref(res), C.byref(err)), err)
    return res


def _assert_same(a, b):
    for f in FIELDS:
        va, vb = getattr(a, f), getattr(b, f)
        if hasattr(va, "__len__"):
            assert list(va) == list(vb), f
        else:
            assert va == vb, (f, va, vb)


@pytest.mark.parametrize("tier", ["fp64", "fp32"])
def test_noalloc_equals_price_put_leg_on_two_calls_in_a_row(tier):
    api.gpu_backend()
    L = N.lib()
    first = W.lv_book(paths=65_536 + 333, steps=756)  # 3 assets, local vol, ragged path count
    second = W.flat_book(paths=4096 + 17, steps=252)  # smaller: fewer paths, steps and tables
    for req in (first, second):
        md = W.max_dimension_for(req)
        a = _call(L.riga_price_put_leg_noalloc, req, md, tier)
        b = _call(L.riga_price_put_leg, req, md, tier)
        _assert_same(a, b)
        assert a.paths == req.config.paths and a.ko_count > 0


def test_noalloc_monte_carlo_lane_equals_price_put_leg():
    # the reference's golden term sheet (ChaCha8 lane): its stream parse in the scratch array
    L = N.lib()
    req = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json")).request
    small = req.clone()
    small.config.paths = 5000
    for q in (req, small):
        a = _call(L.riga_price_put_leg_noalloc, q, 512, "fp64")
        b = _call(L.riga_price_put_leg, q, 512, "fp64")
        _assert_same(a, b)


def test_noalloc_refuses_requests_beyond_its_scratch():
    req = W.flat_book(paths=16_000_000, steps=2)
    with pytest.raises(RigaError) as e:
        _call(N.lib().riga_price_put_leg_noalloc, req, 512, "fp64")
    assert e.value.code == "scratch_exceeded"
