"""The PriceRequest JSON codec and the price_trade front-end's decoding errors
(riga-mcp/src/main.rs:1821-1831, 2713-2719; serde rules of riga-core/src/lib.rs)."""
import copy
import json
import os

import pytest

from riga_flow_amd import api, workloads as W
from riga_flow_amd.fixture import load_termsheet1_fixture
from riga_flow_amd.types import RigaError, request_from_json, request_to_json

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def termsheet():
    return load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json")).request


@pytest.mark.parametrize("make", [termsheet, lambda: W.flat_book(paths=64, steps=20, years=20 / 252)])
def test_round_trip(make):
    req = make()
    doc = json.loads(json.dumps(request_to_json(req)))
    back = request_from_json(doc)
    assert request_to_json(back) == request_to_json(req)
    assert back.trade == req.trade and back.market == req.market and back.config == req.config


def test_malformed_request_has_uniform_error_fields():
    # main.rs:2713-2719
    with pytest.raises(RigaError) as e:
        api.price_trade({})
    assert e.value.code == "invalid_price_request" and isinstance(e.value.message, str)
    with pytest.raises(RigaError) as e:
        api.price_trade("{not json")
    assert e.value.code == "invalid_price_request"


@pytest.mark.parametrize("path,value", [
    (("trade", "steps"), -3), (("trade", "strike"), "high"), (("config", "trace_level"), "loud"),
    (("config", "variance_reduction"), {"kind": "antithetic"}), (("market", "spots"), [1.0, True]),
    (("config", "seed"), 1.5),
])
def test_field_errors(path, value):
    doc = request_to_json(termsheet())
    doc[path[0]][path[1]] = value
    with pytest.raises(RigaError) as e:
        request_from_json(doc)
    assert e.value.code == "invalid_price_request"


def test_optional_and_default_fields():
    doc = request_to_json(termsheet())
    doc["trade"]["knock_out_barrier"] = None
    del doc["trade"]["coupon_rate"]
    del doc["trade"]["fixed_first_coupon"]
    del doc["config"]["variance_reduction"]
    req = request_from_json(doc)
    assert req.trade.knock_out_barrier is None and req.trade.coupon_rate is None
    assert req.trade.fixed_first_coupon is False and req.config.variance_reduction.kind == "none"
    missing = copy.deepcopy(doc)
    del missing["as_of"]
    with pytest.raises(RigaError) as e:
        request_from_json(missing)
    assert e.value.code == "invalid_price_request"
