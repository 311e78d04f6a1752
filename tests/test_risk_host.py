"""Host-side logic of the CRN risk module (riga-risk/src/lib.rs:260-279, 338-385):
configuration checks and the vega bump of the implied grids.  No device."""
import pytest

from riga_flow_amd import risk, workloads as W
from riga_flow_amd.types import LocalVolSpec, RigaError


def implied_request():
    req = W.lv_book(paths=256, steps=21, years=1.0 / 12.0, F=2)
    sig = W.svi_sigmas(2)
    req.local_vol = LocalVolSpec(grids=[W.svi_implied(sig[f]) for f in range(2)], grid_x=96)
    return req


def test_vega_bump_shifts_every_implied_node_of_one_asset():
    req = implied_request()
    out = risk.bump_volatility_request(req, 1, 0.01)
    assert out.market.vols[1] == pytest.approx(req.market.vols[1] + 0.01)
    assert out.market.vols[0] == req.market.vols[0]
    g0, g1 = req.local_vol.grids[1], out.local_vol.grids[1]
    for r0, r1 in zip(g0.vols, g1.vols):
        assert all(b == a + 0.01 for a, b in zip(r0, r1))
    assert out.local_vol.grids[0].vols == req.local_vol.grids[0].vols
    assert req.local_vol.grids[1].vols[0][0] == g0.vols[0][0]  # input untouched


@pytest.mark.parametrize("field,value,code", [
    ("spot_relative_bump", 0.0, "invalid_spot_bump"),
    ("spot_relative_bump", 1.0, "invalid_spot_bump"),
    ("volatility_absolute_bump", float("nan"), "invalid_volatility_bump"),
    ("rate_absolute_bump", -1.0, "invalid_rate_bump"),
    ("correlation_absolute_bump", float("inf"), "invalid_correlation_bump"),
    ("random_number_policy", "no_common_random_numbers", "no_crn_rejected"),
])
def test_configuration_checks_name_the_reference_codes(field, value, code):
    cfg = risk.CentralDifferenceConfig(**{field: value})
    with pytest.raises(RigaError) as e:
        risk._validate(implied_request(), cfg)
    assert e.value.code == code


def test_calibrated_only_surfaces_cannot_take_a_vega_bump():
    req = W.lv_book(paths=256, steps=21, years=1.0 / 12.0, F=2)
    with pytest.raises(RigaError) as e:
        risk.bump_volatility_request(req, 0, 0.01)
    assert e.value.code == "risk_surface_shape"
