"""Pins for the oracle's correlation factor and Dupire surface, and agreement of
the product's host-side FP64 preparation (C++ in libriga_hip.so) with them.

Independent checks: closed-form 2x2 Cholesky, L L^T = C, the reference's
unit-test assertions (correlation.rs:266-337, riga-market/src/lib.rs:400-438),
and the term-sheet reference local-vol points (fixture.rs:507-527).
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from riga_flow_amd import api
from riga_flow_amd.fixture import load_termsheet1_fixture

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_two_factor_closed_form():
    rho = 0.4041
    L = O.cholesky([[1.0, rho], [rho, 1.0]])
    assert abs(L[0, 0] - 1.0) < 1e-15 and abs(L[1, 0] - rho) < 1e-15
    assert abs(L[1, 1] - math.sqrt(1 - rho * rho)) < 1e-15
    y = O.chol_apply(L, [0.3, -1.2])
    assert y[0] == 0.3
    assert abs(y[1] - (rho * 0.3 + math.sqrt(1 - rho * rho) * -1.2)) < 1e-15


def test_factor_reconstructs_and_is_lower_triangular():
    C = np.array([[1.0, 0.5, 0.3], [0.5, 1.0, 0.2], [0.3, 0.2, 1.0]])
    L = O.cholesky(C)
    assert np.all(np.triu(L, 1) == 0.0)
    np.testing.assert_allclose(L @ L.T, C, atol=1e-12)


def test_non_psd_matrix_is_clamped():
    bad = np.array([[1.0, 0.9, -0.9], [0.9, 1.0, 0.9], [-0.9, 0.9, 1.0]])
    assert O.min_eigenvalue(bad) < 0.0
    assert abs(O.min_eigenvalue(bad) - np.linalg.eigvalsh(bad).min()) < 1e-12
    rep = O.psd_clamp(bad, 1e-8)
    assert O.min_eigenvalue(rep) >= -1e-9
    assert np.allclose(np.diag(rep), 1.0, atol=1e-12)
    L = O.cholesky(rep)
    np.testing.assert_allclose(L @ L.T, rep, atol=1e-10)


def test_malformed_matrices_rejected():
    for m in ([[0.5, 0.0], [0.0, 1.0]], [[1.0, 2.0], [2.0, 1.0]], [[1.0, 0.3], [0.2, 1.0]]):
        with pytest.raises(ValueError):
            O.cholesky(m)


def test_jacobi_eigenvalues_match_numpy_on_random_correlations():
    # correlation.rs:222 caps Jacobi at 100 rotations: converged for the <= 5
    # factor matrices a trade carries (larger ones are approximate, as in the reference)
    rng = np.random.default_rng(3)
    for n in (2, 3, 4, 5):
        B = rng.normal(size=(n, 3))
        C = B @ B.T + np.diag(rng.uniform(0.1, 0.5, n))
        s = np.sqrt(np.diag(C))
        C = C / np.outer(s, s)
        np.fill_diagonal(C, 1.0)
        assert abs(O.min_eigenvalue(C) - np.linalg.eigvalsh(C).min()) < 1e-10
        L = O.cholesky(C)
        np.testing.assert_allclose(L @ L.T, C, atol=1e-12)


def flat_surface(vol):
    mats = [0.1, 0.25, 0.5, 1.0]
    strikes = [60.0, 80.0, 100.0, 120.0, 140.0]
    return mats, strikes, [[vol] * len(strikes) for _ in mats]


def test_flat_smile_collapses_to_the_constant_vol():
    # riga-market/src/lib.rs:400-412
    lv = O.LocalVol.calibrate(*flat_surface(0.20), spot=100.0, rate=0.0)
    for t in (0.1, 0.2, 0.3, 0.5):
        for s in (70.0, 90.0, 100.0, 110.0, 130.0):
            assert abs(lv.sigma(t, s) - 0.20) < 1e-9


def test_unsorted_axes_rejected():
    mats, strikes, vols = flat_surface(0.2)
    strikes[0], strikes[1] = strikes[1], strikes[0]
    with pytest.raises(ValueError):
        O.LocalVol.calibrate(mats, strikes, vols, spot=100.0, rate=0.0)


def test_interp_linear_flat_extrapolation():
    xs, ys = [0.0, 1.0, 2.0], [1.0, 3.0, 5.0]
    assert O.interp_linear(xs, ys, -1.0) == 1.0
    assert O.interp_linear(xs, ys, 3.0) == 5.0
    assert O.interp_linear(xs, ys, 0.5) == 2.0


def test_termsheet_local_vol_reference_points():
    # fixture.rs:507-527: upstream calibration points at T = 0.40274, +-3e-3
    req = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json")).request
    lvs = O.local_vols_for(req)
    for factor, s, want in [(0, 180.0, 0.467727), (0, 268.0, 0.430912),
                            (1, 240.0, 0.346527), (1, 268.0, 0.352650)]:
        assert abs(lvs[factor].sigma(0.40274, s) - want) < 3e-3


def test_product_host_calibration_matches_oracle_bit_for_bit():
    req = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json")).request
    for f, g in enumerate(req.local_vol.grids):
        for gx in (96, 128):
            ora = O.LocalVol.calibrate(g.maturities_years, g.strikes, g.vols,
                                       req.market.spots[f], req.market.risk_free_rate, gx)
            prod = api.lv_calibrate(g.maturities_years, g.strikes, g.vols, req.market.spots[f],
                                    req.market.risk_free_rate, gx)
            assert np.array_equal(np.array(prod.t_axis), ora.t_axis)
            assert np.array_equal(np.array(prod.log_s), ora.log_s)
            assert np.array_equal(np.array(prod.v), ora.v)
            for t, s in [(0.0, 100.0), (0.2, 250.0), (0.40274, 268.0), (3.0, 500.0), (0.5, 1.0)]:
                assert api.lv_sigma(prod, t, s) == ora.sigma(t, s)
