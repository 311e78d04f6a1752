"""The reference's published experiment figures, recomputed on the device from
the same inputs (the ChaCha8 Monte Carlo, Halton and Sobol lanes reproduce the
reference's streams):

* CRN variance-reduction factor 140.18x (docs/plan/25-implementation-evidence.md:88;
  riga-risk/src/lib.rs:835-845 on the 512-path term sheet, 40 replicates);
* paired-replicate delta standard error 2.6e-5 at 16 and 1.1e-5 at 64 replicates
  (25-implementation-evidence.md:110; riga-risk/src/lib.rs:919-939);
* put-estimator variance ratio MC / Halton 1.05x and MC / Sobol 1.04x at 4,096 paths
  x 12 seeds (25-implementation-evidence.md:107; riga-engine/src/lib.rs:653-688).

Each is checked to the digits the reference publishes.
"""
import os

import pytest

from riga_flow_amd import api, risk, workloads as W
from riga_flow_amd.fixture import load_termsheet1_fixture

pytestmark = pytest.mark.gpu
FX = os.path.join(os.path.dirname(__file__), "golden", "termsheet1.json")


def small():
    req = load_termsheet1_fixture(FX).request
    req.config.paths = 512
    req.config.path_block_size = 128
    req.config.seed = 81_281
    return req


def test_crn_variance_reduction_factor_is_the_published_140_18():
    f = risk.crn_variance_reduction_factor(small(), replicates=40)
    assert round(f, 2) == 140.18


def test_paired_delta_standard_errors_are_the_published_ones():
    assert f"{risk.crn_delta_convergence(small(), replicates=16).standard_error:.1e}" == "2.6e-05"
    assert f"{risk.crn_delta_convergence(small(), replicates=64).standard_error:.1e}" == "1.1e-05"


def test_qmc_variance_ratios_are_the_published_ones():
    base = load_termsheet1_fixture(FX).request
    dim = base.trade.steps * 2

    def variance(kind):
        reqs = []
        for seed in range(1, 13):
            r = base.clone()
            r.config.paths = 4096
            r.config.path_block_size = 4096
            r.config.seed = seed
            r.config.variance_reduction = W.VarianceReduction(kind, 0 if kind == "none" else dim)
            reqs.append(r)
        s = [x.legs.down_and_in_put for x in api.price_batch(reqs, "fp64")]
        m = sum(s) / len(s)
        return sum((v - m) ** 2 for v in s) / (len(s) - 1)

    mc, halton, sobol = variance("none"), variance("halton"), variance("sobol")
    assert round(mc / halton, 2) == 1.05 and round(mc / sobol, 2) == 1.04
