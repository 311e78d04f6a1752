"""Writes tests/golden/oracle_vectors.json from the CPU oracle.

These are regression fixtures of this build's oracle (Sobol integers for the
reference test's range, per-path KO steps and KI flags of a 256-path flat book):
the reference ships no Sobol-lane vectors, so they pin later changes, while
tests/test_oracle_*.py pin the oracle itself.  Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from riga_flow_amd import workloads as W  # noqa: E402

s = O.Sobol(8, 42)
req = W.flat_book(paths=256, steps=252)
res, recs = O.price(req, max_dim=W.max_dimension_for(req), records=True)
doc = {
    "sobol_seed42_scrambled": [[s.scrambled(n, d) for d in range(8)] for n in range(1, 65)],
    "flat256_ko_step": recs["ko_step"].tolist(),
    "flat256_hit_ki": recs["hit_ki"].tolist(),
    "flat256_pv": res.pv,
}
with open(os.path.join(HERE, "oracle_vectors.json"), "w") as fh:
    json.dump(doc, fh)
print("wrote", os.path.join(HERE, "oracle_vectors.json"))
