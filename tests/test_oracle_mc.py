"""The oracle's Monte Carlo lane (VarianceReduction::None, note.rs:58,124-126):
ChaCha8Rng::seed_from_u64 + rand_distr StandardNormal, restated from the
published algorithms of rand_core 0.9 / rand_chacha 0.9.0 / rand_distr 0.5.1.

Pins: the reference's golden PriceResult (fixtures/termsheet1.golden.json, the
full result within the reference test's 1e-12, lib.rs:489-498) exercises the
whole chain -- seed expansion, ChaCha8, word order, ziggurat tables and both slow
paths -- over 6.3M draws; the ChaCha core against RFC 7539's block known answer
(20 rounds: the same code with more rounds); the sampler against the normal law.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

import oracle as O
from parity import assert_json_close
from riga_flow_amd import api
from riga_flow_amd import _native as N
from riga_flow_amd.fixture import load_termsheet1_fixture
from riga_flow_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_chacha_block_rfc7539_known_answer():
    # RFC 7539 section 2.3.2: key 00..1f, nonce 00:00:00:09:00:00:00:4a:00:00:00:00,
    # block count 1, 20 rounds
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    state = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + key + [1, 0x09000000, 0x4A000000, 0]
    want = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204,
            0x4E6CD4C3, 0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE,
            0xE883D0CB, 0x4E3C50A2]
    assert O.chacha_block(state, 10) == want


def test_stream_words_are_consecutive_blocks_low_word_first():
    seed = 1729
    key = O.seed_from_u64(seed)
    st = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + key + [0, 0, 0, 0]
    for block in (0, 1, 7, 123456):
        st[12], st[13] = block & 0xFFFFFFFF, block >> 32
        out = O.chacha_block(st, 4)
        for j in range(8):
            assert O.rng_u64_at(seed, 8 * block + j) == out[2 * j] | (out[2 * j + 1] << 32)
    assert O.seed_from_u64(1) != O.seed_from_u64(2)


def test_ziggurat_tables_shape():
    x, f = O.zig_tables()
    assert x[1] == 3.6541528853610088 and x[256] == 0.0 and f[256] == 1.0
    assert np.all(np.diff(x) < 0)  # layers shrink towards the mode
    # every layer above the base has area V = x_i (f(x_{i+1}) - f(x_i)), to the
    # 18-decimal rounding the tables are stored with
    v = x[1:255] * (f[2:256] - f[1:255])
    np.testing.assert_allclose(v, 0.00492867323399, rtol=2e-9)
    # the base strip: x_0 f(R) = V
    assert abs(x[0] * math.exp(-x[1] ** 2 / 2) - 0.00492867323399) < 1e-15


def test_standard_normal_law():
    z, end = O.std_normals(2024, 400_000)
    assert abs(z.mean()) < 5 / math.sqrt(len(z))
    assert abs(z.var() - 1.0) < 5 * math.sqrt(2 / len(z))
    assert stats.kstest(z, "norm").pvalue > 1e-4
    # the tail beyond R (layer 0's exponential sampler) has its normal mass
    tail = np.mean(np.abs(z) > 3.6541528853610088)
    want = 2 * stats.norm.sf(3.6541528853610088)
    assert abs(tail - want) < 5 * math.sqrt(want / len(z))
    assert end > len(z)  # some draws took more than one word


def test_draws_restart_from_their_first_word():
    # what the device relies on: a draw depends only on the words from where it starts
    seed = 99
    starts = O.draw_starts(seed, 50_000)
    z, _ = O.std_normals(seed, 50_000)
    w = np.diff(starts.astype(np.int64))
    assert w.min() >= 1 and (w > 1).mean() < 0.03
    for k in (0, 1, 17, 4999, 31337, 49_999):
        zk, _ = O.std_normals(seed, 3, start=int(starts[k]))
        assert zk[0] == z[k]


def _golden_doc(req, o):
    r = N.Result()
    pairs = [("pv", "pv"), ("mc_error", "mc_error"), ("ki_probability", "ki_probability"),
             ("leg_coupon", "leg_coupon"), ("leg_memory_coupon", "leg_memory_coupon"),
             ("leg_put", "leg_put"), ("leg_funding", "leg_funding"), ("leg_total", "leg_total"),
             ("paths", "paths"), ("ko_count", "ko_count"), ("ki_count", "ki_count"),
             ("put_eligible_count", "put_eligible_count"), ("coupon_path_count", "coupon_path_count"),
             ("ko_mean", "ko_mean"), ("ko_variance", "ko_var"), ("ki_mean", "ki_mean"),
             ("ki_variance", "ki_var"), ("coupon_mean", "cpn_mean"), ("coupon_variance", "cpn_var")]
    for a, b in pairs:
        setattr(r, a, getattr(o, b))
    r.n_trace = len(o.trace)
    for i, (p, s) in enumerate(o.trace):
        r.trace_path[i], r.trace_step[i] = p, s
    return api.result_to_json(api._assemble(req, r))


def test_termsheet1_golden_full_result():
    lf = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json"))
    req = lf.request
    assert req.config.variance_reduction.kind == "none" and req.config.seed == 1729
    o = O.price(req)
    golden = json.load(open(os.path.join(GOLDEN, "termsheet1.golden.json")))
    assert_json_close(_golden_doc(req, o), golden)
    assert o.ko_count == 15934 and o.put_eligible_count == 3010


def test_termsheet1_golden_is_sensitive_to_the_stream():
    # the golden is sensitive to the stream: the same request on a neighbouring
    # seed (every draw different) misses it
    lf = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json"))
    req = lf.request
    req.config.seed = 1730
    req.config.paths = 3000
    o = O.price(req)
    golden = json.load(open(os.path.join(GOLDEN, "termsheet1.golden.json")))
    assert abs(o.pv - golden["pv"]) > 1e-9


def test_path_range_skips_the_draws_of_earlier_paths():
    req = W.flat_book(paths=300, steps=40, years=40 / 252)
    req.config.variance_reduction = W.VarianceReduction("none")
    full, frec = O.price(req, records=True)
    part, prec = O.price(req, path_range=(123, 200), records=True)
    assert np.array_equal(prec["ko_step"], frec["ko_step"][123:200])
    assert np.array_equal(prec["pv"], frec["pv"][123:200])
    # shocks of one path: L z on the path's own draws
    F, T = 3, 40
    sh = O.path_shocks(req, 5)
    starts = O.draw_starts(req.config.seed, 6 * F * T)
    z, _ = O.std_normals(req.config.seed, F * T, start=int(starts[5 * F * T]))
    L = O.cholesky(np.asarray(req.market.correlation))
    for t in (0, 17, 39):
        np.testing.assert_array_equal(sh[:, t], O.chol_apply(L, z[t * F:(t + 1) * F]))
