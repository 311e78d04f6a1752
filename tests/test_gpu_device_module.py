"""The reference's device module on gfx950 (crates/riga-engine/src/gpu.rs:27-570):
worst_of, resident_worst_of, lv_terminal_spots and resident_lv, as its examples
(examples/gpu_bench.rs, examples/gpu_lv_bench.rs) call them.

Bars: worst_of bit-exact (including the reference's NaN rule); the FP32 local-vol
propagation against a float32 restatement of the kernel's own arithmetic (KO step
identical on at least 99.99% of paths, terminal spots within 1e-6 relative on at
least 99.99% and within 1e-4 on all) and the KO-gated put within 2e-3 of the FP64
CPU propagation of gpu.rs:37-75's semantics, as the reference's benchmark compares
them; resident runs equal the one-shot call bit for bit.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from riga_flow_amd import gpu
from riga_flow_amd.fixture import load_termsheet1_fixture
from riga_flow_amd.types import RigaError

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_worst_of_and_its_resident_form():
    rng = np.random.default_rng(5)
    a = rng.normal(1.0, 0.2, 100_003).astype(np.float32)
    b = rng.normal(1.0, 0.2, 100_003).astype(np.float32)
    a[7], b[9] = np.nan, np.nan
    want = np.where(a < b, a, b)
    got = gpu.worst_of(a, b)
    assert np.array_equal(got, want, equal_nan=True)
    assert np.isnan(got[9]) and got[7] == b[7]
    assert gpu.worst_of([], []).size == 0
    with pytest.raises(RigaError) as e:
        gpu.worst_of([1.0], [1.0, 2.0])
    assert e.value.code == "shape_mismatch"
    r = gpu.resident_worst_of(a, b)
    for _ in range(3):
        r.reduce()
    r.synchronize()
    assert np.array_equal(r.download(), want, equal_nan=True)
    r.close()


def _lv_inputs(paths):
    # examples/gpu_lv_bench.rs:30-80: the term sheet's calibrated surfaces as per-step
    # rows, its correlated ChaCha8 shocks
    req = load_termsheet1_fixture(os.path.join(GOLDEN, "termsheet1.json")).request
    lvs = O.local_vols_for(req)
    steps = req.trade.steps
    dt = req.trade.maturity_years / steps
    gx = len(lvs[0].log_s)
    rows = np.zeros((2, steps, gx), dtype=np.float32)
    for f in range(2):
        for s in range(steps):
            rows[f, s] = np.asarray(lvs[f].row(s * dt), dtype=np.float32)
    ls = np.stack([np.asarray(lvs[f].log_s, dtype=np.float32) for f in range(2)])
    shocks = O.correlated_block(req.config.seed, paths, steps, req.market.correlation)
    return req, lvs, rows, ls, shocks, steps, dt


def _fma32(a, b, c):
    """fmaf in numpy: the exact product and one rounding of the sum (via float64)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def _propagate_f32(shocks, rows, ls, s0, refs, bar, dt, rate, q):
    """The propagation kernel's float32 arithmetic (gpu.rs:37-75's semantics in the
    kernel's own form): knot position x * scale + offset clamped to [0, gx-1], the
    slope-pair lerp v + w (v' - v), x += (vol (vol (-dt/2) + sqrt(dt) z)) + (r - q) dt
    as two FMAs, KO when both log-spots reach ln(bar * reference)."""
    P, T, _ = shocks.shape
    gx = ls.shape[1]
    f32 = np.float32
    d = lambda v: float(f32(v))  # the value the C ABI receives as a float
    dt32, rate32 = d(dt), d(rate)
    scale, off, drift, lnko, x = [], [], [], [], []
    for f in range(2):
        lo, hi = float(ls[f, 0]), float(ls[f, gx - 1])
        sc = (gx - 1) / (hi - lo)
        scale.append(f32(sc))
        off.append(f32(-lo * sc))
        drift.append(f32((rate32 - d(q[f])) * dt32))
        lnko.append(f32(math.log(d(bar) * d(refs[f]))))
        x.append(np.full(P, f32(math.log(d(s0[f]))), dtype=f32))
    sqdt, nhdt = f32(math.sqrt(dt32)), f32(-0.5 * dt32)
    tmax = f32(gx - 1)
    ko = np.full(P, -1, dtype=np.int32)
    for s in range(T):
        for f in range(2):
            t = np.clip(_fma32(x[f], scale[f], off[f]), f32(0), tmax)
            j = np.minimum(t.astype(np.int32), gx - 2)
            w = t - j.astype(f32)
            v = rows[f, s, j]
            dv = rows[f, s, j + 1] - v
            vol = _fma32(w, dv, v)
            sz = sqdt * shocks[:, s, f]
            x[f] = x[f] + _fma32(vol, _fma32(vol, nhdt, sz), drift[f])
        hit = (ko < 0) & (x[0] >= lnko[0]) & (x[1] >= lnko[1])
        ko[hit] = s
    return np.stack([np.exp(x[0]), np.exp(x[1])], axis=1).reshape(-1), ko


def test_lv_terminal_spots_and_the_resident_sweep():
    paths = 20_000
    req, lvs, rows, ls, shocks, steps, dt = _lv_inputs(paths)
    m, t = req.market, req.trade
    s0 = np.asarray(m.spots, dtype=np.float32)
    refs = np.asarray(t.initial_spots, dtype=np.float32)
    q = np.asarray(m.dividend_yields, dtype=np.float32)
    bar = float(t.knock_out_barrier)
    term, ko = gpu.lv_terminal_spots(shocks, rows, ls, s0, refs, bar, paths, steps, dt, m.risk_free_rate, q)
    want, wko = _propagate_f32(shocks, rows, ls, s0, refs, bar, dt, m.risk_free_rate, q)
    # the restatement rounds like the kernel (FMAs included): only the final expf and
    # the rare double rounding of the float64-emulated FMA may differ
    rel = np.abs(term - want) / np.abs(want)
    assert np.median(rel) <= 2.5e-7 and (rel <= 1e-6).mean() >= 0.9999 and rel.max() <= 1e-4
    assert (ko == wko).mean() >= 0.9999
    # the KO-gated put against the FP64 CPU propagation (gpu_lv_bench.rs:84-130)
    init = t.initial_spots
    df = math.exp(-m.risk_free_rate * t.maturity_years)
    cpu = 0.0
    sh = shocks.astype(np.float64)
    for p in range(2000):
        sp = list(m.spots)
        hit = False
        for s in range(steps):
            for f in range(2):
                vol = lvs[f].sigma(s * dt, sp[f])
                sp[f] *= math.exp((m.risk_free_rate - m.dividend_yields[f] - 0.5 * vol * vol) * dt
                                  + vol * math.sqrt(dt) * sh[p, s, f])
            hit = hit or (sp[0] / init[0] >= bar and sp[1] / init[1] >= bar)
        worst = min(sp[0] / init[0], sp[1] / init[1])
        if worst <= t.knock_in_barrier and not hit:
            cpu += max(t.strike - worst, 0.0)
    cpu = cpu / 2000 * df
    w = np.minimum(term[0:4000:2] / init[0], term[1:4000:2] / init[1])
    gated = np.where((w <= t.knock_in_barrier) & (ko[:2000] < 0), np.maximum(t.strike - w, 0.0), 0.0).mean() * df
    assert abs(gated - cpu) <= 2e-3 * max(cpu, 1e-3)
    # resident: upload once, run a spot-bump sweep
    res = gpu.resident_lv(shocks, rows, ls, paths, steps, dt)
    for bump in (1.0, 1.01, 0.99):
        sb = s0 * np.float32(bump)
        res.run(sb, refs, bar, m.risk_free_rate, q)
        got, gko = res.download()
        one, oko = gpu.lv_terminal_spots(shocks, rows, ls, sb, refs, bar, paths, steps, dt, m.risk_free_rate, q)
        assert np.array_equal(got, one) and np.array_equal(gko, oko)
    res.synchronize()
    res.close()
