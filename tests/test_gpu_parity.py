"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (stated per test): Sobol/Halton integers and KO/KI event indices are
bit-exact; FP64 prices within 1e-10 relative; FP32-tier prices within 1e-4
relative of the FP64 oracle.
"""
import math

import numpy as np
import pytest

import oracle as O
from parity import FP32_REL, assert_fp64_parity, rel
from riga_flow_amd import api, workloads as W

pytestmark = pytest.mark.gpu


def test_device_is_gfx950_and_named():
    api.gpu_backend()
    assert api.arch().startswith("gfx950")
    assert api.device_name()


@pytest.mark.parametrize("seed,ladder_dims,first_point,n_points,first_dim,n_dims", [
    (42, 8, 1, 64, 0, 8),            # the reference test's point/dim range (qmc.rs:479-500)
    (1, 756, 1, 1024, 0, 24),
    (1, 512, 60000, 256, 500, 40),   # reference cap: dims >= 512 wrap their direction row
    (7, 2268, 262100, 44, 2240, 28),
    (3, 756, (1 << 21) + 5, 300, 0, 12),           # 22-bit indices: the 24-bit gray-code sweep
    (3, 2268, (1 << 24) + 77, 300, 2250, 18),      # 25-bit indices: the 32-bit sweep
    (5, 756, (1 << 31) - 150, 300, 100, 10),       # across bit 31
    (9, 756, (1 << 32) - 301, 300, 0, 6),          # the last 32-bit points
])
def test_sobol_integers_bit_exact(seed, ladder_dims, first_point, n_points, first_dim, n_dims):
    u, s, z = api.debug_sobol(seed, ladder_dims, first_point, n_points, first_dim, n_dims)
    sob = O.Sobol(ladder_dims, seed, max_dim=max(512, ladder_dims))
    for p in range(n_points):
        n = first_point + p
        for d in range(n_dims):
            dim = first_dim + d
            assert int(u[p, d]) == sob.unscrambled(n, dim)
            assert int(s[p, d]) == sob.scrambled(n, dim)
            # the inverse normal in FP64 (tails: exp_log.h's log, IEEE sqrt and divide) bit for bit
            assert z[p, d] == sob.normal(n - 1, dim), (n, dim, z[p, d], sob.normal(n - 1, dim))


def test_fp64_exp_and_log_on_the_device_equal_their_host_build_bit_for_bit():
    """exp_log.h compiled for gfx950 and by gcc into the oracle: the same bits for
    Euler increments, spots, inverse-normal tail arguments, the whole range,
    subnormals and special values."""
    rng = np.random.default_rng(31)
    x = np.concatenate([rng.normal(0.0, 0.05, 100_000), rng.uniform(-746.0, 710.0, 30_000),
                        100.0 * np.exp(rng.normal(0.0, 0.4, 50_000)), rng.uniform(1e-12, 0.025, 30_000),
                        2.0 ** rng.uniform(-1074.0, 1023.9, 30_000),
                        [0.0, -0.0, 1.0, -1.0, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                         709.78, 709.79, -745.0, -745.2, np.inf, -np.inf, np.nan]])
    e, lg = api.debug_exp_log(x)
    we = np.array([O.shared_exp(float(v)) for v in x])
    wl = np.array([O.shared_log(float(v)) for v in x])
    assert np.array_equal(e.view(np.uint64)[~np.isnan(we)], we.view(np.uint64)[~np.isnan(we)])
    assert np.array_equal(np.isnan(e), np.isnan(we))
    assert np.array_equal(lg.view(np.uint64)[~np.isnan(wl)], wl.view(np.uint64)[~np.isnan(wl)])
    assert np.array_equal(np.isnan(lg), np.isnan(wl))


@pytest.mark.parametrize("factors,owen_table", [(3, True), (5, True), (3, False)])
def test_fp32_tier_ladder_keeps_24_owen_levels_and_full_tails(factors, owen_table):
    """The FP32 tier's integers, as its path kernels compute them (coordinate table, 24
    hashed levels, the wave's tail passes), against the rule applied to the full
    32-level scramble (whose integers the oracle pins): the top 24 bits with low byte
    0x80, and all 32 bits where the top 10 scrambled bits are all 0 or all 1.  The
    FP32 normals are the inverse normal of that uniform to FP32 accuracy."""
    from scipy.special import ndtri
    seed, ladder, nodes, n_pts = 81_281, 2268, 6, 65_536
    xs, z = api.debug_sobol_f32(seed, ladder, 1, n_pts, nodes, factors, owen_table)
    _, full, _ = api.debug_sobol(seed, ladder, 1, n_pts, 0, nodes * factors)
    p24 = full >> np.uint32(8)
    tail = (p24 < (1 << 14)) | (p24 >= (1 << 24) - (1 << 14))
    want = np.where(tail, full, (p24 << np.uint32(8)) | np.uint32(0x80)).astype(np.uint32)
    assert np.array_equal(xs, want)
    # the tail passes ran, including lanes with two tail coordinates in one node
    per_node = tail.reshape(n_pts, nodes, factors).sum(axis=2)
    assert tail.sum() > 1000 and (per_node >= 2).any()
    # the full integers of a sample of cells, tail cells among them, from the oracle
    sob = O.Sobol(ladder, seed, max_dim=ladder)
    rng = np.random.default_rng(7)
    cells = list(zip(rng.integers(0, n_pts, 200), rng.integers(0, nodes * factors, 200)))
    cells += list(zip(*np.nonzero(tail)))[:100]
    for p, d in cells:
        assert int(full[p, d]) == sob.scrambled(int(p) + 1, int(d))
    zr = ndtri((want.astype(np.float64) + 0.5) / 2.0 ** 32)  # Acklam is within 1.2e-9 of it
    assert np.max(np.abs(z - zr) / np.maximum(1.0, np.abs(zr))) < 5e-6


def test_fp32_ladder_normals_do_not_depend_on_the_kernel_variant():
    """The same point and dimension give the same FP32 normal in the one-, three-
    and five-factor ladders (tail coordinates included), so a trade's paths do not
    depend on the launch -- factor count, scenario bucket, shared table or not --
    it is priced in."""
    seed, ladder, n_pts = 1, 720, 8192
    x1, z1 = api.debug_sobol_f32(seed, ladder, 1, n_pts, 720, 1, True)
    x3, z3 = api.debug_sobol_f32(seed, ladder, 1, n_pts, 240, 3, True)
    x5, z5 = api.debug_sobol_f32(seed, ladder, 1, n_pts, 144, 5, True)
    assert np.array_equal(x1, x3) and np.array_equal(x1, x5)
    assert np.array_equal(z1.view(np.uint32), z3.view(np.uint32))
    assert np.array_equal(z1.view(np.uint32), z5.view(np.uint32))
    assert (np.abs(z1) > 1.97).sum() > 10_000  # the inverse-normal tails are exercised


def test_fp32_full_ladder_option_hashes_all_32_levels():
    """The fp32_full_ladder option gives the FP32 tier the FP64 tier's integers."""
    seed, ladder, nodes, factors, n_pts = 99, 2268, 5, 3, 8192
    xs, z = api.debug_sobol_f32(seed, ladder, 1, n_pts, nodes, factors, True, full_ladder=True)
    _, full, _ = api.debug_sobol(seed, ladder, 1, n_pts, 0, nodes * factors)
    assert np.array_equal(xs, full)
    from scipy.special import ndtri
    zr = ndtri((full.astype(np.float64) + 0.5) / 2.0 ** 32)
    assert np.max(np.abs(z - zr) / np.maximum(1.0, np.abs(zr))) < 5e-6
    # priced: every scenario of a CRN batch within the FP32 bar of the FP64 tier
    names, reqs = W.crn_scenarios(paths=2048, steps=126, years=0.5)
    md = max(W.max_dimension_for(r) for r in reqs)
    f64 = api.price_batch(reqs, "fp64", max_dimension=md)
    full32 = api.price_batch(reqs, "fp32", max_dimension=md, fp32_full_ladder=True)
    for a, b in zip(full32, f64):
        assert rel(a.pv, b.pv) <= FP32_REL


@pytest.mark.parametrize("first_point", [(1 << 21) + 3, (1 << 24) + 11])
def test_fp32_tier_ladder_beyond_21_bit_point_indices(first_point):
    """The FP32 tier's ladder (24 hashed levels, tail passes: the coordinate table holds points below 2^22) at point
    indices that take the 24-bit and the 32-bit gray-code sweeps, against the rule
    applied to the FP64 ladder's full 32-level integers at the same points."""
    seed, ladder, nodes, factors, n_pts = 4_242, 2268, 4, 3, 4096
    xs, _ = api.debug_sobol_f32(seed, ladder, first_point, n_pts, nodes, factors, True)
    _, full, _ = api.debug_sobol(seed, ladder, first_point, n_pts, 0, nodes * factors)
    p24 = full >> np.uint32(8)
    tail = (p24 < (1 << 14)) | (p24 >= (1 << 24) - (1 << 14))
    want = np.where(tail, full, (p24 << np.uint32(8)) | np.uint32(0x80)).astype(np.uint32)
    assert np.array_equal(xs, want)
    sob = O.Sobol(ladder, seed, max_dim=ladder)
    for p in (0, 1, 63, 64, 2047, n_pts - 1):
        for d in range(nodes * factors):
            assert int(full[p, d]) == sob.scrambled(first_point + p, d)


@pytest.mark.parametrize("seed,ladder_dims,first_point,n_points,first_dim,n_dims", [
    (1729, 64, 1, 200, 0, 64),            # the reference's term-sheet lane shape (bases 2..311)
    (7, 512, 60_000, 100, 400, 112),      # large bases, 5-digit indices
    (11, 700, (1 << 31) + 12_345, 64, 400, 300),  # past the reference cap of 512 (bases to 5,279)
])
def test_halton_digits_and_radical_inverses_bit_exact(seed, ladder_dims, first_point, n_points, first_dim, n_dims):
    """The device's Halton ladder (qmc.rs:94-153): the permuted base-b digits of every
    point, its radical inverse bit for bit and its inverse normal, against the
    oracle's bases and Fisher-Yates permutations (qmc.rs:129-140)."""
    dig, u, z = api.debug_halton(seed, ladder_dims, first_point, n_points, first_dim, n_dims)
    hal = O.Halton(ladder_dims, seed, max_dim=max(512, ladder_dims))
    perms = {}
    for d in range(n_dims):
        dim = first_dim + d
        di = dim % hal.dimension
        if di not in perms:
            perms[di] = (hal.base(di), hal.perm(di))
        b, perm = perms[di]
        for p in range(n_points):
            n = first_point + p
            want, m = [], n
            while m > 0:
                want.append(perm[m % b])
                m //= b
            k = len(want)
            assert list(dig[p, d, :k]) == want and np.all(dig[p, d, k:] == 0xFFFFFFFF), (n, dim)
            assert u[p, d] == hal.uniform(n, dim), (n, dim)
            assert z[p, d] == pytest.approx(hal.normal(n - 1, dim), rel=1e-15, abs=1e-15)


def test_path_counts_beyond_the_device_point_index_are_refused():
    from riga_flow_amd.types import RigaError
    req = W.flat_book(paths=1 << 31, steps=4)
    with pytest.raises(RigaError) as e:
        api.price_put_leg(req, "fp64")
    assert e.value.code == "invalid_paths"


def _flat(paths, steps=252, **kw):
    req = W.flat_book(paths=paths, steps=steps)
    for k, v in kw.items():
        setattr(req.trade, k, v)
    return req


def test_flat_book_fp64_parity():
    req = _flat(4096)
    md = W.max_dimension_for(req)
    res = api.price_put_leg(req, "fp64", max_dimension=md)
    ora = O.price(req, max_dim=md)
    assert_fp64_parity(res, ora)
    # trace events: first 64 KO (path, step) in path order, note.rs:174-180
    dt = req.trade.maturity_years / req.trade.steps
    want = [f"daily_global_ko:path={p}:t={(s + 1) * dt:.6f}" for p, s in ora.trace]
    assert res.trace_events == want


def test_flat_book_event_indices_bit_exact():
    req = _flat(2048)
    md = W.max_dimension_for(req)
    recs, res = api.replay_paths(req, "fp64", max_dimension=md)
    ora, orecs = O.price(req, max_dim=md, records=True)
    assert np.array_equal(recs["ko_step"], orecs["ko_step"])
    assert np.array_equal(recs["hit_ki"], orecs["hit_ki"])
    for f in ("final_worst", "funding", "put", "coupon_pv", "memory_pv", "pv"):
        np.testing.assert_allclose(recs[f], orecs[f], rtol=1e-12, atol=1e-15)
    assert_fp64_parity(res, ora)


def test_lv_book_fp64_parity():
    req = W.lv_book(paths=512, steps=756)
    md = W.max_dimension_for(req)
    res = api.price_put_leg(req, "fp64", max_dimension=md)
    ora = O.price(req, max_dim=md)
    assert_fp64_parity(res, ora)


def test_lv_book_fp32_tier_within_1e4():
    req = W.lv_book(paths=2048, steps=756)
    md = W.max_dimension_for(req)
    res = api.price_put_leg(req, "fp32", max_dimension=md)
    ora = O.price(req, max_dim=md)
    assert rel(res.pv, ora.pv) <= FP32_REL
    assert rel(res.legs.funding, ora.leg_funding) <= FP32_REL


def test_lv_fp32_tier_uneven_knots_within_1e4():
    # knots moved off the even spacing: the FP32 tier falls back to the
    # bracketing search instead of the one-multiply knot position
    import dataclasses
    req = W.lv_book(paths=2048, steps=252, years=1.0)
    grids = []
    for g in req.local_vol.calibrated:
        n = len(g.log_s)
        xs = [g.log_s[i] + (0.3 * (g.log_s[1] - g.log_s[0]) * math.sin(1.7 * i) if 0 < i < n - 1 else 0.0)
              for i in range(n)]
        assert all(b > a for a, b in zip(xs, xs[1:]))
        grids.append(dataclasses.replace(g, log_s=xs))
    req = dataclasses.replace(req, local_vol=dataclasses.replace(req.local_vol, calibrated=grids))
    md = W.max_dimension_for(req)
    ora = O.price(req, max_dim=md)
    res = api.price_put_leg(req, "fp32", max_dimension=md)
    assert rel(res.pv, ora.pv) <= FP32_REL
    assert rel(res.legs.down_and_in_put, ora.leg_put) <= 1e-3
    f64 = api.price_put_leg(req, "fp64", max_dimension=md)
    assert_fp64_parity(f64, ora)


def test_fp32_memory_coupons_accumulate_within_1e4():
    # a coupon barrier above spot: most periods are missed and carried as
    # memory coupons until a KO or a paying period (the FP32 tier counts them)
    req = W.flat_book(paths=4096, steps=252)
    req.trade.coupon_barrier = 1.02
    req.trade.knock_out_barrier = 1.06
    md = W.max_dimension_for(req)
    ora = O.price(req, max_dim=md)
    res = api.price_put_leg(req, "fp32", max_dimension=md)
    assert ora.leg_memory_coupon > 0.0
    assert rel(res.pv, ora.pv) <= FP32_REL
    assert rel(res.legs.memory_coupon, ora.leg_memory_coupon) <= 1e-3
    assert rel(res.legs.coupon, ora.leg_coupon) <= 1e-3


def test_crn_scenarios_alone_equal_in_mix():
    names, reqs = W.crn_scenarios(paths=1024, steps=126, years=0.5)
    md = max(W.max_dimension_for(r) for r in reqs)
    mix = api.price_batch(reqs, "fp64", max_dimension=md)
    for i in (0, 1, 7):
        alone = api.price_put_leg(reqs[i], "fp64", max_dimension=md)
        assert alone.pv == mix[i].pv
        assert alone.legs.down_and_in_put == mix[i].legs.down_and_in_put
    ora = O.price(reqs[3], max_dim=md)
    assert_fp64_parity(mix[3], ora)


def test_plan_graph_replays_match_eager():
    # a hipGraph-captured plan replayed several times, with eager work between
    # replays, gives the eager result every time
    names, reqs = W.crn_scenarios(paths=2048, steps=126, years=0.5)
    md = max(W.max_dimension_for(r) for r in reqs)
    eager = api.price_batch(reqs, "fp32", max_dimension=md)
    plan = api.Plan(reqs, "fp32", use_graph=True, max_dimension=md)
    assert plan.info.graph == 1
    for _ in range(3):
        got = plan.run()
        for a, b in zip(got, eager):
            assert a.pv == b.pv and a.legs.down_and_in_put == b.legs.down_and_in_put
        api.price_put_leg(reqs[0], "fp64", max_dimension=md)
    plan.close()


def test_halton_lane_parity():
    req = _flat(1024, steps=64)
    req.config.variance_reduction = W.VarianceReduction("halton", 64 * 3)
    res = api.price_put_leg(req, "fp64")
    ora = O.price(req)
    assert_fp64_parity(res, ora)


def test_reference_dimension_cap_wraps_like_the_reference():
    req = _flat(1024, steps=252)  # 756 dims requested, 512 cap of qmc.rs:9
    res = api.price_put_leg(req, "fp64", max_dimension=512)
    ora = O.price(req, max_dim=512)
    assert_fp64_parity(res, ora)


@pytest.mark.parametrize("tier", ["fp64", "fp32"])
def test_owen_table_and_full_hash_agree(tier):
    # the tabulated coordinates and the hashed ones are identical integers, so
    # identical prices: a flat 1y book and the 3y local-vol trade (2,268 dimensions)
    for req in (_flat(2048, steps=252), W.lv_book(paths=2048, steps=756)):
        md = W.max_dimension_for(req)
        a = api.price_batch([req], tier, max_dimension=md, owen_table=True)[0]
        b = api.price_batch([req], tier, max_dimension=md, owen_table=False)[0]
        c = api.price_batch([req], tier, max_dimension=md)[0]  # the library's choice
        assert a.pv == b.pv == c.pv and a.legs.down_and_in_put == b.legs.down_and_in_put
        assert a.stats[0].output_count == b.stats[0].output_count
        assert a.trace_events == b.trace_events


@pytest.mark.parametrize("tier", ["fp64", "fp32"])
def test_owen_table_entries_past_22_bit_point_indices(tier):
    # below 2^22 points a table entry is the scrambled coordinate itself; an item
    # whose point indices need 23 bits hashes the low levels after its entries' top
    # 22: a book of 2^22 + 1,000 paths prices identically with the table and with
    # every level hashed, inline and through the shared normal table
    req = _flat((1 << 22) + 1000, steps=12)
    md = W.max_dimension_for(req)
    a = api.price_batch([req], tier, max_dimension=md, owen_table=True, normal_table=-1)[0]
    b = api.price_batch([req], tier, max_dimension=md, owen_table=False, normal_table=-1)[0]
    c = api.price_batch([req], tier, max_dimension=md, owen_table=True, normal_table=1)[0]
    assert a.pv == b.pv == c.pv and a.legs.down_and_in_put == b.legs.down_and_in_put
    assert a.stats[0].output_count == b.stats[0].output_count == c.stats[0].output_count


def test_owen_tables_are_built_for_plans_that_read_enough_points():
    # a seed no other test uses: a plan reading fewer than 2^18 points per dimension
    # hashes every level; one reading 2^18 builds the seed's coordinate table; a later
    # small plan of the seed reads that table; a forced table and none at all price
    # identically
    def req(paths):
        q = _flat(paths, steps=21)
        q.config.seed = 987_651
        return q
    md = W.max_dimension_for(req(1024))
    small = api.Plan([req(4096)], "fp64", max_dimension=md)
    assert small.info.owen_tables == 0
    a = small.run()[0]
    small.close()
    big = api.Plan([req(1 << 18)], "fp64", max_dimension=md)
    assert big.info.owen_tables == 1
    big.close()
    again = api.Plan([req(4096)], "fp64", max_dimension=md)
    assert again.info.owen_tables == 1
    b = again.run()[0]
    again.close()
    forced = api.Plan([req(4096)], "fp32", max_dimension=md, owen_table=True)
    off = api.Plan([req(4096)], "fp32", max_dimension=md, owen_table=False)
    assert (forced.info.owen_tables, off.info.owen_tables) == (1, 0)
    c, d = forced.run()[0], off.run()[0]
    forced.close()
    off.close()
    assert a.pv == b.pv and a.stats[0].output_count == b.stats[0].output_count
    assert c.pv == d.pv and c.stats[0].output_count == d.stats[0].output_count
    # a table covering only the first 40 of the 63 dimensions (another fresh seed):
    # the shared normal table reads it for those dimensions and hashes every level
    # of the others, and a path kernel hashes every level of an item it does not
    # cover whole -- the integers are the same either way
    for tier in ("fp64", "fp32"):
        q = req(4096)
        q.config.seed = 987_652 if tier == "fp64" else 987_653
        f = api.price_batch([q], tier, max_dimension=md, owen_table=False, normal_table=-1)[0]
        for table in (1, -1):
            part = api.Plan([q], tier, max_dimension=md, owen_table=40, normal_table=table)
            assert part.info.owen_tables == 1 and part.info.normal_tables == (1 if table == 1 else 0)
            e = part.run()[0]
            part.close()
            assert e.pv == f.pv and e.stats[0].output_count == f.stats[0].output_count
            assert e.trace_events == f.trace_events


@pytest.mark.parametrize("ppt,waves", [(2, 4), (4, 4), (1, 1), (8, 2), (1, 3)])
def test_launch_shape_changes_nothing_but_the_shape(ppt, waves):
    req = W.lv_book(paths=3000, steps=126, years=0.5)
    md = W.max_dimension_for(req)
    a = api.price_batch([req], "fp64", max_dimension=md, paths_per_thread=1, waves_per_block=4)[0]
    b = api.price_batch([req], "fp64", max_dimension=md, paths_per_thread=ppt,
                        waves_per_block=waves)[0]
    assert b.stats[0].output_count == a.stats[0].output_count
    assert b.trace_events == a.trace_events  # first KO events in path order
    assert rel(b.pv, a.pv) <= 1e-14


@pytest.mark.parametrize("tier", ["fp64", "fp32"])
def test_persistent_workgroups_price_as_one_workgroup_per_unit(tier):
    # persistent launches (as many workgroups as the device holds, each claiming
    # units until the queue is empty) against one workgroup per unit: a CRN batch
    # of more units than the device holds workgroups, and a book of 1-5 factor
    # trades (one launch per factor count, FP64 ones with pooled bridge stacks)
    names, crn = W.crn_scenarios(paths=262_144, steps=126, years=0.5)
    book = W.uneven_book(n_trades=10, paths=4096, seed=21)
    for q in book:
        q.trade.steps = min(q.trade.steps, 300)
        q.trade.maturity_years = q.trade.steps / 252.0
        q.trade.coupon_periods = max(1, round(12 * q.trade.maturity_years))
        q.config.variance_reduction = W.VarianceReduction("sobol", q.trade.steps * q.factors)
    for reqs in (crn, book):
        md = max(W.max_dimension_for(r) for r in reqs)
        a = api.price_batch(reqs, tier, max_dimension=md)
        b = api.price_batch(reqs, tier, max_dimension=md, persistent=True)
        for x, y in zip(a, b):
            assert x.pv == y.pv and x.legs.down_and_in_put == y.legs.down_and_in_put
            assert x.stats[0].output_count == y.stats[0].output_count
            assert x.trace_events == y.trace_events


def test_autotuner_picks_a_shape_that_plans_then_use():
    names, reqs = W.crn_scenarios(paths=8192, steps=126, years=0.5)
    md = max(W.max_dimension_for(r) for r in reqs)
    rep = api.autotune(reqs, "fp32", max_dimension=md, reps=1)
    assert len(rep["trials"]) == 16 and rep["device_ms"] > 0
    assert rep["device_ms"] == min(t[2] for t in rep["trials"] if t[2] > 0)
    plan = api.Plan(reqs, "fp32", use_graph=True, max_dimension=md)
    assert (plan.info.paths_per_thread, plan.info.waves_per_block) == \
        (rep["paths_per_thread"], rep["waves_per_block"])
    tuned = plan.run()
    plan.close()
    ref = api.price_batch(reqs, "fp32", max_dimension=md, paths_per_thread=1, waves_per_block=4)
    for a, b in zip(tuned, ref):
        assert rel(a.pv, b.pv) <= 1e-14


def test_uneven_book_shared_normal_table_is_exact():
    # trades of 1-5 underlyings and different maturities in one launch; the
    # shared [dim][path] normal table gives the same integers and normals as
    # per-trade generation, and every trade matches the oracle
    book = W.uneven_book(n_trades=12, paths=1024, seed=3)
    for q in book:  # keep the oracle fast: shorten the horizons
        q.trade.steps = min(q.trade.steps, 160)
        q.trade.maturity_years = q.trade.steps / 252.0
        q.trade.coupon_periods = max(1, round(12 * q.trade.maturity_years))
        q.config.variance_reduction = W.VarianceReduction("sobol", q.trade.steps * q.factors)
    md = 5 * 160
    on = api.price_batch(book, "fp64", max_dimension=md, normal_table=1)
    off = api.price_batch(book, "fp64", max_dimension=md, normal_table=-1)
    for a, b in zip(on, off):
        assert a.pv == b.pv and a.stats[0].output_count == b.stats[0].output_count
    for i in (0, 5, 11):
        assert_fp64_parity(on[i], O.price(book[i], max_dim=md))
    f32 = api.price_batch(book, "fp32", max_dimension=md, normal_table=1)
    for a, b in zip(f32, on):
        assert rel(a.pv, b.pv) <= FP32_REL
