"""Benchmark of the Phoenix path-simulation hot path on one MI355X.

Metric: path-steps/s of the FP32 batched form (10 CRN scenarios of the 3-asset
local-vol trade sharing 1,048,576 Sobol paths x 756 daily steps, one hipGraph
replay per step) beside the plain FP64 form (the same trade, one scenario,
262,144 paths), with the FP32 price/Greek error against FP64 on identical
paths, the roofline of the dominant kernel and the CPU oracle timed on the
host's cores.  Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--quick]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from riga_flow_amd import api, workloads as W  # noqa: E402

# ---------------------------------------------------------------- roofline model
# Algorithmic work of the code path each tier runs, in 32-bit lane operations
# (an FP64 operation counts one: MI355X issues FP64 at the full vector rate).
# Per generated coordinate:
OWEN_ROUND = 21     # hash input xor 2 (three-input), xorshift 2, 64-bit multiply 4,
                    # xorshift 3, top bit of the 2nd multiply 5, scrambled bit 2, prefix*C3 update 3
                    # (the prefix itself is one multiply of P by C3^-1 after the last level)
# Points below 2^OWEN_TABLE_BITS (RIGA_OWEN_PREFIX_BITS in riga_flow_amd/csrc/phoenix_layout.h)
# read their scrambled coordinates from the seed's per-dimension table; past them a
# coordinate is the gray-code XOR and hashed levels: all 32 in the FP64 tier, a float's 24
# in the FP32 tier plus levels 7..0 for coordinates within 2^-10 of 0 or 1 (a fraction 2^-9)
OWEN_TABLE_BITS = 22


def owen_rounds(tier: str, nbits: int) -> float:
    if nbits <= OWEN_TABLE_BITS:
        return 0.0
    return 32 if tier == "fp64" else 24 + 8 * 2.0 ** -9


OWEN_TABLE = 3      # coordinate table address and load
PROBIT = 22         # rational (12) + divide (8) + tails (5% of log + sqrt + rational)
UNIFORM = 4         # int->float, +0.5, scale, clamp
# Per scenario and factor:
LV_LOOKUP = {"fp64": 41,  # log 15, knot bracket 6, divide 8, three lerps 12
             "fp32": 9}   # knot position 1, clamp 1 (med3), index 1, weight 1, address 1,
                          # slope-row lerps 4
EULER = {"fp64": 22,      # drift 4, increment 2, exp 15, multiply 1
         "fp32": 3}       # X += c + vol (dW - vol dt/2): two FMAs and an add (no exp)
# Phoenix scan per scenario step: the FP64 tier's ratios to the initial spots (F divides),
# worst-of, KO test, accrual and payment; the FP32 tier's worst-of on log spots (F - 1),
# accrual 4, KO test 3, payment branch and its rare body (~2)
SCAN = {"fp64": lambda F: 9 * F + 5, "fp32": lambda F: F + 10}
VALU_PEAK = 256 * 4 * 16 * 2.4e9  # CUs x SIMDs x lanes x clock = 3.93e13 lane-ops/s
HBM_PEAK = 8.0e12


def ops_per_path_step(F: int, S: int, nbits: int, local_vol: bool, tier: str) -> float:
    """Lane operations of one path advanced one step for all S scenarios."""
    source = OWEN_TABLE if nbits <= OWEN_TABLE_BITS else 2 * nbits + owen_rounds(tier, nbits) * OWEN_ROUND
    coord = source + UNIFORM + PROBIT
    node = F * coord + F * (F + 1) + 3 * F + 2 * F  # coordinates, L z, bridge node, shock
    scen = F * ((LV_LOOKUP[tier] if local_vol else 0) + EULER[tier]) + SCAN[tier](F)
    return node + S * scen


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "phoenix_kernel_lv3_crn10_fp32_traffic.txt")
BOOK_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "phoenix_kernel_uneven_book_traffic.txt")


def measured_traffic(path: str = TRAFFIC_FILE):
    """HBM bytes per launch (or per book) from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-kernel calibrated) and the commit it was profiled
    at, or (None, None)."""
    try:
        vals, commit = {}, None
        for line in open(path):
            if line.startswith("commit="):
                commit = line.strip().split("=", 1)[1]
            elif "=" in line and not line.startswith("#"):
                k, v = line.strip().split("=", 1)
                vals[k] = float(v)
        return (vals["fetch_kib_per_dispatch"] * 1024 * vals["fetch_scale"] +
                vals["write_kib_per_dispatch"] * 1024 * vals["write_scale"]), commit
    except (OSError, KeyError, ValueError):
        return None, None


def profiled_issue(path: str = TRAFFIC_FILE):
    """VALU lane-operations the kernel issued per (scenario) path-step, from the same
    committed rocprofv3 summary (SQ_INSTS_VALU x 64 / path-steps), or None."""
    try:
        for line in open(path):
            if line.startswith("issued_lane_ops_per_path_step="):
                return float(line.strip().split("=", 1)[1])
    except (OSError, ValueError):
        pass
    return None


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline(seconds_target: float = 12.0):
    """The oracle (single thread, then all host threads) on a bounded sample of the
    3-asset local-vol trade; path-steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker, timed here as the reported CPU baseline
    req = W.lv_book(paths=1024, steps=756)
    md = W.max_dimension_for(req)
    lvs = O.local_vols_for(req)
    t0 = time.perf_counter()
    O.price(req, max_dim=md, lvs=lvs, path_range=(0, 256))
    per_path = (time.perf_counter() - t0) / 256
    n1 = max(256, min(65536, int(seconds_target / per_path)))
    req1 = W.lv_book(paths=n1, steps=756)
    t0 = time.perf_counter()
    O.price(req1, max_dim=md, lvs=lvs)
    t1 = time.perf_counter() - t0
    single = n1 * 756 / t1
    # every core this process may run on, up to the 16-CPU share one GPU's box allots
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, affinity))
    nper = max(64, int(seconds_target / per_path / 2))
    reqm = W.lv_book(paths=nper * threads, steps=756)

    def work(i):
        O.price(reqm, max_dim=md, lvs=lvs, path_range=(i * nper, (i + 1) * nper))

    ts = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    tm = time.perf_counter() - t0
    multi = nper * threads * 756 / tm
    return {
        "value": single, "unit": "path-steps/s", "cores": 1, "kind": "port",
        "sample": f"oracle FP64 on the 3-asset local-vol trade, {n1} paths x 756 steps, "
                  f"1 thread ({t1:.1f} s)",
        "multi_thread": {"value": multi, "cores": threads,
                         "sample": f"{nper * threads} paths x 756 steps over {threads} threads "
                                   f"({tm:.1f} s)"},
        "host": {"os_cpu_count": os.cpu_count(), "affinity_cpus": affinity,
                 "cpu_model": _cpu_model(), "threads_used": threads},
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------- GPU legs
def timed(plan, steps: int, warmup: int):
    for _ in range(warmup):
        plan.run(raw=True)
    dev = []
    t0 = time.perf_counter()
    last = None
    for _ in range(steps):
        last = plan.run(raw=True)
        dev.append(last[0].device_ms)
    wall = time.perf_counter() - t0
    return wall, dev, last


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-book", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-mc", action="store_true", help="skip the Monte Carlo (ChaCha8) lane leg")
    ap.add_argument("--quick", action="store_true", help="smaller path counts (smoke)")
    ap.add_argument("--ppt", type=int, default=0, help="FP32 batch: paths per lane (0 = autotuned)")
    ap.add_argument("--waves", type=int, default=0, help="FP32 batch: waves per workgroup (0 = autotuned)")
    ap.add_argument("--ppt64", type=int, default=0, help="FP64 plain: paths per lane (0 = autotuned)")
    ap.add_argument("--waves64", type=int, default=0, help="FP64 plain: waves per workgroup (0 = autotuned)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (shape 1 x 4 waves)")
    ap.add_argument("--dispatched", action="store_true",
                    help="one workgroup per unit instead of persistent workgroups (as many as the "
                         "device holds, each claiming units until the queue is empty)")
    ap.add_argument("--book-only", action="store_true",
                    help="run only the uneven-book leg and print its JSON (for profiling)")
    ap.add_argument("--mc-only", action="store_true",
                    help="run only the Monte Carlo lane leg and print its JSON (for profiling)")
    ap.add_argument("--legs-only", action="store_true",
                    help="time only the FP32 batch and FP64 plain legs (for profiling: no error "
                         "checks or other legs launch the same kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the hot path shards by contiguous path range "
              f"(Sobol point = path + 1) with one merge of the per-block double-double "
              f"partials, so --gpus {args.gpus} is not run here.",
              file=sys.stderr)
        sys.exit(2)

    api.gpu_backend()
    if args.book_only:
        print(json.dumps({"uneven_book": book_leg(args.quick, not args.dispatched)}), flush=True)
        return
    if args.mc_only:
        print(json.dumps({"mc_lane": mc_leg(args)}), flush=True)
        return
    big = 1_048_576 if not args.quick else 65_536
    plain_paths = 262_144 if not args.quick else 16_384
    names, crn = W.crn_scenarios(paths=big, steps=756, years=3.0)
    md = max(W.max_dimension_for(r) for r in crn)
    tuning = {}
    if not args.no_tune and not (args.ppt and args.waves):
        # the autotuner times every launch shape on these batches (outside the
        # timed region) and the plans below pick up its choice
        tuning["fp32_batched"] = api.autotune(crn, "fp32", use_graph=True, max_dimension=md, reps=1)
        log(f"autotune fp32: {tuning['fp32_batched']['paths_per_thread']} paths/lane x "
            f"{tuning['fp32_batched']['waves_per_block']} waves")
    ppt, waves = args.ppt, args.waves
    fp32 = api.Plan(crn, "fp32", paths_per_thread=ppt, waves_per_block=waves, use_graph=True,
                    max_dimension=md, persistent=not args.dispatched)
    plain = W.lv_book(paths=plain_paths, steps=756)
    if not args.no_tune and not (args.ppt64 and args.waves64):
        tuning["fp64_plain"] = api.autotune([plain], "fp64", use_graph=True, max_dimension=md,
                                            reps=1)
        log(f"autotune fp64: {tuning['fp64_plain']['paths_per_thread']} paths/lane x "
            f"{tuning['fp64_plain']['waves_per_block']} waves")
    fp64 = api.Plan([plain], "fp64", paths_per_thread=args.ppt64, waves_per_block=args.waves64,
                    use_graph=True, max_dimension=md, persistent=not args.dispatched)

    wall32, dev32, res32 = timed(fp32, args.steps, args.warmup)
    log(f"fp32 batched: {wall32 / args.steps * 1e3:.1f} ms/step")
    wall64, dev64, res64 = timed(fp64, args.steps, args.warmup)
    log(f"fp64 plain: {wall64 / args.steps * 1e3:.1f} ms/step")
    if args.legs_only:
        print(json.dumps({"fp32_batched_ms": wall32 / args.steps * 1e3, "fp64_plain_ms": wall64 / args.steps * 1e3,
                          "fp32_device_ms": sum(dev32) / len(dev32), "fp64_device_ms": sum(dev64) / len(dev64)}),
              flush=True)
        fp32.close()
        fp64.close()
        return
    # the same two batches in the other launch mode (same shapes), for the record
    modes = {}
    for key, plan, reqs, tier in (("fp32_batched", fp32, crn, "fp32"), ("fp64_plain", fp64, [plain], "fp64")):
        other = api.Plan(reqs, tier, paths_per_thread=plan.info.paths_per_thread,
                         waves_per_block=plan.info.waves_per_block, use_graph=True, max_dimension=md,
                         persistent=args.dispatched)
        w, d, _ = timed(other, max(3, args.steps // 2), 1)
        other.close()
        this_ms = (sum(dev32) / len(dev32)) if key == "fp32_batched" else (sum(dev64) / len(dev64))
        modes[key] = {"persistent_device_ms": this_ms if not args.dispatched else sum(d) / len(d),
                      "dispatched_device_ms": sum(d) / len(d) if not args.dispatched else this_ms}
    log(f"launch modes: {modes}")
    ps32 = fp32.path_steps
    ps64 = fp64.path_steps
    value = ps32 * args.steps / wall32
    fp64_rate = ps64 * args.steps / wall64

    # FP32 error against FP64 on identical paths and scenarios (the FP64 GPU
    # path is the oracle's arithmetic; tests/test_gpu_parity.py pins it).
    ref = api.Plan(crn, "fp64", max_dimension=md).run()
    got = fp32.run()
    rel_pv = max(abs(a.pv - b.pv) / abs(b.pv) for a, b in zip(got, ref))
    h = [0.01 * s for s in crn[0].market.spots]

    def greeks(rs):
        out = []
        for f in range(3):
            up, dn, base, vg = rs[1 + 2 * f], rs[2 + 2 * f], rs[0], rs[7 + f]
            out.append(((up.pv - dn.pv) / (2 * h[f]),
                        (up.pv - 2 * base.pv + dn.pv) / (h[f] * h[f]),
                        (vg.pv - base.pv) / 0.01))
        return out

    per_scenario_err = [abs(a.pv - b.pv) / abs(b.pv) for a, b in zip(got, ref)]
    g32, g64 = greeks(got), greeks(ref)

    def greek_errors(g):
        e = {}
        for i, k in enumerate(("delta", "gamma", "vega")):
            e[k + "_abs_err"] = max(abs(a[i] - b[i]) for a, b in zip(g, g64))
            e[k + "_rel_err"] = max(abs(a[i] - b[i]) / max(abs(b[i]), 1e-300) for a, b in zip(g, g64))
        return e

    gerr = greek_errors(g32)
    for i, k in enumerate(("delta", "gamma", "vega")):
        gerr[k + "_fp64"] = [b[i] for b in g64]
    # the FP32 tier with all 32 Owen levels hashed (the FP64 tier's integers): what
    # the 24-level rule's sequence change costs, apart from FP32 rounding
    full_plan = api.Plan(crn, "fp32", max_dimension=md, fp32_full_ladder=True)
    full = full_plan.run()
    full_ms = full[0].device_ms
    full_plan.close()
    full_err = {"max_rel_pv_err": max(abs(a.pv - b.pv) / abs(b.pv) for a, b in zip(full, ref)),
                "rel_pv_err_per_scenario": {n: abs(a.pv - b.pv) / abs(b.pv) for n, a, b in zip(names, full, ref)},
                **greek_errors(greeks(full)), "device_ms": full_ms}

    nbits = (big).bit_length()
    ops = ops_per_path_step(3, 10, nbits, True, "fp32") / 10.0  # per (scenario) path-step
    dev_s = sum(dev32) / len(dev32) / 1e3
    achieved = ops * ps32 / dev_s
    traffic, traffic_commit = measured_traffic()
    issued = profiled_issue()
    ops64 = ops_per_path_step(3, 1, plain_paths.bit_length(), True, "fp64")
    dev64_s = sum(dev64) / len(dev64) / 1e3
    achieved64 = ops64 * ps64 / dev64_s
    out = {
        "metric": "path-steps/s FP32-batched vs FP64-plain at 1e-4 price err; p50/p99 latency",
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "Dupire local-vol worst-of Phoenix, 3 underlyings, 3y x 756 daily steps; "
                        "FP32 batched = 10 CRN scenarios (base, spot +-1% x3, vega +1pt x3) on "
                        f"{big} shared Sobol paths; FP64 plain = 1 scenario, {plain_paths} paths",
            "paths": big, "scenarios": 10, "steps": 756, "underlyings": 3,
            "lv_grid": "64 tenors x 128 log-spot knots", "hip_graph": True,
            "launch_shape": {"paths_per_thread": fp32.info.paths_per_thread,
                             "waves_per_block": fp32.info.waves_per_block},
        },
        "fp64_plain": {"value": fp64_rate, "unit": "path-steps/s", "paths": plain_paths,
                       "ms_per_step": wall64 / args.steps * 1e3,
                       "device_ms": sum(dev64) / len(dev64),
                       "launch_shape": {"paths_per_thread": fp64.info.paths_per_thread,
                                        "waves_per_block": fp64.info.waves_per_block},
                       "roofline": {"bound": "valu", "achieved": achieved64 / 1e12,
                                    "peak": VALU_PEAK / 1e12, "unit": "Tops/s",
                                    "frac": achieved64 / VALU_PEAK,
                                    "kernel": "phoenix_kernel<double,3,1>",
                                    "ops_per_path_step": ops64}},
        "fp32_batched_over_fp64_plain": value / fp64_rate,
        "fp32_vs_fp64": {"max_rel_pv_err": rel_pv, "rel_pv_err_per_scenario": dict(zip(names, per_scenario_err)),
                         **gerr,
                         "base_pv_fp32": got[0].pv, "base_pv_fp64": ref[0].pv,
                         "reference": "the FP64 tier on the same 1,048,576 paths (bit-exact to the "
                                      "oracle: tests/test_gpu_configs.py)",
                         "ladder": "24 hashed Owen levels + cell middle (all 32 within 2^-10 of 0 or 1)",
                         "full_32_level_ladder": full_err},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
                     "unit": "Tops/s", "frac": achieved / VALU_PEAK, "traffic": traffic,
                     "traffic_gbps": traffic / dev_s / 1e9 if traffic else None,
                     "hbm_frac": traffic / dev_s / HBM_PEAK if traffic else None,
                     "kernel": "phoenix_kernel<float,3,10>", "device_ms": dev_s * 1e3,
                     "ops_per_path_step": ops,
                     "traffic_profiled_at": traffic_commit,
                     "issued_ops_per_path_step": issued,
                     "issue_frac": issued * ps32 / dev_s / VALU_PEAK if issued else None,
                     "note": "VALU-bound: no MFMA (hashing, scalar FP); achieved counts the "
                             "algorithmic lane-ops of the op model, issue_frac the lane-ops the "
                             "kernel issued (rocprofv3 SQ_INSTS_VALU x 64 per scenario path-step); "
                             "traffic = rocprofv3 FETCH_SIZE+WRITE_SIZE bytes per launch, "
                             "copy-calibrated; both from "
                             "profiles/phoenix_kernel_lv3_crn10_fp32_traffic.txt (profiled at "
                             "commit traffic_profiled_at)"},
        "autotune": tuning,
        "launch_mode": "dispatched" if args.dispatched else "persistent",
        "launch_modes": modes,
    }
    fp32.close()
    fp64.close()
    log("fp32 error check done")
    out["crn_variance_reduction"] = crn_vrf_leg(args.quick)
    out["reference_published"] = published_leg()
    if not args.no_mc:
        out["mc_lane"] = mc_leg(args)
    if not args.no_book:
        out["uneven_book"] = book_leg(args.quick, not args.dispatched)
    if not args.no_serving:
        out["serving"] = serving_leg(args.quick)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        log("cpu baseline done")
    print(json.dumps(out), flush=True)


def mc_leg(args) -> dict:
    """The Monte Carlo lane (VarianceReduction::None: ChaCha8Rng + StandardNormal,
    note.rs:58,124-126) on the same trade and batch shapes as the QMC legs; the
    timed region includes the stream parse that finds each path's first word.
    Also the reference's golden term sheet (30,000 paths) priced on the device."""
    from riga_flow_amd.fixture import load_termsheet1_fixture
    big = 1_048_576 if not args.quick else 65_536
    plain_paths = 262_144 if not args.quick else 16_384
    names, crn = W.crn_scenarios(paths=big, steps=756, years=3.0)
    plain = W.lv_book(paths=plain_paths, steps=756)
    for r in crn + [plain]:
        r.config.variance_reduction = W.VarianceReduction("none")
    shapes = {}
    if not args.no_tune:
        for key, reqs, tier in (("fp32_batched", crn, "fp32"), ("fp64_plain", [plain], "fp64")):
            t = api.autotune(reqs, tier, use_graph=True, reps=1)
            shapes[key] = {"paths_per_thread": t["paths_per_thread"], "waves_per_block": t["waves_per_block"]}
    p32 = api.Plan(crn, "fp32", use_graph=True)
    p64 = api.Plan([plain], "fp64", use_graph=True)
    w32, d32, r32 = timed(p32, args.steps, args.warmup)
    w64, d64, r64 = timed(p64, args.steps, args.warmup)
    ref = api.Plan(crn, "fp64").run()
    err = max(abs(a.pv - b.pv) / abs(b.pv) for a, b in zip(r32, ref))
    v32 = p32.path_steps * args.steps / w32
    v64 = p64.path_steps * args.steps / w64
    log(f"mc lane: fp32 batched {w32 / args.steps * 1e3:.1f} ms, fp64 plain {w64 / args.steps * 1e3:.1f} ms")
    launch = {"fp32_batched": [p32.info.paths_per_thread, p32.info.waves_per_block],
              "fp64_plain": [p64.info.paths_per_thread, p64.info.waves_per_block]}
    p32.close()
    p64.close()
    here = os.path.dirname(os.path.abspath(__file__))
    ts = load_termsheet1_fixture(os.path.join(here, "tests", "golden", "termsheet1.json")).request
    with open(os.path.join(here, "tests", "golden", "termsheet1.golden.json")) as fh:
        golden = json.load(fh)
    t0 = time.perf_counter()
    g = api.price_put_leg(ts, "fp64")
    gsec = time.perf_counter() - t0
    return {"fp32_batched": {"value": v32, "unit": "path-steps/s", "ms_per_step": w32 / args.steps * 1e3,
                             "device_ms": sum(d32) / len(d32), "paths": big, "scenarios": len(crn)},
            "fp64_plain": {"value": v64, "unit": "path-steps/s", "ms_per_step": w64 / args.steps * 1e3,
                           "device_ms": sum(d64) / len(d64), "paths": plain_paths},
            "fp32_batched_over_fp64_plain": v32 / v64,
            "fp32_max_rel_pv_err": err,
            "launch_shapes": launch, "autotuned": bool(shapes),
            "termsheet1_golden": {"pv": g.pv, "golden_pv": golden["pv"], "abs_pv_diff": abs(g.pv - golden["pv"]),
                                  "ko_count": g.stats[0].output_count,
                                  "golden_ko_count": golden["stats"][0]["output_count"],
                                  "seconds": gsec}}


def published_leg() -> dict:
    """The reference's published experiment figures (docs/plan/25-implementation-evidence.md
    :88,:107,:110), recomputed on the device from the same inputs: the CRN variance-reduction
    factor (riga-risk/src/lib.rs:835-845), the paired delta standard errors (lib.rs:919-939)
    and the QMC put-variance ratios (riga-engine/src/lib.rs:653-688)."""
    from riga_flow_amd import risk
    from riga_flow_amd.fixture import load_termsheet1_fixture
    fx = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "termsheet1.json")

    def small():
        r = load_termsheet1_fixture(fx).request
        r.config.paths, r.config.path_block_size, r.config.seed = 512, 128, 81_281
        return r

    t0 = time.perf_counter()
    vrf = risk.crn_variance_reduction_factor(small(), replicates=40)
    se16 = risk.crn_delta_convergence(small(), replicates=16).standard_error
    se64 = risk.crn_delta_convergence(small(), replicates=64).standard_error
    base = load_termsheet1_fixture(fx).request
    dim = base.trade.steps * 2
    var = {}
    for kind in ("none", "halton", "sobol"):
        reqs = []
        for seed in range(1, 13):
            r = base.clone()
            r.config.paths, r.config.path_block_size, r.config.seed = 4096, 4096, seed
            r.config.variance_reduction = W.VarianceReduction(kind, 0 if kind == "none" else dim)
            reqs.append(r)
        s = [x.legs.down_and_in_put for x in api.price_batch(reqs, "fp64")]
        m = sum(s) / len(s)
        var[kind] = sum((v - m) ** 2 for v in s) / (len(s) - 1)
    secs = time.perf_counter() - t0
    log(f"published figures: vrf {vrf:.2f}, se {se16:.2e}/{se64:.2e}, qmc {var['none'] / var['halton']:.2f}/"
        f"{var['none'] / var['sobol']:.2f} ({secs:.2f} s)")
    return {"crn_variance_reduction_factor": {"device": vrf, "published": 140.18},
            "delta_standard_error_16": {"device": se16, "published": 2.6e-5},
            "delta_standard_error_64": {"device": se64, "published": 1.1e-5},
            "put_variance_mc_over_halton": {"device": var["none"] / var["halton"], "published": 1.05},
            "put_variance_mc_over_sobol": {"device": var["none"] / var["sobol"], "published": 1.04},
            "tier": "fp64", "seconds": secs}


def crn_vrf_leg(quick: bool) -> dict:
    """CRN variance-reduction factor of the asset-0 spot delta (riga-risk's
    crn_variance_reduction_factor, lib.rs:296-336) on the 3-asset local-vol trade with
    implied grids recalibrated per scenario; 16 replicates x 4 valuations, one batch."""
    from riga_flow_amd import risk
    from riga_flow_amd.types import LocalVolSpec
    paths = 262_144 if not quick else 16_384
    req = W.lv_book(paths=paths, steps=756, years=3.0)
    sig = W.svi_sigmas(3)
    req.local_vol = LocalVolSpec(grids=[W.svi_implied(sig[f]) for f in range(3)], grid_x=128)
    t0 = time.perf_counter()
    factor, crn, ind = risk.crn_variance_reduction_factor(
        req, replicates=16, precision="fp32", max_dimension=W.max_dimension_for(req), details=True)
    secs = time.perf_counter() - t0
    log(f"crn variance reduction factor {factor:.1f} ({secs:.2f} s)")
    return {"factor": factor, "replicates": 16, "paths": paths, "tier": "fp32",
            "greek": "central spot delta of asset 0, 1% bump",
            "delta_mean_crn": sum(crn) / len(crn), "delta_mean_independent": sum(ind) / len(ind),
            "seconds_including_calibration": secs}


def book_roofline(book, tier: str, device_ms: float) -> dict:
    """Roofline of the uneven book's launches (shared normal table, then the
    table-mode path kernels).  Algorithmic bytes: the table written once
    (dims x paths x real) and read F normals per trade path-step; algorithmic
    lane-ops: per path-step F normals' correlation and bridge node and, per
    factor, the flat-vol Euler step, plus the Phoenix scan."""
    real = 4 if tier == "fp32" else 8
    paths = max(q.config.paths for q in book)
    dims = max(q.factors * q.trade.steps for q in book)
    reads = sum(q.config.paths * q.trade.steps * q.factors for q in book) * real
    writes = dims * paths * real
    ops = sum(q.config.paths * q.trade.steps *
              (q.factors * (q.factors + 1) + 5 * q.factors + q.factors * EULER[tier] + SCAN[tier](q.factors))
              for q in book)
    sec = device_ms / 1e3
    hbm = (reads + writes) / sec
    valu = ops / sec
    traffic, commit = measured_traffic(BOOK_TRAFFIC_FILE) if tier == "fp32" else (None, None)
    if tier == "fp64":
        traffic, commit = measured_traffic(BOOK_TRAFFIC_FILE.replace("_traffic", "_fp64_traffic"))
    return {"bound": "hbm" if hbm / HBM_PEAK >= valu / VALU_PEAK else "valu",
            "achieved": hbm / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": hbm / HBM_PEAK,
            "algorithmic_bytes": reads + writes, "table_bytes_written": writes, "normals_bytes_read": reads,
            "traffic": traffic, "traffic_profiled_at": commit,
            "valu_tops": valu / 1e12, "valu_frac": valu / VALU_PEAK, "device_ms": device_ms,
            "note": "device_ms spans the normal-table kernel and every path launch (events on the "
                    "plan's stream); traffic = rocprofv3 FETCH_SIZE+WRITE_SIZE bytes of one book run "
                    "(normal table + path launches), copy-calibrated, from "
                    "profiles/phoenix_kernel_uneven_book_traffic.txt / _fp64_traffic.txt"}


def book_leg(quick: bool, persistent: bool = False) -> dict:
    """2,048 uneven Phoenix trades (1-5 of 20 assets, 0.5-5y daily, 65,536 paths
    each) in one hipGraph: a shared normal table, then one launch per factor count,
    the launches running concurrently on streams of their own (persistent
    workgroups claiming the trades' path chunks longest first)."""
    n = 2048 if not quick else 128
    book = W.uneven_book(n_trades=n, paths=65_536)
    out = {"trades": n, "paths": 65_536}
    for tier in ("fp32", "fp64"):
        t0 = time.perf_counter()
        plan = api.Plan(book, tier, use_graph=True, max_dimension=7000, persistent=persistent)
        setup = time.perf_counter() - t0
        plan.run(raw=True)
        t0 = time.perf_counter()
        reps = 2
        dev = []
        for _ in range(reps):
            res = plan.run(raw=True)
            dev.append(res[0].device_ms)
        wall = (time.perf_counter() - t0) / reps
        out[tier] = {"path_steps": plan.path_steps, "ms_per_book": wall * 1e3,
                     "path_steps_per_s": plan.path_steps / wall,
                     "normal_tables": plan.info.normal_tables, "launches": plan.info.n_launches,
                     "concurrent_launches": True, "persistent_workgroups": persistent,
                     "plan_setup_s": setup, "device_gb": plan.info.device_bytes / 1e9,
                     "roofline": book_roofline(book, tier, min(dev))}
        log(f"book {tier}: {wall * 1e3:.1f} ms/book")
        plan.close()
    return out


def serving_leg(quick: bool) -> dict:
    """Open-loop Poisson arrivals over a resident 20-asset local-vol market."""
    import numpy as np
    from riga_flow_amd import serving as SV
    u = SV.Universe()
    u.make_resident()  # the market's surfaces live in device memory for the whole run
    rng = np.random.default_rng(42)
    pool = [SV.sample_request(u, rng) for _ in range(24 if not quick else 6)]
    tiers = ["fp32" if rng.random() < 0.5 else "fp64" for _ in pool]
    log(f"serving pool ready: {len(pool)} requests, {sum(len(r) for r in pool)} trade-scenarios")
    sizes = [sum(q.config.paths * q.trade.steps for q in req) for req in pool]

    def run(pool, tiers, qps, mean_size, tag):
        srv = SV.Server(hbm_budget_gib=96.0, max_inflight=3, max_batch_requests=4096,
                        max_queue_seconds=3.0, paths_per_thread=1, max_batch_path_steps=1e11)
        r = SV.poisson_load(srv, pool, tiers, qps, duration_s=5.0 if not quick else 1.0,
                            seed=int(qps))
        r["max_concurrent_plans"] = srv.stats()["max_concurrent"]
        r["offered_path_steps_per_s"] = qps * mean_size
        srv.close()
        log(f"serving {tag} qps={qps:g}: completed {r['completed']}, shed {r['shed']}, "
            f"p50 {r['p50_s']}, p99 {r['p99_s']}, wall {r['wall_s']:.1f} s")
        return r

    runs = [run(pool, tiers, qps, float(np.mean(sizes)), "config") for qps in (50.0, 200.0, 800.0)]
    # the same pool at a rate one GPU sustains: a mean request is ~5.7e10 path-steps,
    # about 0.3 s of FP32 or 4.8 s of FP64 work, so capacity is near 0.4 requests/s;
    # no queue limit here, every ticket completes
    srv = SV.Server(hbm_budget_gib=96.0, max_inflight=3, max_batch_requests=4096,
                    max_queue_seconds=0.0, paths_per_thread=1, max_batch_path_steps=1e11)
    qps_sus = 0.3
    sus = SV.poisson_load(srv, pool, tiers, qps_sus, duration_s=45.0 if not quick else 5.0, seed=3)
    sus["max_concurrent_plans"] = srv.stats()["max_concurrent"]
    sus["offered_path_steps_per_s"] = qps_sus * float(np.mean(sizes))
    srv.close()
    log(f"serving sustainable qps={qps_sus:g}: completed {sus['completed']}/{sus['submitted']}, "
        f"p50 {sus['p50_s']}, p99 {sus['p99_s']}, wall {sus['wall_s']:.1f} s")
    # the same market and tiers with small requests (16k/65k paths, 1-8 trades,
    # 1-4 bumps): the latency of the loop itself at a load one GPU can carry
    small = [SV.sample_request(u, rng, path_choices=(16_384, 65_536), max_trades=8, max_bumps=4)
             for _ in range(24 if not quick else 6)]
    small_tiers = ["fp32" if rng.random() < 0.5 else "fp64" for _ in small]
    small_sizes = [sum(q.config.paths * q.trade.steps for q in req) for req in small]
    small_runs = [run(small, small_tiers, qps, float(np.mean(small_sizes)), "small")
                  for qps in (10.0, 40.0)]
    return {"runs": runs, "mean_request_path_steps": float(np.mean(sizes)),
            "requests_in_pool": len(pool),
            "sustainable": sus,
            "note": "at 50/200/800 req/s the offered load exceeds one GPU by 2-3 orders of "
                    "magnitude; tickets queued longer than 3 s are shed; p50_s/p99_s are over "
                    "completed tickets, 'censored' counts shed tickets as later than all "
                    "(None where a percentile lands on one); one plan holds at most 1e11 "
                    "path-steps beyond its first ticket, so config-size tickets (5.7e10 on "
                    "average) batch one or two to a plan; 'sustainable' runs the same pool "
                    "at 0.3 req/s without shedding",
            "small_requests": {"runs": small_runs, "mean_request_path_steps": float(np.mean(small_sizes)),
                               "pool": "16k/65k paths, 1-8 trades x 1-4 bumps, same market and tiers"}}


if __name__ == "__main__":
    main()
