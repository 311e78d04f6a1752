"""The C-ABI library: it loads, exports every entry point include/riga_hip.h
declares, validates requests with the reference's error codes
(crates/riga-engine/src/lib.rs:209-321) and fails loudly without a device.
No device work happens here."""
import ctypes as C
import os
import re

import pytest

from riga_flow_amd import _native as N
from riga_flow_amd import api, workloads as W
from riga_flow_amd.types import RigaError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "riga_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|double)\s+\**(riga_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    names = declared_functions()
    assert len(names) >= 11
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(N.EXPORTS)


def test_ctypes_struct_sizes_match_the_header_layout():
    # riga_request/riga_result are plain C structs; keep the Python mirror in step
    assert C.sizeof(N.Error) == 512
    assert C.sizeof(N.Result) == 8 * 8 + 5 * 8 + 6 * 8 + 8 + 64 * 8 * 2 + 8
    assert C.sizeof(N.PathRecord) == 56


def _code(req, **trade):
    for k, v in trade.items():
        setattr(req.trade, k, v)
    with pytest.raises(RigaError) as e:
        api.validate(req, max_dimension=W.max_dimension_for(req))
    return e.value.code


def _raw_ladder_code(ladder):
    import ctypes as C
    from riga_flow_amd import _native as N
    p = api._Packed(W.flat_book(paths=16), 512)
    p.r.ladder = ladder
    err = N.Error()
    assert N.lib().riga_validate(C.byref(p.r), C.byref(err)) != 0
    return err.code.decode()


def test_validation_error_codes():
    ok = W.flat_book(paths=16)
    api.validate(ok, max_dimension=W.max_dimension_for(ok))
    req = W.flat_book(paths=16)
    req.market.correlation[0][1] = 2.0
    req.market.correlation[1][0] = 2.0
    assert _code(req) == "invalid_market"  # lib.rs:595-607
    assert _code(W.flat_book(paths=16), steps=0) == "invalid_schedule"
    assert _code(W.flat_book(paths=16), knock_out_barrier=-1.0) == "invalid_trade"
    assert _code(W.flat_book(paths=16), coupon_rate=-0.1) == "invalid_trade"
    assert _code(W.flat_book(paths=16), strike=0.0) == "invalid_trade"
    req = W.flat_book(paths=16)
    req.config.paths = 0
    assert _code(req) == "invalid_paths"
    req = W.flat_book(paths=16)
    req.market.spots[1] = -1.0
    assert _code(req) == "invalid_market"
    req = W.flat_book(paths=16)
    req.config.variance_reduction = W.VarianceReduction("none")
    api.validate(req)  # the ChaCha8 Monte Carlo lane
    assert _raw_ladder_code(7) == "unsupported_variance_reduction"
    req = W.flat_book(paths=16)
    req.market.spots = [100.0] * 6
    req.market.vols = [0.2] * 6
    req.market.dividend_yields = [0.0] * 6
    req.trade.initial_spots = [100.0] * 6
    req.market.correlation = [[1.0 if i == j else 0.0 for j in range(6)] for i in range(6)]
    assert _code(req) == "invalid_market_shape"


def test_pricing_without_a_device_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device node is present")
    with pytest.raises(RigaError) as e:
        api.price_put_leg(W.flat_book(paths=16))
    assert e.value.code == "gpu_unavailable"


def test_ctypes_mirrors_match_the_header_compiled_by_gcc(tmp_path):
    # the C compiler's sizes and offsets of the structs that gained fields,
    # against the ctypes mirror the Python side passes through the ABI
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "riga_hip.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(riga_launch_config),\n'
        "         offsetof(riga_launch_config, persistent), sizeof(riga_ticket_info),\n"
        "         offsetof(riga_ticket_info, device_end_s), sizeof(riga_server_stats),\n"
        "         offsetof(riga_server_stats, overlapped_plans), sizeof(riga_plan_info),\n"
        "         offsetof(riga_plan_info, owen_tables));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = [C.sizeof(N.LaunchConfig), N.LaunchConfig.persistent.offset, C.sizeof(N.TicketInfo),
            N.TicketInfo.device_end_s.offset, C.sizeof(N.ServerStats),
            N.ServerStats.overlapped_plans.offset, C.sizeof(N.PlanInfo), N.PlanInfo.owen_tables.offset]
    assert got == want


def test_bench_op_model_uses_the_built_owen_table_width():
    # the roofline's op model counts the Owen levels the kernels hash: those below
    # the coordinate table's width, which the library is built with
    layout = open(os.path.join(ROOT, "riga_flow_amd", "csrc", "phoenix_layout.h")).read()
    built = int(re.search(r"#define RIGA_OWEN_PREFIX_BITS (\d+)", layout).group(1))
    model = int(re.search(r"^OWEN_TABLE_BITS = (\d+)", open(os.path.join(ROOT, "bench.py")).read(),
                          re.M).group(1))
    assert built == model
