"""ctypes binding of the CPU FP64 oracle (libriga_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The C library
restates the reference engine (see riga_oracle.c for file:line citations).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libriga_oracle.so")
TRACE_CAP = 64
_LIB = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class _Lv(C.Structure):
    _fields_ = [("n_t", C.c_int), ("n_x", C.c_int), ("t_axis", C.POINTER(C.c_double)),
                ("log_s", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double))]


class _Grid(C.Structure):
    _fields_ = [("n_maturities", C.c_int), ("n_strikes", C.c_int),
                ("maturities", C.POINTER(C.c_double)), ("strikes", C.POINTER(C.c_double)),
                ("vols", C.POINTER(C.c_double))]


class _Sobol(C.Structure):
    _fields_ = [("dimension", C.c_int), ("v", C.POINTER(C.c_uint32)), ("seed", C.c_uint64)]


class _Halton(C.Structure):
    _fields_ = [("dimension", C.c_int), ("bases", C.POINTER(C.c_uint32)),
                ("perms", C.POINTER(C.POINTER(C.c_uint32)))]


_D = C.POINTER(C.c_double)


class _Request(C.Structure):
    _fields_ = [
        ("factors", C.c_int),
        ("spots", _D), ("vols", _D), ("dividend_yields", _D), ("correlation", _D),
        ("rate", C.c_double),
        ("initial_spots", _D),
        ("maturity_years", C.c_double),
        ("steps", C.c_int64),
        ("strike", C.c_double), ("knock_in_barrier", C.c_double),
        ("has_ko", C.c_int), ("knock_out_barrier", C.c_double),
        ("has_cb", C.c_int), ("coupon_barrier", C.c_double),
        ("has_cr", C.c_int), ("coupon_rate", C.c_double),
        ("memory_coupon", C.c_int),
        ("coupon_periods", C.c_int64),
        ("fixed_first_coupon", C.c_int),
        ("call_start_step", C.c_int64),
        ("paths", C.c_uint64), ("path_block_size", C.c_uint64), ("seed", C.c_uint64),
        ("trace_level", C.c_int), ("vr_kind", C.c_int), ("vr_dimension", C.c_int),
        ("max_dim", C.c_int),
        ("lv", C.POINTER(_Lv)),
    ]


class _Result(C.Structure):
    _fields_ = [
        ("pv", C.c_double), ("mc_error", C.c_double), ("ki_probability", C.c_double),
        ("leg_coupon", C.c_double), ("leg_memory_coupon", C.c_double), ("leg_put", C.c_double),
        ("leg_funding", C.c_double), ("leg_total", C.c_double),
        ("paths", C.c_uint64), ("ko_count", C.c_uint64), ("ki_count", C.c_uint64),
        ("put_eligible_count", C.c_uint64), ("coupon_path_count", C.c_uint64),
        ("ko_mean", C.c_double), ("ko_var", C.c_double), ("ki_mean", C.c_double),
        ("ki_var", C.c_double), ("cpn_mean", C.c_double), ("cpn_var", C.c_double),
        ("funding_sum", C.c_double), ("put_sum", C.c_double), ("coupon_sum", C.c_double),
        ("memory_sum", C.c_double),
        ("n_trace", C.c_int),
        ("trace_path", C.c_int64 * TRACE_CAP),
        ("trace_step", C.c_int64 * TRACE_CAP),
    ]


class PathRecord(C.Structure):
    _fields_ = [("ko_step", C.c_int32), ("hit_ki", C.c_int32), ("final_worst", C.c_double),
                ("funding", C.c_double), ("put", C.c_double), ("coupon_pv", C.c_double),
                ("memory_pv", C.c_double), ("pv", C.c_double)]


class CouponRow(C.Structure):
    _fields_ = [("period", C.c_int64), ("years", C.c_double), ("qualified_steps", C.c_int64),
                ("period_steps", C.c_int64), ("accrual", C.c_double), ("memory_paid", C.c_double),
                ("paid", C.c_double), ("memory_accrued", C.c_double), ("discount_factor", C.c_double),
                ("pv", C.c_double), ("knocked_out", C.c_int32), ("pad", C.c_int32)]


COUPON_DTYPE = np.dtype([("period", "<i8"), ("years", "<f8"), ("qualified_steps", "<i8"),
                         ("period_steps", "<i8"), ("accrual", "<f8"), ("memory_paid", "<f8"),
                         ("paid", "<f8"), ("memory_accrued", "<f8"), ("discount_factor", "<f8"),
                         ("pv", "<f8"), ("knocked_out", "<i4"), ("pad", "<i4")])

RECORD_DTYPE = np.dtype([("ko_step", "<i4"), ("hit_ki", "<i4"), ("final_worst", "<f8"),
                         ("funding", "<f8"), ("put", "<f8"), ("coupon_pv", "<f8"),
                         ("memory_pv", "<f8"), ("pv", "<f8")])


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ro_probit.restype = C.c_double
        L.ro_probit.argtypes = [C.c_double]
        L.ro_radical_inverse.restype = C.c_double
        L.ro_radical_inverse.argtypes = [C.c_uint64, C.c_uint32]
        L.ro_primitive_polynomials.restype = C.c_int
        L.ro_primitive_polynomials.argtypes = [C.c_int, C.POINTER(C.c_uint32)]
        L.ro_direction_numbers.argtypes = [C.c_uint32, C.POINTER(C.c_uint32)]
        L.ro_mix64.restype = C.c_uint64
        L.ro_mix64.argtypes = [C.c_uint64]
        L.ro_owen_scramble.restype = C.c_uint32
        L.ro_owen_scramble.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64]
        L.ro_sobol_init.argtypes = [C.POINTER(_Sobol), C.c_int, C.c_uint64, C.c_int]
        L.ro_sobol_free.argtypes = [C.POINTER(_Sobol)]
        for name in ("ro_sobol_unscrambled", "ro_sobol_scrambled"):
            getattr(L, name).restype = C.c_uint32
            getattr(L, name).argtypes = [C.POINTER(_Sobol), C.c_uint64, C.c_int]
        for name in ("ro_sobol_uniform", "ro_sobol_normal"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.POINTER(_Sobol), C.c_uint64, C.c_int]
        L.ro_halton_init.argtypes = [C.POINTER(_Halton), C.c_int, C.c_uint64, C.c_int]
        L.ro_halton_free.argtypes = [C.POINTER(_Halton)]
        L.ro_halton_digit_perm.restype = C.c_uint32
        L.ro_halton_digit_perm.argtypes = [C.POINTER(_Halton), C.c_int, C.c_uint32]
        for name in ("ro_halton_uniform", "ro_halton_normal"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.POINTER(_Halton), C.c_uint64, C.c_int]
        L.ro_brownian_increments.argtypes = [_D, C.c_int, C.c_double, _D]
        L.ro_corr_validate.argtypes = [C.c_int, _D, C.c_char_p, C.c_size_t]
        L.ro_min_eigenvalue.restype = C.c_double
        L.ro_min_eigenvalue.argtypes = [C.c_int, _D]
        L.ro_psd_clamp.argtypes = [C.c_int, _D, C.c_double, _D]
        L.ro_cholesky.argtypes = [C.c_int, _D, _D, C.c_char_p, C.c_size_t]
        L.ro_chol_apply.argtypes = [C.c_int, _D, _D, _D]
        L.ro_lv_calibrate.argtypes = [C.POINTER(_Grid), C.c_double, C.c_double, C.c_int,
                                      C.POINTER(_Lv), C.c_char_p, C.c_size_t]
        L.ro_lv_resample.argtypes = [C.POINTER(_Lv), C.c_int, _D, C.POINTER(_Lv)]
        L.ro_lv_free.argtypes = [C.POINTER(_Lv)]
        L.ro_lv_copy.argtypes = [C.c_int, C.c_int, _D, _D, _D, C.POINTER(_Lv)]
        L.ro_lv_sigma.restype = C.c_double
        L.ro_lv_sigma.argtypes = [C.POINTER(_Lv), C.c_double, C.c_double]
        L.ro_lv_row.argtypes = [C.POINTER(_Lv), C.c_double, _D]
        L.ro_interp_linear.restype = C.c_double
        L.ro_interp_linear.argtypes = [C.c_int, _D, _D, C.c_double]
        L.ro_validate.argtypes = [C.POINTER(_Request), C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]
        L.ro_price.argtypes = [C.POINTER(_Request), C.c_uint64, C.c_uint64, C.POINTER(_Result),
                               C.POINTER(PathRecord), C.POINTER(CouponRow), C.c_char_p, C.c_size_t,
                               C.c_char_p, C.c_size_t]
        _F32 = C.POINTER(C.c_float)
        L.ro_correlated_block.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int, _D, _F32]
        L.ro_gbm_evolve.argtypes = [_D, _D, C.c_double, _D, C.c_double, C.c_uint64, C.c_int64, C.c_int,
                                    _F32, _F32]
        L.ro_path_shocks.argtypes = [C.POINTER(_Request), C.c_uint64, _D]
        _U32 = C.POINTER(C.c_uint32)
        L.ro_chacha_block.argtypes = [_U32, C.c_int, _U32]
        L.ro_seed_from_u64.argtypes = [C.c_uint64, _U32]
        L.ro_rng_u64_at.restype = C.c_uint64
        L.ro_rng_u64_at.argtypes = [C.c_uint64, C.c_uint64]
        L.ro_zig_tables.argtypes = [_D, _D]
        L.ro_std_normals.restype = C.c_uint64
        L.ro_std_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, _D]
        L.ro_draw_starts.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
        L.ro_set_device_math.argtypes = [C.c_int]
        L.ro_set_device_math.restype = None
        L.ro_get_device_math.restype = C.c_int
        for name in ("ro_shared_exp", "ro_shared_log"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        _LIB = L
    return _LIB


# exp and log inside the path: False = glibc, the reference's own (the default);
# True = the device FP64 tier's table-driven pair (riga_flow_amd/csrc/exp_log.h),
# for comparing device paths bit for bit.  Applied per call, in the calling thread.
DEVICE_MATH = False


def _math(device_math=None):
    lib().ro_set_device_math(1 if (DEVICE_MATH if device_math is None else device_math) else 0)


def shared_exp(x: float) -> float:
    """The device FP64 tier's exp (exp_log.h), built on the host."""
    return lib().ro_shared_exp(x)


def shared_log(x: float) -> float:
    """The device FP64 tier's log (exp_log.h), built on the host."""
    return lib().ro_shared_log(x)


def _darr(values) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(values, dtype=np.float64))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


# ---------------------------------------------------------------- primitives
def probit(p: float) -> float:
    _math()
    return lib().ro_probit(p)


def radical_inverse(index: int, base: int) -> float:
    return lib().ro_radical_inverse(index, base)


def primitive_polynomials(count: int) -> List[int]:
    out = (C.c_uint32 * max(count, 1))()
    n = lib().ro_primitive_polynomials(count, out)
    return list(out[:n])


def direction_numbers(mask: int) -> List[int]:
    out = (C.c_uint32 * 32)()
    lib().ro_direction_numbers(mask, out)
    return list(out)


def mix64(z: int) -> int:
    return lib().ro_mix64(z)


def owen_scramble(x: int, dim: int, seed: int) -> int:
    return lib().ro_owen_scramble(x, dim, seed)


class Sobol:
    """``Sobol::new(dimension, seed)`` (qmc.rs:248-262); ``max_dim`` 512 = reference."""

    def __init__(self, dimension: int, seed: int, max_dim: int = 512):
        self._s = _Sobol()
        if lib().ro_sobol_init(C.byref(self._s), dimension, seed, max_dim):
            raise MemoryError("sobol init")

    @property
    def dimension(self) -> int:
        return self._s.dimension

    def direction_table(self) -> np.ndarray:
        n = self._s.dimension * 32
        return np.ctypeslib.as_array(self._s.v, shape=(n,)).reshape(self._s.dimension, 32).copy()

    def unscrambled(self, n: int, dim: int) -> int:
        return lib().ro_sobol_unscrambled(C.byref(self._s), n, dim)

    def scrambled(self, n: int, dim: int) -> int:
        return lib().ro_sobol_scrambled(C.byref(self._s), n, dim)

    def uniform(self, n: int, dim: int) -> float:
        return lib().ro_sobol_uniform(C.byref(self._s), n, dim)

    def normal(self, point: int, dim: int) -> float:
        _math()
        return lib().ro_sobol_normal(C.byref(self._s), point, dim)

    def __del__(self):
        try:
            lib().ro_sobol_free(C.byref(self._s))
        except Exception:
            pass


class Halton:
    """``Halton::new(dimension, seed)`` (qmc.rs:104-112)."""

    def __init__(self, dimension: int, seed: int, max_dim: int = 512):
        self._h = _Halton()
        if lib().ro_halton_init(C.byref(self._h), dimension, seed, max_dim):
            raise MemoryError("halton init")

    @property
    def dimension(self) -> int:
        return self._h.dimension

    def base(self, dim_index: int) -> int:
        return self._h.bases[dim_index]

    def perm(self, dim_index: int) -> List[int]:
        b = self._h.bases[dim_index]
        return [lib().ro_halton_digit_perm(C.byref(self._h), dim_index, d) for d in range(b)]

    def uniform(self, n: int, dim: int) -> float:
        return lib().ro_halton_uniform(C.byref(self._h), n, dim)

    def normal(self, point: int, dim: int) -> float:
        _math()
        return lib().ro_halton_normal(C.byref(self._h), point, dim)

    def __del__(self):
        try:
            lib().ro_halton_free(C.byref(self._h))
        except Exception:
            pass


def chacha_block(words: Sequence[int], double_rounds: int) -> List[int]:
    """The ChaCha block function on a 16-word input state (RFC 7539 section 2.3)."""
    a = (C.c_uint32 * 16)(*words)
    out = (C.c_uint32 * 16)()
    lib().ro_chacha_block(a, double_rounds, out)
    return list(out)


def seed_from_u64(seed: int) -> List[int]:
    """rand_core 0.9 ``seed_from_u64``: the eight key words (PCG32 expansion)."""
    out = (C.c_uint32 * 8)()
    lib().ro_seed_from_u64(seed, out)
    return list(out)


def rng_u64_at(seed: int, k: int) -> int:
    """The k-th u64 of ``ChaCha8Rng::seed_from_u64(seed)``."""
    return lib().ro_rng_u64_at(seed, k)


def zig_tables():
    x = np.zeros(257)
    f = np.zeros(257)
    lib().ro_zig_tables(_dptr(x), _dptr(f))
    return x, f


def std_normals(seed: int, n: int, start: int = 0):
    """n ``StandardNormal`` draws from u64 index ``start``; returns (draws, next index)."""
    out = np.zeros(max(n, 1))
    _math()
    end = lib().ro_std_normals(seed, start, n, _dptr(out))
    return out[:n], end


def draw_starts(seed: int, n: int, start: int = 0) -> np.ndarray:
    """u64 index at which each of n consecutive ``StandardNormal`` draws starts."""
    out = np.zeros(max(n, 1), dtype=np.uint64)
    _math()
    lib().ro_draw_starts(seed, start, n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out[:n]


def correlated_block(seed: int, paths: int, times: int, correlation) -> np.ndarray:
    """pathcube.rs:47-76 ``generate_correlated_block``: float32 [paths][times][factors]."""
    L = cholesky(correlation)
    F = L.shape[0]
    _math()
    out = np.zeros((paths, times, F), dtype=np.float32)
    lib().ro_correlated_block(seed, paths, times, F, _dptr(L), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def gbm_evolve(s0, vols, rate: float, dividend_yields, dt: float, block: np.ndarray) -> np.ndarray:
    """pathcube.rs:146-187 ``gbm_evolve``: float32 spot cube of the block's shape."""
    blk = np.ascontiguousarray(block, dtype=np.float32)
    P, T, F = blk.shape
    out = np.zeros_like(blk)
    lib().ro_gbm_evolve(_dptr(_darr(s0)), _dptr(_darr(vols)), rate, _dptr(_darr(dividend_yields)), dt,
                        P, T, F, blk.ctypes.data_as(C.POINTER(C.c_float)),
                        out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def brownian_increments(normals: Sequence[float], dt: float) -> np.ndarray:
    z = _darr(normals)
    out = np.zeros_like(z)
    lib().ro_brownian_increments(_dptr(z), len(z), dt, _dptr(out))
    return out


def min_eigenvalue(c) -> float:
    a = _darr(c)
    return lib().ro_min_eigenvalue(a.shape[0], _dptr(a))


def psd_clamp(c, floor: float) -> np.ndarray:
    a = _darr(c)
    out = np.zeros_like(a)
    lib().ro_psd_clamp(a.shape[0], _dptr(a), floor, _dptr(out))
    return out


def cholesky(c) -> np.ndarray:
    a = _darr(c)
    n = a.shape[0]
    out = np.zeros_like(a)
    err = C.create_string_buffer(256)
    if lib().ro_cholesky(n, _dptr(a), _dptr(out), err, 256):
        raise ValueError(err.value.decode())
    return out


def chol_apply(l, z) -> np.ndarray:
    L = _darr(l)
    zz = _darr(z)
    y = np.zeros_like(zz)
    lib().ro_chol_apply(L.shape[0], _dptr(L), _dptr(zz), _dptr(y))
    return y


# ---------------------------------------------------------------- local vol
class LocalVol:
    """Calibrated Dupire surface (``DupireLocalVol``), owned by the C library."""

    def __init__(self, lv: _Lv):
        self._lv = lv

    @classmethod
    def calibrate(cls, maturities, strikes, vols, spot: float, rate: float, grid_x: int = 96):
        m, k, v = _darr(maturities), _darr(strikes), _darr(vols)
        g = _Grid(len(m), len(k), _dptr(m), _dptr(k), _dptr(v))
        lv = _Lv()
        err = C.create_string_buffer(256)
        if lib().ro_lv_calibrate(C.byref(g), spot, rate, grid_x, C.byref(lv), err, 256):
            raise ValueError(err.value.decode())
        return cls(lv)

    @classmethod
    def from_arrays(cls, t_axis, log_s, v):
        """Copy caller arrays into a C-owned surface."""
        t, x, vv = _darr(t_axis), _darr(log_s), _darr(v)
        out = _Lv()
        if lib().ro_lv_copy(len(t), len(x), _dptr(t), _dptr(x), _dptr(vv), C.byref(out)):
            raise MemoryError("lv copy")
        return cls(out)

    def resample(self, times) -> "LocalVol":
        t = _darr(times)
        out = _Lv()
        lib().ro_lv_resample(C.byref(self._lv), len(t), _dptr(t), C.byref(out))
        return LocalVol(out)

    def sigma(self, t: float, s: float) -> float:
        return lib().ro_lv_sigma(C.byref(self._lv), t, s)

    def row(self, t: float) -> np.ndarray:
        out = np.zeros(self._lv.n_x)
        lib().ro_lv_row(C.byref(self._lv), t, _dptr(out))
        return out

    @property
    def t_axis(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._lv.t_axis, shape=(self._lv.n_t,)).copy()

    @property
    def log_s(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._lv.log_s, shape=(self._lv.n_x,)).copy()

    @property
    def v(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._lv.v, shape=(self._lv.n_t * self._lv.n_x,)).reshape(
            self._lv.n_t, self._lv.n_x).copy()

    def __del__(self):
        try:
            lib().ro_lv_free(C.byref(self._lv))
        except Exception:
            pass


def interp_linear(xs, ys, x: float) -> float:
    a, b = _darr(xs), _darr(ys)
    return lib().ro_interp_linear(len(a), _dptr(a), _dptr(b), x)


# ---------------------------------------------------------------- pricer
@dataclass
class OracleResult:
    pv: float
    mc_error: float
    ki_probability: float
    leg_coupon: float
    leg_memory_coupon: float
    leg_put: float
    leg_funding: float
    leg_total: float
    paths: int
    ko_count: int
    ki_count: int
    put_eligible_count: int
    coupon_path_count: int
    ko_mean: float
    ko_var: float
    ki_mean: float
    ki_var: float
    cpn_mean: float
    cpn_var: float
    funding_sum: float
    put_sum: float
    coupon_sum: float
    memory_sum: float
    trace: List[tuple]


_VR = {"none": 0, "halton": 1, "sobol": 2}
_TL = {"off": 0, "aggregated": 1, "sampled": 2, "full": 3}


class OracleError(Exception):
    def __init__(self, code: str, message: str):
        super().__init__(f"{code}: {message}")
        self.code = code
        self.message = message


def local_vols_for(request) -> Optional[List[LocalVol]]:
    """``build_local_vol`` (riga-engine/src/lib.rs:326-343), or pre-calibrated grids."""
    spec = request.local_vol
    if spec is None:
        return None
    if spec.calibrated is not None:
        return [LocalVol.from_arrays(g.t_axis, g.log_s, g.v) for g in spec.calibrated]
    out = []
    for f, g in enumerate(spec.grids):
        out.append(LocalVol.calibrate(g.maturities_years, g.strikes, g.vols,
                                      request.market.spots[f], request.market.risk_free_rate,
                                      spec.grid_x))
    return out


def _pack(request, max_dim: int, lvs: Optional[List[LocalVol]]):
    t, m, c = request.trade, request.market, request.config
    F = len(m.spots)
    keep = {}

    def arr(name, v):
        keep[name] = _darr(v)
        return _dptr(keep[name])

    r = _Request()
    r.factors = F
    r.spots = arr("spots", m.spots)
    r.vols = arr("vols", m.vols)
    r.dividend_yields = arr("q", m.dividend_yields)
    r.correlation = arr("corr", np.asarray(m.correlation, dtype=np.float64).reshape(-1))
    r.rate = m.risk_free_rate
    r.initial_spots = arr("s0", t.initial_spots)
    r.maturity_years = t.maturity_years
    r.steps = t.steps
    r.strike = t.strike
    r.knock_in_barrier = t.knock_in_barrier
    r.has_ko = t.knock_out_barrier is not None
    r.knock_out_barrier = t.knock_out_barrier if t.knock_out_barrier is not None else math.inf
    r.has_cb = t.coupon_barrier is not None
    r.coupon_barrier = t.coupon_barrier if t.coupon_barrier is not None else math.inf
    r.has_cr = t.coupon_rate is not None
    r.coupon_rate = t.coupon_rate if t.coupon_rate is not None else 0.0
    r.memory_coupon = bool(t.memory_coupon)
    r.coupon_periods = -1 if t.coupon_periods is None else t.coupon_periods
    r.fixed_first_coupon = bool(t.fixed_first_coupon)
    r.call_start_step = -1 if t.call_start_step is None else t.call_start_step
    r.paths = c.paths
    r.path_block_size = c.path_block_size
    r.seed = c.seed
    r.trace_level = _TL[c.trace_level]
    r.vr_kind = _VR[c.variance_reduction.kind]
    r.vr_dimension = c.variance_reduction.dimension
    r.max_dim = max_dim
    if lvs is not None:
        arrs = (_Lv * F)(*[lv._lv for lv in lvs])
        keep["lv"] = arrs
        keep["lvobjs"] = lvs
        r.lv = C.cast(arrs, C.POINTER(_Lv))
    return r, keep


def price(request, max_dim: int = 512, path_range=None, records: bool = False,
          lvs: Optional[List[LocalVol]] = None, ledger: bool = False, device_math=None):
    """``price_full_phoenix`` (note.rs:38-373) on the QMC lanes and the ChaCha8
    Monte Carlo lane (``variance_reduction.kind == "none"``).  ``device_math``
    chooses the path's exp/log (None: the module's ``DEVICE_MATH``).

    Returns ``OracleResult`` (and a numpy record array when ``records``, and the
    coupon strips ``[path][period]`` when ``ledger``, explain.rs:146-262)."""
    if lvs is None:
        lvs = local_vols_for(request)
    r, keep = _pack(request, max_dim, lvs)
    b, e = (0, request.config.paths) if path_range is None else path_range
    res = _Result()
    recs = (PathRecord * max(e - b, 1))() if (records or ledger) else None
    periods = 9 if request.trade.coupon_periods is None else max(1, request.trade.coupon_periods)
    rows = (CouponRow * max((e - b) * periods, 1))() if ledger else None
    code = C.create_string_buffer(64)
    msg = C.create_string_buffer(256)
    _math(device_math)
    rc = lib().ro_price(C.byref(r), b, e, C.byref(res), recs, rows, code, 64, msg, 256)
    if rc:
        raise OracleError(code.value.decode(), msg.value.decode())
    out = OracleResult(
        **{f: getattr(res, f) for f, _ in _Result._fields_ if f not in ("n_trace", "trace_path", "trace_step")},
        trace=[(res.trace_path[i], res.trace_step[i]) for i in range(res.n_trace)],
    )
    if ledger:
        arr = np.frombuffer(recs, dtype=RECORD_DTYPE, count=e - b).copy()
        led = np.frombuffer(rows, dtype=COUPON_DTYPE, count=(e - b) * periods).copy()
        return out, arr, led.reshape(e - b, periods)
    if records:
        arr = np.frombuffer(recs, dtype=RECORD_DTYPE, count=e - b).copy()
        return out, arr
    return out


def validate(request, max_dim: int = 512):
    r, keep = _pack(request, max_dim, None)
    code = C.create_string_buffer(64)
    msg = C.create_string_buffer(256)
    if lib().ro_validate(C.byref(r), code, 64, msg, 256):
        raise OracleError(code.value.decode(), msg.value.decode())


def path_shocks(request, path_id: int, max_dim: int = 512) -> np.ndarray:
    r, keep = _pack(request, max_dim, None)
    F, T = request.factors, request.trade.steps
    _math()
    out = np.zeros(F * T)
    if lib().ro_path_shocks(C.byref(r), path_id, _dptr(out)):
        raise OracleError("invalid_market", "path_shocks failed")
    return out.reshape(F, T)
