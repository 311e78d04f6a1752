"""Pins for the oracle's QMC ladder (qmc.rs restated in oracle/riga_oracle.c).

Each check is independent of the oracle's own formula: textbook tables,
published generator outputs, scipy, net/stratification theorems, closed-form
covariances, or the reference's own unit-test assertions.
"""
import math

import numpy as np
import pytest
from scipy.special import ndtri

import oracle as O

# Primitive polynomials over GF(2), degrees 2..6, in increasing value (e.g. Joe-Kuo's list).
TEXTBOOK_POLYS = [0b111, 0b1011, 0b1101, 0b10011, 0b11001,
                  37, 41, 47, 55, 59, 61,
                  67, 91, 97, 103, 109, 115]


def euler_phi(n):
    out, m, p = n, n, 2
    while p * p <= m:
        if m % p == 0:
            while m % p == 0:
                m //= p
            out -= out // p
        p += 1
    if m > 1:
        out -= out // m
    return out


def test_primitive_polynomials_textbook_prefix():
    assert O.primitive_polynomials(len(TEXTBOOK_POLYS)) == TEXTBOOK_POLYS


def test_primitive_polynomial_counts_per_degree():
    # the number of primitive polynomials of degree d is phi(2^d - 1) / d
    want = [euler_phi(2 ** d - 1) // d for d in range(2, 12)]
    polys = O.primitive_polynomials(sum(want))
    degrees = [p.bit_length() - 1 for p in polys]
    assert [degrees.count(d) for d in range(2, 12)] == want
    assert degrees == sorted(degrees)


def test_direction_numbers_are_odd_and_bounded():
    # Joe-Kuo: m_k is odd and m_k < 2^k; v_k = m_k << (32-k)
    for mask in O.primitive_polynomials(60):
        v = O.direction_numbers(mask)
        for k in range(1, 33):
            m = v[k - 1] >> (32 - k)
            assert m % 2 == 1 and m < 2 ** k
            assert v[k - 1] & ((1 << (32 - k)) - 1) == 0


def bitrev32(x):
    return int(f"{x:032b}"[::-1], 2)


def test_dimension_zero_is_van_der_corput_in_gray_order():
    s = O.Sobol(4, 0)
    for n in range(1, 300):
        assert s.unscrambled(n, 0) == bitrev32(n ^ (n >> 1))
        # and the radical inverse in base 2 of gray(n) (qmc.rs:502-510)
        assert s.unscrambled(n, 0) / 2 ** 32 == O.radical_inverse(n ^ (n >> 1), 2)


def degree_of_dim(d):
    # dim 0 is the identity (degree 1); dim d >= 1 uses the d-th primitive polynomial
    return 1 if d == 0 else O.primitive_polynomials(d)[d - 1].bit_length() - 1


@pytest.mark.parametrize("dims", [(0, 1), (1, 2), (0, 3), (3, 7)])
def test_sobol_pairs_are_nets_with_the_theoretical_t(dims):
    # Sobol's theorem: dims with polynomial degrees s_a, s_b form a (t,m,2)-net
    # with t = (s_a - 1) + (s_b - 1): every elementary box of volume 2^(t-m)
    # holds exactly 2^t points of each aligned block of 2^m points.
    s = O.Sobol(64, 0)
    a, b = dims
    t = degree_of_dim(a) - 1 + degree_of_dim(b) - 1
    m = 8
    for block in range(3):
        pts = [(s.unscrambled(n, a), s.unscrambled(n, b))
               for n in range(block * 2 ** m, (block + 1) * 2 ** m)]
        for i in range(m - t + 1):
            j = m - t - i
            counts = {}
            for x, y in pts:
                key = (x >> (32 - i) if i else 0, y >> (32 - j) if j else 0)
                counts[key] = counts.get(key, 0) + 1
            assert len(counts) == 2 ** (m - t) and set(counts.values()) == {2 ** t}
        for c in (0, 1):  # each coordinate alone is a (0,m,1)-net
            assert len({p[c] >> (32 - m) for p in pts}) == 2 ** m


def test_owen_scramble_is_a_nested_bijection():
    seed, dim = 1729, 37
    # top-k bits of the output depend only on the top-k bits of the input and
    # the induced map on them is a permutation, for every k
    xs = np.random.default_rng(0).integers(0, 2 ** 32, size=4000, dtype=np.uint64)
    outs = [O.owen_scramble(int(x), dim, seed) for x in xs]
    for k in (1, 4, 8, 12):
        f = {}
        for x, y in zip(xs, outs):
            f.setdefault(int(x) >> (32 - k), set()).add(y >> (32 - k))
        assert all(len(v) == 1 for v in f.values())
    top8 = {}
    for t in range(256):
        top8[t] = O.owen_scramble(t << 24, dim, seed) >> 24
    assert sorted(top8.values()) == list(range(256))


def test_owen_keeps_stratification_of_aligned_blocks():
    s = O.Sobol(16, 99)
    m = 10
    for d in (0, 5, 15):
        vals = [s.scrambled(n, d) >> (32 - m) for n in range(2 ** m, 2 ** (m + 1))]
        assert len(set(vals)) == 2 ** m


def test_mix64_is_the_splitmix64_finaliser():
    # splitmix64 seeded with 0: published first outputs
    x, outs = 0, []
    for _ in range(3):
        x = (x + 0x9E3779B97F4A7C15) % 2 ** 64
        outs.append(O.mix64(x))
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_sobol_reference_unit_test():
    # qmc.rs:479-500: reproducible, unbiased per dimension, seed-decorrelated
    a, b, c = O.Sobol(8, 42), O.Sobol(8, 42), O.Sobol(8, 43)
    for point in range(1, 65):
        for d in range(8):
            assert a.normal(point, d) == b.normal(point, d)
    for d in range(8):
        mean = sum(a.uniform(p, d) for p in range(1, 4097)) / 4096
        assert abs(mean - 0.5) < 0.01
    assert any(a.normal(1, d) != c.normal(1, d) for d in range(8))


def test_dimension_cap_wraps_direction_rows_but_not_the_scramble():
    # qmc.rs:249,271: rows wrap modulo the clamped dimension; owen uses the raw dim
    s = O.Sobol(2000, 5)  # clamped to 512
    assert s.dimension == 512
    for n in (1, 77, 4096):
        assert s.unscrambled(n, 600) == s.unscrambled(n, 600 - 512)
        assert s.scrambled(n, 600) != s.scrambled(n, 600 - 512) or n == 0


def test_probit_against_scipy_within_acklam_bound():
    ps = np.concatenate([np.geomspace(1e-12, 0.02425, 500), np.linspace(0.02425, 0.97575, 2001),
                         1.0 - np.geomspace(1e-12, 0.02425, 500)])
    for p in ps:
        want = ndtri(p)
        assert abs(O.probit(p) - want) <= 1.16e-9 * max(abs(want), 1e-9) + 1e-15
    assert abs(O.probit(0.5)) < 1e-9  # qmc.rs:512-517
    assert abs(O.probit(0.975) - 1.959963984540054) < 1e-6
    assert abs(O.probit(0.025) + 1.959963984540054) < 1e-6
    assert O.probit(0.0) == -math.inf and O.probit(1.0) == math.inf


def test_radical_inverse_reference_values():
    # qmc.rs:502-510
    assert [O.radical_inverse(i, 2) for i in range(1, 5)] == [0.5, 0.25, 0.75, 0.125]
    assert abs(O.radical_inverse(1, 3) - 1 / 3) < 1e-15
    assert abs(O.radical_inverse(2, 3) - 2 / 3) < 1e-15


def test_halton_digit_permutations_and_first_digit():
    h = O.Halton(20, 42)
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71]
    for d in range(20):
        assert h.base(d) == primes[d]
        perm = h.perm(d)
        assert sorted(perm) == list(range(primes[d]))
        for n in range(1, primes[d]):  # one digit: u = perm[n] / base
            assert h.uniform(n, d) == perm[n] * (1.0 / primes[d])
    h2 = O.Halton(20, 43)
    assert any(h.perm(d) != h2.perm(d) for d in range(2, 20))
    # qmc.rs:519-533
    for point in range(1, 50):
        for d in range(8):
            assert math.isfinite(h.normal(point, d))


def bridge_matrix(T, dt):
    cols = []
    for i in range(T):
        e = np.zeros(T)
        e[i] = 1.0
        cols.append(np.cumsum(O.brownian_increments(e, dt)))
    return np.array(cols).T  # W = B z


@pytest.mark.parametrize("T", [1, 2, 4, 64, 256])
def test_bridge_is_a_brownian_motion_for_dyadic_steps(T):
    # with iid normals, W_t must have covariance min(t_i, t_j) dt exactly when
    # every split is a true midpoint (T a power of two)
    dt = 1.0 / 252
    B = bridge_matrix(T, dt)
    t = np.arange(1, T + 1)
    np.testing.assert_allclose(B @ B.T, np.minimum.outer(t, t) * dt, rtol=0, atol=1e-13)


def test_bridge_uses_half_half_mean_at_uneven_splits():
    # qmc.rs:184-187 conditions on 0.5 (W_lo + W_hi) even when mid is not the
    # time midpoint (T = 3 splits at 1); worked by hand:
    #   W3 = sqrt(3) z0, W1 = W3/2 + sqrt(2/3) z1, W2 = (W1 + W3)/2 + sqrt(1/2) z2
    B = bridge_matrix(3, 1.0)
    r3, a, c = math.sqrt(3.0), math.sqrt(2.0 / 3.0), math.sqrt(0.5)
    want = np.array([[r3 / 2, a, 0.0],
                     [0.75 * r3, a / 2, c],
                     [r3, 0.0, 0.0]])
    np.testing.assert_allclose(B, want, rtol=0, atol=1e-15)


def test_bridge_preorder_node_numbering():
    # T = 4: z0 terminal; z1 -> W2 (root midpoint); z2 -> W1 (left); z3 -> W3 (right)
    dt = 1.0
    B = bridge_matrix(4, dt)
    assert np.allclose(B[:, 0], [0.5, 1.0, 1.5, 2.0])          # terminal sqrt(4) * t/4
    assert np.allclose(B[:, 1], [0.5, 1.0, 0.5, 0.0])          # W2 = sqrt(1), W1,W3 half of it
    assert np.allclose(B[:, 2], [math.sqrt(0.5), 0, 0, 0])      # only W1 moves
    assert np.allclose(B[:, 3], [0, 0, math.sqrt(0.5), 0])      # only W3 moves


def test_bridge_variance_reference_test():
    # qmc.rs:436-477 (statistical)
    n, dt, reps = 64, 1.0 / 252, 4000
    rng = np.random.default_rng(12345)
    inc2, term = 0.0, []
    for _ in range(reps):
        z = rng.standard_normal(n)
        inc = O.brownian_increments(z, dt)
        inc2 += float(inc @ inc)
        term.append(inc.sum())
    assert abs(inc2 / (reps * n) - dt) < 0.05 * dt
    assert abs(np.mean(np.square(term)) - n * dt) < 0.1 * n * dt
