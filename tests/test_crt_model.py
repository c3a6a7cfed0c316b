"""CPU checks of the Ozaki scheme II (CRT) arithmetic, against exact rational arithmetic.

The device path (TMM_OPT_OZAKI_MODULI) follows crt_model.py step by step; these tests pin the
claims the kernels rely on: the moduli are pairwise coprime and their balanced residues fit int8,
the bit budget keeps |A'B'| inside the reconstructible range, the CRT sum's constants are exact and
it rebuilds the exact integer (exact when representable, otherwise correctly rounded on the
samples and within one unit in the last place), and the emulated product is within the
scaling/rounding error bound of the exact product.
"""
import math
from fractions import Fraction
from itertools import combinations

import numpy as np
import pytest

import crt_model as cm


def test_moduli_pairwise_coprime_and_int8():
    for a, b in combinations(cm.MODULI, 2):
        assert math.gcd(a, b) == 1, (a, b)
    for m in cm.MODULI:
        lo, hi = min(cm.balanced(r, m) for r in range(m)), max(cm.balanced(r, m) for r in range(m))
        assert -128 <= lo and hi <= 127 and hi - lo == m - 1
        # the device bias in the residue epilogue and Garner steps
        assert m >= 167 and (m << 23) >= 2 ** 30 and (m << 23) < 2 ** 32


@pytest.mark.parametrize("nmod", range(6, 21))
@pytest.mark.parametrize("k", [1, 2, 63, 64, 1000, 8000, 40000, 65536])
def test_bit_budget_stays_in_range(nmod, k):
    bits = cm.bits_for(nmod, k)
    M = math.prod(cm.MODULI[:nmod])
    assert 1 <= bits <= 62
    # worst case |sum_k a' b'| <= k 2^(2 bits) must stay below M/4 (balanced range is ~ +-M/2)
    assert k * 2 ** (2 * bits) < M // 4
    # and the budget is tight: with k rounded up to a power of two, one more bit per operand
    # would break it (below the 62-bit cap)
    if bits < 62:
        assert 2 ** math.ceil(math.log2(k)) * 2 ** (2 * bits + 2) >= M // 4


def test_int32_products_exact():
    # per-modulus int8 products: |sum| <= k * 128^2 < 2^30 for k <= MAX_K
    assert cm.MAX_K * 128 * 128 <= 2 ** 30


@pytest.mark.parametrize("nmod", [6, 9, 14, 16, 17, 20])
def test_crt_sum_table(nmod):
    """The constants of the device's CRT sum: W_t = (M/m_t) y_t is 1 modulo m_t and 0 modulo every
    other modulus, its limbs (and M's) are exact 26-bit integers, and f_t is y_t / m_t rounded."""
    f, w, mm, L = cm.direct_table(nmod)
    mods = cm.MODULI[:nmod]
    M = math.prod(mods)
    assert L * cm.LIMB >= M.bit_length() and sum(int(x) << (cm.LIMB * i) for i, x in enumerate(mm)) == M
    for t, m in enumerate(mods):
        assert all(x == int(x) and 0 <= x < 2 ** cm.LIMB for x in w[t])
        W = sum(int(x) << (cm.LIMB * i) for i, x in enumerate(w[t]))
        assert 0 < W < M
        for s, ms in enumerate(mods):
            assert W % ms == (1 if s == t else 0), (t, s)
        assert f[t] == float(Fraction(W, M))  # = y_t / m_t correctly rounded


@pytest.mark.parametrize("nmod", [6, 9, 14, 16, 17, 20])
def test_crt_sum_rebuilds_exact_integers(nmod):
    """The device's CRT sum (crt_direct) rebuilds every |x| < M/4 from its residues: exactly when x
    is representable in FP64, correctly rounded on the samples otherwise, and its limb sums stay
    exact integers (below 2^40)."""
    rng = np.random.default_rng(nmod)
    M = math.prod(cm.MODULI[:nmod])
    q4 = M // 4
    samples = [0, 1, -1, q4 - 1, -(q4 - 1), 123456789, -(2 ** 40 + 17), 2 ** 53 + 1, 2 ** 54 + 2, -(2 ** 60 + 2 ** 7)]
    samples += [int(x) for x in rng.integers(-2 ** 62, 2 ** 62, 50)]
    samples += [int.from_bytes(rng.bytes(24), "little") % (2 * q4 - 1) - (q4 - 1) for _ in range(100)]
    # representable values: 53-bit odd significands at every scale
    samples += [int(s) << e for s in rng.integers(2 ** 52, 2 ** 53, 20) for e in (0, 30, 60)]
    for x in samples:
        if abs(x) >= q4:
            continue
        res = [x % m for m in cm.MODULI[:nmod]]
        y = cm.crt_direct(res, nmod)
        assert y == float(x), (x, y)
        # the limb sums before the carries: sum_t r_t W_ti - q M_i, integers below 2^40
        f, w, mm, L = cm.direct_table(nmod)
        q = round(sum(Fraction(r) * Fraction(ft) for r, ft in zip(res, f)))
        for i in range(L):
            S = sum(r * int(w[t][i]) for t, r in enumerate(res)) - q * int(mm[i])
            assert abs(S) < 2 ** 40


def test_crt_sum_near_midpoint_within_one_ulp():
    """The documented limit of the final double-double step: a value within 2^-53 ulp of a rounding
    midpoint (here 2^106 + 2^53 + 1, just above the midpoint 2^106 + 2^53) can come out one unit
    in the last place off; everything else in this module is exact or correctly rounded."""
    for x in (2 ** 106 + 2 ** 53 + 1, -(2 ** 106 + 2 ** 53 + 1)):
        y = cm.crt_direct([x % m for m in cm.MODULI[:16]], 16)
        assert abs(Fraction(y) - x) <= Fraction(math.ulp(float(x)))


@pytest.mark.parametrize("nmod,shape", [(8, (3, 4, 5)), (12, (4, 3, 9)), (16, (5, 5, 7)), (20, (2, 3, 4))])
def test_emulated_product_within_rounding_bound(nmod, shape):
    m, n, k = shape
    rng = np.random.default_rng(sum(shape) + nmod)
    A = rng.uniform(-1, 1, (m, k)) * np.exp(2.0 * rng.standard_normal((m, k)))
    B = rng.uniform(-1, 1, (k, n)) * np.exp(2.0 * rng.standard_normal((k, n)))
    A[0, 0] = 0.0
    C, exact, bits, ea, eb = cm.emulate(A, B, nmod)
    for i in range(m):
        for j in range(n):
            true = sum(Fraction(float(A[i, kk])) * Fraction(float(B[kk, j])) for kk in range(k))
            # truncation: each a, b off by at most half a unit of 2^(e - bits)
            bound = Fraction(0)
            for kk in range(k):
                bound += Fraction(2) ** (ea[i] - bits - 1) * abs(Fraction(float(B[kk, j])))
                bound += Fraction(2) ** (eb[j] - bits - 1) * abs(Fraction(float(A[i, kk])))
                bound += Fraction(2) ** (ea[i] + eb[j] - 2 * bits - 2)
            # reconstruction: the scaled integer product, correctly rounded
            scaled = Fraction(exact[i][j]) * Fraction(2) ** (ea[i] + eb[j] - 2 * bits)
            assert C[i, j] == float(scaled)
            assert abs(scaled - true) <= bound


def test_integers_are_exact():
    rng = np.random.default_rng(5)
    A = rng.integers(-9, 10, (4, 6)).astype(float)
    B = rng.integers(-9, 10, (6, 3)).astype(float)
    C, *_ = cm.emulate(A, B, 8)
    np.testing.assert_array_equal(C, A @ B)
