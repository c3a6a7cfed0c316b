"""Exact-integer model of the Ozaki scheme II (CRT) DGEMM emulation (TMM_OPT_OZAKI_MODULI).

Mirrors the device algorithm step by step with Python integers, so that its arithmetic claims can be
checked against exact rational arithmetic on the CPU:
  * scaling: a'(i,k) = sign(a) * round(|a| * 2^(bits - ea[i])) with |a(i,k)| < 2^ea[i];
  * residues: balanced residues of a', b' modulo pairwise coprime moduli <= 256 (int8 range);
  * one integer product per modulus, reduced to [0, m_t);
  * Garner's mixed-radix digits (balanced) and their double-double Horner evaluation.
"""
import math
from fractions import Fraction

import numpy as np

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191, 181, 179, 173, 167]
MAX_K = 65536


def bits_for(nmod, k):
    l = sum(math.log2(m) for m in MODULI[:nmod])
    lk = math.ceil(math.log2(max(k, 1)))
    return max(1, min(62, math.floor((l - 2.0 - lk) / 2.0)))


def frexp_exp(x):
    """e with |x| < 2^e (frexp convention); None for 0."""
    if x == 0:
        return None
    return math.frexp(x)[1]


def row_exponents(X):
    out = []
    for row in X:
        es = [frexp_exp(float(v)) for v in row if v != 0]
        out.append(max(es) if es else None)
    return out


def scale_round(x, bits, e):
    """sign(x) * round-half-even(|x| * 2^(bits - e)) as a Python int (the device uses rint)."""
    if e is None or x == 0:
        return 0
    v = Fraction(abs(float(x))) * Fraction(2) ** (bits - e)
    r = round(v)  # Fraction rounds half to even, like rint
    return -r if x < 0 else r


def balanced(r, m):
    r %= m
    return r - m if 2 * r >= m else r


def garner(residues, nmod):
    """Balanced mixed-radix digits of the integer with the given residues (each in [0, m_t))."""
    v = []
    for t in range(nmod):
        m = MODULI[t]
        x = residues[t]
        for s in range(t):
            x = ((x - v[s]) * pow(MODULI[s], -1, m)) % m
        v.append(balanced(x, m))
    return v


def fma(a, b, c):
    """a*b + c with one rounding (IEEE fused multiply-add)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def garner_sum(residues, nmod):
    """The same digits in the form the device computes them: v_t = (r_t - sum_{s<t} v_s P_s) P_t^-1
    mod m_t with P_s = m_0 ... m_(s-1), reduced once per digit."""
    v = []
    for t in range(nmod):
        m = MODULI[t]
        P = [math.prod(MODULI[:s]) % m for s in range(t + 1)]
        a = residues[t] - sum(v[s] * P[s] for s in range(t))
        v.append(balanced(a * pow(P[t], -1, m), m))
    return v


def horner(v):
    """Double-double evaluation from the top digit down, operation for operation as the device:
    y <- y * m_t + v_t with the product split exactly by an fma and the sum by TwoSum."""
    yh, yl = float(v[-1]), 0.0
    for t in range(len(v) - 2, -1, -1):
        mt = float(MODULI[t])
        p = yh * mt
        pe = fma(yh, mt, -p)
        q = fma(yl, mt, pe + float(v[t]))
        s = p + q
        bb = s - p
        e = (p - (s - bb)) + (q - bb)
        yh, yl = s, e
    return yh + yl


def emulate(A, B, nmod):
    """C = A @ B through the scheme (A: m x k, B: k x n, float64 arrays). Returns (C, exact scaled
    integer product, bits, ea, eb)."""
    m, k = A.shape
    n = B.shape[1]
    bits = bits_for(nmod, k)
    ea = row_exponents(A)
    eb = row_exponents(B.T)
    Ai = [[scale_round(A[i, kk], bits, ea[i]) for kk in range(k)] for i in range(m)]
    Bi = [[scale_round(B[kk, j], bits, eb[j]) for j in range(n)] for kk in range(k)]
    exact = [[sum(Ai[i][kk] * Bi[kk][j] for kk in range(k)) for j in range(n)] for i in range(m)]
    C = np.zeros((m, n))
    for i in range(m):
        for j in range(n):
            # per-modulus products of balanced residues, reduced to [0, m_t)
            res = []
            for t in range(nmod):
                mt = MODULI[t]
                s = sum(balanced(Ai[i][kk], mt) * balanced(Bi[kk][j], mt) for kk in range(k))
                res.append(s % mt)
            v = garner(res, nmod)
            y = horner(v)
            if ea[i] is None or eb[j] is None:
                C[i, j] = 0.0
            else:
                C[i, j] = math.ldexp(y, ea[i] + eb[j] - 2 * bits)
    return C, exact, bits, ea, eb
