"""Exact-integer model of the Ozaki scheme II (CRT) DGEMM emulation (TMM_OPT_OZAKI_MODULI).

Mirrors the device algorithm step by step with Python integers, so that its arithmetic claims can be
checked against exact rational arithmetic on the CPU:
  * scaling: a'(i,k) = sign(a) * round(|a| * 2^(bits - ea[i])) with |a(i,k)| < 2^ea[i];
  * residues: balanced residues of a', b' modulo pairwise coprime moduli <= 256 (int8 range);
  * one integer product per modulus, reduced to [0, m_t);
  * the CRT sum Y = sum_t r_t W_t - q M, evaluated in 26-bit FP64 limbs as the device does
    (crt_direct).
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


def fma(a, b, c):
    """a*b + c with one rounding (IEEE fused multiply-add)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


LIMB = 26


def direct_table(nmod):
    """Constants of the direct CRT sum for the first nmod moduli: f_t = y_t / m_t rounded to FP64,
    the 26-bit limbs of W_t = (M / m_t) y_t with y_t = (M / m_t)^-1 mod m_t, and of M."""
    mods = MODULI[:nmod]
    M = math.prod(mods)
    L = (8 * nmod + LIMB - 1) // LIMB
    f, w = [], []
    for m in mods:
        Mt = M // m
        y = pow(Mt % m, -1, m)
        f.append(float(Fraction(y, m)))
        w.append([float(((Mt * y) >> (LIMB * i)) & ((1 << LIMB) - 1)) for i in range(L)])
    mm = [float((M >> (LIMB * i)) & ((1 << LIMB) - 1)) for i in range(L)]
    assert sum(int(x) << (LIMB * i) for i, x in enumerate(mm)) == M
    return f, w, mm, L


def _carry(S):
    """S_i -> [0, 2^26) for every limb but the top, carries upward (each step exact)."""
    for i in range(len(S) - 1):
        c = math.floor(S[i] * 2.0 ** -LIMB)
        S[i] = fma(-c, 2.0 ** LIMB, S[i])
        S[i + 1] = S[i + 1] + c
    return S


def crt_direct(residues, nmod):
    """FP64 value of the integer Y in (-M/4, M/4) with Y = residues[t] mod m_t, operation for
    operation as the device: q = rint(sum r_t f_t); limb sums S_i = sum_t r_t W_ti - q M_i (exact
    integers below 2^40); carries; sign; limb pairs below 2^52; top-down sum in double-double."""
    f, w, mm, L = direct_table(nmod)
    qs = 0.0
    for t in range(nmod):
        qs = fma(float(residues[t]), f[t], qs)
    q = float(round(qs))  # rint: |qs - integer| <= 1/4 + 2^-35, never a tie
    S = [-q * mm[i] for i in range(L)]
    for t in range(nmod):
        r = float(residues[t])
        for i in range(L):
            S[i] = fma(r, w[t][i], S[i])
    _carry(S)
    sgn = -1.0 if S[L - 1] < 0 else 1.0
    S = _carry([sgn * s for s in S])
    P = [fma(S[2 * j + 1], 2.0 ** LIMB, S[2 * j]) if 2 * j + 1 < L else S[2 * j] for j in range((L + 1) // 2)]
    h, lo = math.ldexp(P[-1], 2 * LIMB * (len(P) - 1)), 0.0
    for j in range(len(P) - 2, -1, -1):
        T = math.ldexp(P[j], 2 * LIMB * j)
        s = h + T
        bb = s - h
        e = (h - (s - bb)) + (T - bb)
        h = s
        lo = lo + e
    return sgn * (h + lo)


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
            y = crt_direct(res, nmod)
            if ea[i] is None or eb[j] is None:
                C[i, j] = 0.0
            else:
                C[i, j] = math.ldexp(y, ea[i] + eb[j] - 2 * bits)
    return C, exact, bits, ea, eb
