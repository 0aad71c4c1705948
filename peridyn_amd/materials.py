"""Host-side material and contact constants, as the reference constructors
compute them before anything reaches the hot path.

  make_matrix            src/misc/utils.jl:12-25
  bond stiffness         src/materials/bond_based.jl:58   18 * K ./ pi ./ horizon^4
  EP yield value psi     src/materials/elasto_plastic.jl:59-71 (VonMises, :364-366)
  short-range stiffness  src/contacts/shortrange.jl:25    18K/(pi*horizon^4) * mfactor
  particle size          src/materials/general_material.jl:83  volume[1]^(1/3)
  contact distances      src/contacts/contacts.jl:76-91 (ContactModel12_gcal)

`horizon^n` with an integer n is Julia's `^(::Float64, ::Integer)`
(base/math.jl pow_body): power by squaring carried in double-double with fused
multiply-adds.  Python 3.10 has no math.fma, so the fused operations are
evaluated exactly with Fractions and rounded once.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np


def _fma(a: float, b: float, c: float) -> float:
    if not (math.isfinite(a) and math.isfinite(b) and math.isfinite(c)):
        return a * b + c
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def julia_pow_int(x: float, n: int) -> float:
    x = float(x)
    if n == 0:
        return 1.0
    if n == 3:
        return x * x * x
    y, xnlo, ynlo = 1.0, 0.0, 0.0
    if n < 0:
        rx = 1.0 / x
        if n == -2:
            return rx * rx
        if math.isfinite(x):
            xnlo = -_fma(x, rx, -1.0) * rx
        x, n = rx, -n
    while n > 1:
        if n & 1:
            err = _fma(y, xnlo, x * ynlo)
            p = x * y
            plo = _fma(x, y, -p)
            y, ynlo = p, plo + err
        err = x * 2.0 * xnlo
        p = x * x
        plo = _fma(x, x, -p)
        x, xnlo = p, plo + err
        n >>= 1
    err = _fma(y, xnlo, x * ynlo)
    return _fma(x, y, err) if (math.isfinite(x) and math.isfinite(err)) else x * y


def make_matrix(S) -> np.ndarray:
    S = np.asarray(S, dtype=np.float64)
    if S.ndim == 2:
        return S.copy()
    S = S.ravel()
    n = S.size
    N = int(math.sqrt(2 * n + 0.5 ** 2) - 0.5)
    if N * (N + 1) // 2 != n:
        raise ValueError("packed parameter vector must hold N(N+1)/2 entries")
    M = np.zeros((N, N))
    a = 0
    for i in range(N):
        for j in range(i, N):
            M[i, j] = S[a]
            M[j, i] = S[a]
            a += 1
    return M


def bond_stiffness(K: np.ndarray, horizon: float) -> np.ndarray:
    return ((18.0 * np.asarray(K, dtype=np.float64)) / math.pi) / julia_pow_int(horizon, 4)


def ep_psi(sigma_y: np.ndarray, horizon: float) -> np.ndarray:
    s = np.asarray(sigma_y, dtype=np.float64)
    se2 = ((s * s + 0.0 * 0.0) + s * s) / 2.0
    return ((75.0 / (8.0 * math.pi)) * se2) / julia_pow_int(horizon, 5)


def shortrange_stiffness(K: float, horizon: float, mfactor: float = 1.0) -> float:
    return ((18.0 * K) / (math.pi * julia_pow_int(horizon, 4))) * mfactor


def particle_size(vol) -> float:
    v = float(np.asarray(vol).ravel()[0])
    r = round(v ** (1.0 / 3.0))
    if r > 0 and float(r) ** 3 == v:
        return float(r)
    return math.pow(v, 1.0 / 3.0)
