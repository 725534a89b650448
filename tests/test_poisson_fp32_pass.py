"""Error bounds of the path kernels' FP32 passes over two branches of Giles'
inverse Poisson CDF: the tail branch and the Temme correction.

Giles' inverse Poisson CDF (reference math/Numerics.hpp:49-272) takes, for
lam > 4 and |W| >= 3 (W = Phi^-1(U)), a Newton iteration for R with
eta(R) = W / sqrt(lam) and the continuous quantile
S = lam R + log(sqrt(2 R (1 - R + R log R)) / |R - 1|) / log R,
S -= 0.0218 / (S + 0.065 lam), S += Del = 0.01 / S; the count is floor(S)
outside the correction band [X, X + 2 Del) and for S >= 10. The device runs
that branch in FP32 (giles_tail32 in dev_math.hpp) and returns floor(S) only
where S clears 10, the integers and the band by tol = 1e-3 + 1e-5 lam, so its
counts are the FP64 method's wherever |S32 - S64| < tol.

This test restates both arithmetics -- FP64 as the method and the oracle
form them, FP32 as the device does, with every FP32 logarithm, square root
and reciprocal perturbed by up to 2 ulp (the device's transcendental units
are not correctly rounded) and W from the FP32 form of the normal quantile --
and checks |S32 - S64| < tol / 8 wherever S64 >= 9, and that the two cases
the device sends to the bottom-up sum without iterating (W / sqrt(lam) <
-1.3; W < 0 with lam <= 9.4) have S64 < 10 - tol. The counts themselves are
compared with the oracle's inversion on the GPU (test_gpu_parity.py).

The Temme correction (S > 10 within [X, X + 2 Del) of an integer X >= 10,
0.5 lam < X < 2 lam) chooses X - 1 where Temme's expansion Sm of the CDF at X - 1
exceeds U (for X < lam; Sm - 1 against U - 1 above). The device forms Sm in
FP32 (giles_temme32) and decides only where it clears U by 4e-6; the second
test checks |Sm32 - Sm64| < 4e-6 / 4 with the same perturbations (erfc by 4
ulp).
"""
import math

import numpy as np
from scipy.special import erfc

C = [-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00, -2.549732539343734e+00,
     4.374664141464968e+00, 2.938163982698783e+00]
D = [7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00, 3.754408661907416e+00]
F32 = np.float32


def _noisy(x, rng, ulps=2):
    x = np.asarray(x, dtype=F32)
    k = rng.integers(-ulps, ulps + 1, x.shape)
    out = x.copy()
    for step in range(1, ulps + 1):
        up = k >= step
        dn = k <= -step
        out[up] = np.nextafter(out[up], F32(np.inf))
        out[dn] = np.nextafter(out[dn], F32(-np.inf))
    return out


def _w64(q):
    """Acklam's lower-tail quantile in FP64 (q < 0.02425)."""
    r = np.sqrt(-2.0 * np.log(q))
    num = ((((C[0] * r + C[1]) * r + C[2]) * r + C[3]) * r + C[4]) * r + C[5]
    den = (((D[0] * r + D[1]) * r + D[2]) * r + D[3]) * r + 1.0
    return num / den


def _w32(q, rng):
    """The device's FP32 form: q rounded to FP32, r = sqrt(-2 ln2 log2 q), the
    rational in FP32 with a reciprocal."""
    qf = q.astype(F32)
    lg = _noisy(np.log(qf), rng)
    r = _noisy(np.sqrt(F32(-2.0) * lg), rng)
    num = (((((F32(C[0]) * r + F32(C[1])) * r + F32(C[2])) * r + F32(C[3])) * r + F32(C[4])) * r + F32(C[5]))
    den = ((((F32(D[0]) * r + F32(D[1])) * r + F32(D[2])) * r + F32(D[3])) * r + F32(1.0))
    return (num * _noisy(F32(1.0) / den, rng)).astype(F32)


def _s64(lam, W):
    S = W / np.sqrt(lam)
    R = np.maximum(1.0 + S, 0.1)
    live = np.ones(lam.shape, bool)
    for _ in range(200):
        Tl = np.log(R)
        S2 = np.sqrt(np.maximum(2.0 * ((1.0 - R) + R * Tl), 0.0))
        S2 = np.where(R < 1.0, -S2, S2)
        Rn = R - (S2 - S) * S2 / Tl
        Rn = np.maximum(Rn, 0.1 * R)
        done = np.abs(Rn - R) <= 1e-8
        R = np.where(live, Rn, R)
        live &= ~done
        if not live.any():
            break
    assert not live.any()
    Tl = np.log(R)
    S = lam * R + np.log(np.sqrt(2.0 * R * ((1.0 - R) + R * Tl)) / np.abs(R - 1.0)) / Tl
    S = S - 0.0218 / (S + 0.065 * lam)
    return S + 0.01 / S


def _s32(lam, W, rng):
    lf = lam.astype(F32)
    s = W * _noisy(F32(1.0) / _noisy(np.sqrt(lf), rng), rng)
    R = np.maximum(F32(1.0) + s, F32(0.1)).astype(F32)
    live = np.ones(lam.shape, bool)
    conv = np.zeros(lam.shape, bool)
    with np.errstate(invalid="ignore", divide="ignore"):
        for _ in range(12):
            Tl = _noisy(np.log(R), rng)
            S2 = _noisy(np.sqrt(F32(2.0) * ((F32(1.0) - R) + R * Tl)), rng)
            S2 = np.where(R < 1, -S2, S2).astype(F32)
            Rn = (R - (S2 - s) * S2 * _noisy(F32(1.0) / Tl, rng)).astype(F32)
            Rn = np.where(Rn < F32(0.1) * R, F32(0.1) * R, Rn).astype(F32)
            done = np.abs(Rn - R) <= F32(1e-6) * Rn
            R = np.where(live, Rn, R).astype(F32)
            conv |= live & done
            live &= ~done
        Tl = _noisy(np.log(R), rng)
        a = _noisy(np.sqrt(F32(2.0) * R * ((F32(1.0) - R) + R * Tl)), rng)
        S = lf * R + _noisy(np.log(a * _noisy(F32(1.0) / np.abs(R - F32(1.0)), rng)), rng) * \
            _noisy(F32(1.0) / Tl, rng)
        S = S - F32(0.0218) * _noisy(F32(1.0) / (S + F32(0.065) * lf), rng)
        S = S + F32(0.01) * _noisy(F32(1.0) / S, rng)
    return S.astype(F32), conv


def _grid(rng):
    lam = np.concatenate([rng.uniform(4.0, 20.0, 60000), 10 ** rng.uniform(1.3, 3.5, 60000)])
    logq = rng.uniform(np.log(1e-19), np.log(1.3e-3), lam.size)  # |W| in [3, 9]
    q = np.exp(logq)
    sign = np.where(rng.uniform(size=lam.size) < 0.5, -1.0, 1.0)
    return lam, q, sign


def test_tail_pass_error_below_an_eighth_of_tol():
    rng = np.random.default_rng(7)
    lam, q, sign = _grid(rng)
    W64 = sign * _w64(q)
    W32 = (sign * _w32(q, rng)).astype(F32)
    assert np.max(np.abs(W32 - W64) / np.abs(W64)) < 2e-6
    tail = np.abs(W32) > 3.001
    lam, W64, W32 = lam[tail], W64[tail], W32[tail]
    skip = (W32 / np.sqrt(lam) < -1.3) | ((W32 < 0) & (lam <= 9.4))
    S64 = _s64(lam, W64)
    tol = 1e-3 + 1e-5 * lam
    # the cases sent to the sum without iterating: the method sums there too
    assert np.all(S64[skip] < 10.0 - tol[skip])
    S32, conv = _s32(lam[~skip], W32[~skip], rng)
    S64k, tolk = S64[~skip], tol[~skip]
    big = S64k >= 9.0
    assert big.sum() > 50000
    # every case the floor decides has converged
    assert np.all(conv[big])
    err = np.abs(S32[big].astype(np.float64) - S64k[big])
    assert np.max(err / tolk[big]) < 0.125, np.max(err / tolk[big])


B1 = [8.0995211567045583e-16, -5.1391118342426808e-16, -1.3923887224181616e-13, 1.1004392031956284e-13,
      2.4361948020667402e-11, -2.5514193994946487e-11, -4.3820360184533521e-09, 6.7078535434015332e-09,
      8.2967113409530833e-07, -2.1854485106799979e-06, -0.00017875514403292177, 0.0011574074074074078,
      0.083333333333333329]
B0 = [-1.9752288294349411e-15, 2.8534893807047458e-14, 3.3717632624009806e-13, -5.0276692801141763e-12,
      -5.8307721325504166e-11, 9.1476995822367933e-10, 1.0261809784240299e-08, -1.7665952736826086e-07,
      -1.8540622107151585e-06, 3.9192631785224383e-05, 0.00035273368606701921, -0.014814814814814815,
      -0.33333333333333331]


def _temme64(lam, X):
    """The method's Sm (+ or - the erfc term) in FP64."""
    ilam, Xi = 1.0 / lam, 1.0 / X
    eta = X * ilam
    eta = np.sqrt(2.0 * (1.0 - eta + eta * np.log(eta)) / eta)
    eta = np.where(X > lam, -eta, eta)
    b1, b0 = B1[0], B0[0]
    sm = b0 + b1 * eta
    mult = 25.0
    for j in range(1, 13):
        b1 = B1[j] + mult * b1 * Xi
        sm = b1 + sm * eta
        mult -= 1.0
        b0 = B0[j] + mult * b0 * Xi
        sm = b0 + sm * eta
        mult -= 1.0
    sm = sm / (1.0 + b1 * Xi) * np.exp(-0.5 * X * eta * eta) / np.sqrt(2.0 * math.pi * X)
    return np.where(X < lam, sm + 0.5 * erfc(eta * np.sqrt(0.5 * X)), sm - 0.5 * erfc(-eta * np.sqrt(0.5 * X)))


def _temme32(lam, X, rng):
    d = ((X - lam) * (1.0 / lam)).astype(F32)
    r = (F32(1.0) + d).astype(F32)
    p = F32(1.0 / 182.0)
    for k in range(13, 1, -1):
        p = (p * d + F32((-1) ** k / (k * (k - 1)))).astype(F32)
    g = np.where(np.abs(d) < F32(0.25), d * d * p, r * _noisy(np.log(r), rng) - d).astype(F32)
    eta = _noisy(np.sqrt(F32(2.0) * g * _noisy(F32(1.0) / r, rng)), rng)
    eta = np.where(X > lam, -eta, eta).astype(F32)
    Xf = X.astype(F32)
    xi = _noisy(F32(1.0) / Xf, rng)
    b1, b0 = F32(B1[0]), F32(B0[0])
    sm = (b0 + b1 * eta).astype(F32)
    mult = F32(25.0)
    for j in range(1, 13):
        b1 = (F32(B1[j]) + mult * b1 * xi).astype(F32)
        sm = (b1 + sm * eta).astype(F32)
        mult = F32(mult - 1)
        b0 = (F32(B0[j]) + mult * b0 * xi).astype(F32)
        sm = (b0 + sm * eta).astype(F32)
        mult = F32(mult - 1)
    sm = (sm * _noisy(F32(1.0) / (F32(1.0) + b1 * xi), rng)).astype(F32)
    sm = (sm * _noisy(np.exp(F32(-0.5) * Xf * eta * eta), rng) *
          _noisy(F32(1.0) / _noisy(np.sqrt(F32(2.0 * math.pi) * Xf), rng), rng)).astype(F32)
    z = (eta * _noisy(np.sqrt(F32(0.5) * Xf), rng)).astype(F32)
    ep = _noisy(erfc(z.astype(np.float64)).astype(F32), rng, 4)
    em = _noisy(erfc(-z.astype(np.float64)).astype(F32), rng, 4)
    return np.where(X < lam, sm + F32(0.5) * ep, sm - F32(0.5) * em).astype(F32)


def test_temme_pass_error_below_a_quarter_of_tol():
    rng = np.random.default_rng(5)
    lam = np.concatenate([rng.uniform(5.0, 30.0, 100000), 10 ** rng.uniform(1.4, 3.5, 100000)])
    X = np.floor(lam + rng.uniform(-3.2, 3.2, lam.size) * np.sqrt(lam))
    keep = (X >= 10) & (X > 0.5 * lam) & (X < 2.0 * lam)
    lam, X = lam[keep], X[keep]
    assert lam.size > 150000
    err = np.abs(_temme32(lam, X, rng).astype(np.float64) - _temme64(lam, X))
    assert np.max(err) < 1e-6, np.max(err)
