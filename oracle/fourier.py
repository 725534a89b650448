"""ORACLE (test infrastructure only): Gil-Pelaez Fourier prices of European
options under the ASVJ family, the reference's analytic check for every MC
strip (reference include/myql/pricers/fourier/FourierPricer.hpp:34-407,
engines/fourier/kernels/GilPelaezKernel.hpp:32-119).

Characteristic exponent of the martingale part Y_T of log(S_T/S0)
(AffineTraits.hpp:18-36, CFPolicies.hpp:9-60):
  Heston factor (Albrecher's "little trap" form)
    d = sqrt((rho s i u - kappa)^2 + s^2 (i u + u^2))
    g = (kappa - rho s i u - d) / (kappa - rho s i u + d)
    psi = kappa theta / s^2 [(kappa - rho s i u - d) T - 2 log((1 - g e^{-dT})/(1 - g))]
          + v0 / s^2 (kappa - rho s i u - d)(1 - e^{-dT}) / (1 - g e^{-dT})
  Black-Scholes: -vol^2 T (i u + u^2) / 2
  Merton:  lambda T [exp(i u mu - u^2 delta^2/2) - 1 - i u k],  k = e^{mu+delta^2/2} - 1
  Kou:     lambda T [p eta1/(eta1 - i u) + (1-p) eta2/(eta2 + i u) - 1 - i u k]
Factors add (independent CIR factors, DoubleFactorModel).

Gil-Pelaez, with k = log(K/S0) and phi_X(u) = exp(i u (r-q) T + psi(u)):
  P2 = 1/2 + 1/pi int_0^inf Re[e^{-iuk} phi_X(u) / (iu)] du
  P1 = 1/2 + 1/pi int_0^inf Re[e^{-iuk} phi_X(u - i) / (iu phi_X(-i))] du
  call = S0 e^{-qT} P1 - K e^{-rT} P2; cash-or-nothing call = e^{-rT} P2;
  asset-or-nothing call = S0 e^{-qT} P1; puts by parity.
The integrals use scipy's adaptive Gauss-Kronrod (QUADPACK) instead of the
reference's adaptive Simpson: the oracle targets the integral itself.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate


def _heston_psi(u: complex, T: float, h) -> complex:
    kappa, theta, s, rho, v0 = h.kappa, h.theta, h.sigma, h.rho, h.v0
    iu = 1j * u
    s2 = s * s
    dp = rho * s * iu
    D = np.sqrt((dp - kappa) ** 2 + s2 * (iu + u * u))
    g = (kappa - dp - D) / (kappa - dp + D)
    e = np.exp(-D * T)
    A = (kappa * theta / s2) * ((kappa - dp - D) * T - 2.0 * np.log((1.0 - g * e) / (1.0 - g)))
    B = (v0 / s2) * (kappa - dp - D) * (1.0 - e) / (1.0 - g * e)
    return A + B


def _jump_psi(u: complex, T: float, jump_kind: int, jump) -> complex:
    iu = 1j * u
    if jump_kind == 1:
        lam, mu, delta = jump[0], jump[1], jump[2]
        k = math.exp(mu + 0.5 * delta * delta) - 1.0
        return lam * T * (np.exp(iu * mu - 0.5 * delta * delta * u * u) - 1.0 - iu * k)
    if jump_kind == 2:
        lam, p, e1, e2 = jump
        k = p * e1 / (e1 - 1.0) + (1.0 - p) * e2 / (e2 + 1.0) - 1.0
        return lam * T * (p * e1 / (e1 - iu) + (1.0 - p) * e2 / (e2 + iu) - 1.0 - iu * k)
    return 0.0


class AffineModel:
    """n_factors: 0 (vol), 1 or 2 Heston factors; jump_kind 0/1/2 as the C-ABI."""

    def __init__(self, n_factors: int, heston=(), vol: float = 0.0, jump_kind: int = 0,
                 jump=(0.0, 0.0, 0.0, 0.0)):
        self.n_factors = n_factors
        self.heston = list(heston)
        self.vol = vol
        self.jump_kind = jump_kind
        self.jump = tuple(jump) + (0.0,) * (4 - len(jump))

    def psi(self, u: complex, T: float) -> complex:
        if T < 1e-8:
            return 0.0
        if self.n_factors == 0:
            out = -0.5 * self.vol ** 2 * T * (1j * u + u * u)
        else:
            out = 0.0
            for h in self.heston[: self.n_factors]:
                out = out + _heston_psi(u, T, h)
        return out + _jump_psi(u, T, self.jump_kind, self.jump)


def _probabilities(model: AffineModel, S0: float, K: float, T: float, r: float, q: float,
                   tol: float = 1e-12):
    k = math.log(K / S0)
    drift = (r - q) * T

    def f2(u):
        if u == 0.0:
            u = 1e-12
        v = np.exp(1j * u * drift + model.psi(u, T) - 1j * u * k) / (1j * u)
        return v.real

    def f1(u):
        if u == 0.0:
            u = 1e-12
        v = np.exp(1j * u * drift + model.psi(u - 1j, T) - 1j * u * k) / (1j * u)
        return v.real

    def integral(f):
        total, a, width = 0.0, 0.0, 5.0
        # integrate in growing panels until the panel contribution is negligible
        for _ in range(400):
            val, _err = integrate.quad(f, a, a + width, epsabs=tol * 1e-2, epsrel=1e-13,
                                       limit=400)
            total += val
            a += width
            if abs(val) < tol * 1e-3 and a > 50.0:
                break
            width = min(width * 1.5, 200.0)
        return total

    P1 = 0.5 + integral(f1) / math.pi
    P2 = 0.5 + integral(f2) / math.pi
    return P1, P2


def price(model: AffineModel, S0: float, K: float, T: float, r: float, q: float,
          payoff: int = 0, option_type: int = 0) -> float:
    """payoff 0 vanilla, 1 cash-or-nothing, 2 asset-or-nothing; option 0 call, 1 put."""
    P1, P2 = _probabilities(model, S0, K, T, r, q)
    dfr, dfq = math.exp(-r * T), math.exp(-q * T)
    if payoff == 0:
        call = S0 * dfq * P1 - K * dfr * P2
        return call if option_type == 0 else call - S0 * dfq + K * dfr
    if payoff == 1:
        return dfr * P2 if option_type == 0 else dfr * (1.0 - P2)
    return S0 * dfq * P1 if option_type == 0 else S0 * dfq * (1.0 - P1)


def greeks_fd(model: AffineModel, S0: float, K: float, T: float, r: float, q: float,
              payoff: int = 0, option_type: int = 0, hS: float = 1e-2) -> dict:
    """Delta and gamma by central differences of the semi-analytic price."""
    p0 = price(model, S0, K, T, r, q, payoff, option_type)
    pu = price(model, S0 + hS, K, T, r, q, payoff, option_type)
    pd = price(model, S0 - hS, K, T, r, q, payoff, option_type)
    return dict(price=p0, delta=(pu - pd) / (2 * hS), gamma=(pu - 2 * p0 + pd) / hS ** 2)


def black_scholes(S0, K, T, r, q, vol, option_type=0):
    from scipy.stats import norm
    d1 = (math.log(S0 / K) + (r - q + 0.5 * vol * vol) * T) / (vol * math.sqrt(T))
    d2 = d1 - vol * math.sqrt(T)
    call = S0 * math.exp(-q * T) * norm.cdf(d1) - K * math.exp(-r * T) * norm.cdf(d2)
    return call if option_type == 0 else call - S0 * math.exp(-q * T) + K * math.exp(-r * T)
