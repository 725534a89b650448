"""ORACLE (test infrastructure only): the terminal-jump expectation that the
learned control variate subtracts (myquantlib_amd/csrc/cv_kernel.hpp,
rb_expect / rb_sum / kou_sum), restated by a different route.

European paths whose variance does not depend on the jumps take the whole
maturity's jump at the end (JumpPolicies.hpp:55-135): J = m0 + sum of n
jumps, n ~ Poisson(lambda T), m0 = -lambda kappa T the compensation. The
estimator uses E[P(S_T- e^J) | S_T-] for vanilla and sharp digital payoffs
(Payoffs.hpp:14-69: vanilla (S - K)^+ / (K - S)^+, cash 1{S > K}, asset
S 1{S > K}).

  terminal_expectation(...)  E[P(S e^J)] (include_zero=False: the n >= 1
                             part only)
  mc_terminal_expectation    Monte Carlo of the same (also for digitals
                             smoothed with bandwidth eps, which the device
                             integrates over the payoff's transition window
                             and the GPU test checks against this)

Kou jumps (up with probability p, Exp(eta1); down Exp(eta2)): the n >= 1
part is summed over the independent up and down counts (j, m) with the
expectation over the up side A ~ Gamma(j, eta1) in closed form (the upper
regularised incomplete gamma function) and the down side B ~ Gamma(m,
eta2) integrated numerically (composite Gauss-Legendre, split at the
payoff's kink). The
device sums a closed form of both sides instead. Merton jumps: the n-jump
sum is normal and the expectation is integrated numerically over it.

Pinned in tests/test_cv_jump.py by Monte Carlo with jumps drawn as the
reference draws them (a Poisson count, then each jump's side and size),
and by put-call parity (call - put = S - K, since E[e^J] = 1).
"""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate, special, stats

VANILLA, CASH, ASSET = 0, 1, 2


def _payoff(payoff: int, call: bool, S, K, eps: float | None = None):
    """Payoffs.hpp:14-69; eps > 0: digitals smoothed to Phi(+-(S - K) / eps)
    (0.5 erfc(-diff / (sqrt(2) eps)), as the pricer smooths them in its Greek
    modes)."""
    S = np.asarray(S, dtype=np.float64)
    d = S - K if call else K - S
    if payoff == VANILLA:
        return np.maximum(d, 0.0)
    w = special.ndtr(d / eps) if eps else (d > 0.0).astype(np.float64)
    return w if payoff == CASH else S * w


def kou_kappa(p: float, e1: float, e2: float) -> float:
    return p * e1 / (e1 - 1.0) + (1.0 - p) * e2 / (e2 + 1.0) - 1.0


def _up_side(payoff: int, call: bool, X, K: float, j: int, e1: float):
    """E[P(X e^A)] for A ~ Gamma(j, e1), j >= 1 (X an array)."""
    X = np.asarray(X, dtype=np.float64)
    c = np.log(K / X)
    cp = np.maximum(c, 0.0)
    g1 = (e1 / (e1 - 1.0)) ** j                            # E[e^A]
    t0 = np.where(c > 0, special.gammaincc(j, e1 * cp), 1.0)  # P(A > c)
    t1 = g1 * np.where(c > 0, special.gammaincc(j, (e1 - 1.0) * cp), 1.0)  # E[e^A 1{A > c}]
    if payoff == VANILLA:
        callv = X * t1 - K * t0
        return callv if call else callv - (X * g1 - K)
    if payoff == CASH:
        return t0 if call else 1.0 - t0
    return X * t1 if call else X * (g1 - t1)


_GL_X, _GL_W = np.polynomial.legendre.leggauss(48)


def _kou_term(payoff, call, Sm, K, j, m, e1, e2) -> float:
    """E[P(Sm e^{A - B})], A ~ Gamma(j, e1), B ~ Gamma(m, e2), j + m >= 1:
    the down side by composite Gauss-Legendre quadrature (panels split at
    the payoff's kink, where Sm e^{-B} = K)."""
    def h(b):
        X = Sm * np.exp(-b)
        if j == 0:
            return _payoff(payoff, call, X, K)
        return _up_side(payoff, call, X, K, j, e1)

    if m == 0:
        return float(h(np.array([0.0]))[0])
    g = stats.gamma(a=m, scale=1.0 / e2)
    kink = math.log(Sm / K)
    hi = g.isf(1e-18)
    edges = [0.0] + ([kink] if 0.0 < kink < hi else []) + [hi]
    tot = 0.0
    for a, b in zip(edges[:-1], edges[1:]):
        p = np.linspace(a, b, 13)
        lo, up = p[:-1, None], p[1:, None]
        x = 0.5 * (up - lo) * _GL_X[None, :] + 0.5 * (up + lo)
        w = 0.5 * (up - lo) * _GL_W[None, :]
        tot += float(np.sum(w * g.pdf(x) * h(x)))
    return tot


def _poisson_weights(mean: float, tol: float = 1e-17):
    w, pn, mass, n = [], math.exp(-mean), 0.0, 0
    while mass < 1.0 - tol and n < 60:
        w.append(pn)
        mass += pn
        n += 1
        pn *= mean / n
    return w


def terminal_expectation(payoff: int, call: bool, K: float, S: float, T: float, jump_kind: str,
                         params, include_zero: bool = True) -> float:
    """E[P(S e^{m0 + J})] over the terminal jump of a maturity-T path.
    jump_kind "kou": params (lambda, p, eta1, eta2); "merton": (lambda, mu, delta)."""
    if jump_kind == "kou":
        lam, p, e1, e2 = params
        m0 = -lam * kou_kappa(p, e1, e2) * T
        Sm = S * math.exp(m0)
        wu = _poisson_weights(lam * T * p)
        wd = _poisson_weights(lam * T * (1.0 - p))
        tot = 0.0
        for j, a in enumerate(wu):
            for m, b in enumerate(wd):
                if j == 0 and m == 0:
                    continue
                if a * b < 1e-19:
                    continue
                tot += a * b * _kou_term(payoff, call, Sm, K, j, m, e1, e2)
    elif jump_kind == "merton":
        lam, mu, dl = params
        m0 = -lam * (math.exp(mu + 0.5 * dl * dl) - 1.0) * T
        Sm = S * math.exp(m0)
        w = _poisson_weights(lam * T)
        tot = 0.0
        for n in range(1, len(w)):
            s = math.sqrt(n) * dl
            f = lambda z: stats.norm.pdf(z) * float(_payoff(payoff, call, Sm * math.exp(n * mu + s * z), K))
            z0 = (math.log(K / Sm) - n * mu) / s
            tot += w[n] * sum(integrate.quad(f, a, b, epsabs=1e-14, epsrel=1e-12, limit=200)[0]
                              for a, b in ((-12.0, min(max(z0, -12.0), 12.0)), (min(max(z0, -12.0), 12.0), 12.0)))
    else:
        raise ValueError(jump_kind)
    if include_zero:
        tot += math.exp(-lam * T) * float(_payoff(payoff, call, Sm, K))
    return tot


def mc_terminal_expectation(payoff: int, call: bool, K: float, S: float, T: float, jump_kind: str, params,
                            n: int, seed: int, eps: float | None = None):
    """Monte Carlo estimate (mean, standard error) of E[P(S e^{m0 + J})]
    with the jumps drawn as the reference draws them: a Poisson count for
    the maturity, then per jump its side and size (Kou) or its normal size
    (Merton). eps: smoothed digitals."""
    rng = np.random.default_rng(seed)
    lam = params[0]
    cnt = rng.poisson(lam * T, size=n)
    idx = np.repeat(np.arange(n), cnt)
    if jump_kind == "kou":
        _, p, e1, e2 = params
        m0 = -lam * kou_kappa(p, e1, e2) * T
        up = rng.random(idx.size) < p
        size = np.where(up, rng.exponential(1.0 / e1, idx.size), -rng.exponential(1.0 / e2, idx.size))
    else:
        _, mu, dl = params
        m0 = -lam * (math.exp(mu + 0.5 * dl * dl) - 1.0) * T
        size = rng.normal(mu, dl, idx.size)
    J = np.bincount(idx, weights=size, minlength=n)
    v = _payoff(payoff, call, S * np.exp(m0 + J), K, eps)
    return float(v.mean()), float(v.std(ddof=1) / math.sqrt(n))
