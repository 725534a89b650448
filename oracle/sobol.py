"""Sobol' sequence (Bratley & Fox, ACM TOMS Alg. 659) as the engine builds
it for randomized quasi-Monte Carlo -- CPU checker only (test
infrastructure; the product never imports it).

The reference has no QMC (its README.md:332-335 lists it as future work), so
this restates the engine's construction independently, in plain Python
integers:
* dimension 0: the van der Corput sequence, V_k = 2^(32-k);
* dimension d >= 1: the d-th primitive polynomial over GF(2) in order of
  degree, then of its integer encoding (x^s + a_1 x^(s-1) + ... + 1 has the
  bits a_1..a_(s-1) between the two ends), found by testing that x has
  order 2^s - 1 modulo the polynomial; the initial direction integers
  m_1..m_s are odd, m_k < 2^k, drawn from splitmix64 seeded with the
  dimension; m_k for k > s by the Bratley-Fox recurrence
  m_k = 2 a_1 m_(k-1) ^ 4 a_2 m_(k-2) ^ ... ^ 2^s m_(k-s) ^ m_(k-s);
  V_k = m_k 2^(32-k), k = 1..32;
* point c, dimension d: x = XOR of V_k over the set bits (k-1) of c, as a
  32-bit integer; u = (x + 1/2) 2^-32.
The engine walks c in a wavefront order -- c = (gray(h) << 6) | lane with
gray(h) = h ^ (h >> 1) -- and randomizes by a digital shift (XOR of one
32-bit integer per dimension); `engine_point` restates that mapping.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(state: int):
    """splitmix64 step: (new state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def _polymulmod(a: int, b: int, p: int, s: int) -> int:
    """a * b mod p over GF(2), p of degree s."""
    r = 0
    while b:
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if a >> s & 1:
            a ^= p
    return r


def _x_pow(e: int, p: int, s: int) -> int:
    r, base = 1, 2  # the polynomial x
    while e:
        if e & 1:
            r = _polymulmod(r, base, p, s)
        base = _polymulmod(base, base, p, s)
        e >>= 1
    return r


def _prime_factors(n: int):
    out, f = [], 2
    while f * f <= n:
        if n % f == 0:
            out.append(f)
            while n % f == 0:
                n //= f
        f += 1
    if n > 1:
        out.append(n)
    return out


def is_primitive(p: int, s: int) -> bool:
    """p (bit i = coefficient of x^i, degree s) is primitive: x has
    multiplicative order exactly 2^s - 1 modulo p."""
    if not (p >> s & 1) or not (p & 1):
        return False
    order = (1 << s) - 1
    if _x_pow(order, p, s) != 1:
        return False
    return all(_x_pow(order // q, p, s) != 1 for q in _prime_factors(order))


def primitive_polynomials(count: int):
    """The first `count` primitive polynomials, as (degree, bits)."""
    out, s = [], 1
    while len(out) < count:
        for mid in range(1 << (s - 1)):
            p = (1 << s) | (mid << 1) | 1
            if is_primitive(p, s):
                out.append((s, p))
                if len(out) == count:
                    break
        s += 1
    return out


def direction_numbers(dims: int) -> np.ndarray:
    """V[d, k-1] = V_k of dimension d, k = 1..32, as uint32."""
    V = np.zeros((dims, 32), dtype=np.uint64)
    for k in range(1, 33):
        V[0, k - 1] = 1 << (32 - k)
    polys = primitive_polynomials(max(dims - 1, 0))
    for d in range(1, dims):
        s, p = polys[d - 1]
        st = d
        m = [0] * 33
        for k in range(1, s + 1):
            st, z = splitmix64(st)
            m[k] = (2 * (z % (1 << (k - 1))) + 1) if k > 1 else 1
        for k in range(s + 1, 33):
            v = m[k - s] ^ (m[k - s] << s)
            for i in range(1, s):
                if p >> (s - i) & 1:
                    v ^= m[k - i] << i
            m[k] = v
        for k in range(1, 33):
            V[d, k - 1] = (m[k] << (32 - k)) & 0xFFFFFFFF
    return V.astype(np.uint32)


def point(V: np.ndarray, c: int, d: int) -> int:
    x, k = 0, 0
    while c:
        if c & 1:
            x ^= int(V[d, k])
        c >>= 1
        k += 1
    return x


def engine_point(V: np.ndarray, shift: np.ndarray, h: int, lane: int, d: int) -> int:
    """The engine's coordinate d for wave step h and lane: the Sobol' point
    (gray(h) << 6) | lane, digitally shifted."""
    g = h ^ (h >> 1)
    return point(V, (g << 6) | lane, d) ^ int(shift[d])


def replicate_shift(seed: int, rep: int, dims: int) -> np.ndarray:
    """Digital shift of replicate `rep`: splitmix64 seeded with
    seed ^ (rep * 0x9E3779B97F4A7C15), one output (low 32 bits) per
    dimension."""
    st = (seed ^ ((rep * 0x9E3779B97F4A7C15) & MASK64)) & MASK64
    out = np.zeros(dims, dtype=np.uint32)
    for d in range(dims):
        st, z = splitmix64(st)
        out[d] = z & 0xFFFFFFFF
    return out
