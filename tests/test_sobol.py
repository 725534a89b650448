"""The Sobol' construction behind the engine's randomized quasi-Monte Carlo
layout (MYQL_STREAMS_SOBOL), restated in oracle/sobol.py and pinned here by
what the mathematics fixes. The reference has no QMC (README.md:332-335
lists it as future work); the device's coordinates are compared with this
restatement bit for bit in tests/test_gpu_qmc.py.
"""
import math

import numpy as np
import pytest

from oracle import sobol


def _order_of_x(p: int, s: int) -> int:
    """Multiplicative order of x modulo p by repeated multiplication."""
    r, n = 2, 1
    if r >> s & 1:
        r ^= p
    while r != 1:
        r <<= 1
        if r >> s & 1:
            r ^= p
        n += 1
        if n > (1 << s):
            return -1
    return n


def _euler_phi(n: int) -> int:
    out, m, f = n, n, 2
    while f * f <= m:
        if m % f == 0:
            while m % f == 0:
                m //= f
            out -= out // f
        f += 1
    if m > 1:
        out -= out // m
    return out


@pytest.mark.parametrize("s", range(1, 11))
def test_primitive_polynomials_by_brute_force(s):
    """Every polynomial of degree s the test accepts has x of order exactly
    2^s - 1 (direct multiplication), and there are phi(2^s - 1)/s of them."""
    found = [p for (deg, p) in sobol.primitive_polynomials(400) if deg == s]
    assert len(found) == _euler_phi((1 << s) - 1) // s
    for p in found:
        assert _order_of_x(p, s) == (1 << s) - 1
    # and a reducible or non-primitive one is rejected
    if s >= 2:
        rejected = [p for p in range(1 << s, 1 << (s + 1)) if p & 1 and p not in found]
        for p in rejected[:20]:
            assert not sobol.is_primitive(p, s)
            assert _order_of_x(p, s) != (1 << s) - 1


def test_first_dimensions_are_the_classical_ones():
    V = sobol.direction_numbers(3)
    # van der Corput, and the x + 1 dimension: m = 1, 3, 5, 15, 17, 51, ...
    assert [int(V[0, k]) for k in range(5)] == [1 << (31 - k) for k in range(5)]
    m = [int(V[1, k]) >> (31 - k) for k in range(8)]
    assert m == [1, 3, 5, 15, 17, 51, 85, 255]


def test_direction_integers_have_unit_diagonal():
    V = sobol.direction_numbers(256).astype(np.uint64)
    for k in range(32):
        # m_k odd, m_k < 2^k: the lowest set bit of V_k is bit 31 - k and
        # nothing sits above bit 31
        assert np.all((V[:, k] >> (31 - k)) & 1 == 1)
        assert np.all(V[:, k] < (1 << 32))


def test_one_dimensional_stratification():
    """The first 2^m points of every coordinate fall one in each interval
    [j 2^-m, (j + 1) 2^-m)."""
    V = sobol.direction_numbers(64)
    m = 9
    for d in range(64):
        top = sorted(sobol.point(V, c, d) >> (32 - m) for c in range(1 << m))
        assert top == list(range(1 << m)), d


def test_first_two_coordinates_form_a_0_m_2_net():
    """(x_0, x_1) of the first 2^m points: every elementary interval of area
    2^-m holds exactly one point."""
    V = sobol.direction_numbers(2)
    m = 10
    pts = [(sobol.point(V, c, 0), sobol.point(V, c, 1)) for c in range(1 << m)]
    for i in range(m + 1):
        j = m - i
        cells = {(a >> (32 - i) if i else 0, b >> (32 - j) if j else 0) for a, b in pts}
        assert len(cells) == 1 << m, (i, j)


def test_engine_walk_covers_the_net():
    """The engine's walk -- point (gray(h) << 6) | lane for h < 2^k -- visits
    exactly the first 2^(k + 6) points."""
    V = sobol.direction_numbers(4)
    shift = np.zeros(4, dtype=np.uint32)
    k = 4
    walk = sorted(sobol.engine_point(V, shift, h, lane, 2) for h in range(1 << k) for lane in range(64))
    direct = sorted(sobol.point(V, c, 2) for c in range(1 << (k + 6)))
    assert walk == direct


def test_digital_shift_is_a_uniform_xor():
    s1 = sobol.replicate_shift(42, 0, 16)
    s2 = sobol.replicate_shift(42, 1, 16)
    assert not np.array_equal(s1, s2)
    V = sobol.direction_numbers(16)
    assert sobol.engine_point(V, s1, 5, 7, 3) == sobol.engine_point(V, np.zeros(16, np.uint32), 5, 7, 3) ^ int(s1[3])
    # the shift's bits look fair over many coordinates
    s = sobol.replicate_shift(7, 3, 4096).astype(np.uint64)
    frac = np.mean([(int(x) >> b) & 1 for x in s for b in range(32)])
    assert abs(frac - 0.5) < 0.01
