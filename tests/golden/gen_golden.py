"""Writes tests/golden/golden.json: known-answer vectors for the oracle,
computed here independently of it.

  xoshiro256++ and splitmix64 outputs from a pure-Python big-integer
    restatement of the published algorithms (Blackman & Vigna; Steele et al.);
  the published xoshiro256 JUMP (2^128) and LONG_JUMP (2^192) polynomials;
  chi-squared quantiles and regularised incomplete gamma values from scipy;
  inverse normal values from scipy.special.ndtri.
Run: python tests/golden/gen_golden.py
"""
import json
import os

import numpy as np
from scipy import special, stats

M64 = (1 << 64) - 1


def splitmix(seed):
    x = seed & M64
    while True:
        x = (x + 0x9E3779B97F4A7C15) & M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        yield z ^ (z >> 31)


def rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


def xoshiro_pp(state, n):
    s = list(state)
    out = []
    for _ in range(n):
        out.append((rotl((s[0] + s[3]) & M64, 23) + s[0]) & M64)
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
    return out


def main():
    g = {}
    for seed in (0, 42, 999983, 2 ** 63 + 5):
        sm = splitmix(seed)
        st = [next(sm) for _ in range(4)]
        g[f"xoshiro_seed_{seed}"] = {"state": [str(v) for v in st],
                                     "outputs": [str(v) for v in xoshiro_pp(st, 64)]}
    g["jump_2_128"] = [str(v) for v in (0x180ec6d33cfd0aba, 0xd5a61266f0c9392c,
                                        0xa9582618e03fc9aa, 0x39abdc4529b1661c)]
    g["jump_2_192"] = [str(v) for v in (0x76e15d3efefdcbbf, 0xc5004e441c522fb3,
                                        0x77710069854ee241, 0x39109bb02acbe635)]
    rng = np.random.default_rng(20240607)
    dof = np.concatenate([[0.32, 0.32, 0.32, 2.32, 10.32, 44.32, 88.32, 0.05, 1.0, 500.0],
                          rng.uniform(0.05, 120.0, 60)])
    p = np.concatenate([[1e-4, 0.5, 0.999, 0.3, 0.7, 0.9999999, 0.999, 1e-9, 0.5, 0.5],
                        rng.uniform(0.0, 1.0, 60)])
    g["chi2"] = {"dof": dof.tolist(), "p": p.tolist(), "q": stats.chi2.ppf(p, dof).tolist()}
    a = rng.uniform(0.05, 60, 80)
    x = a * rng.uniform(0.05, 3.0, 80)
    g["igamma"] = {"a": a.tolist(), "x": x.tolist(), "P": special.gammainc(a, x).tolist(),
                   "Q": special.gammaincc(a, x).tolist()}
    pn = np.concatenate([10.0 ** -np.arange(1, 16), np.linspace(0.001, 0.999, 99),
                         1 - 10.0 ** -np.arange(1, 16)])
    g["ndtri"] = {"p": pn.tolist(), "z": special.ndtri(pn).tolist()}
    # degrees of freedom d + 2N over the whole NCI table (N up to 1023)
    rt = np.random.default_rng(20261203)
    dof_t = np.concatenate([[849.0313428193025, 1264.3770645306565, 2046.32], rt.uniform(120.0, 2100.0, 397)])
    p_t = rt.uniform(0.0, 1.0, dof_t.size)
    p_t[:3] = [0.7674501184822347, 0.5044304231979618, 0.5]
    p_t[3:23] = 0.999
    p_t[23:43] = 1e-6
    p_t[43:63] = 1.0 - 1e-9
    g["chi2_table_range"] = {"dof": dof_t.tolist(), "p": p_t.tolist(), "q": stats.chi2.ppf(p_t, dof_t).tolist()}
    # the first NCI grid interval, inverted exactly for rows N >= 2: p from
    # the smallest nonzero uniform (2^-64) up to the grid step 0.999 / 9999
    rl = np.random.default_rng(20261204)
    dof_l = 0.32 + 2.0 * np.concatenate([[2, 5, 1023], rl.integers(2, 1024, 197)])
    p_l = np.exp(rl.uniform(np.log(2.0 ** -64), np.log(0.999 / 9999), dof_l.size))
    p_l[:3] = [2.0 ** -64, 0.999 / 9999, 2.0 ** -64]
    g["chi2_low_tail"] = {"dof": dof_l.tolist(), "p": p_l.tolist(), "q": stats.chi2.ppf(p_l, dof_l).tolist()}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(out, "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
