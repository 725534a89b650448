"""Writes tests/golden/smoothing_ref.json: the expected average variance and
the automatic digital-smoothing bandwidth (scalar strike K = 110 and strike
strip, S0 = 100) of ten models at four maturities, from the reference's own
Boost-free variance helpers (models/asvj/data/ModelParams.hpp:54-80) on its
model containers, compiled by oracle/Makefile's `ref` target
(oracle/_ref/smoothing_ref, caller tests/cpp/smoothing_ref.cpp). Run here,
where the reference is present:

    python tests/golden/gen_smoothing_ref.py

The model parameters are stored with the cases so the tests can rebuild
each call.
"""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BIN = os.path.join(ROOT, "oracle", "_ref", "smoothing_ref")

# the parameters tests/cpp/smoothing_ref.cpp uses (kept in step with it)
H1 = dict(kappa=2.0, theta=0.04, sigma=0.3, rho=-0.7, v0=0.04)
H_HI = dict(kappa=2.0, theta=0.04, sigma=1.0, rho=-0.7, v0=0.04)
H2 = dict(kappa=1.0, theta=0.03, sigma=0.25, rho=-0.5, v0=0.02)
H_BK = dict(kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7, v0=0.04)
H_FLAT = dict(kappa=1e-9, theta=0.05, sigma=0.2, rho=-0.3, v0=0.03)
MERTON = dict(kind="merton", p=[0.5, -0.1, 0.15, 0.0])
KOU = dict(kind="kou", p=[1.0, 0.4, 10.0, 5.0])
MODELS = {
    "black_scholes": dict(n_factors=0, heston=[], vol=0.25, jump=None),
    "merton": dict(n_factors=0, heston=[], vol=0.2, jump=MERTON),
    "kou": dict(n_factors=0, heston=[], vol=0.2, jump=KOU),
    "heston": dict(n_factors=1, heston=[H1], vol=0.0, jump=None),
    "bates": dict(n_factors=1, heston=[H_HI], vol=0.0, jump=MERTON),
    "bates_kou": dict(n_factors=1, heston=[H_BK], vol=0.0, jump=KOU),
    "double_heston": dict(n_factors=2, heston=[H1, H2], vol=0.0, jump=None),
    "double_bates": dict(n_factors=2, heston=[H1, H2], vol=0.0, jump=MERTON),
    "double_bates_kou": dict(n_factors=2, heston=[H1, H2], vol=0.0, jump=KOU),
    "heston_flat": dict(n_factors=1, heston=[H_FLAT], vol=0.0, jump=None),
}


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([BIN], capture_output=True, text=True, check=True)
    cases = [json.loads(line) for line in out.stdout.splitlines()]
    fixture = {"S0": 100.0, "K_scalar": 110.0, "models": MODELS, "cases": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "smoothing_ref.json")
    with open(path, "w") as f:
        json.dump(fixture, f, indent=0)
    print("wrote", path, len(cases), "cases")


if __name__ == "__main__":
    main()
