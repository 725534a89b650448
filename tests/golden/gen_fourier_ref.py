"""Writes tests/golden/fourier_ref.json: prices and Greeks of European strips
from the reference's own Gil-Pelaez pricer (FourierPricer.hpp, whose
closure needs no Boost), compiled from the reference's headers where they
lie by oracle/Makefile's `ref` target (oracle/_ref/fourier_ref, caller
tests/cpp/fourier_ref.cpp). Run here, where the reference is present:

    python tests/golden/gen_fourier_ref.py

The reference's engine cuts the integration range into
max(8, 4 * omp_get_max_threads()) chunks, so the program runs with
OMP_NUM_THREADS=16 (64 chunks). The fixture stores the model parameters,
spot and rates with every case so the tests can rebuild each call.
"""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BIN = os.path.join(ROOT, "oracle", "_ref", "fourier_ref")

# the parameters tests/cpp/fourier_ref.cpp prices (kept in step with it)
H1 = dict(kappa=2.0, theta=0.04, sigma=0.3, rho=-0.7, v0=0.04)
H_HI = dict(kappa=2.0, theta=0.04, sigma=1.0, rho=-0.7, v0=0.04)
H2 = dict(kappa=1.0, theta=0.03, sigma=0.25, rho=-0.5, v0=0.02)
H_BK = dict(kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7, v0=0.04)
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
}


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS="16")
    out = subprocess.run([BIN], capture_output=True, text=True, env=env, check=True)
    cases = [json.loads(line) for line in out.stdout.splitlines()]
    fixture = {"S0": 100.0, "r": 0.03, "q": 0.01, "omp_threads": 16, "models": MODELS, "cases": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fourier_ref.json")
    with open(path, "w") as f:
        json.dump(fixture, f)
    print("wrote", path, len(cases), "cases")


if __name__ == "__main__":
    main()
