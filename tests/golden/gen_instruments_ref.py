"""Writes tests/golden/instruments_ref.json: base / spot-up / spot-down
payoffs of 21 instruments (European strips with vanilla and smoothed
digital payoffs, geometric and arithmetic Asians, lookbacks, barriers with
and without rebates, in-flight variants) on 8 fixed log-price paths, from
the reference's own trackers and instruments compiled where they lie by
oracle/Makefile's `ref` target (oracle/_ref/instruments_ref, caller
tests/cpp/instruments_ref.cpp). Run here, where the reference is present:

    python tests/golden/gen_instruments_ref.py
"""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BIN = os.path.join(ROOT, "oracle", "_ref", "instruments_ref")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([BIN], capture_output=True, text=True, check=True)
    cases = [json.loads(line) for line in out.stdout.splitlines()]
    fixture = {"S0": 100.0, "fd_bump": 1e-4, "greek_mode": "essential", "n_steps": 50, "T": 1.0, "cases": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "instruments_ref.json")
    with open(path, "w") as f:
        json.dump(fixture, f)
    print("wrote", path, len(cases), "cases")


if __name__ == "__main__":
    main()
