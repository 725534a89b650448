"""Writes tests/golden/numerics_ref.json: the reference's own Acklam inverse
normal (Numerics::normcdfinv_as241) and Giles inverse Poisson CDF
(Numerics::poissinvcdf) on a fixed grid, from math/Numerics.hpp compiled
where it lies by oracle/Makefile's `ref` target (oracle/_ref/numerics_ref,
caller tests/cpp/numerics_ref.cpp). Run here, where the reference is
present:

    python tests/golden/gen_numerics_ref.py
"""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BIN = os.path.join(ROOT, "oracle", "_ref", "numerics_ref")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([BIN], capture_output=True, text=True, check=True)
    data = json.loads(out.stdout)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "numerics_ref.json")
    with open(path, "w") as f:
        json.dump(data, f)
    print("wrote", path, len(data["p"]), "inverse-normal and", len(data["lam"]), "inverse-Poisson points")


if __name__ == "__main__":
    main()
