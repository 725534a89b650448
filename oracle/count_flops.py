"""Algorithmic floating-point operations per path-step of the Bates strip,
counted by the oracle's op-counting build (lib/libmyql_oracle_ops.so).

The count follows the reference algorithm as written: every Greek variant
(base, v0 +-, T +-, r +-) replays its own path, an FMA counts 2 and an add,
mul, div, sqrt, exp, log or erfc counts 1; integer RNG work is not counted.
It covers the whole pricing call -- stepping, Poisson and chi-squared
inversion, jumps, payoffs, Greek estimators and the per-strike sums -- and is
divided by the path-steps of the call (paths x steps x 7 variants).

    python oracle/count_flops.py [--paths-log2 16] [--reference-nci-rows]

prints one JSON line with the per-maturity and overall flops per path-step;
bench.py's FLOP_PER_PATH_STEP is this script's overall figure at 2^16 paths,
FLOP_PER_PATH_STEP_REF_ROWS the same under the reference's NCI rows and
fallback (MYQL_FLAG_REFERENCE_NCI_ROWS).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def count(paths: int, maturities=None, reference_rows: bool = False) -> dict:
    import bench
    from myquantlib_amd import capi
    from myquantlib_amd.pricing import InstrumentSpec, mc_config
    from oracle.oracle import Oracle

    o = Oracle(ops=True)
    m = bench.model()
    per, tot_ops, tot_steps = {}, 0, 0
    for T in (maturities or bench.MATURITIES):
        steps = int(round(100 * T))
        inst = InstrumentSpec(strikes=bench.STRIKES, maturity=T)
        cfg = mc_config(paths, steps, seed=42, stream_layout=capi.STREAMS_SEEDED, num_streams=1,
                        flags=capi.FLAG_REFERENCE_NCI_ROWS if reference_rows else 0)
        o.op_count(reset=True)
        o.mc(m, [capi.SCHEME_NCI], inst, capi.GREEK_FULL, cfg, bench.S0, bench.R, bench.Q,
             nthreads=1)
        ops = o.op_count(reset=True)
        ps = paths * steps * bench.VARIANTS
        per[str(T)] = ops / ps
        tot_ops += ops
        tot_steps += ps
    return {"paths_per_maturity": paths, "reference_nci_rows": reference_rows,
            "flop_per_path_step": tot_ops / tot_steps,
            "per_maturity": per}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--paths-log2", type=int, default=16)
    ap.add_argument("--reference-nci-rows", action="store_true")
    args = ap.parse_args()
    print(json.dumps(count(1 << args.paths_log2, reference_rows=args.reference_nci_rows)))


if __name__ == "__main__":
    main()
