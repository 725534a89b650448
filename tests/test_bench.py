"""bench.py's workload definitions and CLI contract, on the CPU: the strip's
path-step count, the batched Double-Heston parameter sets, and the
one-GPU refusal of --gpus above 1."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from myquantlib_amd import capi  # noqa: E402


def test_strip_path_steps():
    """8 maturities x 2^24 paths x round(100 T) steps x 7 Greek variants =
    8.63e10 path-steps (735 steps summed over the maturities)."""
    assert sum(int(round(100 * T)) for T in bench.MATURITIES) == 735
    assert bench.path_steps(1 << 24) == 735 * (1 << 24) * 7 == 86318776320
    assert len(bench.STRIKES) == 64 and min(bench.STRIKES) == 70.0 and max(bench.STRIKES) == 130.0


def test_strip_jobs_shape():
    J = bench.jobs(1 << 10, capi.FP32)
    assert len(J) == 8
    for (m, schemes, inst, mode, cfg, S0, r, q), T in zip(J, bench.MATURITIES):
        assert schemes == [capi.SCHEME_NCI] and mode == capi.GREEK_FULL
        assert inst.maturity == T and cfg.time_steps == int(round(100 * T)) and cfg.seed == 42
        assert m.heston[0].sigma == 1.0 and m.jump_kind == capi.JUMP_MERTON


def test_double_heston_sets_are_fixed():
    """The batched workload's 1,024 sets come from the xoshiro stream seeded
    42 (the seed the workload is frozen with): 729 of the 2,048 factors
    violate Feller and take SchemeExact, and the path counts sum to
    263,915,406 (the figures the bench line reports)."""
    sets = bench.double_heston_sets(1024)
    assert len(sets) == 1024
    assert sum(s.count(capi.SCHEME_EXACT) for _h, s, _p in sets) == 729
    assert sum(p for _h, _s, p in sets) == 263915406
    for hs, schemes, paths in sets:
        assert 1 << 14 <= paths <= 1 << 20
        for h, sch in zip(hs, schemes):
            assert (sch == capi.SCHEME_NV) == (h.sigma ** 2 <= 4.0 * h.kappa * h.theta)


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in json.loads(out.stdout.strip().splitlines()[-1])["error"]


def test_flop_counts_of_the_batched_and_mlmc_legs():
    """bench.py's flops per path-step of the batched Double-Heston and MLMC
    legs are what the oracle's op-counting build gives for those workloads
    (oracle/count_flops.py --workload batched / mlmc)."""
    import bench
    from oracle.count_flops import count_batched, count_mlmc
    assert abs(count_mlmc()["flop_per_path_step"] - bench.FLOP_PER_PATH_STEP_MLMC) < 0.01
    got = count_batched(n_sets=1024, paths=1 << 8)["flop_per_path_step"]
    assert abs(got - bench.FLOP_PER_PATH_STEP_BATCHED) < 0.01
