"""The drop-in boundary without a GPU: the C-ABI library loads and exports
every entry point include/myql_capi.h declares, the ctypes mirror has the C
header's layout, calls on a null context fail with MYQL_E_INVALID_ARG
(no device work), and the C++20 headers under include/myql compile and link
as reference-style user code (tests/cpp/dropin_check.cpp)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from myquantlib_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "myql_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(myql_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_exactly_the_exported_list():
    assert declared_functions() == sorted(capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    defined = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in defined]
    assert not missing, missing


STRUCTS = {
    "myql_heston_params": capi.HestonParams,
    "myql_model": capi.Model,
    "myql_instrument": capi.Instrument,
    "myql_mc_config": capi.MCConfig,
    "myql_mc_result": capi.MCResult,
    "myql_mc_job": capi.MCJob,
    "myql_mlmc_config": capi.MLMCConfig,
    "myql_mlmc_result": capi.MLMCResult,
    "myql_fourier_config": capi.FourierConfig,
    "myql_fourier_job": capi.FourierJob,
    "myql_cv_config": capi.CVConfig,
    "myql_cv_net": capi.CVNet,
    "myql_cv_result": capi.CVResult,
}


def test_ctypes_layout_matches_c_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "myql_capi.h"',
             "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_null_context_is_rejected_without_device_work():
    lib = capi.load_library()
    assert lib.myql_last_error(None) == b"null context"
    m, inst, cfg, res = capi.Model(), capi.Instrument(), capi.MCConfig(), capi.MCResult()
    schemes = (C.c_int32 * 2)(0, 0)
    st = lib.myql_mc_calculate(None, C.byref(m), schemes, C.byref(inst), 0, C.byref(cfg), 100.0,
                               0.0, 0.0, C.byref(res))
    assert st == capi.E_INVALID_ARG
    assert lib.myql_mc_calculate_batch(None, None, 0) == capi.E_INVALID_ARG
    mres = capi.MLMCResult()
    mcfg = capi.MLMCConfig()
    st = lib.myql_mlmc_calculate(None, C.byref(m), schemes, C.byref(inst), 0, C.byref(mcfg), 100.0,
                                 0.0, 0.0, C.byref(mres))
    assert st == capi.E_INVALID_ARG
    assert lib.myql_fourier_calculate_batch(None, None, 0) == capi.E_INVALID_ARG
    buf = C.create_string_buffer(64)
    assert lib.myql_tuning_report(None, buf, len(buf)) == capi.E_INVALID_ARG
    assert lib.myql_debug_set_launch_shape(None, capi.FP32, 1, 5, 0, 256, 4, 4, 0) == capi.E_INVALID_ARG


def test_cpp_dropin_headers_compile_and_link(tmp_path):
    """The reference-style program builds against include/myql with the
    reference's own header paths and class templates, and links to the
    engine library (it is run on the GPU by test_gpu_dropin.py)."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp")
    exe = tmp_path / "dropin_check"
    out = subprocess.run(["g++", "-std=c++20", "-O0", "-Wall", "-Wextra", "-Werror", "-I",
                          os.path.join(ROOT, "include"), src, "-L",
                          os.path.dirname(capi.LIB_PATH), "-lmyql_hip", "-o", str(exe)],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert exe.exists()


REF_INCLUDE = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF_INCLUDE, "myql", "dispatcher")),
                    reason="needs the reference's dispatcher headers")
def test_reference_dispatcher_needs_only_dropin_headers(tmp_path):
    """The reference's runtime dispatcher (myql/dispatcher/*.hpp), reached
    through links placed first on the include path (as oracle/Makefile
    builds oracle/_ref/dispatch_check), takes every other myql/ header it
    includes from the drop-in: the reference supplies only the dispatcher."""
    stage = tmp_path / "myql" / "dispatcher"
    stage.mkdir(parents=True)
    ref_disp = os.path.join(REF_INCLUDE, "myql", "dispatcher")
    for h in os.listdir(ref_disp):
        os.symlink(os.path.join(ref_disp, h), stage / h)
    src = os.path.join(ROOT, "tests", "cpp", "dispatch_check.cpp")
    out = subprocess.run(["g++", "-std=c++20", "-M", "-I", str(tmp_path), "-I", os.path.join(ROOT, "include"),
                          "-I", REF_INCLUDE, src], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    deps = out.stdout.replace("\\\n", " ").split()
    staged = sorted(os.path.basename(d) for d in deps if d.startswith(str(tmp_path)))
    assert staged == sorted(os.listdir(ref_disp)), staged
    assert not [d for d in deps if d.startswith(REF_INCLUDE)]


def test_dropin_dispatcher_is_self_contained():
    """The drop-in's own dispatcher (include/myql/dispatcher) compiles from
    include/ alone: every myql/ header the caller reaches is the drop-in's."""
    src = os.path.join(ROOT, "tests", "cpp", "dispatch_check.cpp")
    out = subprocess.run(["g++", "-std=c++20", "-M", "-I", os.path.join(ROOT, "include"), src],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    deps = [d for d in out.stdout.replace("\\\n", " ").split() if "/myql/" in d]
    assert any(d.endswith("myql/dispatcher/PricingDispatch.hpp") for d in deps)
    assert all(d.startswith(os.path.join(ROOT, "include")) for d in deps), deps


@pytest.mark.parametrize("snippet,needle", [
    # a stepper whose tracker differs from the instrument's
    ("MonteCarloPricer<HestonModel, HestonStepper<SchemeNV>, UpAndOutCall<>> p(m, c);",
     "tracker"),
    # a two-factor stepper on a one-factor model
    ("MonteCarloPricer<HestonModel, DoubleHestonStepper<SchemeNV, SchemeNV>, "
     "EuropeanOption<PayoffVanilla<OptionType::Call>>> p(m, c);"
     " (void)p.calculate(100, 0, 0, EuropeanOption<PayoffVanilla<OptionType::Call>>(100.0, 1.0));",
     "factor count"),
])
def test_cpp_type_mismatches_fail_to_compile(tmp_path, snippet, needle):
    src = tmp_path / "bad.cpp"
    src.write_text("#include <myql/instruments/options/Barrier.hpp>\n"
                   "#include <myql/instruments/options/European.hpp>\n"
                   "#include <myql/pricers/montecarlo/MonteCarloPricer.hpp>\n"
                   "int main() { HestonModel m(HestonParams{2, .04, .3, -.7, .04});"
                   " MonteCarloConfig c; " + snippet + " return 0; }\n")
    out = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                          str(src)], capture_output=True, text=True)
    assert out.returncode != 0
    assert needle in out.stderr


def test_instrument_structs_keep_their_strikes_alive():
    """One InstrumentSpec backing several jobs of a batch: every struct it
    returned must keep pointing at live strikes."""
    import gc

    from myquantlib_amd.pricing import InstrumentSpec
    spec = InstrumentSpec(strikes=[90.0, 100.0, 110.0])
    structs = [spec.struct() for _ in range(4)]
    spec.strikes = [1.0, 2.0, 3.0]
    spec.struct()
    gc.collect()
    junk = [bytearray(1 << 16) for _ in range(64)]  # reuse freed memory if any
    for s in structs:
        assert [s.strikes[i] for i in range(3)] == [90.0, 100.0, 110.0]
    del junk
