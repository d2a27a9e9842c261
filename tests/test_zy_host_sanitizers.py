"""The host compiler, VCD reader, C API and CPU oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (gcc, host code only, no
GPU code involved), running the CPU tests that exercise them: netlist
generation, partitioning and multi-level mapping, partition merge, flatten,
the validator, features, gather tables, the lane layout, timing words,
the input VCD reader, the output VCD writers, and the oracle's
interpreter, AIG simulator and planted-mistake builds.  Any invalid access or undefined behaviour aborts
the run."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the host library's sources, as jacquard_amd/build.py lists them
HOST_SRC = ["aig.cpp", "netgen.cpp", "compile.cpp", "verilog.cpp", "capi.cpp", "vcd.cpp"]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
TESTS = ["test_compile_direct.py", "test_batched_encoding.py", "test_timing_pins.py",
         "test_vcd_reader.py", "test_vcd_writer.py", "test_golden_vcd.py",
         "test_capi_vcd_events.py", "test_oracle_kats.py", "test_word_symbols.py"]


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True,
                         text=True, check=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def _asan_runtime():
    # the C++ runtime is preloaded with the sanitizer runtime: Python does
    # not link it, and ASan's exception interceptor must find it at start
    a, c = _runtime("libasan.so"), _runtime("libstdc++.so")
    return f"{a} {c}" if a and c else None


@pytest.mark.skipif(_asan_runtime() is None, reason="gcc's AddressSanitizer runtime is absent")
def test_host_libraries_under_asan_and_ubsan(tmp_path):
    d = str(tmp_path)
    src = [os.path.join(ROOT, "jacquard_amd", "csrc", "host", f) for f in HOST_SRC]
    src = [f for f in src if os.path.exists(f)]
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
                           *SAN, "-o", os.path.join(d, "libjacquard_host.so"), *src])
    subprocess.check_call(["gcc", "-O1", "-g", "-std=c11", "-fPIC", "-shared", "-fopenmp",
                           *SAN, "-o", os.path.join(d, "libjq_oracle.so"),
                           os.path.join(ROOT, "oracle", "jq_oracle.c")])
    env = dict(os.environ,
               LD_PRELOAD=_asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               JACQUARD_HOST_LIB_DIR=d, JACQUARD_ORACLE_LIB_DIR=d, OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider",
                        *[os.path.join(ROOT, "tests", t) for t in TESTS]],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert " passed" in r.stdout, tail
