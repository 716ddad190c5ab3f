"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds the library's host code (model.cpp, engine.cpp, server.cpp) and the oracle with -fsanitize=address,undefined
(no recovery from undefined behaviour) into a temporary directory, then runs the CPU tests of the plan builder, the
device-LU emulation, the fresh host plans, option checks and the oracle against those builds in a child process with
the ASan runtime preloaded. engine.cpp holds a __device__ array, so hipcc compiles it with the sanitizers on its host
pass only (-Xarch_host). The kernels' launchers are replaced by stubs that report no device: there is no GPU here, and
the device code is not host code. Leak checking is off (the Python interpreter itself does not free everything).
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gridsim_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = "/opt/rocm/bin/hipcc"

STUBS = r'''
#include "engine_internal.h"
namespace gs {
hipError_t launch_nr_batch(int, int, int, const InstDesc*, const DevPlan&, int, const NrParams&, size_t, hipStream_t) {
  return hipErrorNoDevice;
}
hipError_t launch_group(int, int, int, const InstDesc*, const DevPlan&, int, const NrParams&, size_t, hipStream_t) {
  return hipErrorNoDevice;
}
hipError_t launch_seam(const SeamDesc*, const DevPlan&, int, size_t, hipStream_t) { return hipErrorNoDevice; }
hipError_t launch_sincos(const double*, double*, double*, int64_t, hipStream_t) { return hipErrorNoDevice; }
hipError_t launch_iterative_step(const InstDesc*, const DevPlan&, const NrParams&, double*, int*, hipStream_t) {
  return hipErrorNoDevice;
}
hipError_t launch_screen(const InstDesc*, const DevPlan&, int, gs_screen_summary*, double*, double*, hipStream_t) {
  return hipErrorNoDevice;
}
hipError_t launch_wide_start(const InstDesc*, const DevPlan&, int, WideCtl*, hipStream_t) { return hipErrorNoDevice; }
hipError_t launch_wide_step(int, const InstDesc*, const DevPlan&, int, const NrParams&, WideCtl*,
                            const std::vector<int>&, const std::vector<int>&, const std::vector<int>&, size_t,
                            const WideSplit&, hipStream_t) {
  return hipErrorNoDevice;
}
hipError_t launch_fp64_probe(double*, hipStream_t) { return hipErrorNoDevice; }
}  // namespace gs
'''

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
HOST = ["-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", INCLUDE, "-I", CSRC]


def _gcc_library(name):
    try:
        path = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True,
                              check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return path if os.path.isabs(path) and os.path.exists(path) else None


def _asan_runtime():
    """The ASan runtime, then libstdc++ (preloaded with it: the interpreter does not link libstdc++, and ASan's
    exception interceptor needs the real __cxa_throw at start-up)."""
    asan, stdcxx = _gcc_library("libasan.so"), _gcc_library("libstdc++.so")
    return f"{asan}:{stdcxx}" if asan and stdcxx else None


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    runtime = _asan_runtime()
    if runtime is None or not os.path.exists(HIPCC) or shutil.which("g++") is None:
        pytest.skip("no ASan runtime or compilers")
    out = str(tmp_path_factory.mktemp("sanitized"))
    with open(os.path.join(out, "stubs.cpp"), "w") as f:
        f.write(STUBS)
    hip_host = ["-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include"]
    jobs = {
        "model.o": ["g++"] + HOST + SAN + ["-c", os.path.join(CSRC, "model.cpp")],
        "server.o": ["g++"] + HOST + SAN + hip_host + ["-c", os.path.join(CSRC, "server.cpp")],
        "stubs.o": ["g++"] + HOST + SAN + hip_host + ["-c", os.path.join(out, "stubs.cpp")],
        # the sanitizers on the host pass only
        "engine.o": [HIPCC, "-x", "hip"] + HOST + ["-Xarch_host", "-fsanitize=address", "-Xarch_host",
                                                   "-fsanitize=undefined", "-Xarch_host",
                                                   "-fno-sanitize-recover=undefined", "--offload-arch=gfx950",
                                                   "-fno-gpu-rdc", "-c", os.path.join(CSRC, "engine.cpp")],
        "liboracle.so": ["g++"] + HOST + SAN + ["-shared", os.path.join(ROOT, "oracle", "oracle.cpp"), "-lpthread"],
    }
    procs = {k: subprocess.Popen(cmd + ["-o", os.path.join(out, k)], stdout=subprocess.PIPE,
                                 stderr=subprocess.STDOUT, text=True) for k, cmd in jobs.items()}
    for k, p in procs.items():
        log = p.communicate()[0]
        assert p.returncode == 0, f"{k}: {log[-3000:]}"
    lib = os.path.join(out, "libgridsim_hip.so")
    subprocess.run(["g++", "-shared"] + SAN + ["-o", lib] + [os.path.join(out, o) for o in
                                                            ("model.o", "engine.o", "server.o", "stubs.o")] +
                   ["-L/opt/rocm/lib", "-lamdhip64", "-lpthread"], check=True)
    return runtime, lib, os.path.join(out, "liboracle.so")


def test_host_code_is_clean_under_address_and_undefined_behaviour_sanitizers(sanitized):
    runtime, lib, orc = sanitized
    env = dict(os.environ, LD_PRELOAD=runtime, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", GRIDSIM_HIP_LIB=lib, GRIDSIM_ORACLE_LIB=orc)
    tests = [os.path.join(ROOT, "tests", t) for t in ("test_engine_cpu.py", "test_oracle.py", "test_oracle_retry.py",
                                                      "test_nk_screen.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + tests,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    report = r.stdout[-4000:] + r.stderr[-8000:]
    assert r.returncode == 0, report
    assert "runtime error" not in report and "AddressSanitizer" not in report, report
