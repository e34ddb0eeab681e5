"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the engine's host code: validation
of caller-supplied programs and observables, lowering and every planner mode (lower.cpp,
planner.cpp), driven by tests/cpp/sanitize_driver.cpp on valid random programs and on malformed
ones. The sanitizers instrument the host code only; no GPU is used."""
import os
import subprocess

import pytest

import helpers as H

BIN = os.path.join(H.ROOT, "tests", "_build", "sanitize_driver")


def build():
    srcs = [os.path.join(H.ROOT, "tests", "cpp", "sanitize_driver.cpp"),
            os.path.join(H.ROOT, "quantr_amd", "csrc", "lower.cpp"),
            os.path.join(H.ROOT, "quantr_amd", "csrc", "planner.cpp")]
    hdrs = [os.path.join(H.ROOT, "quantr_amd", "csrc", f) for f in ("engine.hpp", "kernels.hpp")]
    newest = max(os.path.getmtime(f) for f in srcs + hdrs)
    if os.path.exists(BIN) and os.path.getmtime(BIN) >= newest:
        return
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
    cmd = (["/opt/rocm/bin/hipcc", "-O1", "-g", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
            "-I" + os.path.join(H.ROOT, "include"), "-I" + os.path.join(H.ROOT, "quantr_amd", "csrc")]
           + san + srcs + ["-o", BIN])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-4000:]


@pytest.mark.timeout(900)
def test_host_code_under_sanitizers():
    build()
    env = dict(os.environ)
    # leak checking stays off: the HIP runtime linked into the binary keeps allocations that
    # are released at exit by design; bounds and undefined-behaviour checks stay on
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=800)
    assert r.returncode == 0, r.stdout[-6000:]
    assert "sanitize driver: ok" in r.stdout
