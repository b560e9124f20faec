"""The host structure library under AddressSanitizer and UndefinedBehaviorSanitizer.

build() also produces lib/libtenet_host_san.so: the same sources with
-fsanitize=address,undefined and -fno-sanitize-recover=undefined.  This test
reruns the host-side CPU tests (structure compile, descriptor replay, batched
descriptors, narrow fusion, snapshots, the CPU port over compiled plans) in a
child process that loads that library (TENET_AMD_HOST_LIB) with the ASan
runtime preloaded, so an out-of-bounds access, use after free or undefined
operation anywhere in the host compiler fails the run.  Python's own
allocations are not instrumented; leak checking is off because the
interpreter keeps memory until exit.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN_LIB = os.path.join(ROOT, "tenet_amd", "lib", "libtenet_host_san.so")
HOST_TESTS = ["test_host_structure.py", "test_batch_desc.py", "test_narrow_fusion.py", "test_snapshot.py",
              "test_abi_layout.py", "test_cpu_port.py"]


def _gcc_lib(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if out.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.timeout(900)
def test_host_tests_pass_under_asan_ubsan():
    if not os.path.exists(SAN_LIB):
        from tenet_amd import build
        build.build_host_sanitized()
    asan, cxx = _gcc_lib("libasan.so"), _gcc_lib("libstdc++.so")
    if asan is None or cxx is None:
        pytest.skip("no ASan runtime next to g++")
    env = dict(os.environ)
    # libstdc++ right after the ASan runtime: its __cxa_throw interceptor binds
    # the real function at start-up, before the interpreter loads any C++
    env.update(LD_PRELOAD=f"{asan} {cxx}", TENET_AMD_HOST_LIB="libtenet_host_san.so",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"]
    cmd += [os.path.join(HERE, t) for t in HOST_TESTS]
    cmd += ["-k", "not huge"]
    p = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True)
    tail = (p.stdout[-4000:] + "\n" + p.stderr[-4000:])
    assert p.returncode == 0, tail
    assert "passed" in p.stdout, tail
    # the child really ran the sanitized build
    assert "AddressSanitizer" not in p.stderr, tail
