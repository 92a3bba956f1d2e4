"""The host-side C under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): the
library's host layer (text and binary readers and writers, the %lf-exact printer, argument
checks, the serve stream's request generator, the phantom generators) and the oracle are built
with -fsanitize=address,undefined (wrt-project_amd/Makefile and oracle/Makefile, target
`sanitize`), and the CPU tests that exercise them run against those builds in a child process
with the ASan runtime preloaded; any report fails the child (halt on error).  Leak detection is
off: the Python interpreter itself holds memory to exit."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wrt-project_amd")
SAN_LIB = os.path.join(PKG, "lib", "san", "libwrt.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
# the CPU tests that run the host C and the oracle
FILES = ["test_io.py", "test_serve_host.py", "test_abi.py", "test_oracle.py",
         "test_oracle_pins.py", "test_oracle_reduction_rays.py", "test_cli.py"]


def _asan_runtime():
    try:
        out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_host_c_under_asan_ubsan():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("gcc's ASan runtime is not installed")
    for d in (PKG, os.path.join(ROOT, "oracle")):
        r = subprocess.run(["make", "-s", "-j8", "sanitize"], cwd=d, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, WRT_LIB=SAN_LIB, ORACLE_LIB=SAN_ORACLE, LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    # the child loads the sanitizer builds (and the ASan runtime), nothing else
    probe = ("import sys; sys.path[:0] = [%r, %r]; import wrt, orc; wrt.lib(); orc.L(); "
             "print(open('/proc/self/maps').read())" % (PKG, os.path.join(ROOT, "tests")))
    maps = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True)
    assert maps.returncode == 0, maps.stderr[-3000:]
    assert SAN_LIB in maps.stdout and SAN_ORACLE in maps.stdout and "libasan" in maps.stdout
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider"] + FILES,
                       cwd=os.path.join(ROOT, "tests"), env=env, capture_output=True, text=True)
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    assert "passed" in r.stdout
