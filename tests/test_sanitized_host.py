"""The host code of libgfx (mesh generation, topology, HDF5 input/output, the
solver's host plan) under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C gf-only-sem_amd sanitize` (part of __graft_entry__.build()) builds
libgfx_san.so: every host translation unit instrumented (device code as
usual).  The CPU test files that drive that code run again in a child process
that loads it through GFX_LIB with the clang ASan runtime preloaded; any
out-of-bounds access, use after free or undefined behaviour aborts the child.
No GPU is used (the host plan stops at its first device call)."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "gf-only-sem_amd", "libgfx_san.so")
FILES = ["tests/test_host_plan.py", "tests/test_meshgen.py", "tests/test_topology.py", "tests/test_io.py",
         "tests/test_cpml_complete.py"]


def _asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def test_host_code_is_clean_under_asan_and_ubsan():
    rt = _asan_runtime()
    if not os.path.exists(SAN_LIB) or rt is None:
        pytest.skip("sanitizer build not present (make -C gf-only-sem_amd sanitize)")
    syms = subprocess.run(["nm", "-D", SAN_LIB], capture_output=True, text=True, check=True).stdout
    assert "__asan_report" in syms and "__ubsan_handle" in syms  # really instrumented
    env = dict(os.environ, GFX_LIB=SAN_LIB, LD_PRELOAD=rt, HIP_VISIBLE_DEVICES="",
               ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        *FILES], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
