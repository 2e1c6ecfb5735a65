"""The host code under its checked build (CPU only).

libmercury_hip_checked.so is the same library with bounds-checked standard
containers (_GLIBCXX_ASSERTIONS) and trapping undefined-behaviour checks in
every host translation unit (mpcns_mercury_amd/build.py, CHECK_FLAGS): an
out-of-range container index, a signed overflow, a bad shift or a null
dereference in the plan builders, the CASE reader, the geometry or the restart
code stops the process.  This runs the host-side tests against it in a child
process (MERCURY_HIP_LIB selects the library at load time).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECKED = os.path.join(ROOT, "mpcns_mercury_amd", "libmercury_hip_checked.so")
HOST_TESTS = ["tests/test_abi_cpu.py", "tests/test_case_reader.py", "tests/test_restart_format.py"]


@pytest.mark.skipif(not os.path.exists(CHECKED), reason="checked library not built (__graft_entry__.build builds it)")
def test_host_code_passes_its_tests_under_the_checked_build():
    env = dict(os.environ, MERCURY_HIP_LIB=CHECKED)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider"] + HOST_TESTS,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert " passed" in r.stdout


def test_checked_flags_reach_the_host_pass_only():
    from mpcns_mercury_amd import build
    flags = build.CHECK_FLAGS
    for i, f in enumerate(flags):
        if f.startswith("-fsanitize") or f.startswith("-fno-sanitize"):
            assert i > 0 and flags[i - 1] == "-Xarch_host", f
    assert "-fsanitize-trap=undefined" in flags and "-D_GLIBCXX_ASSERTIONS" in flags
