"""The host-only C++ (structure assembly, cell list, pair search, visiting
order, MATLAB dump reader/writer) under AddressSanitizer and
UndefinedBehaviorSanitizer: largetb_amd/csrc/host_check.cpp drives every host
path and the sanitizers abort on any invalid access, leak or undefined
operation."""
import os
import subprocess

from largetb_amd import build


def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = build.build_host_check(sanitize=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host_check ok" in r.stdout
