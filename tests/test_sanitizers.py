"""Host-side code under AddressSanitizer + UndefinedBehaviorSanitizer and ThreadSanitizer, with
no GPU (binaries built by `make -C tests/cpp`):

- tests/cpp/server_fake_test: the serving loop (fecl_amd/csrc/ldpc_server.cpp) on a CPU fake of
  the HIP runtime and of the engine (tests/cpp/fakedev/), whose streams are real asynchronous
  queues: requests from several threads in pageable, page-locked and device memory, admission
  by device memory, device out-of-memory at creation and while running, destroy with requests
  queued.  Every result must equal the fake decode.
- tests/cpp/host_units_test: code construction and the generator-form encoder, the archive
  reader under truncation and random damage, the autotune cache from many threads, and the
  host conversion threads shared by concurrent callers.

A sanitizer report makes the binary exit non-zero (ASan/UBSan abort; TSan exits 66)."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")


@pytest.mark.parametrize("name", ["server_fake_test", "server_fake_test_asan", "server_fake_test_tsan",
                                  "host_units_test_asan", "host_units_test_tsan"])
def test_sanitized_binary_runs_clean(name):
    exe = os.path.join(HERE, name)
    assert os.path.exists(exe), "build it with `make -C tests/cpp`"
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66"
    env["ASAN_OPTIONS"] = "detect_leaks=1"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "0 failures" in r.stdout
    assert "ThreadSanitizer" not in r.stderr and "AddressSanitizer" not in r.stderr
