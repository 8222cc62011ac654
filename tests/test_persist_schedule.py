"""The persistent multi-code decode's work queue (fecl_amd/csrc/ldpc_persist.h), checked on the
CPU by tests/cpp/persist_schedule_test: walked in queue order, every item's wait is already
satisfied by earlier items (so the launch cannot deadlock whatever workgroups run), and the
queue lists each (kind, code, tile, chunk, step) once."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_persistent_queue_order_satisfies_every_wait():
    exe = os.path.join(HERE, "cpp", "persist_schedule_test")
    assert os.path.exists(exe), "build it with `make -C tests/cpp`"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
