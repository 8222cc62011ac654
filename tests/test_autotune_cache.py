"""Autotune cache (fecl_amd/csrc/ldpc_tune_cache.cpp): configuration, the file format and its
error handling run on the CPU; the GPU test checks that a second autotune of the same key pins
the kept pick without probing."""
import os

import numpy as np
import pytest

import fecl_amd as F

HEADER = "fecl-autotune 1\n"


@pytest.fixture(autouse=True)
def _restore():
    yield
    F.autotune_cache(1)
    F.autotune_cache_clear()


def test_file_is_created_with_its_header(tmp_path):
    p = tmp_path / "tune.txt"
    F.autotune_cache(2, str(p))
    assert p.read_text() == HEADER


def test_an_empty_file_gets_its_header(tmp_path):
    p = tmp_path / "tune.txt"
    p.write_text("")
    F.autotune_cache(2, str(p))
    assert p.read_text() == HEADER
    st = F.autotune_cache_stats()
    assert st["entries"] == 0 and st["write_failures"] == 0


def test_file_entries_load_and_malformed_lines_are_skipped(tmp_path):
    p = tmp_path / "tune.txt"
    p.write_text(HEADER
                 + "gfx950/256|n64800 m32400 e226799 h0123456789abcdef|alg0|fmt1|f1024\t1 1 3.5 3.6 -1 -1 -1 -1\n"
                 + "gfx950/256|n64800 m32400 e226799 h0123456789abcdef|alg0|fmt0|f1024\t4 1 3.8 3.7 3.4 3.4 3.43 -1\n"
                 + "no tab on this line\n"
                 + "gfx950/256|x\t3 1 1 1 1 1 1 1\n"      # 3 words is no layout
                 + "gfx950/256|y\t1 2 1 1\n")             # too few times
    F.autotune_cache_clear()
    F.autotune_cache(2, str(p))
    st = F.autotune_cache_stats()
    assert st["entries"] == 2 and st["skipped_lines"] == 3 and st["hits"] == 0


def test_bad_header_and_bad_arguments_are_errors(tmp_path):
    p = tmp_path / "tune.txt"
    p.write_text("something else\n")
    with pytest.raises(ValueError, match="not a cache file"):
        F.autotune_cache(2, str(p))
    with pytest.raises(ValueError, match="needs a file path"):
        F.autotune_cache(2)
    with pytest.raises(ValueError, match="mode must be"):
        F.autotune_cache(3)
    with pytest.raises(ValueError, match="cannot create"):
        F.autotune_cache(2, str(tmp_path / "missing_dir" / "tune.txt"))


@pytest.mark.gpu
def test_second_autotune_pins_the_kept_pick(tmp_path):
    import torch
    p = tmp_path / "tune.txt"
    F.autotune_cache_clear()
    F.autotune_cache(2, str(p))
    H = F.DvbS2.matrix(64800, 0.5)
    opts = F.DecoderOptions().algorithm(F.Approximate).iterations(50).msgFormat(F.MsgFormat.F16)
    a = F.Ldpc(H, opts)
    best_a, times_a = a.autotune(256, probe_iterations=3)
    s1 = F.autotune_cache_stats()
    assert s1["misses"] == 1 and s1["entries"] == 1
    lines = p.read_text().splitlines()
    assert lines[0] == HEADER.strip() and len(lines) == 2 and "|f256" in lines[1]
    # another handle of the same code and batch-size bucket: no probing, same pick
    b = F.Ldpc(H, opts)
    best_b, times_b = b.autotune(200, probe_iterations=3)
    s2 = F.autotune_cache_stats()
    assert s2["hits"] == 1 and (best_b, b.varsPerWave()) == (best_a, a.varsPerWave())
    assert times_b == pytest.approx(times_a, rel=1e-5)
    # a fresh process view: forget the picks, read the file again
    F.autotune_cache_clear()
    F.autotune_cache(2, str(p))
    c = F.Ldpc(H, opts)
    assert c.autotune(256, probe_iterations=3)[0] == best_a and F.autotune_cache_stats()["hits"] == 1
    # another batch-size bucket or message format is probed
    c.autotune(1024, probe_iterations=2)
    d = F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(50))
    d.autotune(256, probe_iterations=2)
    assert F.autotune_cache_stats()["misses"] == 2 and len(p.read_text().splitlines()) == 4
    # decoding with a kept pick is unchanged
    x = torch.randn(64800 * 8, device="cuda") * 3
    torch.cuda.synchronize()
    assert torch.equal(a.decode(x), b.decode(x))
