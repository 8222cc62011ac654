"""Pins the CPU oracle to stock FeCl through the golden vectors in tests/golden/ (written by
tests/golden/make_golden.py from FeCl built out of the reference sources)."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.vectors import dvbs2_llr, gallager_llr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "fecl_golden.npz"))
CASES = json.load(open(os.path.join(HERE, "golden", "fecl_golden.json")))["cases"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def gallager_code():
    N, K = (int(x) for x in GOLD["gallager_nk"])
    return O.Code(GOLD["gallager_row_ptr"], GOLD["gallager_col"], N, K)


_INPUTS = {}


def inputs(name):
    if name not in _INPUTS:
        code = gallager_code()
        gen = {"g_conv": lambda: gallager_llr(code.N, code.K, frames=3, ebn0=2.0, seed=11),
               "g_hard": lambda: gallager_llr(code.N, code.K, frames=2, ebn0=-1.0, seed=12),
               "d_10": lambda: dvbs2_llr(frames=3, ebn0=1.0, seed=21),
               "d_20": lambda: dvbs2_llr(frames=2, ebn0=2.0, seed=22)}[name]
        llr = gen()
        assert sha(llr) == GOLD["llr_sha_" + name].tobytes().decode(), "input generator changed"
        _INPUTS[name] = llr
    return _INPUTS[name]


def sf_table(code, sf):
    if sf.startswith("v:"):
        return O.scaling_table(code, [float(x) for x in sf[2:].split(",")])
    return O.scaling_table(code, float(sf))


def test_dvbs2_r12_matrix_matches_fecl():
    c = O.dvbs2(64800, 1, 2)
    assert sha(c.row_ptr) + sha(c.col) == GOLD["dvbs2_r12_csr_sha"].tobytes().decode()


def _run(case):
    code = gallager_code() if case["code"] == "gallager" else O.dvbs2(64800, 1, 2)
    llr = inputs(case["input"])
    kw = {}
    if case.get("state_in") == "chain7":
        prev = O.decode(code, llr, case["alg"], 7, prec=64, want=("state",))
        kw["state_in"] = prev["state"]
        kw["syst_in"] = np.random.default_rng(case["syst_in_seed"]).normal(0.0, 1.0, (llr.shape[0], code.K))
    return O.decode(code, llr, case["alg"], case["iters"], sf=sf_table(code, case["sf"]), prec=case["prec"], **kw)


@pytest.mark.parametrize("i", [i for i, c in enumerate(CASES) if "sha" in c])
def test_outputs_bit_identical_to_fecl(i):
    case = CASES[i]
    out = _run(case)
    for k, h in case["sha"].items():
        assert sha(out[k]) == h, f"{k} differs from FeCl for {case}"


@pytest.mark.parametrize("i", [i for i, c in enumerate(CASES) if "counts" in c])
def test_iteration_counts_match_fecl(i):
    case = CASES[i]
    out = _run(case)
    assert out["iters"].tolist() == case["counts"], case
