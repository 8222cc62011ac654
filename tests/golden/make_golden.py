"""Writes tests/golden/fecl_golden.npz from stock FeCl (oracle/_ref/fecl_golden, built from
/root/reference by oracle/ref.mk).  Run from the repo root:  python tests/golden/make_golden.py

Inputs are regenerated from seeds in tests (numpy PCG64) and their SHA-256 is stored so a
change of generator is caught.  Outputs are stored as SHA-256 of the raw little-endian bytes
(bit-exact comparison) plus packed hard bits and per-frame iteration counts.  Iteration counts
are derived from FeCl itself: the smallest budget I0 for which soDecode's state output equals
the output at the full budget for every budget in [I0, full] (FeCl does not expose the count).
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from tests.golden.vectors import gallager_llr, dvbs2_llr  # noqa: E402

DRV = os.path.join(ROOT, "oracle", "_ref", "fecl_golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def read_out(path, dtype):
    raw = np.fromfile(path, dtype=np.uint8)
    F, L = np.frombuffer(raw[:16].tobytes(), dtype=np.int64)
    return np.frombuffer(raw[16:].tobytes(), dtype=dtype).reshape(F, L)


def read_csr(path):
    a = np.fromfile(path, dtype=np.int64)
    M, N, K, E = (int(x) for x in a[:4])
    return a[4:4 + M + 1].astype(np.int32), a[4 + M + 1:].astype(np.int32), N, K


def write_llr(path, llr):
    with open(path, "wb") as f:
        np.array(llr.shape, np.int64).tofile(f)
        np.ascontiguousarray(llr, np.float64).tofile(f)


def run(tmp, code, llr_path, alg, iters, sf, prec, tag, state_in="-", syst_in="-"):
    pre = os.path.join(tmp, tag)
    subprocess.check_call([DRV, "decode", code, llr_path, str(alg), str(iters), sf, str(prec), pre,
                           state_in, syst_in])
    return {k: read_out(pre + "." + k, np.uint8 if k == "hard" else np.float64)
            for k in ("hard", "msg", "parity", "state")}


def counts(tmp, code, llr_path, alg, iters, sf, prec):
    out = os.path.join(tmp, "iters.bin")
    subprocess.check_call([DRV, "iters", code, llr_path, str(alg), str(iters), sf, str(prec), out])
    return read_out(out, np.int32).ravel()


def main():
    if not os.path.exists(DRV):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    arrays, cases = {}, []
    with tempfile.TemporaryDirectory() as tmp:
        # --- Gallager(2048, 3, 5, seed 0): the reference's own test code (test/Cpp/Ldpc.cpp:95).
        subprocess.check_call([DRV, "csr-gallager", "2048", "3", "5", "0", os.path.join(tmp, "g.csr")])
        rp, col, N, K = read_csr(os.path.join(tmp, "g.csr"))
        arrays["gallager_row_ptr"], arrays["gallager_col"] = rp, col
        arrays["gallager_nk"] = np.array([N, K], np.int64)
        gcode = "gallager:2048:3:5:0"
        inputs = {"g_conv": gallager_llr(N, K, frames=3, ebn0=2.0, seed=11),
                  "g_hard": gallager_llr(N, K, frames=2, ebn0=-1.0, seed=12)}
        for name, llr in inputs.items():
            write_llr(os.path.join(tmp, name), llr)
            arrays["llr_sha_" + name] = np.frombuffer(sha(llr).encode(), np.uint8)
        sf_vec5 = "v:0.5,0.625,0.75,0.875,1.0"
        for inp in inputs:
            for alg in (0, 1, 2):
                for prec in (64, 32):
                    for iters, sf in [(1, "1.0"), (2, "1.0"), (5, "1.0"), (30, "1.0"), (50, "1.0"),
                                      (30, "0.75"), (30, "0.9"), (5, sf_vec5)]:
                        o = run(tmp, gcode, os.path.join(tmp, inp), alg, iters, sf, prec, "x")
                        cases.append({"code": "gallager", "input": inp, "alg": alg, "prec": prec,
                                      "iters": iters, "sf": sf,
                                      "sha": {k: sha(v) for k, v in o.items()}})
                    c = counts(tmp, gcode, os.path.join(tmp, inp), alg, 50, "1.0", prec)
                    cases.append({"code": "gallager", "input": inp, "alg": alg, "prec": prec,
                                  "iters": 50, "sf": "1.0", "counts": c.tolist()})
        # soDecode with a-priori state and systematic inputs (BpDecoderImpl.cpp:73-93).
        llr = inputs["g_hard"]
        for alg in (0, 1, 2):
            o = run(tmp, gcode, os.path.join(tmp, "g_hard"), alg, 7, "1.0", 64, "s1")
            write_llr(os.path.join(tmp, "state_in"), o["state"])
            syst = np.random.default_rng(99).normal(0.0, 1.0, (llr.shape[0], K))
            write_llr(os.path.join(tmp, "syst_in"), syst)
            o2 = run(tmp, gcode, os.path.join(tmp, "g_hard"), alg, 7, "1.0", 64, "s2",
                     os.path.join(tmp, "state_in"), os.path.join(tmp, "syst_in"))
            # (the driver's hard bits come from decode(), which takes no a-priori input)
            cases.append({"code": "gallager", "input": "g_hard", "alg": alg, "prec": 64, "iters": 7,
                          "sf": "1.0", "state_in": "chain7", "syst_in_seed": 99,
                          "sha": {k: sha(v) for k, v in o2.items() if k != "hard"}})

        # --- DVB-S2 normal frame, rate 1/2: the matrix and a few frames.
        subprocess.check_call([DRV, "csr-dvbs2", "64800", "0.5", os.path.join(tmp, "d.csr")])
        rp, col, N, K = read_csr(os.path.join(tmp, "d.csr"))
        arrays["dvbs2_r12_csr_sha"] = np.frombuffer((sha(rp) + sha(col)).encode(), np.uint8)
        dcode = "dvbs2:64800:0.5"
        dinputs = {"d_10": dvbs2_llr(frames=3, ebn0=1.0, seed=21),
                   "d_20": dvbs2_llr(frames=2, ebn0=2.0, seed=22)}
        for name, llr in dinputs.items():
            write_llr(os.path.join(tmp, name), llr)
            arrays["llr_sha_" + name] = np.frombuffer(sha(llr).encode(), np.uint8)
        for inp in dinputs:
            for alg in (0, 1, 2):
                for prec in (64, 32):
                    o = run(tmp, dcode, os.path.join(tmp, inp), alg, 50, "1.0", prec, "d")
                    cases.append({"code": "dvbs2_r12", "input": inp, "alg": alg, "prec": prec,
                                  "iters": 50, "sf": "1.0", "sha": {k: sha(v) for k, v in o.items()},
                                  "hard_ones": int(o["hard"].sum())})
            o = run(tmp, dcode, os.path.join(tmp, inp), 2, 50, "0.75", 32, "d")
            cases.append({"code": "dvbs2_r12", "input": inp, "alg": 2, "prec": 32, "iters": 50,
                          "sf": "0.75", "sha": {k: sha(v) for k, v in o.items()}})
        for inp in dinputs:
            for alg, prec in ((2, 32), (2, 64), (1, 64), (0, 64)):
                c = counts(tmp, dcode, os.path.join(tmp, inp), alg, 50, "1.0", prec)
                cases.append({"code": "dvbs2_r12", "input": inp, "alg": alg, "prec": prec,
                              "iters": 50, "sf": "1.0", "counts": c.tolist()})
    out = os.path.join(ROOT, "tests", "golden")
    np.savez_compressed(os.path.join(out, "fecl_golden.npz"), **arrays)
    with open(os.path.join(out, "fecl_golden.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py (stock FeCl via oracle/ref.mk)",
                   "cases": cases}, f, indent=0)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
