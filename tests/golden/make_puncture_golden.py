"""Writes tests/golden/fecl_puncture.json: FeCl's Ldpc::Structure::puncturing
(src/detail/Ldpc.cpp:333-353) for a set of masks, computed by stock FeCl built from the
reference sources (oracle/ref.mk -> oracle/_ref/fecl_golden).  Each case stores the
permutation's input size, output size and the SHA-256 of its int64 index sequence."""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
DRV = os.path.join(ROOT, "oracle", "_ref", "fecl_golden")

CODES = ["gallager:2048:3:5:0", "dvbs2:64800:0.5"]
MASKS = [("-", "-"), ("1,0", "-"), ("1,0", "1"), ("-", "1,0"), ("1,1,0", "1,0,1"), ("0,1", "0"), ("1,0,0,1", "-"),
         ("1", "0,1,1"), ("0,0,1", "1,1,0,1")]


def main():
    if not os.path.exists(DRV):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    cases = []
    with tempfile.TemporaryDirectory() as tmp:
        for code in CODES:
            csr = os.path.join(tmp, "h.csr")
            if code.startswith("gallager"):
                subprocess.check_call([DRV, "csr-gallager", *code.split(":")[1:], csr])
            else:
                n, rate = code.split(":")[1:]
                subprocess.check_call([DRV, "csr-dvbs2", n, rate, csr])
            M, N, K, E = (int(x) for x in np.fromfile(csr, np.int64, 4))
            for mask, syst in MASKS:
                out = os.path.join(tmp, "p.bin")
                subprocess.check_call([DRV, "puncture", code, mask, syst, out])
                d = np.fromfile(out, np.int64)
                seq = d[2:]
                cases.append({"code": code, "K": K, "N": N, "mask": mask, "syst_mask": syst,
                              "input_size": int(d[0]), "output_size": int(d[1]),
                              "sha256": hashlib.sha256(seq.astype("<i8").tobytes()).hexdigest(),
                              "head": [int(x) for x in seq[:8]], "tail": [int(x) for x in seq[-8:]]})
    with open(os.path.join(HERE, "fecl_puncture.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_puncture_golden.py (stock FeCl via oracle/ref.mk)",
                   "cases": cases}, f, indent=1)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
