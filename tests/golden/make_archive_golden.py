"""Writes tests/golden/fecl_archive.json and fecl_archive.npz: codecs saved by stock FeCl
(fec::detail::save of a fec::Ldpc through a `const fec::Codec*`, types Convolutional, Turbo,
Ldpc registered as FeCl's Matlab wrapper does; src/Serialization.h:81-101, wrap/wrap.h:35),
and what stock FeCl's loader reads back from them (fec::detail::load, src/Serialization.h:103-115).
FeCl is built from the reference sources by oracle/ref.mk (oracle/_ref/fecl_golden).

For each case: the code (a Gallager shape, a hand-made matrix from fecl_structure.npz, or a
DVB-S2 code), the options and work group size, the archive (full bytes for small ones, length
and SHA-256 for all), and the loaded sizes, work group size, algorithm, iterations and the
SHA-256 of the loaded checks()."""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
DRV = os.path.join(ROOT, "oracle", "_ref", "fecl_golden")

# name, code, algorithm, iterations, scaling ("f" or "v:f0,f1,..."), work group size
CASES = [("gallager30_minsum", {"gallager": [30, 3, 5, 0]}, 2, 10, "1", 8),
         ("gallager30_linear_vector", {"gallager": [30, 3, 5, 0]}, 1, 5, "v:0.5,0.625,0.75,0.875,1", 4),
         ("gallager2048_exact", {"gallager": [2048, 3, 5, 0]}, 0, 30, "1", 8),
         ("gallager1000_dependent", {"gallager": [1000, 3, 5, 1]}, 2, 20, "0.75", 16),
         ("dependent_rows", {"csr": "dependent_rows_30x80", "N": 80}, 2, 3, "0.9", 8),
         ("staircase_repeat", {"csr": "staircase_repeat_21x50", "N": 50}, 1, 7, "1", 2),
         ("dvbs2_16200_r12", {"dvbs2": [16200, 1, 2]}, 2, 50, "1", 8)]
FULL_BYTES = 8192


def matrix(code):
    import fecl_amd as F
    if "gallager" in code:
        return F.Gallager.matrix(*code["gallager"])
    if "dvbs2" in code:
        n, a, b = code["dvbs2"]
        return F.DvbS2.matrix(n, (a, b))
    z = np.load(os.path.join(HERE, "fecl_structure.npz"))
    return F.CheckMatrix(z[code["csr"] + "_row_ptr"], z[code["csr"] + "_col"], code["N"])


def csr_sha(rp, col):
    h = hashlib.sha256()
    h.update(np.asarray(rp, "<i4").tobytes())
    h.update(np.asarray(col, "<i4").tobytes())
    return h.hexdigest()


def main():
    if not os.path.exists(DRV):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    cases, arrays = [], {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, code, alg, iters, sf, wgs in CASES:
            H = matrix(code)
            csr = os.path.join(tmp, "h.csr")
            with open(csr, "wb") as f:
                np.asarray([H.rows(), H.cols(), H.cols() - H.rows(), H.size()], "<i8").tofile(f)
                np.asarray(H.row_ptr, "<i8").tofile(f)
                np.asarray(H.col, "<i8").tofile(f)
            arc = os.path.join(tmp, "a.bin")
            subprocess.check_call([DRV, "save", csr, str(alg), str(iters), sf, str(wgs), arc])
            data = open(arc, "rb").read()
            ld = os.path.join(tmp, "ld.bin")
            subprocess.check_call([DRV, "load", arc, ld])
            hdr = np.fromfile(ld, "<i8")
            d = np.fromfile(ld + ".csr", "<i8")
            M, N, K, E = (int(x) for x in d[:4])
            rp, col = d[4:4 + M + 1], d[4 + M + 1:4 + M + 1 + E]
            if len(data) <= FULL_BYTES:
                arrays[name] = np.frombuffer(data, np.uint8)
            cases.append({"name": name, "code": code, "algorithm": alg, "iterations": iters, "scaling": sf,
                          "work_group_size": wgs, "archive_bytes": len(data),
                          "archive_sha256": hashlib.sha256(data).hexdigest(), "full": len(data) <= FULL_BYTES,
                          "loaded": {"msg": int(hdr[0]), "syst": int(hdr[1]), "parity": int(hdr[2]),
                                     "state": int(hdr[3]), "work_group_size": int(hdr[4]),
                                     "algorithm": int(hdr[5]), "iterations": int(hdr[6]),
                                     "checks_sha256": csr_sha(rp, col), "M": M}})
            print(name, len(data))
    np.savez_compressed(os.path.join(HERE, "fecl_archive.npz"), **arrays)
    with open(os.path.join(HERE, "fecl_archive.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_archive_golden.py (stock FeCl via oracle/ref.mk)",
                   "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
