"""Writes tests/golden/fecl_structure.json and fecl_structure.npz: stock FeCl's
Ldpc::Gallager::matrix (src/Ldpc.cpp:77-100), the parity-check form its Structure builds
(Structure::computeGeneratorMatrix, src/detail/Ldpc.cpp:216-331) and its encoder
(Structure::encode, src/detail/Ldpc.cpp:180-209), computed by stock FeCl built from the
reference sources (oracle/ref.mk -> oracle/_ref/fecl_golden).

Cases: Gallager matrices (FeCl's generator; some with dependent checks, some with n not a
multiple of wr, one with a seed above 2^32), and hand-made matrices that reach the rarer steps
of the transform (duplicate and summed rows, unused columns, a wide gap between the triangular
part and the rest, a staircase code with a repeated check).  For each: the input CSR (full
arrays, or a SHA-256 for the Gallager ones, whose input comes from the generator under test),
FeCl's form (SHA-256 of row_ptr and col as int32, K), and FeCl's codewords for 3 random
messages (bit-packed)."""
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

GALLAGER = [(2048, 3, 5, 0), (1000, 3, 5, 1), (999, 4, 7, 123), (512, 3, 6, 2**40 + 5), (96, 2, 4, 3),
            (30, 3, 5, 0), (1200, 4, 8, 9)]


def csr_hash(rp, col):
    h = hashlib.sha256()
    h.update(np.asarray(rp, "<i4").tobytes())
    h.update(np.asarray(col, "<i4").tobytes())
    return h.hexdigest()


def rows_to_csr(rows):
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate([np.asarray(sorted(r), np.int32) for r in rows]).astype(np.int32)
    return rp, col


def custom_codes():
    rng = np.random.default_rng(2024)

    def rand_rows(M, N, lo, hi, ncols=None):
        ncols = ncols or N
        return [sorted(rng.choice(ncols, size=int(rng.integers(lo, hi + 1)), replace=False).tolist())
                for _ in range(M)]

    out = []
    out.append(("random_40x100", 100, rand_rows(40, 100, 3, 8)))
    r = rand_rows(30, 80, 3, 7)
    r[17] = list(r[5])
    r[20] = sorted(set(r[1]) ^ set(r[2]))
    r[25] = sorted(set(r[3]) ^ set(r[4]) ^ set(r[6]))
    out.append(("dependent_rows_30x80", 80, r))
    out.append(("unused_columns_25x70", 70, rand_rows(25, 70, 3, 6, ncols=60)))
    out.append(("wide_gap_30x60", 60, rand_rows(30, 60, 12, 20)))
    # staircase over 20 checks with info degree 3, plus a repeat of check 7
    K, M = 30, 20
    st = []
    for j in range(M):
        info = sorted(rng.choice(K, size=3, replace=False).tolist())
        st.append(info + ([K + j - 1] if j else []) + [K + j])
    st.append(list(st[7]))
    out.append(("staircase_repeat_21x50", 50, st))
    return out


def run(*args):
    subprocess.check_call([DRV, *[str(a) for a in args]])


def write_csr(path, rp, col, N):
    M = len(rp) - 1
    with open(path, "wb") as f:
        np.asarray([M, N, N - M, rp[-1]], "<i8").tofile(f)
        np.asarray(rp, "<i8").tofile(f)
        np.asarray(col, "<i8").tofile(f)


def read_csr(path):
    d = np.fromfile(path, "<i8")
    M, N, K, E = (int(x) for x in d[:4])
    rp = d[4:4 + M + 1].astype(np.int32)
    col = d[4 + M + 1:4 + M + 1 + E].astype(np.int32)
    return rp, col, N, K


def main():
    if not os.path.exists(DRV):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    rng = np.random.default_rng(77)
    cases, arrays = [], {}
    with tempfile.TemporaryDirectory() as tmp:
        inputs = []
        for n, wc, wr, seed in GALLAGER:
            raw = os.path.join(tmp, "raw.csr")
            run("gallager-raw", n, wc, wr, seed, raw)
            rp, col, N, _ = read_csr(raw)
            inputs.append((f"gallager_{n}_{wc}_{wr}_{seed}", {"gallager": [n, wc, wr, seed]}, rp, col, N))
        for name, N, rows in custom_codes():
            rp, col = rows_to_csr(rows)
            arrays[f"{name}_row_ptr"] = rp
            arrays[f"{name}_col"] = col
            inputs.append((name, {"csr": name}, rp, col, N))
        for name, src, rp, col, N in inputs:
            hin = os.path.join(tmp, "in.csr")
            write_csr(hin, rp, col, N)
            hout = os.path.join(tmp, "out.csr")
            run("structure", hin, hout)
            srp, scol, _, K = read_csr(hout)
            F = 3
            msg = rng.integers(0, 2, size=(F, K), dtype=np.uint8)
            mpath = os.path.join(tmp, "msg.bin")
            with open(mpath, "wb") as f:
                np.asarray([F, K], "<i8").tofile(f)
                msg.tofile(f)
            cpath = os.path.join(tmp, "cw.bin")
            run("encode", hin, mpath, cpath)
            cw = np.fromfile(cpath, np.uint8, offset=16).reshape(F, N)
            arrays[f"{name}_msg"] = np.packbits(msg, axis=1)
            arrays[f"{name}_cw"] = np.packbits(cw, axis=1)
            cases.append({"name": name, **src, "M": len(rp) - 1, "N": N, "E": int(rp[-1]), "K": int(K),
                          "input_sha256": csr_hash(rp, col), "checks_sha256": csr_hash(srp, scol),
                          "identity": bool(np.array_equal(rp, srp) and np.array_equal(col, scol)),
                          "frames": F})
    np.savez_compressed(os.path.join(HERE, "fecl_structure.npz"), **arrays)
    with open(os.path.join(HERE, "fecl_structure.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_structure_golden.py (stock FeCl via oracle/ref.mk)",
                   "cases": cases}, f, indent=1)
    for c in cases:
        print(c["name"], "K", c["K"], "N-M", c["N"] - c["M"], "identity", c["identity"])


if __name__ == "__main__":
    main()
