"""Writes tests/golden/fecl_structure_random.npz: stock FeCl's parity-check form
(Structure::computeGeneratorMatrix, src/detail/Ldpc.cpp:216-331) and encoder
(Structure::encode, src/detail/Ldpc.cpp:180-209) for 40 seeded random sparse matrices, computed
by stock FeCl built from the reference sources (oracle/ref.mk -> oracle/_ref/fecl_golden).

The matrices vary the shape (M x N from 6 x 10 to 120 x 300), the row weight (2 to 14), the
share of dependent rows (copies and sums of other rows) and unused columns, so the greedy
triangularisation meets ties, count-2+ pivots, wide gaps and added message columns.  Stored per
case: the input CSR, FeCl's K, the SHA-256 of FeCl's form (row_ptr and col as int32) and FeCl's
codewords for 2 random messages (bit-packed)."""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
DRV = os.path.join(ROOT, "oracle", "_ref", "fecl_golden")
CASES = 40


def random_matrix(rng):
    M = int(rng.integers(6, 121))
    N = int(M + rng.integers(4, 2 * M + 20))
    lo = int(rng.integers(2, 5))
    hi = int(min(N - 1, lo + rng.integers(0, 10)))
    used = N - int(rng.integers(0, max(1, N // 8))) if rng.random() < 0.3 else N
    rows = []
    for _ in range(M):
        d = int(rng.integers(lo, hi + 1))
        rows.append(sorted(rng.choice(used, size=min(d, used), replace=False).tolist()))
    for _ in range(int(rng.integers(0, 4))):          # dependent rows: copies and sums
        i, j = rng.integers(0, M, size=2)
        s = sorted(set(rows[i]) ^ set(rows[j])) if rng.random() < 0.5 else list(rows[i])
        if len(s) >= 2:
            rows[int(rng.integers(0, M))] = s
    return rows, N


def main():
    if not os.path.exists(DRV):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    rng = np.random.default_rng(99)
    arrays = {}
    meta = []
    with tempfile.TemporaryDirectory() as tmp:
        for c in range(CASES):
            rows, N = random_matrix(rng)
            rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
            col = np.concatenate([np.asarray(r, np.int32) for r in rows])
            hin = os.path.join(tmp, "in.csr")
            with open(hin, "wb") as f:
                np.asarray([len(rows), N, N - len(rows), rp[-1]], "<i8").tofile(f)
                np.asarray(rp, "<i8").tofile(f)
                np.asarray(col, "<i8").tofile(f)
            hout = os.path.join(tmp, "out.csr")
            subprocess.check_call([DRV, "structure", hin, hout])
            d = np.fromfile(hout, "<i8")
            M, _, K, E = (int(x) for x in d[:4])
            srp, scol = d[4:4 + M + 1], d[4 + M + 1:4 + M + 1 + E]
            h = hashlib.sha256()
            h.update(np.asarray(srp, "<i4").tobytes())
            h.update(np.asarray(scol, "<i4").tobytes())
            msg = rng.integers(0, 2, size=(2, K), dtype=np.uint8)
            mpath = os.path.join(tmp, "msg.bin")
            with open(mpath, "wb") as f:
                np.asarray([2, K], "<i8").tofile(f)
                msg.tofile(f)
            cpath = os.path.join(tmp, "cw.bin")
            subprocess.check_call([DRV, "encode", hin, mpath, cpath])
            cw = np.fromfile(cpath, np.uint8, offset=16).reshape(2, N)
            arrays[f"c{c}_row_ptr"] = rp
            arrays[f"c{c}_col"] = col
            arrays[f"c{c}_msg"] = np.packbits(msg, axis=1)
            arrays[f"c{c}_cw"] = np.packbits(cw, axis=1)
            meta.append((N, K, h.hexdigest()))
    arrays["meta_n_k"] = np.asarray([(n, k) for n, k, _ in meta], np.int64)
    arrays["meta_sha"] = np.frombuffer("".join(s for _, _, s in meta).encode(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "fecl_structure_random.npz"), **arrays)
    grew = sum(1 for i, (n, k, _) in enumerate(meta) if k > n - (len(arrays[f"c{i}_row_ptr"]) - 1))
    print(CASES, "cases;", grew, "with dependent checks")


if __name__ == "__main__":
    main()
