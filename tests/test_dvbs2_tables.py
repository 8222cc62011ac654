"""The library's DVB-S2 matrices against the reference's own data.

tests/golden/dvbs2_tables.json holds, for each of the 21 codes, the parameter q and a SHA-256
of the address table as the reference lists it (src/DvbS2.cpp:26-1593), written by
tests/golden/make_dvbs2_tables_golden.py from that file.  Here the table is read back out of
the matrix the library builds (information bit 360 g joins exactly the checks of table row g;
bit 360 g + i joins (a + i q) mod M for every address a of the row), so a change to the
library's tables or to its construction fails one of these checks.  The edge counts and degree
structure of the seven normal rates the batching and serving workloads use are also checked
against figures counted from stock FeCl's tables."""
import hashlib
import json
import os

import numpy as np
import pytest

import fecl_amd as F

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = json.load(open(os.path.join(HERE, "golden", "dvbs2_tables.json")))["codes"]

# rate: (E, max row degree, {information-bit degrees}) for n = 64800
NORMAL = {(1, 4): (194399, 4, {12, 3}), (1, 2): (226799, 7, {8, 3}), (3, 5): (285119, 11, {12, 3}),
          (2, 3): (215999, 10, {13, 3}), (3, 4): (226799, 14, {12, 3}), (5, 6): (237599, 22, {13, 3}),
          (9, 10): (194399, 30, {4, 3})}


def columns(H):
    """Rows of each column, ascending: (ptr, rows) in CSC form."""
    rows = np.repeat(np.arange(H.rows(), dtype=np.int64), np.diff(H.row_ptr))
    order = np.lexsort((rows, H.col))
    ptr = np.concatenate([[0], np.cumsum(np.bincount(H.col, minlength=H.cols()))])
    return ptr, rows[order]


@pytest.mark.parametrize("key", sorted(TABLES))
def test_table_and_construction_match_the_reference(key):
    n, rate = key.split()
    n = int(n)
    num, den = (int(x) for x in rate.split("/"))
    ref = TABLES[key]
    H = F.DvbS2.matrix(n, (num, den))
    q = ref["q"]
    M = 360 * q
    K = n - M
    assert (H.rows(), H.cols()) == (M, n) and K == 360 * ref["groups"]
    ptr, rows = columns(H)
    deg = np.diff(ptr)
    table = []
    for g, length in enumerate(ref["row_lengths"]):
        c0 = 360 * g
        assert np.all(deg[c0:c0 + 360] == length), g
        base = rows[ptr[c0]:ptr[c0 + 1]]
        table.append(base)
        got = rows[ptr[c0]:ptr[c0 + 360]].reshape(360, length)
        want = np.sort((base[None, :] + np.arange(360)[:, None] * q) % M, axis=1)
        assert np.array_equal(got, want), g
    canon = ";".join(",".join(str(int(a)) for a in r) for r in table)
    assert hashlib.sha256(canon.encode()).hexdigest() == ref["table_sha256"]
    # staircase parity part: parity bit K + j on checks j and j + 1, the last one on M - 1 only
    assert np.all(deg[K:n - 1] == 2) and deg[n - 1] == 1
    par = rows[ptr[K]:ptr[n - 1]].reshape(-1, 2)
    assert np.array_equal(par, np.stack([np.arange(M - 1), np.arange(1, M)], axis=1))
    assert rows[ptr[n - 1]] == M - 1


@pytest.mark.parametrize("rate", sorted(NORMAL))
def test_normal_rates_edge_counts_and_degrees(rate):
    E, dc, info_dv = NORMAL[rate]
    H = F.DvbS2.matrix(64800, rate)
    K = 64800 - H.rows()
    col_deg = np.bincount(H.col, minlength=H.cols())
    assert H.size() == E and int(np.diff(H.row_ptr).max()) == dc
    assert set(col_deg[:K].tolist()) == info_dv
    assert np.bincount(col_deg[K:]).tolist()[1:] == [1, H.rows() - 1]
