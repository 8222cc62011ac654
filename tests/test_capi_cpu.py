"""CPU-side checks of the product library: it loads, exports every C-ABI symbol the header
declares, and its host logic (DVB-S2 construction, scaling-table rules, option builder,
argument validation) behaves as FeCl's.  No decode runs here (no GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import fecl_amd as F
from fecl_amd._lib import LIB_PATH, SIGNATURES, lib
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "include", "fecl_ldpc.h")
GOLD = np.load(os.path.join(HERE, "golden", "fecl_golden.npz"))


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(fecl_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 17
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(SIGNATURES), "binding table out of sync with the header"


def test_dvbs2_r12_is_fecl_matrix():
    import hashlib
    H = F.DvbS2.matrix(64800, 0.5)
    h = hashlib.sha256(H.row_ptr.tobytes()).hexdigest() + hashlib.sha256(H.col.tobytes()).hexdigest()
    assert h == GOLD["dvbs2_r12_csr_sha"].tobytes().decode()


@pytest.mark.parametrize("n,rate", [(n, r) for n, rs in F.ldpc._DVBS2_RATES.items() for r in rs])
def test_dvbs2_all_rates_match_independent_construction(n, rate):
    H = F.DvbS2.matrix(n, rate)
    oc = O.dvbs2(n, *rate)
    assert np.array_equal(H.row_ptr, oc.row_ptr) and np.array_equal(H.col, oc.col)
    # K info bits, M = N - K checks, staircase parity, codewords satisfy H c = 0
    K = H.cols() - H.rows()
    rng = np.random.default_rng(n + rate[0])
    info = rng.integers(0, 2, (2, K), dtype=np.uint8)
    cw = O.encode(oc, info)
    assert O.syndrome_weight(oc, cw[0]) == 0 and O.syndrome_weight(oc, cw[1]) == 0


def test_dvbs2_invalid():
    with pytest.raises(ValueError, match="Invalid length"):
        F.DvbS2.matrix(1000, 0.5)
    with pytest.raises(ValueError, match="Invalid rate"):
        F.DvbS2.matrix(64800, 0.7)
    with pytest.raises(ValueError, match="Invalid rate"):
        F.DvbS2.matrix(16200, 0.9)  # 9/10 exists only for normal frames


def test_scaling_table_rules():
    """Structure::scalingMapToVector (src/detail/Ldpc.cpp:64-112)."""
    H = F.DvbS2.matrix(64800, 0.5)  # row degrees 6 and 7 -> 6 columns
    t = F.scaling_table(H, F.DecoderOptions().iterations(3).scalingFactor(0.75))
    assert t.shape == (1, 6) and np.all(t == 0.75)
    t = F.scaling_table(H, F.DecoderOptions().iterations(3).scalingFactor({0: [0.5, 0.6, 0.7], 7: [1, 2, 3]}))
    assert t.shape == (3, 6)
    assert np.allclose(t[:, 0], [0.5, 0.6, 0.7]) and np.allclose(t[:, 5], [1, 2, 3])
    t = F.scaling_table(H, F.DecoderOptions().iterations(3).scalingFactor({6: [0.25], 7: [0.5]}))
    assert t.shape == (1, 6) and t[0, 4] == 0.25 and t[0, 5] == 0.5
    with pytest.raises(ValueError, match="Wrong size for scaling factor"):
        F.scaling_table(H, F.DecoderOptions().iterations(3).scalingFactor({0: [0.5, 0.6]}))
    with pytest.raises(ValueError, match="Wrong size for scaling factor"):
        F.scaling_table(H, F.DecoderOptions().iterations(3).scalingFactor({0: [0.5], 7: [1, 2]}))
    with pytest.raises(ValueError, match="Scaling factor not defined and no default"):
        F.scaling_table(H, F.DecoderOptions().iterations(3).scalingFactor({7: [0.5]}))


def test_decoder_options_defaults():
    """Ldpc::DecoderOptions defaults (src/detail/Ldpc.h:56-74)."""
    o = F.DecoderOptions()
    assert o.algorithm() == F.Approximate
    assert o.iterations() is None
    assert o.scalingFactor() == {0: [1.0]}
    assert o.msgFormat() == F.MsgFormat.F32
    assert int(F.Exact) == 0 and int(F.Linear) == 1 and int(F.Approximate) == 2


def test_create_reports_errors_without_gpu():
    """No GPU here: handle creation reports a device error, never crashes; bad graphs are
    rejected before any device call."""
    H = F.DvbS2.matrix(16200, 0.5)
    opts = F.DecoderOptions().iterations(10)
    bad = F.CheckMatrix(np.array([0, 1], np.int32), np.array([0], np.int32), 4)
    with pytest.raises(ValueError, match="at least two edges"):
        F.Ldpc(bad, opts)
    oob = F.CheckMatrix(np.array([0, 2], np.int32), np.array([0, 9], np.int32), 4)
    with pytest.raises(ValueError, match="out of range"):
        F.Ldpc(oob, opts)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(Exception):
            F.Ldpc(H, opts)


def test_last_error_is_a_string():
    assert isinstance(lib().fecl_last_error(), bytes)


PUNCTURE_GOLD = os.path.join(HERE, "golden", "fecl_puncture.json")


def _mask(s):
    return [] if s == "-" else [int(x) for x in s.split(",")]


def test_puncturing_matches_stock_fecl():
    """Ldpc::Structure::puncturing (src/detail/Ldpc.cpp:333-353), index for index, against
    sequences written by stock FeCl (tests/golden/make_puncture_golden.py) for the Gallager and
    DVB-S2 r1/2 codes and mask combinations, including FeCl's index rule for systematic masks."""
    import hashlib
    import json
    cases = json.load(open(PUNCTURE_GOLD))["cases"]
    assert len(cases) >= 18
    for c in cases:
        p = F.puncturing(c["K"], c["N"], F.PunctureOptions().mask(_mask(c["mask"])).systMask(_mask(c["syst_mask"])))
        assert p.inputSize() == c["input_size"] and p.outputSize() == c["output_size"], c
        seq = np.array([p[i] for i in range(p.outputSize())], dtype="<i8")
        assert hashlib.sha256(seq.tobytes()).hexdigest() == c["sha256"], c
        assert list(seq[:8]) == c["head"] and list(seq[-8:]) == c["tail"]


def test_permutation_host_semantics():
    """Permutation::permute / dePermute (src/Permutation.h): out[i] = in[seq[i]] per block, and
    the inverse scatter into zeroed blocks, over several blocks."""
    p = F.Permutation([4, 0, 2, 6], 8)
    assert p.inputSize() == 8 and p.outputSize() == 4
    x = np.arange(16, dtype=np.float64) + 1.0
    y = p.permute(x)
    assert np.array_equal(y, [5, 1, 3, 7, 13, 9, 11, 15])
    z = p.dePermute(y)
    assert np.array_equal(z, [1, 0, 3, 0, 5, 0, 7, 0, 9, 0, 11, 0, 13, 0, 15, 0])
    assert F.Permutation([3, 1]).inputSize() == 4


def test_puncturing_argument_errors():
    with pytest.raises(ValueError):
        F.puncturing(10, 5, F.PunctureOptions())


# ---- Gallager matrices, FeCl's parity-check form and its encoder --------------------------
STRUCT_GOLD = os.path.join(HERE, "golden", "fecl_structure.json")
STRUCT_NPZ = os.path.join(HERE, "golden", "fecl_structure.npz")


def _structure_cases():
    import json
    return json.load(open(STRUCT_GOLD))["cases"]


def _csr_sha(rp, col):
    import hashlib
    h = hashlib.sha256()
    h.update(np.asarray(rp, "<i4").tobytes())
    h.update(np.asarray(col, "<i4").tobytes())
    return h.hexdigest()


def _case_matrix(c):
    if "gallager" in c:
        return F.Gallager.matrix(*c["gallager"])
    z = np.load(STRUCT_NPZ)
    return F.CheckMatrix(z[c["name"] + "_row_ptr"], z[c["name"] + "_col"], c["N"])


def _gf2_rank(H):
    """Rank of H over GF(2), rows as Python integers (Gaussian elimination)."""
    pivots = {}
    rank = 0
    for r in range(H.rows()):
        x = 0
        for c in H.col[H.row_ptr[r]:H.row_ptr[r + 1]]:
            x ^= 1 << int(c)
        while x:
            top = x.bit_length() - 1
            if top not in pivots:
                pivots[top] = x
                rank += 1
                break
            x ^= pivots[top]
    return rank


def _col_degrees(H):
    return np.bincount(H.col, minlength=H.cols())


@pytest.mark.parametrize("case", [c for c in _structure_cases() if "gallager" in c], ids=lambda c: c["name"])
def test_gallager_matrix_matches_stock_fecl(case):
    """Gallager::matrix (src/Ldpc.cpp:77-100) gives FeCl's matrix for the same arguments,
    including n not a multiple of wr and a seed above 2^32 (FeCl seeds with (int)seed).  What
    the construction fixes independently: n//wr*wc rows of wr ones, block 0 in order, and each
    block a partition of (a permutation of) the covered columns."""
    n, wc, wr, seed = case["gallager"]
    H = F.Gallager.matrix(n, wc, wr, seed)
    assert _csr_sha(H.row_ptr, H.col) == case["input_sha256"]
    per = n // wr
    assert H.rows() == per * wc and H.cols() == n and np.all(np.diff(H.row_ptr) == wr)
    assert np.array_equal(H.col[:per * wr], np.arange(per * wr))
    for b in range(1, wc):
        block = H.col[b * per * wr:(b + 1) * per * wr]
        assert len(np.unique(block)) == per * wr
    assert _col_degrees(H).max() <= wc


@pytest.mark.parametrize("case", _structure_cases(), ids=lambda c: c["name"])
def test_structure_matches_stock_fecl(case):
    """Structure(EncoderOptions(H)).checks() and msgSize() (src/detail/Ldpc.cpp:216-331) equal
    stock FeCl's.  What the mathematics fixes without FeCl: the form is H with its rows and
    columns permuted (same row and column degree multisets and edge count), and the message
    size is N - rank(H) over GF(2), so dependent checks add message bits."""
    H = _case_matrix(case)
    assert _csr_sha(H.row_ptr, H.col) == case["input_sha256"]
    S, K = F.structure(H)
    assert K == case["K"]
    assert _csr_sha(S.row_ptr, S.col) == case["checks_sha256"]
    assert (S.rows(), S.cols(), S.size()) == (H.rows(), H.cols(), H.size())
    assert sorted(np.diff(S.row_ptr)) == sorted(np.diff(H.row_ptr))
    assert sorted(_col_degrees(S)) == sorted(_col_degrees(H))
    assert K == H.cols() - _gf2_rank(H)
    assert _gf2_rank(S) == _gf2_rank(H)
    assert all(np.all(np.diff(S.col[a:b]) > 0) for a, b in zip(S.row_ptr[:-1], S.row_ptr[1:]))


@pytest.mark.parametrize("case", _structure_cases(), ids=lambda c: c["name"])
def test_encoder_matches_stock_fecl(case):
    """Structure::encode (src/detail/Ldpc.cpp:180-209): FeCl's codewords for the same messages,
    systematic in the form's column order, and every one a codeword of that form."""
    H = _case_matrix(case)
    z = np.load(STRUCT_NPZ)
    K, N, frames = case["K"], case["N"], case["frames"]
    msg = np.unpackbits(z[case["name"] + "_msg"], axis=1)[:, :K]
    cw = np.unpackbits(z[case["name"] + "_cw"], axis=1)[:, :N]
    out = F.encode(H, msg).reshape(frames, N)
    assert np.array_equal(out, cw)
    assert np.array_equal(out[:, :K], msg)
    S, _ = F.structure(H)
    oc = O.Code(S.row_ptr, S.col, N, K)
    assert all(O.syndrome_weight(oc, out[f]) == 0 for f in range(frames))


@pytest.mark.parametrize("n,rate", [(64800, (1, 2)), (64800, (1, 4)), (64800, (9, 10)), (16200, (1, 3)),
                                    (16200, (8, 9))])
def test_dvbs2_form_is_the_matrix_itself_and_the_encoder_the_accumulator(n, rate):
    """For staircase codes FeCl's form leaves H unchanged (so decoding a DVB-S2 matrix uses its
    own column order), and its encoder reduces to the parity accumulator of the oracle."""
    H = F.DvbS2.matrix(n, rate)
    S, K = F.structure(H)
    assert S == H and K == H.cols() - H.rows()
    oc = O.dvbs2(n, *rate)
    info = np.random.default_rng(5).integers(0, 2, size=(3, K), dtype=np.uint8)
    assert np.array_equal(F.encode(H, info).reshape(3, -1), O.encode(oc, info))


def test_structure_argument_errors():
    unsorted = F.CheckMatrix(np.array([0, 3, 6], np.int32), np.array([0, 2, 1, 1, 3, 4], np.int32), 6)
    with pytest.raises(ValueError, match="ascending"):
        F.structure(unsorted)
    short = F.CheckMatrix(np.array([0, 3, 4], np.int32), np.array([0, 1, 2, 3], np.int32), 6)
    with pytest.raises(ValueError, match="two edges"):
        F.structure(short)
    with pytest.raises(ValueError):
        F.Gallager.matrix(10, 3, 11)
    with pytest.raises(ValueError, match="Invalid size"):
        F.encode(F.Gallager.matrix(30, 3, 5), np.zeros(5, np.uint8))


# ---- saved codecs: FeCl's archive format ------------------------------------------------
ARCH_GOLD = os.path.join(HERE, "golden", "fecl_archive.json")
ARCH_NPZ = os.path.join(HERE, "golden", "fecl_archive.npz")


def _archive_cases():
    import json
    return json.load(open(ARCH_GOLD))["cases"]


def _archive_matrix(code):
    if "gallager" in code:
        return F.Gallager.matrix(*code["gallager"])
    if "dvbs2" in code:
        n, a, b = code["dvbs2"]
        return F.DvbS2.matrix(n, (a, b))
    z = np.load(STRUCT_NPZ)
    return F.CheckMatrix(z[code["csr"] + "_row_ptr"], z[code["csr"] + "_col"], code["N"])


def _archive_options(c):
    sf = c["scaling"]
    f = {0: [float(x) for x in sf[2:].split(",")]} if sf.startswith("v:") else float(sf)
    return F.DecoderOptions().algorithm(c["algorithm"]).iterations(c["iterations"]).scalingFactor(f)


@pytest.mark.parametrize("case", _archive_cases(), ids=lambda c: c["name"])
def test_archive_save_is_byte_identical_to_stock_fecl(case):
    """fec::detail::save of fec::Ldpc (src/Serialization.h:92-101, Matlab wrapper registration):
    the same bytes as stock FeCl for Gallager codes (with dependent checks too), hand-made
    matrices, a DVB-S2 code, scalar and per-iteration scaling, each algorithm."""
    import hashlib
    H = _archive_matrix(case["code"])
    data = F.archive_save(H, _archive_options(case), case["work_group_size"])
    assert len(data) == case["archive_bytes"]
    assert hashlib.sha256(data).hexdigest() == case["archive_sha256"]
    if case["full"]:
        assert data == np.load(ARCH_NPZ)[case["name"]].tobytes()


@pytest.mark.parametrize("case", [c for c in _archive_cases() if c["full"]], ids=lambda c: c["name"])
def test_archive_load_reads_stock_fecl_archives(case):
    """fec::detail::load (src/Serialization.h:103-115) of archives written by stock FeCl gives
    what FeCl's own loader gives: sizes, work group size, algorithm, iterations, checks(); the
    options and the structure are those the codec was saved with.  Bytes after the archive (the
    Matlab wrapper saves into a buffer with 4096 bytes of slack) are ignored."""
    data = np.load(ARCH_NPZ)[case["name"]].tobytes()
    ld = case["loaded"]
    for blob in (data, data + bytes(4096)):
        S, K, opts, wgs = F.archive_load(blob)
        assert (K, S.cols(), S.size(), wgs) == (ld["msg"], ld["parity"], ld["state"], ld["work_group_size"])
        assert (int(opts._algorithm), opts._iterations) == (ld["algorithm"], ld["iterations"])
        assert _csr_sha(S.row_ptr, S.col) == ld["checks_sha256"]
        H = _archive_matrix(case["code"])
        assert (S, K) == F.structure(H)
        want = F.scaling_table(H, _archive_options(case))
        for d in np.unique(np.diff(S.row_ptr)):
            assert np.array_equal(opts._scaling[int(d)], want[:, d - 2])


def test_archive_load_rejects_damaged_bytes():
    """Truncated or altered archives are reported as invalid arguments, never read past."""
    data = np.load(ARCH_NPZ)["gallager30_minsum"].tobytes()
    for cut in list(range(0, 80)) + list(range(80, len(data), 97)):
        with pytest.raises(ValueError):
            F.archive_load(data[:cut])
    wrong_class = bytearray(data)
    wrong_class[40] = 1            # the pointer's class id: Turbo instead of Ldpc
    with pytest.raises(ValueError, match="Ldpc codec"):
        F.archive_load(bytes(wrong_class))
    wrong_version = bytearray(data)
    wrong_version[30] = 11         # Boost library version
    with pytest.raises(ValueError, match="version"):
        F.archive_load(bytes(wrong_version))
    bad_size = bytearray(data)
    bad_size[97] = 0x1f            # paritySize_ no longer the matrix's column count
    with pytest.raises(ValueError, match="sizes"):
        F.archive_load(bytes(bad_size))
    # rows are consecutive {begin, end} ranges of the element list: a row that starts anywhere
    # else (here row 1 spans rows 0 and 1, so the ranges overlap) is rejected, which also caps
    # the expanded matrix at the stored element count
    rows = np.array([0, 5, 5, 10], dtype="<u8").tobytes()
    at = data.find(rows)
    assert at > 0 and data.count(rows) == 1
    overlap = bytearray(data)
    overlap[at + 16:at + 24] = np.array([0], dtype="<u8").tobytes()
    with pytest.raises(ValueError, match="consecutive"):
        F.archive_load(bytes(overlap))
    # random byte changes: each archive is either rejected or read into a consistent structure
    rng = np.random.default_rng(8)
    for _ in range(1500):
        d = bytearray(data)
        for _ in range(int(rng.integers(1, 4))):
            d[int(rng.integers(0, len(d)))] = int(rng.integers(0, 256))
        try:
            S, K, opts, _ = F.archive_load(bytes(d))
        except ValueError:
            continue
        assert S.rows() < S.cols() and S.cols() - S.rows() <= K < S.cols() and S.size() == S.row_ptr[-1]


def test_matlab_gateway_host_calls():
    """The MATLAB gateway (fecl_amd/matlab/fecl_wrap.cpp) through the MEX test double, host-only
    calls: FeCl's function ids, the DvbS2 matrix as a sparse logical array with the C++ API's
    pattern, errors raised through mexErrMsgIdAndTxt, no arrays leaked."""
    import subprocess
    exe = os.path.join(HERE, "cpp", "mex_gateway_test")
    assert os.path.exists(exe), "build it with `make -C tests/cpp`"
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


RANDOM_NPZ = os.path.join(HERE, "golden", "fecl_structure_random.npz")


def test_structure_and_encoder_match_stock_fecl_on_random_matrices():
    """40 seeded random sparse matrices (tests/golden/make_structure_random_golden.py): shapes
    from 6 x 10 to 120 x 300, row weights 2-14, copied and summed rows (25 cases with dependent
    checks), unused columns.  FeCl's form (SHA-256), K and codewords are reproduced exactly, and
    K = N - rank(H) over GF(2) holds for each."""
    z = np.load(RANDOM_NPZ)
    nk = z["meta_n_k"]
    shas = z["meta_sha"].tobytes().decode()
    for c in range(len(nk)):
        N, K = (int(x) for x in nk[c])
        H = F.CheckMatrix(z[f"c{c}_row_ptr"], z[f"c{c}_col"], N)
        S, k = F.structure(H)
        assert k == K, c
        assert _csr_sha(S.row_ptr, S.col) == shas[64 * c:64 * (c + 1)], c
        assert K == N - _gf2_rank(H), c
        msg = np.unpackbits(z[f"c{c}_msg"], axis=1)[:, :K]
        cw = np.unpackbits(z[f"c{c}_cw"], axis=1)[:, :N]
        assert np.array_equal(F.encode(H, msg).reshape(2, N), cw), c
