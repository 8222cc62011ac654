"""ctypes front-end of the CPU oracle (liboracle.so, built by `make -C oracle`).

Test infrastructure only: tests, the smoke check and the benchmark's CPU-baseline leg call it
as the checker; the product never imports it.  See fecl_oracle.h for what pins each function.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

EXACT, LINEAR, APPROXIMATE, EXACT_STABLE = 0, 1, 2, 3
MSG_F32, MSG_F16, MSG_BF16 = 0, 1, 2


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64, i32 = ctypes.c_int64, ctypes.c_int
        L.orc_dvbs2_build.argtypes = [i32, i32, i32, P, P, P]
        L.orc_dvbs2_encode.argtypes = [P, P, i64, i64, P, P, i64]
        L.orc_syndrome_weight.argtypes = [P, P, i64, P]
        L.orc_syndrome_weight.restype = i64
        L.orc_quantize.argtypes = [ctypes.c_float, i32]
        L.orc_quantize.restype = ctypes.c_float
        L.orc_bp_decode.argtypes = [P, P, i64, i64, i64, P, i64, i32, i32, P, i32, i32, i32, i32,
                                    P, P, P, P, P, P, P, P, i32]
        L.orc_bp_decode_diag.argtypes = [P, P, i64, i64, i64, P, i64, i32, i32, P, i32, i32, i32, i32,
                                         P, P, P, P, P, P, P, P, P, i32]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Code:
    """A parity-check matrix in CSR form: rows are checks, edges in row-major order."""

    def __init__(self, row_ptr, col, n, k):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.M = len(self.row_ptr) - 1
        self.N = int(n)
        self.K = int(k)
        self.E = int(self.row_ptr[-1])

    @property
    def max_degree(self):
        return int(np.max(np.diff(self.row_ptr)))


def dvbs2(n, rate_num, rate_den):
    dims = np.zeros(4, np.int64)
    if lib().orc_dvbs2_build(n, rate_num, rate_den, _p(dims), None, None) != 0:
        raise ValueError("unknown DVB-S2 code")
    M, N, K, E = (int(x) for x in dims)
    rp = np.zeros(M + 1, np.int32)
    col = np.zeros(E, np.int32)
    lib().orc_dvbs2_build(n, rate_num, rate_den, _p(dims), _p(rp), _p(col))
    return Code(rp, col, N, K)


def encode(code, info):
    info = np.ascontiguousarray(info, dtype=np.uint8).reshape(-1, code.K)
    cw = np.zeros((info.shape[0], code.N), np.uint8)
    lib().orc_dvbs2_encode(_p(code.row_ptr), _p(code.col), code.M, code.N, _p(info), _p(cw), info.shape[0])
    return cw


def syndrome_weight(code, cw):
    cw = np.ascontiguousarray(cw, dtype=np.uint8)
    return int(lib().orc_syndrome_weight(_p(code.row_ptr), _p(code.col), code.M, _p(cw)))


def quantize(x, fmt):
    return float(lib().orc_quantize(float(x), fmt))


def scaling_table(code, factor=1.0, rows=None):
    """[L][maxdeg-1] table; `factor` is a scalar or a per-iteration sequence."""
    f = np.atleast_1d(np.asarray(factor, dtype=np.float64))
    return np.repeat(f[:, None], code.max_degree - 1, axis=1).copy()


def decode(code, llr, alg, iterations, sf=None, prec=64, msg_fmt=MSG_F32, syst_in=None,
           state_in=None, want=("hard", "msg", "parity", "syst", "state", "iters"), threads=None):
    llr = np.ascontiguousarray(llr, dtype=np.float64).reshape(-1, code.N)
    F = llr.shape[0]
    if sf is None:
        sf = scaling_table(code, 1.0)
    sf = np.ascontiguousarray(sf, dtype=np.float64)
    out = {}
    if "hard" in want:
        out["hard"] = np.zeros((F, code.K), np.uint8)
    if "msg" in want:
        out["msg"] = np.zeros((F, code.K))
    if "parity" in want:
        out["parity"] = np.zeros((F, code.N))
    if "syst" in want:
        out["syst"] = np.zeros((F, code.K))
    if "state" in want:
        out["state"] = np.zeros((F, code.E))
    if "iters" in want:
        out["iters"] = np.zeros(F, np.int32)
    if "lin_side" in want:   # Linear only: hash of the cutoff side of every table lookup
        out["lin_side"] = np.zeros(F, np.uint64)
    if syst_in is not None:
        syst_in = np.ascontiguousarray(syst_in, dtype=np.float64).reshape(F, code.K)
    if state_in is not None:
        state_in = np.ascontiguousarray(state_in, dtype=np.float64).reshape(F, code.E)
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib().orc_bp_decode_diag(_p(code.row_ptr), _p(code.col), code.M, code.N, code.K, _p(llr), F,
                                  alg, iterations, _p(sf), sf.shape[0], sf.shape[1], prec, msg_fmt,
                                  _p(syst_in), _p(state_in), _p(out.get("hard")), _p(out.get("msg")),
                                  _p(out.get("parity")), _p(out.get("syst")), _p(out.get("state")),
                                  _p(out.get("iters")), _p(out.get("lin_side")), threads)
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    return out
