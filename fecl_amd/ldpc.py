"""Python mirror of FeCl's LDPC API over the HIP decoder's C ABI (include/fecl_ldpc.h).

Names, argument meaning and errors follow the reference:
    fec::DecoderAlgorithm            src/DecoderAlgorithm.h:30-34
    fec::Ldpc::DecoderOptions        src/detail/Ldpc.h:56-74   (builder methods)
    fec::Ldpc::DvbS2::matrix         src/Ldpc.h:66-75, src/DvbS2.cpp:1595-1645
    fec::Ldpc / Codec decode, soDecode, encode, check, sizes
                                     src/Ldpc.h:59-108, src/Codec.h:60-82, 210-307
Errors raise ValueError where FeCl throws std::invalid_argument.

Host arrays (numpy float64, FeCl's type) go through the host entry points; torch CUDA float32
tensors stay on the device and run on the current stream.
"""
import ctypes
import os
import enum
from fractions import Fraction

import numpy as np

from ._lib import lib, check


class DecoderAlgorithm(enum.IntEnum):
    Exact = 0
    Linear = 1
    Approximate = 2


Exact = DecoderAlgorithm.Exact
Linear = DecoderAlgorithm.Linear
Approximate = DecoderAlgorithm.Approximate


class MsgFormat(enum.IntEnum):
    """Device storage of edge messages (arithmetic is FP32 in all formats)."""
    F32 = 0
    F16 = 1
    BF16 = 2


class CheckMatrix:
    """Parity-check matrix in CSR form (rows are checks, row-major edge order), the role of
    FeCl's SparseBitMatrix (src/BitMatrix.h:42-625)."""

    def __init__(self, row_ptr, col, cols):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self._cols = int(cols)

    def rows(self):
        return len(self.row_ptr) - 1

    def cols(self):
        return self._cols

    def size(self):
        return int(self.row_ptr[-1])

    def rowSizes(self):
        return np.diff(self.row_ptr)

    def __eq__(self, other):
        return (isinstance(other, CheckMatrix) and self._cols == other._cols
                and np.array_equal(self.row_ptr, other.row_ptr) and np.array_equal(self.col, other.col))


_DVBS2_RATES = {64800: [(1, 4), (1, 3), (2, 5), (1, 2), (3, 5), (2, 3), (3, 4), (4, 5), (5, 6), (8, 9), (9, 10)],
                16200: [(1, 4), (1, 3), (2, 5), (1, 2), (3, 5), (2, 3), (3, 4), (4, 5), (5, 6), (8, 9)]}


class DvbS2:
    @staticmethod
    def matrix(n, rate):
        """DVB-S2 parity-check matrix for frame length n (64800 or 16200) and nominal rate
        (a float such as 0.5 as in FeCl, or a (num, den) pair / Fraction)."""
        if n not in _DVBS2_RATES:
            raise ValueError("Invalid length")
        if isinstance(rate, tuple):
            num, den = rate
        elif isinstance(rate, Fraction):
            num, den = rate.numerator, rate.denominator
        else:
            match = [(a, b) for (a, b) in _DVBS2_RATES[n] if float(rate) == a / b]
            if not match:
                raise ValueError("Invalid rate")
            num, den = match[0]
        dims = np.zeros(4, np.int64)
        check(lib().fecl_dvbs2_matrix(n, num, den, _p(dims), None, None))
        M, N, K, E = (int(x) for x in dims)
        rp = np.zeros(M + 1, np.int32)
        col = np.zeros(E, np.int32)
        check(lib().fecl_dvbs2_matrix(n, num, den, _p(dims), _p(rp), _p(col)))
        return CheckMatrix(rp, col, N)


class Gallager:
    @staticmethod
    def matrix(n, wc, wr, seed=0):
        """Ldpc::Gallager::matrix (src/Ldpc.cpp:77-100): the regular (wc, wr) matrix FeCl builds
        for the same arguments, n columns and n//wr*wc rows."""
        dims = np.zeros(3, np.int64)
        check(lib().fecl_gallager_matrix(int(n), int(wc), int(wr), int(seed) & (2**64 - 1), _p(dims), None, None))
        M, N, E = (int(x) for x in dims)
        rp = np.zeros(M + 1, np.int32)
        col = np.zeros(E, np.int32)
        check(lib().fecl_gallager_matrix(int(n), int(wc), int(wr), int(seed) & (2**64 - 1), _p(dims), _p(rp), _p(col)))
        return CheckMatrix(rp, col, N)


def structure(checkMatrix):
    """(H_, K): the matrix FeCl decodes and encodes with for checkMatrix, and its message size
    (Structure(EncoderOptions(H)).checks() / msgSize(), src/detail/Ldpc.cpp:46-55, 216-331).
    Equal to checkMatrix for staircase codes such as DVB-S2."""
    H = checkMatrix
    dims = np.zeros(4, np.int64)
    check(lib().fecl_ldpc_structure(_p(H.row_ptr), _p(H.col), H.rows(), H.cols(), _p(dims), None, None))
    M, N, K, E = (int(x) for x in dims)
    rp = np.zeros(M + 1, np.int32)
    col = np.zeros(E, np.int32)
    check(lib().fecl_ldpc_structure(_p(H.row_ptr), _p(H.col), H.rows(), H.cols(), _p(dims), _p(rp), _p(col)))
    return CheckMatrix(rp, col, N), K


def encode(checkMatrix, msg):
    """Codec::encode for the code of checkMatrix on the host, without a decoder or a device:
    msg holds frames of K = structure(checkMatrix)[1] bits; returns frames of N codeword bits in
    the order of structure(checkMatrix)[0] (src/detail/Ldpc.cpp:180-209)."""
    H = checkMatrix
    _, K = structure(H)
    m = np.ascontiguousarray(np.asarray(msg).reshape(-1), dtype=np.uint8)
    if K == 0 or m.size % K:
        raise ValueError("Invalid size for message")
    frames = m.size // K
    out = np.zeros(frames * H.cols(), np.uint8)
    check(lib().fecl_encode(_p(H.row_ptr), _p(H.col), H.rows(), H.cols(), _p(m), frames, _p(out)))
    return out


class DecoderOptions:
    """Ldpc::DecoderOptions builder.  Defaults as FeCl's C++ API: Approximate, scaling 1.0;
    iterations has no default and must be set before use."""

    def __init__(self):
        self._algorithm = DecoderAlgorithm.Approximate
        self._iterations = None
        self._scaling = {0: [1.0]}
        self._msg_format = MsgFormat.F32

    def algorithm(self, a=None):
        if a is None:
            return self._algorithm
        self._algorithm = DecoderAlgorithm(a)
        return self

    def iterations(self, n=None):
        if n is None:
            return self._iterations
        self._iterations = int(n)
        return self

    def scalingFactor(self, factor=None):
        if factor is None:
            return {k: list(v) for k, v in self._scaling.items()}
        if isinstance(factor, dict):
            self._scaling = {int(k): [float(x) for x in np.atleast_1d(v)] for k, v in factor.items()}
        else:
            self._scaling = {0: [float(factor)]}
        return self

    def msgFormat(self, fmt=None):
        if fmt is None:
            return self._msg_format
        self._msg_format = MsgFormat(fmt)
        return self


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def scaling_table(H, options):
    """Structure::scalingMapToVector: {degree: factors} -> [rows][max degree - 1]."""
    keys = np.array(sorted(options._scaling), np.int64)
    vals = [np.asarray(options._scaling[k], np.float64) for k in keys]
    lengths = np.array([len(v) for v in vals], np.int64)
    offsets = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
    values = np.concatenate(vals) if vals else np.zeros(0)
    cap = max(1, int(options._iterations or 1)) * int(np.max(np.diff(H.row_ptr)))
    table = np.zeros(cap, np.float64)
    rows, cols = ctypes.c_int32(), ctypes.c_int32()
    check(lib().fecl_ldpc_scaling_table(_p(H.row_ptr), H.rows(), int(options._iterations or 0), _p(keys),
                                        _p(lengths), _p(offsets), _p(values), len(keys), _p(table), cap,
                                        ctypes.byref(rows), ctypes.byref(cols)))
    return table[:rows.value * cols.value].reshape(rows.value, cols.value).copy()


class _COptions(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("scaling", ctypes.c_void_p), ("scaling_rows", ctypes.c_int32),
                ("scaling_cols", ctypes.c_int32), ("msg_format", ctypes.c_int32)]


def _options_from_table(algorithm, iterations, table, msg_format=MsgFormat.F32):
    scaling = {d + 2: [float(x) for x in table[:, d]] for d in range(table.shape[1])}
    return DecoderOptions().algorithm(algorithm).iterations(iterations).scalingFactor(scaling).msgFormat(msg_format)


def archive_save(checkMatrix, decoderOptions, workGroupSize=8):
    """FeCl's saved form of fec::Ldpc(EncoderOptions(checkMatrix), decoderOptions,
    workGroupSize): the bytes fec::detail::save writes (src/Serialization.h:92-101) with the
    Matlab wrapper's type registration.  Host only; no device is needed."""
    H = checkMatrix
    o = decoderOptions
    if o._iterations is None or o._iterations < 1:
        raise ValueError("iterations must be >= 1")
    table = scaling_table(H, o)
    opts = _COptions(int(o._algorithm), int(o._iterations), table.ctypes.data, table.shape[0], table.shape[1],
                     int(o._msg_format))
    n = ctypes.c_int64()
    check(lib().fecl_archive_save(_p(H.row_ptr), _p(H.col), H.rows(), H.cols(), ctypes.byref(opts),
                                  int(workGroupSize), None, 0, ctypes.byref(n)))
    buf = np.zeros(n.value, np.uint8)
    check(lib().fecl_archive_save(_p(H.row_ptr), _p(H.col), H.rows(), H.cols(), ctypes.byref(opts),
                                  int(workGroupSize), _p(buf), len(buf), ctypes.byref(n)))
    return buf.tobytes()


def archive_load(data):
    """Reads a saved fec::Ldpc (fec::detail::load, src/Serialization.h:103-115) without a
    device: returns (checks, K, DecoderOptions, workGroupSize), checks being the stored matrix
    in FeCl's generator form."""
    buf = np.frombuffer(bytes(data), np.uint8)
    dims = np.zeros(4, np.int64)
    o = _COptions()
    check(lib().fecl_archive_load(_p(buf), len(buf), _p(dims), None, None, ctypes.byref(o), None, 0, None))
    M, N, K, E = (int(x) for x in dims)
    rp = np.zeros(M + 1, np.int32)
    col = np.zeros(E, np.int32)
    table = np.zeros(max(1, o.scaling_rows * o.scaling_cols), np.float64)
    wgs = ctypes.c_int32()
    check(lib().fecl_archive_load(_p(buf), len(buf), _p(dims), _p(rp), _p(col), ctypes.byref(o), _p(table),
                                  len(table), ctypes.byref(wgs)))
    table = table[:o.scaling_rows * o.scaling_cols].reshape(o.scaling_rows, o.scaling_cols)
    return CheckMatrix(rp, col, N), K, _options_from_table(o.algorithm, o.iterations, table), wgs.value


class Ldpc:
    """fec::Ldpc over the HIP decoder.  workGroupSize is accepted for signature parity with
    FeCl (src/Ldpc.h:82-84); frames run as one device batch regardless of it."""

    def __init__(self, checkMatrix, decoderOptions=None, workGroupSize=8):
        self._H = checkMatrix
        self._wgs = int(workGroupSize)
        self._h = ctypes.c_void_p()
        self._options = None
        opts, keep = self._c_options(decoderOptions or DecoderOptions().iterations(1))
        check(lib().fecl_ldpc_create(_p(checkMatrix.row_ptr), _p(checkMatrix.col), checkMatrix.rows(),
                                     checkMatrix.cols(), ctypes.byref(opts), ctypes.byref(self._h)))
        self._options = decoderOptions
        sizes = [ctypes.c_int64() for _ in range(4)]
        check(lib().fecl_ldpc_sizes(self._h, *[ctypes.byref(s) for s in sizes]))
        self._msg, self._syst, self._parity, self._state = (s.value for s in sizes)

    def _c_options(self, o):
        if o._iterations is None or o._iterations < 1:
            raise ValueError("iterations must be >= 1")
        table = scaling_table(self._H, o)
        opts = _COptions(int(o._algorithm), int(o._iterations), table.ctypes.data, table.shape[0],
                         table.shape[1], int(o._msg_format))
        return opts, table

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().fecl_ldpc_destroy(h)
            except Exception:
                pass
            self._h = None

    # sizes (src/Codec.h:60-63)
    def msgSize(self):
        return self._msg

    def systSize(self):
        return self._syst

    def paritySize(self):
        return self._parity

    def stateSize(self):
        return self._state

    def getWorkGroupSize(self):
        return self._wgs

    def setWorkGroupSize(self, n):
        self._wgs = int(n)

    def checkMatrix(self):
        """The matrix the decoder was built from, as given."""
        return self._H

    def checks(self):
        """Structure::checks(): the matrix in FeCl's generator form, which the decoder and the
        encoder use (message bit k is codeword bit k; the state follows its edge order)."""
        dims = np.zeros(4, np.int64)
        check(lib().fecl_ldpc_checks(self._h, _p(dims), None, None))
        M, N, K, E = (int(x) for x in dims)
        rp = np.zeros(M + 1, np.int32)
        col = np.zeros(E, np.int32)
        check(lib().fecl_ldpc_checks(self._h, _p(dims), _p(rp), _p(col)))
        return CheckMatrix(rp, col, N)

    def setDecoderOptions(self, decoderOptions):
        opts, keep = self._c_options(decoderOptions)
        check(lib().fecl_ldpc_set_options(self._h, ctypes.byref(opts)))
        self._options = decoderOptions

    def getDecoderOptions(self):
        o = _COptions()
        check(lib().fecl_ldpc_get_options(self._h, ctypes.byref(o), None, 0))
        buf = np.zeros(max(1, o.scaling_rows * o.scaling_cols), np.float64)
        check(lib().fecl_ldpc_get_options(self._h, ctypes.byref(o), _p(buf), len(buf)))
        table = buf[:o.scaling_rows * o.scaling_cols].reshape(o.scaling_rows, o.scaling_cols)
        return _options_from_table(o.algorithm, o.iterations, table, o.msg_format)

    @classmethod
    def fromStructure(cls, checks, msgSize, decoderOptions=None, workGroupSize=8):
        """fec::Ldpc(const Structure&) (src/Ldpc.h:83): a decoder for a matrix already in FeCl's
        generator form (Structure::checks(), msgSize()), used as given.  It encodes and saves
        only when the matrix is its own form (staircase codes)."""
        self = cls.__new__(cls)
        self._H = checks
        self._wgs = int(workGroupSize)
        self._h = ctypes.c_void_p()
        self._options = None
        opts, keep = self._c_options(decoderOptions or DecoderOptions().iterations(1))
        check(lib().fecl_ldpc_create_checks(_p(checks.row_ptr), _p(checks.col), checks.rows(), checks.cols(),
                                            int(msgSize), ctypes.byref(opts), ctypes.byref(self._h)))
        self._options = decoderOptions
        sizes = [ctypes.c_int64() for _ in range(4)]
        check(lib().fecl_ldpc_sizes(self._h, *[ctypes.byref(s) for s in sizes]))
        self._msg, self._syst, self._parity, self._state = (s.value for s in sizes)
        return self

    # ---- saved codecs (src/Serialization.h:81-115)
    def save(self):
        """FeCl's archive of this codec (fec::detail::save with the Matlab wrapper's type
        registration), as bytes; the work group size is stored too."""
        n = ctypes.c_int64()
        check(lib().fecl_ldpc_save(self._h, self._wgs, None, 0, ctypes.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        check(lib().fecl_ldpc_save(self._h, self._wgs, _p(buf), len(buf), ctypes.byref(n)))
        return buf.tobytes()

    @classmethod
    def load(cls, data):
        """A decoder for a codec saved by FeCl or by save() (fec::detail::load): the stored
        structure as is (no transform), options and work group size; FP32 messages."""
        buf = np.frombuffer(bytes(data), np.uint8)
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        wgs = ctypes.c_int32()
        check(lib().fecl_ldpc_load(_p(buf), len(buf), ctypes.byref(self._h), ctypes.byref(wgs)))
        self._wgs = wgs.value
        self._H = self.checks()
        self._options = self.getDecoderOptions()
        sizes = [ctypes.c_int64() for _ in range(4)]
        check(lib().fecl_ldpc_sizes(self._h, *[ctypes.byref(s) for s in sizes]))
        self._msg, self._syst, self._parity, self._state = (s.value for s in sizes)
        return self

    # ---- device layout (no counterpart in FeCl)
    def setLayout(self, words):
        """Pin the lane layout: 32-bit words per lane per edge (1, 2, 4; 0 = automatic)."""
        check(lib().fecl_ldpc_set_layout(self._h, int(words)))

    def setFusedSyndrome(self, enable):
        """Syndrome in the next check kernel (True) or in its own kernel (False, default)."""
        check(lib().fecl_ldpc_set_fused_syndrome(self._h, int(bool(enable))))

    def setCachePolicy(self, policy):
        """Non-temporal message rows: bit 0 check kernel, bit 1 variable kernel; -1 automatic."""
        check(lib().fecl_ldpc_set_cache_policy(self._h, int(policy)))

    def setPassFrames(self, frames):
        """Decode batches in passes of at most `frames` frames (0 = 1024); results are the same."""
        check(lib().fecl_ldpc_set_pass_frames(self._h, int(frames)))

    def setCompaction(self, enable):
        """Move live frames into fewer tiles as frames converge (default on); same results."""
        check(lib().fecl_ldpc_set_compaction(self._h, int(bool(enable))))

    def setGraphPersistent(self, enable):
        """Whether BatchGraphs made with this decoder may use the persistent kernel (default
        off; used when every decoder of the graph asks for it and is eligible); same results."""
        check(lib().fecl_ldpc_set_graph_persistent(self._h, int(bool(enable))))

    def layout(self):
        """(words per lane, frames per tile) of the current device layout."""
        w, tw = ctypes.c_int32(), ctypes.c_int32()
        check(lib().fecl_ldpc_get_layout(self._h, ctypes.byref(w), ctypes.byref(tw)))
        return w.value, tw.value

    def setVarsPerWave(self, n):
        """Variables one wavefront of the variable kernel serves (1, 2; 0 = automatic)."""
        check(lib().fecl_ldpc_set_vars_per_wave(self._h, int(n)))

    def varsPerWave(self):
        v = ctypes.c_int32()
        check(lib().fecl_ldpc_get_vars_per_wave(self._h, ctypes.byref(v)))
        return v.value

    def autotune(self, frames, probe_iterations=10):
        """Times each eligible (words, variables per wave) pair on a frames-frame batch and pins
        the fastest.  Returns (words, {(words, vars_per_wave): ms})."""
        best = ctypes.c_int32()
        ms = (ctypes.c_double * 6)()
        check(lib().fecl_ldpc_autotune(self._h, int(frames), int(probe_iterations), ctypes.byref(best), ms))
        keys = [(w, v) for w in (1, 2, 4) for v in (1, 2)]
        return best.value, {k: ms[i] for i, k in enumerate(keys) if ms[i] >= 0}

    # ---- encode / check (host)
    def puncturing(self, options):
        """Ldpc::puncturing (src/Ldpc.h:94): FeCl's Permutation for these PunctureOptions."""
        return puncturing(self._syst, self._parity, options)

    def encode(self, msg, stream=None):
        """Codec::encode: message bits [F*K] (0/1) -> codeword [F*N].  A CUDA uint8 tensor is
        encoded on the device (on the current stream) and the codeword returned as one."""
        if _is_cuda_tensor(msg):
            import torch
            x = msg.contiguous().view(-1)
            if x.dtype != torch.uint8:
                raise ValueError("message bits must be uint8")
            frames = x.numel() // self._msg
            if x.numel() != frames * self._msg:
                raise ValueError("Invalid size for message")
            out = torch.empty(frames * self._parity, dtype=torch.uint8, device=x.device)
            st = stream if stream is not None else torch.cuda.current_stream(x.device).cuda_stream
            check(lib().fecl_ldpc_encode_device(self._h, ctypes.c_void_p(x.data_ptr()), frames,
                                                ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(st)))
            return out
        msg = np.ascontiguousarray(msg, dtype=np.uint8).ravel()
        frames = msg.size // self._msg
        if msg.size != frames * self._msg:
            raise ValueError("Invalid size for message")
        parity = np.zeros(frames * self._parity, np.uint8)
        check(lib().fecl_ldpc_encode(self._h, _p(msg), frames, _p(parity)))
        return parity

    def check(self, parity):
        parity = np.ascontiguousarray(parity, dtype=np.uint8).ravel()
        frames = parity.size // self._parity
        if parity.size != frames * self._parity:
            raise ValueError("Invalid size for parity")
        ok = ctypes.c_int32()
        check(lib().fecl_ldpc_check(self._h, _p(parity), frames, ctypes.byref(ok)))
        return bool(ok.value)

    # ---- decode
    def _frames(self, n, what="parity"):
        frames = n // self._parity
        if n != frames * self._parity:
            raise ValueError("Invalid size for " + what)
        return frames

    def decode(self, parity, iterations=False, stream=None):
        """Hard-decision decode.  numpy input: float64 LLRs (FeCl), returns uint8 [F*K].
        torch CUDA input: float32 LLRs on the device, returns a CUDA uint8 tensor [F*K].
        With iterations=True also returns the per-frame check-update counts."""
        if _is_cuda_tensor(parity):
            import torch
            x = parity.contiguous().view(-1)
            if x.dtype != torch.float32:
                raise ValueError("device LLRs must be float32")
            frames = self._frames(x.numel())
            out = torch.empty(frames * self._msg, dtype=torch.uint8, device=x.device)
            it = torch.empty(frames, dtype=torch.int32, device=x.device)
            st = stream if stream is not None else torch.cuda.current_stream(x.device).cuda_stream
            check(lib().fecl_ldpc_decode_device(self._h, ctypes.c_void_p(x.data_ptr()), frames,
                                                ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(it.data_ptr()),
                                                ctypes.c_void_p(st)))
            return (out, it) if iterations else out
        llr = np.ascontiguousarray(parity, dtype=np.float64).ravel()
        frames = self._frames(llr.size)
        out = np.zeros(frames * self._msg, np.uint8)
        it = np.zeros(frames, np.int32)
        check(lib().fecl_ldpc_decode(self._h, _p(llr), frames, _p(out), _p(it)))
        return (out, it) if iterations else out

    def soDecode(self, parity, syst=None, state=None,
                 outputs=("msg", "syst", "parity", "state", "iterations"), stream=None):
        """Soft-output decode (Codec::soDecode).  Returns a dict of the requested outputs:
        msg [F*K] a-posteriori, syst [F*K] / parity [F*N] extrinsic, state [F*E]."""
        if parity is None:
            raise ValueError("Input must contains parity")
        if _is_cuda_tensor(parity):
            import torch
            x = parity.contiguous().view(-1)
            frames = self._frames(x.numel())
            dev = x.device

            def opt_in(a, n, what):
                if a is None:
                    return None, None
                a = a.contiguous().view(-1).to(torch.float32)
                if a.numel() != frames * n:
                    raise ValueError("Invalid size for " + what)
                return a, ctypes.c_void_p(a.data_ptr())
            s_in, s_ptr = opt_in(syst, self._syst, "msg")
            st_in, st_ptr = opt_in(state, self._state, "state")
            res, ptrs = {}, {}
            for name, n, dt in (("msg", self._msg, torch.float32), ("syst", self._syst, torch.float32),
                                ("parity", self._parity, torch.float32), ("state", self._state, torch.float32),
                                ("iterations", 1, torch.int32)):
                if name in outputs:
                    res[name] = torch.empty(frames * n, dtype=dt, device=dev)
                    ptrs[name] = ctypes.c_void_p(res[name].data_ptr())
                else:
                    ptrs[name] = None
            st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
            check(lib().fecl_ldpc_sodecode_device(self._h, ctypes.c_void_p(x.data_ptr()), s_ptr, st_ptr, frames,
                                                  ptrs["msg"], ptrs["syst"], ptrs["parity"], ptrs["state"],
                                                  ptrs["iterations"], ctypes.c_void_p(st)))
            return res
        llr = np.ascontiguousarray(parity, dtype=np.float64).ravel()
        frames = self._frames(llr.size)
        if syst is not None:
            syst = np.ascontiguousarray(syst, dtype=np.float64).ravel()
            if syst.size != frames * self._syst:
                raise ValueError("Invalid size for msg")
        if state is not None:
            state = np.ascontiguousarray(state, dtype=np.float64).ravel()
            if state.size != frames * self._state:
                raise ValueError("Invalid size for state")
        res = {}
        for name, n, dt in (("msg", self._msg, np.float64), ("syst", self._syst, np.float64),
                            ("parity", self._parity, np.float64), ("state", self._state, np.float64),
                            ("iterations", 1, np.int32)):
            if name in outputs:
                res[name] = np.zeros(frames * n, dt)
        check(lib().fecl_ldpc_sodecode(self._h, _p(llr), _p(syst), _p(state), frames, _p(res.get("msg")),
                                       _p(res.get("syst")), _p(res.get("parity")), _p(res.get("state")),
                                       _p(res.get("iterations"))))
        return res


def _is_cuda_tensor(x):
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


class PunctureOptions:
    """Ldpc::PunctureOptions (src/detail/Ldpc.h:76-82): mask and systMask, lists of 0/1."""

    def __init__(self):
        self._mask, self._syst = [], []

    def mask(self, m=None):
        if m is None:
            return list(self._mask)
        self._mask = [1 if x else 0 for x in m]
        return self

    def systMask(self, m=None):
        if m is None:
            return list(self._syst)
        self._syst = [1 if x else 0 for x in m]
        return self


class Permutation:
    """fec::Permutation (src/Permutation.h): output element i is input element sequence[i].
    permute: blocks of inputSize -> outputSize; dePermute: blocks of outputSize -> inputSize
    with the entries no index names left 0.  numpy arrays are handled on the host; CUDA float32
    tensors on the device (fecl_permute_device, on the current stream)."""

    def __init__(self, sequence, inputSize=None):
        self._seq = np.ascontiguousarray(sequence, dtype=np.int64).ravel()
        self._in = int(inputSize) if inputSize is not None else (int(self._seq.max()) + 1 if self._seq.size else 0)
        self._dev = {}

    def inputSize(self):
        return self._in

    def outputSize(self):
        return int(self._seq.size)

    def __getitem__(self, i):
        return int(self._seq[i])

    def _device(self, x, inverse, stream):
        import torch
        v = x.contiguous().view(-1)
        if v.dtype != torch.float32:
            raise ValueError("device permutation takes float32 tensors")
        n_from = self.outputSize() if inverse else self._in
        n_to = self._in if inverse else self.outputSize()
        if n_from == 0 or v.numel() % n_from:
            raise ValueError("Invalid input size")
        frames = v.numel() // n_from
        key = v.device.index
        if key not in self._dev:
            self._dev[key] = torch.from_numpy(self._seq).to(v.device)
        out = torch.empty(frames * n_to, dtype=torch.float32, device=v.device)
        st = stream if stream is not None else torch.cuda.current_stream(v.device).cuda_stream
        check(lib().fecl_permute_device(ctypes.c_void_p(self._dev[key].data_ptr()), self.outputSize(), self._in,
                                        ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(out.data_ptr()), frames,
                                        int(inverse), ctypes.c_void_p(st)))
        return out

    def permute(self, x, stream=None):
        if _is_cuda_tensor(x):
            return self._device(x, False, stream)
        x = np.asarray(x).reshape(-1, self._in)
        return x[:, self._seq].ravel()

    def dePermute(self, x, stream=None):
        if _is_cuda_tensor(x):
            return self._device(x, True, stream)
        x = np.asarray(x).reshape(-1, self.outputSize())
        out = np.zeros((x.shape[0], self._in), dtype=x.dtype)
        out[:, self._seq] = x
        return out.ravel()


def puncturing(systSize, paritySize, options):
    """Ldpc::Structure::puncturing (src/detail/Ldpc.cpp:333-353) for a code of systSize
    systematic bits in paritySize: FeCl's Permutation for these PunctureOptions."""
    m = np.asarray(options.mask(), np.uint8)
    sm = np.asarray(options.systMask(), np.uint8)
    n = ctypes.c_int64()
    check(lib().fecl_ldpc_puncturing(int(systSize), int(paritySize), _p(m) if m.size else None, m.size,
                                     _p(sm) if sm.size else None, sm.size, None, 0, ctypes.byref(n)))
    seq = np.zeros(max(n.value, 1), np.int64)
    check(lib().fecl_ldpc_puncturing(int(systSize), int(paritySize), _p(m) if m.size else None, m.size,
                                     _p(sm) if sm.size else None, sm.size, _p(seq), seq.size, ctypes.byref(n)))
    return Permutation(seq[:n.value], paritySize)


def autotune_cache(mode=1, path=None):
    """Where Ldpc.autotune keeps its picks (no counterpart in the reference): 0 = nowhere, 1 = this
    process (the default), 2 = this process and the text file `path` (read now, appended with
    every new pick).  Picks are keyed by GPU architecture, code, algorithm, message format and
    batch size (rounded up to a multiple of 64 up to 8192 frames, to a power of two above)."""
    check(lib().fecl_autotune_cache_config(int(mode), None if path is None else os.fsencode(path)))


def autotune_cache_stats():
    """{"hits", "misses", "entries", "skipped_lines", "write_failures"} of the autotune cache."""
    v = [ctypes.c_int64() for _ in range(5)]
    check(lib().fecl_autotune_cache_stats(*[ctypes.byref(x) for x in v]))
    return dict(zip(("hits", "misses", "entries", "skipped_lines", "write_failures"), (x.value for x in v)))


def autotune_cache_clear():
    """Forgets the picks held in this process (a cache file is left as it is)."""
    check(lib().fecl_autotune_cache_clear())


class BatchGraph:
    """Several codes' batches decoded by one hipGraph launch (fecl_ldpc_graph_*).

    decoders: Ldpc handles; llr: CUDA float32 tensors [frames_i * N_i] that stay alive and are
    re-read at every launch().  Outputs live in .msg[i] (uint8 [frames_i * K_i]) and
    .iterations[i] (int32 [frames_i])."""

    def __init__(self, decoders, llr):
        import torch
        if len(decoders) != len(llr) or not decoders:
            raise ValueError("one LLR tensor per decoder")
        self._decoders = list(decoders)
        for x in llr:
            if not x.is_cuda or x.dtype != torch.float32:
                raise ValueError("LLRs must be CUDA float32 tensors")
            # the graph reads the tensor's memory at every launch: a private contiguous copy
            # would never see the caller's later writes
            if not x.is_contiguous():
                raise ValueError("LLR tensors must be contiguous (the graph re-reads their memory)")
        self.llr = [x.view(-1) for x in llr]
        self.frames = [d._frames(x.numel()) for d, x in zip(self._decoders, self.llr)]
        self.msg = [torch.empty(f * d.msgSize(), dtype=torch.uint8, device=x.device)
                    for f, d, x in zip(self.frames, self._decoders, self.llr)]
        self.iterations = [torch.empty(f, dtype=torch.int32, device=x.device)
                           for f, x in zip(self.frames, self.llr)]
        n = len(decoders)
        P = ctypes.c_void_p
        hs = (P * n)(*[d._h.value for d in self._decoders])
        fr = (ctypes.c_int64 * n)(*self.frames)
        ll = (P * n)(*[x.data_ptr() for x in self.llr])
        ms = (P * n)(*[m.data_ptr() for m in self.msg])
        it = (P * n)(*[i.data_ptr() for i in self.iterations])
        self._g = P()
        check(lib().fecl_ldpc_graph_create(hs, fr, n, ll, ms, it, ctypes.byref(self._g)))

    def launch(self, stream=None):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream(self.llr[0].device).cuda_stream
        check(lib().fecl_ldpc_graph_launch(self._g, ctypes.c_void_p(st)))

    @property
    def persistent(self):
        """True when the graph runs the persistent multi-code kernel."""
        v = ctypes.c_int32()
        check(lib().fecl_ldpc_graph_is_persistent(self._g, ctypes.byref(v)))
        return bool(v.value)

    def check(self):
        """Raises if the last finished launch of a persistent graph stopped on its watchdog
        (call after synchronising the launch stream)."""
        check(lib().fecl_ldpc_graph_check(self._g))

    def __del__(self):
        g = getattr(self, "_g", None)
        if g is not None and g.value:
            try:
                lib().fecl_ldpc_graph_destroy(g)
            except Exception:
                pass
            self._g = None


class Server:
    """Continuous-batching decode server (fecl_ldpc_server_*): one device slot pool per code,
    requests admitted as slots free up, message state resident on the device.

    decoders: Ldpc handles (kept alive by the server); slots: frames per pool (None = share
    memory_bytes, by default half of the free device memory)."""

    def __init__(self, decoders, slots=None, memory_bytes=0):
        self._decoders = list(decoders)
        n = len(self._decoders)
        P = ctypes.c_void_p
        hs = (P * n)(*[d._h.value for d in self._decoders])
        sl = (ctypes.c_int64 * n)(*(slots if slots is not None else [0] * n))
        self._s = P()
        self._pending = {}
        check(lib().fecl_ldpc_server_create(hs, n, sl, int(memory_bytes), ctypes.byref(self._s)))

    def slots(self, code):
        v = ctypes.c_int64()
        check(lib().fecl_ldpc_server_slots(self._s, int(code), ctypes.byref(v)))
        return v.value

    def setRoundSteps(self, steps):
        """Steps each pool takes per round between the host's admissions and retirements
        (1..64, default 1): more steps cost fewer host round trips, but a finished frame waits
        up to steps - 1 more steps for its retirement and admissions come less often."""
        check(lib().fecl_ldpc_server_set_round_steps(self._s, int(steps)))

    def submit(self, code, llr):
        """Queue frames of code index `code`: float32 [frames * N], a host array or a
        contiguous CUDA tensor on the server's device (read in place, no host copy; kept alive
        until wait() returns it).  Returns a ticket."""
        d = self._decoders[code]
        if _is_cuda_tensor(llr):
            import torch
            if llr.dtype != torch.float32 or not llr.is_contiguous():
                raise ValueError("device LLRs must be a contiguous float32 tensor")
            x = llr
            frames = d._frames(x.numel())
            # the server reads the frames on its own stream: they must be complete now
            torch.cuda.current_stream(x.device).synchronize()
            ptr = ctypes.c_void_p(x.data_ptr())
        else:
            x = np.ascontiguousarray(llr, dtype=np.float32).ravel()
            frames = d._frames(x.size)
            ptr = _p(x)
        msg = np.zeros(frames * d.msgSize(), np.uint8)
        it = np.zeros(frames, np.int32)
        t = ctypes.c_uint64()
        check(lib().fecl_ldpc_server_submit(self._s, int(code), ptr, frames, _p(msg), _p(it), ctypes.byref(t)))
        self._pending[t.value] = (x, msg, it)
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """(msg uint8 [frames*K], iterations int32 [frames], latency seconds), or None on timeout."""
        done, lat = ctypes.c_int32(), ctypes.c_double()
        check(lib().fecl_ldpc_server_wait(self._s, ctypes.c_uint64(ticket), float(timeout), ctypes.byref(done),
                                          ctypes.byref(lat)))
        if not done.value:
            return None
        _, msg, it = self._pending.pop(ticket)
        return msg, it, lat.value

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().fecl_ldpc_server_stats(self._s, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"frames_done": a.value, "steps": b.value, "queued_frames": c.value}

    def close(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib().fecl_ldpc_server_destroy(s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
