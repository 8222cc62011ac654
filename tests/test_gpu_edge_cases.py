"""Inputs at the edges of the arithmetic, through the HIP decoder, against the oracle.

  * 16-bit message storage saturates at the format's largest finite value (+-65504 for FP16,
    +-0x1.fep127 for BF16) before rounding to nearest even: LLRs of +-1e5 (FP16) and LLRs beyond
    the FP32 range (BF16) decode bit for bit like the oracle, which stores messages the same way.
  * FP64 LLRs beyond the FP32 range (+-1e39) become +-inf when the host path rounds them, as in
    FeCl's decoder instantiated with a float LLR type: infinite channel values propagate the
    same way (hard bits, iteration counts, a-posteriori values and state equal the oracle's).
  * NaN channel values (FP32 message storage): a NaN message acts in a min-sum check row as
    a magnitude above every other (it never becomes the row's minimum) carrying its sign bit,
    and its variable's a-posteriori value is NaN (hard decision 0, since NaN >= 0 is false).
    FeCl's std::min folds (min(a, b) = b < a ? b : a) agree with that at a row's middle
    positions, but a NaN first in a row takes over FeCl's forward prefixes (outputs 1 .. d-1
    turn NaN) and a NaN last in a row its backward suffix (output 0 turns NaN), from where NaN
    spreads through the frame.  NaN on variables that are neither first nor last in any row
    decodes bit for bit like the oracle; a NaN first or last in a row leaves every GPU output
    but the variable's own a-posteriori value finite.
  * Exact with |LLR| above ~37, where FeCl's tanh(-x/2) rounds to +-1 and its post returns
    +-inf (src/detail/LlrMetrics.h:137-138): the GPU's complement form stays finite.  Check-to-
    variable messages agree with the exact box-plus value within 1e-4 where that value is
    below 30, and saturate near 87 (|x| is clamped to 87 in the prior) where every other input
    of the row is larger; hard decisions and iteration counts equal FeCl's.
"""
import numpy as np
import pytest

import fecl_amd as F
from oracle import oracle as O
from tests.golden.vectors import dvbs2_llr

pytestmark = pytest.mark.gpu


def short12():
    oc = O.dvbs2(16200, 1, 2)
    return oc, F.CheckMatrix(oc.row_ptr, oc.col, oc.N)


def u32(a):
    return np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32)).view(np.uint32)


def assert_same_bits(gpu, ref, frames, oc):
    assert np.array_equal(gpu["iterations"], ref["iters"]), "iteration counts"
    for name, n in (("msg", oc.K), ("syst", oc.K), ("parity", oc.N), ("state", oc.E)):
        g = u32(np.asarray(gpu[name]).reshape(frames, n))
        r = u32(ref[name].reshape(frames, n))
        gn = np.isnan(np.asarray(gpu[name]).reshape(frames, n))
        rn = np.isnan(ref[name].reshape(frames, n))
        assert np.array_equal(gn, rn), name + " NaN positions"
        assert np.array_equal(g[~gn], r[~rn]), name


@pytest.mark.parametrize("fmt,ofmt,big", [(F.MsgFormat.F16, O.MSG_F16, 1e5), (F.MsgFormat.BF16, O.MSG_BF16, 1e39),
                                          (F.MsgFormat.F32, O.MSG_F32, 1e39)])
def test_saturating_and_infinite_llrs(fmt, ofmt, big):
    """Every variable of a few check rows, and a few hundred scattered positions, get LLRs of
    +-big, most with the sign of the transmitted bit and some against it; FP64 inputs beyond
    the FP32 range round to +-inf on the way in.  In the rows made entirely of such variables
    the check-to-variable messages are the saturated (or infinite) values themselves."""
    oc, H = short12()
    frames = 6
    llr, info = dvbs2_llr(frames, 2.2, seed=1800, rate=(1, 2), n=16200, return_info=True)
    cw = O.encode(oc, info)
    rng = np.random.default_rng(1801)
    for f in range(frames):
        rows = rng.choice(oc.M, size=5, replace=False)
        pos = np.concatenate([oc.col[oc.row_ptr[r]:oc.row_ptr[r + 1]] for r in rows] +
                             [rng.choice(oc.N, size=300, replace=True)])
        sign = np.where(cw[f, pos] == 1, 1.0, -1.0)      # FeCl's LLR is positive for a 1
        sign[-20:] = -sign[-20:]                          # a few confident errors
        llr[f, pos] = big * sign
    ref = O.decode(oc, llr, O.APPROXIMATE, 20, prec=32, msg_fmt=ofmt, threads=16)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(20).msgFormat(fmt))
    out = dec.soDecode(llr)
    assert_same_bits(out, ref, frames, oc)
    hard, it = dec.decode(llr, iterations=True)
    assert np.array_equal(hard.reshape(frames, oc.K), ref["hard"]) and np.array_equal(it, ref["iters"])
    if big > 3.5e38:
        assert np.isinf(out["msg"]).any()                 # the case is exercised
    if fmt != F.MsgFormat.F32:                            # stored messages saturate
        lim = 65504.0 if fmt == F.MsgFormat.F16 else 3.38953138925153547590e38
        st = np.asarray(out["state"])
        assert np.isfinite(st).all() and np.abs(st).max() == lim


def row_ends(oc):
    return set(oc.col[oc.row_ptr[:-1]].tolist()) | set(oc.col[oc.row_ptr[1:] - 1].tolist())


def test_nan_llrs_in_the_middle_of_rows():
    """NaN channel values on information bits that are neither first nor last in any of their
    rows: every output equals the oracle's (NaN exactly where it has NaN: the NaN variables'
    own a-posteriori values), min-sum with FP32 messages (NaN inputs are specified for FP32
    message storage; 16-bit storage saturates on conversion and is not specified for NaN)."""
    oc, H = short12()
    frames = 4
    llr = dvbs2_llr(frames, 2.5, seed=1900, rate=(1, 2), n=16200)
    rng = np.random.default_rng(1901)
    ends = row_ends(oc)
    cand = np.array([v for v in range(oc.K) if v not in ends])
    assert cand.size > 1000
    for f in range(frames):
        llr[f, rng.choice(cand, size=25, replace=False)] = np.nan
    ref = O.decode(oc, llr, O.APPROXIMATE, 20, prec=32, threads=16)
    out = F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(20)).soDecode(llr)
    assert_same_bits(out, ref, frames, oc)
    assert np.isnan(np.asarray(out["msg"])).sum() == 25 * frames


@pytest.mark.parametrize("where", ["first", "last"])
def test_nan_llr_first_or_last_in_a_row(where):
    """A NaN on the first column of row 0 (an information bit) or on its last column (parity
    bit K): FeCl's forward (first) or backward (last) fold keeps the NaN and it spreads; the
    GPU treats it as the row's largest magnitude, so every output of the frame but that
    variable's own a-posteriori value stays finite, its hard decision is 0, and the other
    frame of the batch is untouched."""
    oc, H = short12()
    llr, info = dvbs2_llr(2, 3.0, seed=2000, rate=(1, 2), n=16200, return_info=True)
    v = int(oc.col[oc.row_ptr[0]] if where == "first" else oc.col[oc.row_ptr[1] - 1])
    llr[0, v] = np.nan
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(20))
    out = dec.soDecode(llr)
    app = np.asarray(out["msg"]).reshape(2, oc.K)
    par = np.asarray(out["parity"]).reshape(2, oc.N)
    nan_app = np.isnan(app[0])
    assert nan_app.sum() == (1 if v < oc.K else 0) and (v >= oc.K or nan_app[v])
    assert np.isfinite(par).all() and np.isfinite(np.asarray(out["state"])).all() and np.isfinite(app[1]).all()
    hard = dec.decode(llr).reshape(2, oc.K)
    assert np.array_equal(hard[1], info[1])
    if v < oc.K:
        assert hard[0, v] == 0
    ref = O.decode(oc, llr, O.APPROXIMATE, 20, prec=32, want=("msg",), threads=16)
    assert np.isnan(ref["msg"][0]).sum() > 1                # FeCl's NaN spreads
    ref1 = O.decode(oc, llr[1:], O.APPROXIMATE, 20, prec=32, want=("hard",), threads=16)
    assert np.array_equal(hard[1], ref1["hard"][0])


def test_exact_beyond_tanh_saturation():
    """Eb/N0 12 dB: 12 % of the channel LLRs exceed 37 and messages grow to hundreds.  FeCl FP64
    returns +-inf check-to-variable messages there; the GPU's stay finite, equal the exact
    box-plus value (the oracle's stable FP64 evaluation) within 1e-4 where that is below 30,
    and sit between min(|exact|, 80) and 88 with its sign elsewhere.  Hard decisions equal the
    transmitted bits and FeCl's; iteration counts equal FeCl's."""
    oc, H = short12()
    frames = 8
    llr, info = dvbs2_llr(frames, 12.0, seed=1700, rate=(1, 2), n=16200, return_info=True)
    assert (np.abs(llr) > 37).mean() > 0.05
    fecl = O.decode(oc, llr, O.EXACT, 3, prec=64, threads=16)
    exact = O.decode(oc, llr, O.EXACT_STABLE, 3, prec=64, threads=16)
    assert not np.isfinite(fecl["state"]).all()             # FeCl saturates to +-inf
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Exact).iterations(3))
    out = dec.soDecode(llr)
    g = np.asarray(out["state"], np.float64).reshape(frames, oc.E)
    x = exact["state"]
    assert np.isfinite(g).all() and np.isfinite(np.asarray(out["msg"])).all()
    small = np.abs(x) < 30
    assert np.all(np.abs(g[small] - x[small]) <= 1e-4 * np.maximum(1.0, np.abs(x[small])))
    big = ~small
    assert np.all(np.sign(g[big]) == np.sign(x[big]))
    assert np.all(np.abs(g[big]) >= np.minimum(np.abs(x[big]), 80.0) - 1e-3) and np.abs(g).max() <= 88.0
    assert np.array_equal(out["iterations"], fecl["iters"])
    hard = dec.decode(llr).reshape(frames, oc.K)
    assert np.array_equal(hard, info) and np.array_equal(hard, fecl["hard"])
