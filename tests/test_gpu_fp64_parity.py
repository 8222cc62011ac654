"""The HIP decoder against FeCl's FP64 decoder (the oracle at prec 64, bit-identical to stock
FeCl's `double` build) on identical seeded LLRs: the contract stated against the reference's
own precision, beside the bit-exact contract against FeCl's float instantiation that
test_gpu_decoder.py holds.

Bars and why they are what they are:
  * Min-sum FP32 on dyadic LLRs (multiples of 1/8): identical hard bits, iteration counts,
    a-posteriori values and state on every frame, converged or not.  Min-sum's check update
    only selects and negates stored values, and every variable-node sum of multiples of 1/8
    below 2^21 in magnitude is exact in FP32 (24-bit significand), so nothing rounds in either
    precision; the test asserts that premise on FeCl's own outputs.
  * Min-sum FP32 on real-valued FP32 LLRs: identical hard bits and iteration counts on every
    frame at 1.6 dB and above.  In the waterfall (1.45 dB) a few frames that BP decodes only
    after a long transient (34 or more check updates in FeCl FP64) take a different number of
    updates in any FP32 evaluation: the GPU equals FeCl's own float instantiation on every one
    of them, so the difference is FeCl's float-vs-double difference.  Bar: every frame that
    FeCl FP64 decodes within 30 updates is identical, and at most 8 % of the decoded frames
    differ (FeCl's own float build measured 3-10 % at 1.4-1.55 dB over several seeds).
  * Linear FP32: a-posteriori LLRs within 1e-4 * max(1, |FP64|) on every frame whose table
    lookups all fall on the same side of the table's cutoff as in FP64 (|a +- b| >= 3.5, where
    FeCl's correction drops from 0.0298 to 0, src/detail/LinearTable.h:68-74); the oracle's
    lookup-side hash finds the frames where FP32 rounding carried some |a +- b| across it.
    Those frames also keep FP64's iteration counts, and converged frames FP64's hard
    decisions.
  * Exact FP32: a-posteriori LLRs within 1e-4 * max(1, |FP64|) wherever FeCl's value is
    finite and accurate; FeCl's own FP64 evaluation loses up to ~6e-4 relative where a
    check-to-variable message passes |x| ~ 37 (1 - tanh(|x|/2) cancels in double), and those
    entries are held to the exact box-plus value instead.  Hard decisions and iteration counts
    identical (tests in test_gpu_decoder.py, and here after one iteration, where no entry
    needs the exception, and on converged frames).
  * The reference's own CPU case (16 frames of FP64 LLRs at 1.0 dB, Exact, 50 iterations)
    through the host path: hard decisions and iteration counts identical to FeCl FP64, APP
    as for Exact above.
"""
import numpy as np
import pytest

import fecl_amd as F
from oracle import oracle as O
from tests.golden.vectors import dvbs2_llr

pytestmark = pytest.mark.gpu


def r12():
    oc = O.dvbs2(64800, 1, 2)
    return oc, F.CheckMatrix(oc.row_ptr, oc.col, oc.N)


def gpu_decode_device(H, alg, iters, llr32):
    import torch
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(alg).iterations(iters))
    hard, it = dec.decode(torch.from_numpy(np.ascontiguousarray(llr32, np.float32)).cuda(), iterations=True)
    torch.cuda.synchronize()
    return dec, hard.cpu().numpy().reshape(llr32.shape[0], -1), it.cpu().numpy()


def test_minsum_fp32_equals_fp64_on_dyadic_llrs():
    """The config-2 shape (1024 frames of DVB-S2 r1/2 at Eb/N0 0.6 ... 1.4 dB, 50 iterations),
    LLRs rounded to multiples of 1/8: every frame's hard bits and iteration count equal FeCl
    FP64's, and for a sample the soDecode outputs (a-posteriori LLRs, extrinsic values and
    state) are equal too."""
    oc, H = r12()
    points = (0.6, 0.8, 1.0, 1.2, 1.4)
    llr = np.concatenate([dvbs2_llr(1024 // 5 + (1 if i < 1024 % 5 else 0), e, seed=1100 + i)
                          for i, e in enumerate(points)])
    llr = np.round(llr * 8.0) / 8.0
    assert llr.shape[0] == 1024 and np.array_equal(llr.astype(np.float32).astype(np.float64), llr)
    ref = O.decode(oc, llr, O.APPROXIMATE, 50, prec=64, want=("hard", "iters", "msg"), threads=16)
    # the premise: every value FeCl forms stays a multiple of 1/8 far below 2^21
    assert np.abs(ref["msg"]).max() < 2.0 ** 21
    dec, hard, it = gpu_decode_device(H, F.Approximate, 50, llr)
    assert np.array_equal(it, ref["iters"])
    assert np.array_equal(hard, ref["hard"])
    assert (ref["iters"] < 50).sum() >= 1          # the 1.4 dB frames include decoded ones
    pick = np.concatenate([np.arange(0, 1024, 128), np.flatnonzero(ref["iters"] < 50)[:8]])
    sub = llr[pick]
    r = O.decode(oc, sub, O.APPROXIMATE, 50, prec=64, threads=16)
    g = dec.soDecode(sub)
    for name in ("msg", "syst", "parity", "state"):
        assert np.array_equal(np.asarray(g[name]).reshape(r[name].shape), r[name]), name
    assert np.array_equal(g["iterations"], r["iters"])


@pytest.mark.parametrize("ebn0,frames", [(1.6, 512), (2.0, 512)])
def test_minsum_fp32_equals_fp64_above_the_waterfall(ebn0, frames):
    """Real-valued FP32 LLRs (config 2's input format) at 1.6 and 2.0 dB: hard bits and
    iteration counts identical to FeCl FP64 on every frame."""
    oc, H = r12()
    llr = dvbs2_llr(frames, ebn0, seed=1200 + int(ebn0 * 10)).astype(np.float32)
    ref = O.decode(oc, llr.astype(np.float64), O.APPROXIMATE, 50, prec=64, want=("hard", "iters"), threads=16)
    _, hard, it = gpu_decode_device(H, F.Approximate, 50, llr)
    assert np.array_equal(it, ref["iters"])
    assert np.array_equal(hard, ref["hard"])


def test_minsum_fp32_against_fp64_in_the_waterfall():
    """1.45 dB, where about half the frames decode: frames FeCl FP64 decodes within 30 check
    updates are identical; the frames that differ are decoded ones after a long transient, at
    most 8 % of the decoded frames, and on each of them the GPU equals FeCl's float
    instantiation."""
    oc, H = r12()
    frames, I = 512, 50
    llr = dvbs2_llr(frames, 1.45, seed=1300).astype(np.float32)
    ref = O.decode(oc, llr.astype(np.float64), O.APPROXIMATE, I, prec=64, want=("hard", "iters"), threads=16)
    _, hard, it = gpu_decode_device(H, F.Approximate, I, llr)
    same = (it == ref["iters"]) & np.all(hard == ref["hard"], axis=1)
    decoded = ref["iters"] < I
    early = ref["iters"] <= 30
    assert decoded.sum() >= frames // 4
    assert np.all(same[early]), np.flatnonzero(early & ~same)
    assert (decoded & ~same).sum() <= 0.08 * decoded.sum()
    diff = np.flatnonzero(~same)
    if diff.size:
        f32 = O.decode(oc, llr[diff].astype(np.float64), O.APPROXIMATE, I, prec=32, want=("hard", "iters"),
                       threads=16)
        assert np.array_equal(it[diff], f32["iters"]) and np.array_equal(hard[diff], f32["hard"])
    print(f"1.45 dB: {decoded.sum()} of {frames} decoded by FP64, {(~same).sum()} differ "
          f"(FP64 updates of those decoded: {sorted(ref['iters'][diff][decoded[diff]].tolist())})", flush=True)


@pytest.mark.parametrize("iters", [1, 50])
def test_linear_fp32_app_against_fp64(iters):
    """Linear (LinearBoxSum with FeCl's 8-knot table) on 128 frames at 1.0 dB: after one
    iteration and at convergence, the a-posteriori LLRs are within 1e-4 * max(1, |FP64|) of
    FeCl FP64 on every frame whose table lookups fell on the same side of the cutoff as
    FP64's, with the same iteration count; the remaining frames (FP32 rounding carried some
    |a +- b| across 3.5, so one correction of 0.0298 went the other way) are few at one
    iteration, and at convergence every frame keeps FP64's hard decisions."""
    oc, H = r12()
    llr = dvbs2_llr(128, 1.0, seed=1400).astype(np.float32).astype(np.float64)
    want = ("msg", "iters", "hard", "lin_side")
    ref = O.decode(oc, llr, O.LINEAR, iters, prec=64, want=want, threads=16)
    f32 = O.decode(oc, llr, O.LINEAR, iters, prec=32, want=want, threads=16)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Linear).iterations(iters))
    out = dec.soDecode(llr)
    g = np.asarray(out["msg"], np.float64).reshape(128, oc.K)
    # the GPU is FeCl's float instantiation bit for bit, so the oracle's float hashes are its own
    assert np.array_equal(g.astype(np.float32), f32["msg"].astype(np.float32))
    rel = (np.abs(g - ref["msg"]) / np.maximum(1.0, np.abs(ref["msg"]))).max(axis=1)
    same_side = ref["lin_side"] == f32["lin_side"]
    assert np.all(rel[same_side] <= 1e-4), rel[same_side].max()
    its = np.asarray(out["iterations"])
    assert np.array_equal(its[same_side], ref["iters"][same_side])
    hard = dec.decode(llr).reshape(128, oc.K)
    if iters == 1:
        assert (~same_side).sum() <= 8
    else:
        # a crossing can move the syndrome's success by one update (one frame of 128 here)
        conv = ref["iters"] < iters
        assert conv.sum() >= 120
        assert np.array_equal(hard[conv], ref["hard"][conv])
        assert (its != ref["iters"]).sum() <= 4
    print(f"Linear I={iters}: {(~same_side).sum()} of 128 frames crossed the cutoff; max relative "
          f"APP error {rel.max():.3g} over all, {rel[same_side].max():.3g} over the others", flush=True)


@pytest.mark.parametrize("iters", [1, 50])
def test_exact_fp32_app_against_fp64(iters):
    """Exact on 64 frames at 1.0 dB, after one iteration and at convergence: APP within
    1e-4 * max(1, |FP64|) wherever FeCl's value is finite, except entries where FeCl's own FP64
    evaluation has lost accuracy (|APP| >= 36: 1 - tanh(|x|/2) cancels in double), which are
    checked against the exact box-plus value instead; identical hard decisions and iteration
    counts.  After one iteration no entry needs the exception."""
    from tests.test_gpu_decoder import assert_app_against_fecl_fp64
    oc, H = r12()
    llr = dvbs2_llr(64, 1.0, seed=1500).astype(np.float32).astype(np.float64)
    ref = O.decode(oc, llr, O.EXACT, iters, prec=64, want=("msg", "iters", "hard"), threads=16)
    exact = O.decode(oc, llr, O.EXACT_STABLE, iters, prec=64, want=("msg",), threads=16)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Exact).iterations(iters))
    out = dec.soDecode(llr)
    g = np.asarray(out["msg"], np.float64).reshape(64, oc.K)
    x = exact["msg"]
    assert np.all(np.isfinite(g))
    assert (np.abs(g - x) / np.maximum(1.0, np.abs(x))).max() <= 1e-5
    over = assert_app_against_fecl_fp64(g.ravel(), ref["msg"].ravel(), x.ravel())
    if iters == 1:
        assert over == 0
    assert np.array_equal(out["iterations"], ref["iters"])
    assert np.array_equal(dec.decode(llr).reshape(64, oc.K), ref["hard"])
    if iters == 50:
        assert (ref["iters"] < 50).sum() >= 60
    print(f"Exact I={iters}: {over} of {g.size} APP entries beyond 1e-4 of FeCl FP64, all with |APP| >= 36",
          flush=True)


def test_reference_cpu_case_fp64_llrs():
    """The reference's own CPU-runnable case: DVB-S2 n=64800 r1/2, 16 frames, Eb/N0 1.0 dB,
    FP64 LLRs as the channel gives them (not rounded to FP32 first), Exact, 50 iterations,
    through the host path FeCl's decode hook uses: hard decisions and iteration counts equal
    FeCl FP64's, and the a-posteriori LLRs (host soDecode, FP64 out) are within 1e-4 of them,
    or of the exact value where FeCl's FP64 evaluation has lost accuracy."""
    oc, H = r12()
    llr, info = dvbs2_llr(16, 1.0, seed=1600, return_info=True)
    assert np.any(llr.astype(np.float32).astype(np.float64) != llr)
    ref = O.decode(oc, llr, O.EXACT, 50, prec=64, want=("hard", "iters", "msg"), threads=16)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Exact).iterations(50))
    hard, it = dec.decode(llr, iterations=True)
    assert np.array_equal(it, ref["iters"])
    assert np.array_equal(hard.reshape(16, oc.K), ref["hard"])
    assert np.array_equal(ref["hard"], info)      # every frame decodes at this Eb/N0
    from tests.test_gpu_decoder import assert_app_against_fecl_fp64
    g = np.asarray(dec.soDecode(llr, outputs=("msg",))["msg"]).reshape(16, oc.K)
    x = O.decode(oc, llr, O.EXACT_STABLE, 50, prec=64, want=("msg",), threads=16)["msg"]
    assert (np.abs(g - x) / np.maximum(1.0, np.abs(x))).max() <= 1e-5
    assert_app_against_fecl_fp64(g.ravel(), ref["msg"].ravel(), x.ravel())
