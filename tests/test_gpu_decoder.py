"""GPU parity: the HIP decoder (through the C ABI) against the CPU oracle on the same LLRs.

Bar, per path:
  * min-sum FP32 and Linear FP32: bit-identical to the oracle's FP32 restatement (which is
    bit-identical to FeCl's own decoder instantiated with a float LLR type): hard bits,
    a-posteriori/extrinsic values, state and iteration counts.
  * min-sum / Linear with FP16 or BF16 message storage: bit-identical to the oracle's FP32
    restatement with the same storage rounding.
  * Exact FP32 (BoxSum's tanh product carried as q = 1 - |tanh(x/2)|, which neither cancels
    nor saturates) against FeCl's FP64 decoder: identical
    hard decisions and iteration counts on these inputs, a-posteriori LLRs within 1e-4
    relative (|d| <= 1e-4 * max(1, |ref|)) where FeCl's value is finite.
"""
import numpy as np
import pytest

import fecl_amd as F
from oracle import oracle as O
from tests.golden.vectors import dvbs2_llr, gallager_llr

pytestmark = pytest.mark.gpu

GOLD = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "fecl_golden.npz"))


def gallager():
    """FeCl's test code: the decoder is built from Gallager::matrix(2048, 3, 5, 0) as generated;
    the oracle runs on the form stock FeCl decodes with (its Structure's checks(), a fixture)."""
    N, K = (int(x) for x in GOLD["gallager_nk"])
    oc = O.Code(GOLD["gallager_row_ptr"], GOLD["gallager_col"], N, K)
    return oc, F.Gallager.matrix(2048, 3, 5, 0)


def dvbs2(n=64800, rate=(1, 2)):
    oc = O.dvbs2(n, *rate)
    return oc, F.CheckMatrix(oc.row_ptr, oc.col, oc.N)


def gpu_decoder(H, alg, iters, sf=1.0, fmt=F.MsgFormat.F32, words=0, fused=False):
    opt = F.DecoderOptions().algorithm(alg).iterations(iters).scalingFactor(sf).msgFormat(fmt)
    dec = F.Ldpc(H, opt)
    if words:
        dec.setLayout(words)
    if fused:
        dec.setFusedSyndrome(True)
    return dec


def oracle_run(oc, llr, alg, iters, sf=1.0, prec=32, fmt=O.MSG_F32, **kw):
    return O.decode(oc, llr, int(alg), iters, sf=O.scaling_table(oc, sf), prec=prec, msg_fmt=int(fmt), **kw)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64 if a.dtype == np.float64 else np.uint32)


def assert_bit_exact(gpu, ref, frames, K, N, E):
    assert np.array_equal(gpu["iterations"], ref["iters"]), "iteration counts"
    for name, n in (("msg", K), ("syst", K), ("parity", N), ("state", E)):
        g = np.asarray(gpu[name]).reshape(frames, n)
        r = ref[name].reshape(frames, n)
        # the oracle's FP32 values are exact in float64
        assert np.array_equal(bits(g.astype(np.float32)), bits(r.astype(np.float32))), name


@pytest.mark.parametrize("alg", [F.Approximate, F.Linear])
@pytest.mark.parametrize("iters,sf,words,fused", [(1, 1.0, 1, False), (2, 1.0, 4, True), (5, 1.0, 2, False),
                                                  (30, 1.0, 1, True), (30, 1.0, 4, False), (30, 0.75, 2, True),
                                                  (30, 0.9, 4, False), (5, [0.5, 0.625, 0.75, 0.875, 1.0], 1, True),
                                                  (5, [0.5, 0.625, 0.75, 0.875, 1.0], 4, False)])
def test_gallager_fp32_bit_exact(alg, iters, sf, words, fused):
    oc, H = gallager()
    llr = np.concatenate([gallager_llr(oc.N, oc.K, 3, 2.0, 11), gallager_llr(oc.N, oc.K, 2, -1.0, 12)])
    llr32 = llr.astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr32, alg, iters, sf)
    dec = gpu_decoder(H, alg, iters, {0: sf} if isinstance(sf, list) else sf, words=words, fused=fused)
    out = dec.soDecode(llr32)
    assert_bit_exact(out, ref, llr.shape[0], oc.K, oc.N, oc.E)
    hard, it = dec.decode(llr32, iterations=True)
    assert np.array_equal(hard.reshape(-1, oc.K), ref["hard"])
    assert np.array_equal(it, ref["iters"])


@pytest.mark.parametrize("frames,words,fused", [(1, 4, False), (63, 1, True), (65, 2, False), (130, 4, True),
                                                (300, 4, False)])
def test_dvbs2_minsum_fp32_ragged_frames(frames, words, fused):
    oc, H = dvbs2()
    llr = dvbs2_llr(frames, 1.4, seed=100 + frames).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, 50)
    dec = gpu_decoder(H, F.Approximate, 50, words=words, fused=fused)
    assert dec.layout()[0] == words
    out = dec.soDecode(llr)
    assert_bit_exact(out, ref, frames, oc.K, oc.N, oc.E)
    hard, it = dec.decode(llr, iterations=True)
    assert np.array_equal(hard.reshape(frames, oc.K), ref["hard"])


@pytest.mark.parametrize("fmt,ofmt,words", [(F.MsgFormat.F32, O.MSG_F32, 1), (F.MsgFormat.F32, O.MSG_F32, 2),
                                             (F.MsgFormat.F16, O.MSG_F16, 1), (F.MsgFormat.F16, O.MSG_F16, 2)])
def test_two_variables_per_wave(fmt, ofmt, words):
    """The variable kernel serving two variables per wavefront gives the same bits, including
    a code with an odd variable count (the Gallager code without its last column), whose last
    wavefront serves one variable."""
    oc, H = dvbs2(16200, (1, 2))
    llr = dvbs2_llr(70, 1.6, seed=17, rate=(1, 2), n=16200).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, 30, fmt=ofmt)
    dec = gpu_decoder(H, F.Approximate, 30, fmt=fmt, words=words)
    dec.setVarsPerWave(2)
    assert dec.varsPerWave() == 2
    assert_bit_exact(dec.soDecode(llr), ref, 70, oc.K, oc.N, oc.E)
    g, _ = gallager()
    keep = g.col < g.N - 1
    row_ptr = np.concatenate([[0], np.cumsum([int(keep[a:b].sum()) for a, b in zip(g.row_ptr[:-1], g.row_ptr[1:])])])
    H2 = F.CheckMatrix(row_ptr, g.col[keep], g.N - 1)
    dec2 = gpu_decoder(H2, F.Linear, 20, fmt=fmt, words=words)
    S2 = dec2.checks()
    oc2 = O.Code(S2.row_ptr, S2.col, S2.cols(), dec2.msgSize())
    llr2 = gallager_llr(oc2.N, oc2.K, 9, 2.0, seed=5)
    ref2 = oracle_run(oc2, llr2, F.Linear, 20, fmt=ofmt)
    dec2.setVarsPerWave(2)
    assert_bit_exact(dec2.soDecode(llr2), ref2, 9, oc2.K, oc2.N, oc2.E)


@pytest.mark.parametrize("shape", [(1000, 3, 5, 1), (1200, 4, 8, 9)])
def test_codes_with_dependent_checks(shape):
    """A Gallager code whose checks are dependent: FeCl's form carries K > N - M message bits
    (src/detail/Ldpc.cpp:307) and the decoder follows that form: soDecode is bit-exact to the
    oracle run on it, and messages from the host encoder come back through the channel."""
    from tests.golden.vectors import awgn_llr
    H = F.Gallager.matrix(*shape)
    dec = gpu_decoder(H, F.Approximate, 30)
    S, K = dec.checks(), dec.msgSize()
    assert K > H.cols() - H.rows() and (S, K) == F.structure(H)
    oc = O.Code(S.row_ptr, S.col, S.cols(), K)
    llr = gallager_llr(oc.N, oc.K, 6, 2.0, seed=8)
    assert_bit_exact(dec.soDecode(llr), oracle_run(oc, llr, F.Approximate, 30), 6, oc.K, oc.N, oc.E)
    rng = np.random.default_rng(4)
    msg = rng.integers(0, 2, size=(5, K), dtype=np.uint8)
    cw = dec.encode(msg).reshape(5, oc.N)
    assert dec.check(cw)
    assert np.array_equal(dec.decode(awgn_llr(cw, K / oc.N, 4.0, rng)).reshape(5, K), msg)


@pytest.mark.parametrize("pass_frames", [0, 256, 300])
def test_dvbs2_minsum_fp32_1024_frames_device_path(pass_frames):
    """The config-2 shape: 1024 frames of DVB-S2 r1/2, FP32 LLRs in HBM, min-sum, 50 iterations;
    whole batch in one pass, or in passes of 256 frames and of 300 frames (two tiles)."""
    import torch
    oc, H = dvbs2()
    rng_frames = [dvbs2_llr(256, e, seed=300 + i) for i, e in enumerate((0.6, 1.0, 1.4, 2.0))]
    llr = np.concatenate(rng_frames).astype(np.float32)
    ref = oracle_run(oc, llr.astype(np.float64), F.Approximate, 50, want=("hard", "iters"), threads=16)
    dec = gpu_decoder(H, F.Approximate, 50, words=4)
    dec.setPassFrames(pass_frames)
    d_llr = torch.from_numpy(llr).cuda()
    hard, it = dec.decode(d_llr, iterations=True)
    torch.cuda.synchronize()
    assert np.array_equal(hard.cpu().numpy().reshape(-1, oc.K), ref["hard"])
    assert np.array_equal(it.cpu().numpy(), ref["iters"])


@pytest.mark.parametrize("fmt,ofmt", [(F.MsgFormat.F16, O.MSG_F16), (F.MsgFormat.BF16, O.MSG_BF16)])
@pytest.mark.parametrize("alg,words", [(F.Approximate, 1), (F.Approximate, 4), (F.Linear, 2), (F.Linear, 4)])
def test_16bit_messages_bit_exact(fmt, ofmt, alg, words):
    oc, H = dvbs2()
    llr = dvbs2_llr(5, 1.2, seed=7).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, alg, 50, fmt=ofmt)
    out = gpu_decoder(H, alg, 50, fmt=fmt, words=words).soDecode(llr)
    assert_bit_exact(out, ref, 5, oc.K, oc.N, oc.E)


@pytest.mark.parametrize("rate", [(1, 4), (3, 5), (2, 3), (3, 4), (5, 6), (9, 10)])
def test_all_normal_rates_minsum_f16(rate):
    """Every DVB-S2 normal rate (row degrees up to 30, variable degrees up to 13) in the
    widest layout the register budget allows."""
    oc, H = dvbs2(64800, rate)
    llr = dvbs2_llr(3, 2.5 + 2 * rate[0] / rate[1], seed=8, rate=rate).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, 30, fmt=O.MSG_F16)
    out = gpu_decoder(H, F.Approximate, 30, fmt=F.MsgFormat.F16).soDecode(llr)
    assert_bit_exact(out, ref, 3, oc.K, oc.N, oc.E)


@pytest.mark.parametrize("rate,words,sf,fmt,ofmt", [
    ((9, 10), 4, 1.0, F.MsgFormat.F16, O.MSG_F16),     # rows of 30 over 8 frames a lane: streamed check kernel
    ((9, 10), 4, 0.75, F.MsgFormat.F16, O.MSG_F16),    # streamed, with a scaling factor
    ((5, 6), 2, 0.75, F.MsgFormat.BF16, O.MSG_BF16),   # rows of 22 over 4 frames a lane: streamed, scaled
    ((2, 3), 1, 0.75, F.MsgFormat.BF16, O.MSG_BF16),   # register-staged float form (rows of 10), scaled
    ((9, 10), 1, 1.0, F.MsgFormat.F32, O.MSG_F32),     # raw FP32 form, rows of 30
    ((1, 4), 4, 0.875, F.MsgFormat.F32, O.MSG_F32)])   # scaled float form, variable degree 12
def test_check_kernel_forms_on_high_degree_codes(rate, words, sf, fmt, ofmt):
    """Each min-sum check-kernel form (raw encodings, float with scaling, streamed rows) on DVB-S2
    codes whose rows exceed the register-staged bound for the layout."""
    oc, H = dvbs2(64800, rate)
    llr = dvbs2_llr(4, 2.5 + 2 * rate[0] / rate[1], seed=21, rate=rate).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, 20, sf=sf, fmt=ofmt)
    dec = gpu_decoder(H, F.Approximate, 20, sf=sf, fmt=fmt, words=words)
    assert dec.layout()[0] == words
    assert_bit_exact(dec.soDecode(llr), ref, 4, oc.K, oc.N, oc.E)


# Eb/N0 (dB) near each normal-frame rate's waterfall, so that some frames stop early and some
# run out their iterations; short frames get 0.4 dB more
_WATERFALL = {(1, 4): 1.1, (1, 3): 1.3, (2, 5): 1.5, (1, 2): 1.7, (3, 5): 2.4, (2, 3): 2.3, (3, 4): 2.6,
              (4, 5): 2.9, (5, 6): 3.2, (8, 9): 3.8, (9, 10): 4.1}


def assert_app_against_fecl_fp64(g, r, x):
    """APP g (GPU) against FeCl FP64 r: within 1e-4 * max(1, |r|) wherever r is finite, except
    where FeCl's own FP64 evaluation has lost accuracy.  FeCl's post -log((1+p)/(1-p)) cancels
    in 1 - p once a check-to-variable message passes |x| ~ 37 (tanh(x/2) within 1e-16 of 1), so
    its APP there sits up to ~6e-4 relative from the exact box-plus value x.  Each entry beyond
    the bar must lie in that zone (|x| >= 36) and FeCl's distance from x must account for the
    difference; the same signs where FeCl saturates to +-inf."""
    fin = np.isfinite(r)
    scale = np.maximum(1.0, np.abs(r))
    over = fin & (np.abs(g - r) > 1e-4 * scale)
    assert np.all(np.abs(x[over]) >= 36.0), np.abs(x[over]).min()
    assert np.all(np.abs(g[over] - r[over]) <= np.abs(r[over] - x[over]) + 1e-4 * scale[over])
    assert np.all(np.sign(g[~fin]) == np.sign(r[~fin]))
    return int(over.sum())


def assert_exact_close(oc, llr, H, iters, sf, bar):
    """Exact FP32 on the GPU against FeCl FP64 (`ref`) and the exact box-plus value (`exact`,
    the oracle's FP64 stable evaluation): identical hard decisions and iteration counts; APP
    within `bar` relative (to max(1, |exact|)) of the exact value, and within 1e-4 relative of
    FeCl FP64 as assert_app_against_fecl_fp64 states it."""
    ref = oracle_run(oc, llr, F.Exact, iters, sf=sf, prec=64, threads=16)
    exact = oracle_run(oc, llr, O.EXACT_STABLE, iters, sf=sf, prec=64, threads=16)
    out = gpu_decoder(H, F.Exact, iters, sf=sf).soDecode(llr)
    hard = gpu_decoder(H, F.Exact, iters, sf=sf).decode(llr)
    assert np.array_equal(hard.reshape(-1, oc.K), ref["hard"])
    assert np.array_equal(out["iterations"], ref["iters"])
    r, x = ref["msg"].ravel(), exact["msg"].ravel()
    g = out["msg"].astype(np.float64).ravel()
    scale = np.maximum(1.0, np.abs(x))
    assert (np.abs(g - x) / scale).max() <= bar
    assert_app_against_fecl_fp64(g, r, x)


@pytest.mark.parametrize("code", ["gallager", "dvbs2"])
def test_exact_fp32_against_fp64(code):
    """Exact: FeCl evaluates BoxSum as -log((1+p)/(1-p)) with p a product of tanh(-x/2)
    (src/detail/LlrMetrics.h:126-140).  In FP64 that form loses relative accuracy as |p| -> 1
    (1-p cancels), so FeCl's own APP deviates from the exact box-plus value by up to ~4e-4 at
    |APP| ~ 60.  The GPU carries the complement q = 1 - |p| (product 1-q = (1-q_a)(1-q_b),
    post log((2-q)/q)), which evaluates the same value without the cancellation.  Bars: hard
    decisions and iteration counts identical to FeCl FP64; APP within 2e-6 relative of the
    exact value, and within 1e-4 relative of FeCl FP64 plus FeCl's own distance from it."""
    oc, H = gallager() if code == "gallager" else dvbs2()
    if code == "gallager":
        llr = gallager_llr(oc.N, oc.K, 4, 2.0, 11)
    else:
        llr = dvbs2_llr(3, 1.0, seed=21)
    assert_exact_close(oc, llr.astype(np.float32).astype(np.float64), H, 50, 1.0, 2e-6)


@pytest.mark.parametrize("sf", [1.0, 0.75])
def test_every_dvbs2_code_exact_against_fp64(sf):
    """Exact FP32 on all 21 DVB-S2 codes, 3 frames each near the code's waterfall (the inputs
    of test_every_dvbs2_code_bit_exact), 30 iterations, with and without a scaling factor:
    hard decisions and iteration counts identical to FeCl FP64, APP within 1e-5 relative of
    the exact box-plus value."""
    from fecl_amd.ldpc import _DVBS2_RATES
    for n, rates in sorted(_DVBS2_RATES.items()):
        for rate in rates:
            oc, H = dvbs2(n, rate)
            ebn0 = _WATERFALL[rate] + (0.4 if n == 16200 else 0.0)
            llr = dvbs2_llr(3, ebn0, seed=31 + rate[0] * 7 + rate[1], rate=rate, n=n)
            print(f"code n={n} rate={rate[0]}/{rate[1]} Eb/N0={ebn0} sf={sf}", flush=True)
            assert_exact_close(oc, llr.astype(np.float32).astype(np.float64), H, 30, sf, 1e-5)


@pytest.mark.parametrize("words", [1, 4])
def test_sodecode_state_and_syst_inputs(words):
    """soDecode with a-priori state and systematic inputs (BpDecoderImpl.cpp:73-93)."""
    oc, H = gallager()
    llr = gallager_llr(oc.N, oc.K, 2, -1.0, 12).astype(np.float32).astype(np.float64)
    first = oracle_run(oc, llr, F.Approximate, 7)
    state = first["state"].astype(np.float32).astype(np.float64)
    syst = np.random.default_rng(99).normal(0, 1, (2, oc.K)).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, 7, state_in=state, syst_in=syst)
    out = gpu_decoder(H, F.Approximate, 7, words=words).soDecode(llr, syst=syst, state=state)
    assert_bit_exact(out, ref, 2, oc.K, oc.N, oc.E)


def test_host_sodecode_adds_fp64_apriori_once():
    """FeCl's decoder with a float LLR type stores the channel value as float, adds the double
    a-priori value and rounds once (src/detail/BpDecoder/BpDecoderImpl.cpp:75-80).  The host
    soDecode path gets FP64 syst values that are not FP32-representable, among them 2^-24 +
    2^-50 on a channel value of 1.0 (one rounding gives 1 + 2^-23, two roundings give 1.0):
    every output equals the oracle's, which adds them the same way."""
    oc, H = gallager()
    llr = gallager_llr(oc.N, oc.K, 3, 1.0, 14).astype(np.float32).astype(np.float64)
    syst = np.random.default_rng(97).normal(0, 1, (3, oc.K))
    llr[0, :8] = 1.0
    syst[0, :8] = 2.0 ** -24 + 2.0 ** -50
    assert np.any(syst.astype(np.float32).astype(np.float64) != syst)
    for iters in (1, 6):
        ref = oracle_run(oc, llr, F.Approximate, iters, syst_in=syst)
        out = gpu_decoder(H, F.Approximate, iters).soDecode(llr, syst=syst)
        assert_bit_exact(out, ref, 3, oc.K, oc.N, oc.E)
        # the first 8 a-posteriori values with one iteration carry 1 + 2^-23, not 1.0
        if iters == 1:
            twice = oracle_run(oc, llr, F.Approximate, 1, syst_in=syst.astype(np.float32).astype(np.float64))
            assert not np.array_equal(twice["msg"][0, :8], ref["msg"][0, :8])


def test_host_buffer_path_in_several_chunks():
    """FeCl's hooks pass host FP64 buffers: 1300 frames go through device staging in chunks of
    512 whose input copies overlap the previous chunk's decode; with a-priori state and
    systematic inputs, every output equals the oracle's."""
    oc, H = gallager()
    frames = 1300
    llr = np.concatenate([gallager_llr(oc.N, oc.K, frames // 2, 2.0, 31),
                          gallager_llr(oc.N, oc.K, frames - frames // 2, 0.5, 32)]).astype(np.float32).astype(np.float64)
    first = oracle_run(oc, llr, F.Approximate, 4, threads=16)
    state = first["state"].astype(np.float32).astype(np.float64)
    syst = np.random.default_rng(98).normal(0, 1, (frames, oc.K)).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, 12, state_in=state, syst_in=syst, threads=16)
    dec = gpu_decoder(H, F.Approximate, 12, words=2)
    assert dec.layout()[1] < 512
    out = dec.soDecode(llr, syst=syst, state=state)
    assert_bit_exact(out, ref, frames, oc.K, oc.N, oc.E)
    hard, it = dec.decode(llr, iterations=True)
    ref2 = oracle_run(oc, llr, F.Approximate, 12, want=("hard", "iters"), threads=16)
    assert np.array_equal(hard.reshape(frames, oc.K), ref2["hard"]) and np.array_equal(it, ref2["iters"])


def test_autotune_picks_an_eligible_layout():
    oc, H = dvbs2()
    dec = gpu_decoder(H, F.Approximate, 50)
    best, times = dec.autotune(256, probe_iterations=3)
    assert (best, dec.varsPerWave()) in times and dec.layout()[0] == best
    llr = dvbs2_llr(2, 1.4, seed=3).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, 50)
    hard, it = dec.decode(llr, iterations=True)
    assert np.array_equal(hard.reshape(2, oc.K), ref["hard"]) and np.array_equal(it, ref["iters"])


def test_two_phase_resume_equals_single_run():
    """test_soDecode_2phases (test/Cpp/operations.h:344-383): I iterations, state out, state
    back in with I more == one 2I run, at a low SNR where nothing converges."""
    oc, H = gallager()
    llr = gallager_llr(oc.N, oc.K, 2, -5.0, 5).astype(np.float32).astype(np.float64)
    d1 = gpu_decoder(H, F.Approximate, 10)
    a = d1.soDecode(llr, outputs=("state",))
    b = d1.soDecode(llr, state=a["state"], outputs=("msg",))
    c = gpu_decoder(H, F.Approximate, 20).soDecode(llr, outputs=("msg",))
    assert np.array_equal(b["msg"], c["msg"])


def test_zero_frames_and_errors():
    oc, H = gallager()
    dec = gpu_decoder(H, F.Approximate, 5)
    assert dec.decode(np.zeros(0)).size == 0
    with pytest.raises(ValueError):
        dec.decode(np.zeros(oc.N + 1))
    with pytest.raises(ValueError):
        dec.soDecode(np.zeros(oc.N), syst=np.zeros(3))
    with pytest.raises(ValueError):
        F.Ldpc(H, F.DecoderOptions().iterations(0))
    with pytest.raises(ValueError):
        dec.setPassFrames(-1)


def test_batch_graph_multirate_replays():
    """Several codes in one hipGraph launch; replayed with new inputs copied in eagerly between
    launches, every replay checked against the oracle and against the eager decode."""
    import torch
    specs = [(64800, (1, 2), 40, 1.6), (64800, (3, 4), 70, 3.2), (16200, (1, 2), 130, 2.0)]
    decs, tensors, codes = [], [], []
    for n, rate, frames, _ in specs:
        oc, H = dvbs2(n, rate)
        codes.append(oc)
        decs.append(gpu_decoder(H, F.Approximate, 50, fmt=F.MsgFormat.F16))
        if n == 16200:  # two passes of one 128-frame tile and two frames
            decs[-1].setLayout(1)
            decs[-1].setPassFrames(100)
        tensors.append(torch.zeros(frames * oc.N, dtype=torch.float32, device="cuda"))
    g = F.BatchGraph(decs, tensors)
    for rep in range(3):
        refs = []
        for i, (n, rate, frames, ebn0) in enumerate(specs):
            llr = dvbs2_llr(frames, ebn0 + 0.2 * rep, seed=40 + 10 * rep + i, rate=rate, n=n).astype(np.float32)
            tensors[i].copy_(torch.from_numpy(llr.ravel()))
            refs.append(oracle_run(codes[i], llr.astype(np.float64), F.Approximate, 50, fmt=O.MSG_F16,
                                   want=("hard", "iters")))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for i, (n, rate, frames, _) in enumerate(specs):
            K = codes[i].K
            got = g.msg[i].cpu().numpy().reshape(frames, K)
            assert np.array_equal(got, refs[i]["hard"]), (rep, i)
            assert np.array_equal(g.iterations[i].cpu().numpy(), refs[i]["iters"]), (rep, i)
            eager, it = decs[i].decode(tensors[i], iterations=True)
            torch.cuda.synchronize()
            assert np.array_equal(eager.cpu().numpy().reshape(frames, K), got)


def test_batch_graph_keeps_its_options_after_set_options():
    """A captured graph decodes with the options its handles had at creation: changing a
    handle's options afterwards (here from one scaling row to a per-iteration table, which
    reallocates the handle's device table) leaves every replay equal to the oracle run with the
    original options, while eager decodes on the handle follow the new ones."""
    import torch
    oc, H = dvbs2(16200, (1, 2))
    frames = 70
    dec = gpu_decoder(H, F.Approximate, 20, sf=0.75)
    x = torch.zeros(frames * oc.N, dtype=torch.float32, device="cuda")
    g = F.BatchGraph([dec], [x])
    new_sf = [0.5, 0.625, 0.75, 0.875, 1.0] * 4
    dec.setDecoderOptions(F.DecoderOptions().algorithm(F.Approximate).iterations(20).scalingFactor({0: new_sf}))
    for rep in range(2):
        llr = dvbs2_llr(frames, 2.0 + 0.2 * rep, seed=60 + rep, rate=(1, 2), n=16200).astype(np.float32)
        x.copy_(torch.from_numpy(llr.ravel()))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        ref = oracle_run(oc, llr.astype(np.float64), F.Approximate, 20, sf=0.75, want=("hard", "iters"))
        assert np.array_equal(g.msg[0].cpu().numpy().reshape(frames, oc.K), ref["hard"]), rep
        assert np.array_equal(g.iterations[0].cpu().numpy(), ref["iters"]), rep
        now = oracle_run(oc, llr.astype(np.float64), F.Approximate, 20, sf=new_sf,
                         want=("hard", "iters"))
        _, it = dec.decode(llr.astype(np.float64), iterations=True)
        assert np.array_equal(it, now["iters"]), rep


def test_device_encoder_matches_the_oracle_for_all_dvbs2_codes():
    """Codec::encode on the device (staircase accumulator, one workgroup per frame) equals the
    oracle's accumulator encoder (written separately, on the oracle's matrix) and the host
    encoder bit for bit on every DVB-S2 code, and its codewords satisfy every check."""
    import torch
    rng = np.random.default_rng(11)
    normal = [(1, 4), (1, 3), (2, 5), (1, 2), (3, 5), (2, 3), (3, 4), (4, 5), (5, 6), (8, 9), (9, 10)]
    short = [(1, 4), (1, 3), (2, 5), (1, 2), (3, 5), (2, 3), (3, 4), (4, 5), (5, 6), (8, 9)]
    for n, rates in ((64800, normal), (16200, short)):
        for rate in rates:
            oc = O.dvbs2(n, *rate)
            dec = F.Ldpc(F.DvbS2.matrix(n, rate), F.DecoderOptions().iterations(1))
            frames = 5
            msg = rng.integers(0, 2, frames * dec.msgSize(), dtype=np.uint8)
            dev = dec.encode(torch.from_numpy(msg).cuda())
            torch.cuda.synchronize()
            dev = dev.cpu().numpy()
            ref = O.encode(oc, msg.reshape(frames, oc.K)).ravel()
            assert np.array_equal(dev, ref), (n, rate)
            assert np.array_equal(dec.encode(msg), ref), (n, rate)
            assert all(O.syndrome_weight(oc, c) == 0 for c in dev.reshape(frames, oc.N)), (n, rate)


def test_concurrent_calls_on_one_handle():
    """FeCl's Codec::decode calls decodeBlocks from several worker threads at once on disjoint
    frames (src/Codec.h:224-237): concurrent host-path calls on one handle each get their own
    workspace and the results equal the single-threaded ones."""
    import threading
    oc, H = dvbs2()
    dec = gpu_decoder(H, F.Approximate, 50)
    parts = [dvbs2_llr(24, 1.2 + 0.2 * i, seed=200 + i).astype(np.float32).astype(np.float64) for i in range(4)]
    alone = [dec.decode(x, iterations=True) for x in parts]
    got = [None] * len(parts)

    def work(i):
        for _ in range(3):
            got[i] = dec.decode(parts[i], iterations=True)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(parts))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for (h1, i1), (h2, i2) in zip(alone, got):
        assert np.array_equal(h1, h2) and np.array_equal(i1, i2)


def test_punctured_decode_device_permutation():
    """Puncturing (src/detail/Ldpc.cpp:333-353) as FeCl's tests use it (test/Cpp/operations.h:
    130-145): transmit perm.permute(codeword), receive, dePermute (punctured positions enter
    as 0-LLR), decode.  The device permutation equals the host one bit for bit, and so do the
    decodes that follow; at this Eb/N0 every frame decodes."""
    import torch
    oc, H = dvbs2()
    dec = gpu_decoder(H, F.Approximate, 50)
    perm = dec.puncturing(F.PunctureOptions().mask([1, 0]).systMask([1]))
    assert perm.inputSize() == oc.N and oc.K < perm.outputSize() < oc.N
    rng = np.random.default_rng(41)
    frames = 6
    info = rng.integers(0, 2, (frames, oc.K), dtype=np.uint8)
    cw = dec.encode(info).reshape(frames, oc.N)
    tx = perm.permute(cw.astype(np.float64))
    rate = oc.K / perm.outputSize()
    sigma2 = 1.0 / (2 * rate * 10 ** (4.0 / 10))
    rx = (-2.0 * ((1.0 - 2.0 * tx) + rng.normal(0, np.sqrt(sigma2), tx.shape)) / sigma2).astype(np.float32)
    host_full = perm.dePermute(rx.astype(np.float64)).astype(np.float32)
    dev_full = perm.dePermute(torch.from_numpy(rx).cuda())
    dev_tx = perm.permute(dev_full)
    torch.cuda.synchronize()
    assert np.array_equal(dev_full.cpu().numpy(), host_full)
    assert np.array_equal(dev_tx.cpu().numpy(), rx)
    bits_dev, _ = dec.decode(dev_full, iterations=True)
    torch.cuda.synchronize()
    bits_host = dec.decode(host_full.astype(np.float64))
    assert np.array_equal(bits_dev.cpu().numpy(), bits_host)
    assert np.array_equal(bits_host.reshape(frames, oc.K), info)


def test_saved_codecs_round_trip_and_decode_like_the_original():
    """fec::detail::save / load (src/Serialization.h:81-115) through decoder handles.  A handle
    built from FeCl's test code saves the archive stock FeCl writes (Exact, 30 iterations; and
    Linear with per-iteration scaling on a small Gallager code); a handle loaded from FeCl's
    archive keeps the stored structure, saves the same bytes again and decodes bit for bit like
    the handle built from the matrix."""
    import hashlib
    import json
    here = __import__("os").path.dirname(__file__)
    cases = {c["name"]: c for c in json.load(open(__import__("os").path.join(here, "golden", "fecl_archive.json")))["cases"]}
    c = cases["gallager2048_exact"]
    dec = F.Ldpc(F.Gallager.matrix(2048, 3, 5, 0), F.DecoderOptions().algorithm(F.Exact).iterations(30), 8)
    assert hashlib.sha256(dec.save()).hexdigest() == c["archive_sha256"]
    blob = np.load(__import__("os").path.join(here, "golden", "fecl_archive.npz"))["gallager30_linear_vector"].tobytes()
    opts = (F.DecoderOptions().algorithm(F.Linear).iterations(5)
            .scalingFactor({0: [0.5, 0.625, 0.75, 0.875, 1.0]}))
    built = F.Ldpc(F.Gallager.matrix(30, 3, 5, 0), opts, 4)
    assert built.save() == blob
    loaded = F.Ldpc.load(blob + bytes(4096))
    assert loaded.getWorkGroupSize() == 4 and loaded.save() == blob
    assert loaded.checks() == built.checks() and loaded.msgSize() == built.msgSize()
    llr = gallager_llr(30, built.msgSize(), 7, 1.0, seed=3)
    a, b = built.soDecode(llr), loaded.soDecode(llr)
    for k in ("msg", "syst", "parity", "state", "iterations"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_decoder_from_a_structure_given_in_form():
    """fec::Ldpc(const Structure&) (src/Ldpc.h:83), the entry a binding inside FeCl uses: the
    matrix FeCl's Structure holds (a fixture) with its msgSize, taken as given, decodes bit for
    bit like the decoder built from Gallager::matrix; it has no encoder unless the matrix is its
    own form (DVB-S2), where encoding and saving match the ordinary handle."""
    oc, H = gallager()
    S = F.CheckMatrix(oc.row_ptr, oc.col, oc.N)
    opts = F.DecoderOptions().algorithm(F.Approximate).iterations(30)
    a = F.Ldpc(H, opts)
    b = F.Ldpc.fromStructure(S, oc.K, opts)
    assert b.checks() == a.checks() and b.msgSize() == a.msgSize()
    llr = gallager_llr(oc.N, oc.K, 5, 2.0, seed=12)
    ra, rb = a.soDecode(llr), b.soDecode(llr)
    for k in ("msg", "syst", "parity", "state", "iterations"):
        assert np.array_equal(np.asarray(ra[k]), np.asarray(rb[k])), k
    with pytest.raises(Exception, match="encoder"):
        b.encode(np.zeros(oc.K, np.uint8))
    D = F.DvbS2.matrix(16200, 0.5)
    d1, d2 = F.Ldpc(D, opts), F.Ldpc.fromStructure(D, 16200 - D.rows(), opts)
    msg = np.random.default_rng(2).integers(0, 2, size=2 * d1.msgSize(), dtype=np.uint8)
    assert np.array_equal(d1.encode(msg), d2.encode(msg)) and d1.save() == d2.save()


@pytest.mark.parametrize("alg,fmt,ofmt", [(F.Approximate, F.MsgFormat.F32, O.MSG_F32),
                                          (F.Linear, F.MsgFormat.F32, O.MSG_F32),
                                          (F.Approximate, F.MsgFormat.F16, O.MSG_F16)])
def test_every_dvbs2_code_bit_exact(alg, fmt, ofmt):
    """All 21 DVB-S2 codes (11 normal, 10 short rates), 3 frames each near the code's
    waterfall, 30 iterations: every soDecode output and iteration count equal to the oracle's
    (min-sum and Linear FP32; min-sum with FP16 message storage)."""
    from fecl_amd.ldpc import _DVBS2_RATES
    for n, rates in sorted(_DVBS2_RATES.items()):
        for rate in rates:
            oc, H = dvbs2(n, rate)
            ebn0 = _WATERFALL[rate] + (0.4 if n == 16200 else 0.0)
            llr = dvbs2_llr(3, ebn0, seed=31 + rate[0] * 7 + rate[1], rate=rate, n=n)
            llr = llr.astype(np.float32).astype(np.float64)
            print(f"code n={n} rate={rate[0]}/{rate[1]} Eb/N0={ebn0}", flush=True)
            ref = oracle_run(oc, llr, alg, 30, fmt=ofmt, threads=16)
            out = gpu_decoder(H, alg, 30, fmt=fmt).soDecode(llr)
            assert_bit_exact(out, ref, 3, oc.K, oc.N, oc.E)


def test_large_batch_across_the_pass_boundary():
    """8193 frames in one device call run as two passes (8192 + 1 frames): every frame's bits
    and iteration count equal those of decoding the frames in separate calls (frames share no
    state), and a sample equals the oracle."""
    import torch
    oc, H = dvbs2()
    base = dvbs2_llr(64, 1.5, seed=77).astype(np.float32)
    llr = np.concatenate([base] * 128 + [base[:1]])       # 8193 frames
    assert llr.shape[0] == 8193
    dec = gpu_decoder(H, F.Approximate, 50, fmt=F.MsgFormat.F16)
    x = torch.from_numpy(llr).cuda()
    bits, it = dec.decode(x, iterations=True)
    b1, i1 = dec.decode(x[:4096], iterations=True)
    b2, i2 = dec.decode(x[4096:], iterations=True)
    torch.cuda.synchronize()
    assert torch.equal(bits, torch.cat([b1, b2])) and torch.equal(it, torch.cat([i1, i2]))
    K = oc.K
    hb = bits.view(8193, K).cpu().numpy()
    assert all(np.array_equal(hb[f], hb[f % 64]) for f in (64, 4095, 4096, 8191, 8192))
    ref = oracle_run(oc, llr[:8].astype(np.float64), F.Approximate, 50, fmt=O.MSG_F16, want=("hard", "iters"))
    assert np.array_equal(hb[:8], ref["hard"]) and np.array_equal(it[:8].cpu().numpy(), ref["iters"])


def test_one_pass_beyond_the_grid_y_limit():
    """70000 frames of the Gallager code in one pass: the hard-bit unpack, whose launch grid
    counts frames in y, is split at 65535 rows (the device reports 65536 as its grid limit in
    y).  Bits and iteration counts equal the same batch in the default 1024-frame passes, and
    the oracle on frames either side of the split."""
    import torch
    oc, H = gallager()
    frames, distinct = 70000, 4096
    base = gallager_llr(oc.N, oc.K, distinct, 2.5, 31).astype(np.float32)
    idx = np.arange(frames) % distinct
    x = torch.from_numpy(base).cuda()[torch.from_numpy(idx).cuda()].contiguous().view(-1)
    one = gpu_decoder(H, F.Approximate, 20)
    one.setPassFrames(frames)
    passes = gpu_decoder(H, F.Approximate, 20)
    b1, i1 = one.decode(x, iterations=True)
    b2, i2 = passes.decode(x, iterations=True)
    torch.cuda.synchronize()
    assert torch.equal(b1, b2) and torch.equal(i1, i2)
    sample = np.array([0, 1, 65534, 65535, 65536, 69999])
    ref = oracle_run(oc, base[idx[sample]].astype(np.float64), F.Approximate, 20, want=("hard", "iters"))
    assert np.array_equal(b1.view(frames, oc.K).cpu().numpy()[sample], ref["hard"])
    assert np.array_equal(i1.cpu().numpy()[sample], ref["iters"])
