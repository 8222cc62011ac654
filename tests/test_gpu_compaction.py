"""GPU: compaction of live frames into fewer tiles (run_batch in fecl_amd/csrc/ldpc_capi.cpp).

Frames are independent (FeCl decodes each block on its own, BpDecoderImpl.cpp:38-130), so
moving a frame's messages, channel values and state to another slot must change nothing: every
output stays bit-identical to the oracle and to the decode without compaction.  The inputs mix
frames that converge at different iterations with frames that never do, shuffled over the
tiles, so that the live frames shrink to a half, a quarter, ... of the slots while every tile
still holds one: the decoder then compacts more than once, and the later regions reuse the
first region's buffers.
"""
import numpy as np
import pytest

import fecl_amd as F
from oracle import oracle as O
from tests.golden.vectors import dvbs2_llr, gallager_llr
from tests.test_gpu_decoder import assert_bit_exact, gallager, gpu_decoder, oracle_run

pytestmark = pytest.mark.gpu

ITERS = 30


def staggered_llr(oc, frames, seed):
    """All-zero codewords at Eb/N0 from 6 dB (converge in 2-3 updates) down to -3 dB (never),
    shuffled over the batch."""
    rng = np.random.default_rng(seed)
    points = [6.0, 4.0, 3.0, 2.5, 2.0, 1.5, -3.0]
    share = [0.25, 0.2, 0.15, 0.1, 0.1, 0.1, 0.1]
    counts = np.floor(np.array(share) * frames).astype(int)
    counts[0] += frames - counts.sum()
    parts = [gallager_llr(oc.N, oc.K, int(n), e, seed + 7 * i) for i, (n, e) in enumerate(zip(counts, points))]
    llr = np.concatenate(parts)
    return llr[rng.permutation(frames)].astype(np.float32).astype(np.float64)


def compactions(iters, tw, iterations):
    """Compactions the host's rule performs for these per-frame update counts.  A frame with c
    check updates finishes in step c-1, so after step s the host knows the frames live after
    step s-2 (c > s-1).  It compacts when they fit in three quarters of the current region's
    tiles, the freed tile-steps outweigh 1.5 frame-iterations per live frame, and at least 4
    steps are left."""
    tiles = (len(iters) + tw - 1) // tw
    n = 0
    for step in range(2, iterations):
        live = int(np.sum(iters > step - 1))
        if live == 0:
            break
        need = (live + tw - 1) // tw
        remaining = iterations - 1 - step
        if need <= 3 * tiles // 4 and 2 * (tiles - need) * remaining * tw > 3 * live and step + 4 < iterations:
            tiles, n = need, n + 1
    return n


@pytest.mark.parametrize("fmt,ofmt,words", [(F.MsgFormat.F32, O.MSG_F32, 1), (F.MsgFormat.F32, O.MSG_F32, 4),
                                            (F.MsgFormat.F16, O.MSG_F16, 1), (F.MsgFormat.BF16, O.MSG_BF16, 2)])
def test_compacted_decode_matches_the_oracle(fmt, ofmt, words):
    """soDecode (APP, extrinsic, state, iteration counts) and decode bits bit-identical to the
    oracle while the live frames move through several regions."""
    oc, H = gallager()
    frames = 1024
    llr = staggered_llr(oc, frames, 40 + words)
    ref = oracle_run(oc, llr, F.Approximate, ITERS, fmt=ofmt, threads=16)
    dec = gpu_decoder(H, F.Approximate, ITERS, fmt=fmt, words=words)
    tw = dec.layout()[1]
    assert compactions(ref["iters"], tw, ITERS) >= 2
    print(f"fmt={fmt} words={words} tw={tw} compactions={compactions(ref['iters'], tw, ITERS)}", flush=True)
    assert_bit_exact(dec.soDecode(llr), ref, frames, oc.K, oc.N, oc.E)
    bits = dec.decode(llr)
    assert np.array_equal(np.asarray(bits).reshape(frames, oc.K), ref["hard"])


@pytest.mark.parametrize("alg", [F.Linear, F.Exact])
def test_compaction_changes_nothing(alg):
    """Linear and Exact with scaling factors: every output with compaction equals the output
    without it (Exact: the oracle is not bit-exact, so this is the check)."""
    oc, H = gallager()
    frames = 768
    llr = staggered_llr(oc, frames, 50)
    outs = []
    for on in (True, False):
        dec = gpu_decoder(H, alg, ITERS, sf=0.75, words=1)
        dec.setCompaction(on)
        o = dec.soDecode(llr)
        o["bits"] = np.asarray(dec.decode(llr))
        outs.append(o)
    assert compactions(outs[0]["iterations"], 64, ITERS) >= 2
    for name in ("msg", "syst", "parity", "state", "iterations", "bits"):
        a, b = np.asarray(outs[0][name]), np.asarray(outs[1][name])
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name


def test_compaction_with_state_and_syst_inputs():
    """A-priori state and systematic inputs (the state-seeded first variable pass) on a batch
    that compacts: bit-identical to the oracle."""
    oc, H = gallager()
    frames = 512
    llr = staggered_llr(oc, frames, 60)
    rng = np.random.default_rng(61)
    syst = rng.normal(0.0, 0.5, (frames, oc.K)).astype(np.float32).astype(np.float64)
    state = rng.normal(0.0, 0.5, (frames, oc.E)).astype(np.float32).astype(np.float64)
    ref = oracle_run(oc, llr, F.Approximate, ITERS, syst_in=syst, state_in=state, threads=16)
    dec = gpu_decoder(H, F.Approximate, ITERS, words=1)
    assert compactions(ref["iters"], 64, ITERS) >= 1
    assert_bit_exact(dec.soDecode(llr, syst=syst, state=state), ref, frames, oc.K, oc.N, oc.E)


def test_device_path_dvbs2_compaction_on_and_off():
    """DVB-S2 r1/2 near the waterfall through the device path (FP16 messages): bits and
    iteration counts with compaction equal those without it and the oracle's on a sample."""
    import torch
    from tests.test_gpu_decoder import dvbs2
    oc, H = dvbs2()
    frames = 512
    llr = np.concatenate([dvbs2_llr(256, 1.7, seed=71), dvbs2_llr(256, 0.5, seed=72)])
    llr = llr[np.random.default_rng(73).permutation(frames)].astype(np.float32)
    x = torch.from_numpy(llr).cuda()
    res = []
    for on in (True, False):
        dec = gpu_decoder(H, F.Approximate, 50, fmt=F.MsgFormat.F16, words=1)
        dec.setCompaction(on)
        bits, it = dec.decode(x, iterations=True)
        torch.cuda.synchronize()
        res.append((bits.cpu().numpy(), it.cpu().numpy()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert compactions(res[0][1], 128, 50) >= 1
    sample = np.arange(0, frames, 37)
    ref = oracle_run(oc, llr[sample].astype(np.float64), F.Approximate, 50, fmt=O.MSG_F16, threads=16)
    assert np.array_equal(res[0][1][sample], ref["iters"])
    assert np.array_equal(res[0][0].reshape(frames, oc.K)[sample], ref["hard"])


def graph_compactions(iters, tw, iterations):
    """In-place compactions a captured decode makes: after every fourth step (while 4 steps
    remain) the device counts the frames still live (c > s+1 updates) and the tiles holding
    them, and moves them into the first slots (in place, one move) when that frees a tile and
    the freed tile-steps outweigh 1.5 frame-iterations per live frame."""
    slots = list(range(len(iters)))
    n = 0
    for step in range(3, iterations, 4):
        if step + 4 >= iterations:
            break
        live = [f for f in slots if f >= 0 and iters[f] > step + 1]
        tiles_live = len({j // tw for j, f in enumerate(slots) if f >= 0 and iters[f] > step + 1})
        freed = tiles_live - (len(live) + tw - 1) // tw
        if freed > 0 and 2 * freed * (iterations - 1 - step) * tw > 3 * len(live):
            slots, n = live + [-1] * (len(slots) - len(live)), n + 1
    return n


def test_graph_compacts_on_the_device():
    """One hipGraph over two batches that converge at staggered iterations (FP32 and FP16
    messages, 64 and 128 frames per tile): the device decides at every fourth step and compacts
    in place.  Replayed with new inputs copied in eagerly between launches, every replay's bits
    and iteration counts equal the oracle's and the eager decode's."""
    import torch
    oc, H = gallager()
    specs = [(F.MsgFormat.F32, O.MSG_F32, 1024), (F.MsgFormat.F16, O.MSG_F16, 768)]
    decs = [gpu_decoder(H, F.Approximate, ITERS, fmt=fmt, words=1) for fmt, _, _ in specs]
    eager = [gpu_decoder(H, F.Approximate, ITERS, fmt=fmt, words=1) for fmt, _, _ in specs]
    tensors = [torch.zeros(frames * oc.N, dtype=torch.float32, device="cuda") for _, _, frames in specs]
    g = F.BatchGraph(decs, tensors)
    for rep in range(3):
        refs = []
        for i, (fmt, ofmt, frames) in enumerate(specs):
            llr = staggered_llr(oc, frames, 80 + 10 * rep + i)
            tensors[i].copy_(torch.from_numpy(llr.astype(np.float32).ravel()))
            refs.append(oracle_run(oc, llr, F.Approximate, ITERS, fmt=ofmt, want=("hard", "iters"), threads=16))
            n = graph_compactions(refs[-1]["iters"], decs[i].layout()[1], ITERS)
            print(f"replay {rep} batch {i}: {frames} frames, in-place compactions {n}", flush=True)
            assert n >= 1
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for i, (fmt, ofmt, frames) in enumerate(specs):
            got = g.msg[i].cpu().numpy().reshape(frames, oc.K)
            assert np.array_equal(got, refs[i]["hard"]), (rep, i)
            assert np.array_equal(g.iterations[i].cpu().numpy(), refs[i]["iters"]), (rep, i)
            bits, it = eager[i].decode(tensors[i], iterations=True)
            torch.cuda.synchronize()
            assert np.array_equal(bits.cpu().numpy().reshape(frames, oc.K), got), (rep, i)
            assert np.array_equal(it.cpu().numpy(), refs[i]["iters"]), (rep, i)


@pytest.mark.parametrize("fmt,ofmt,words,frames", [(F.MsgFormat.F16, O.MSG_F16, 4, 4096),
                                                   (F.MsgFormat.BF16, O.MSG_BF16, 2, 2048),
                                                   (F.MsgFormat.F32, O.MSG_F32, 2, 1536)])
def test_graph_in_place_compaction_wide_tiles(fmt, ofmt, words, frames):
    """In-place compaction with 512-, 256- and 128-frame tiles (2 and 4 words per lane), over
    several 1024-frame passes: a workgroup of the move then owns fewer edges or variables, and a
    destination tile's sources more often lie in that same tile.  Two replays: bits and
    iteration counts equal the oracle's and the eager decode's."""
    import torch
    oc, H = gallager()
    dec = gpu_decoder(H, F.Approximate, ITERS, fmt=fmt, words=words)
    eager = gpu_decoder(H, F.Approximate, ITERS, fmt=fmt, words=words)
    x = torch.zeros(frames * oc.N, dtype=torch.float32, device="cuda")
    g = F.BatchGraph([dec], [x])
    tw = dec.layout()[1]
    pass_frames = 1024   # the default pass; each pass is a region of its own
    for rep in range(2):
        llr = staggered_llr(oc, frames, 120 + rep)
        ref = oracle_run(oc, llr, F.Approximate, ITERS, fmt=ofmt, want=("hard", "iters"), threads=16)
        n = sum(graph_compactions(ref["iters"][p:p + pass_frames], tw, ITERS) for p in range(0, frames, pass_frames))
        print(f"fmt={fmt} words={words} tile={tw} replay {rep}: in-place compactions {n}", flush=True)
        assert n >= 1
        x.copy_(torch.from_numpy(llr.astype(np.float32).ravel()))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        got = g.msg[0].cpu().numpy().reshape(frames, oc.K)
        assert np.array_equal(got, ref["hard"]), rep
        assert np.array_equal(g.iterations[0].cpu().numpy(), ref["iters"]), rep
        bits, it = eager.decode(x, iterations=True)
        torch.cuda.synchronize()
        assert np.array_equal(bits.cpu().numpy().reshape(frames, oc.K), got), rep
        assert np.array_equal(it.cpu().numpy(), ref["iters"]), rep


@pytest.mark.parametrize("alg,fmt,words", [(F.Linear, F.MsgFormat.F32, 1), (F.Exact, F.MsgFormat.F32, 1),
                                           (F.Linear, F.MsgFormat.F16, 2), (F.Exact, F.MsgFormat.BF16, 2)])
def test_graph_compaction_linear_exact_equal_eager(alg, fmt, words):
    """Linear and Exact with scaling factor 0.75 in one hipGraph (in-place compaction on the
    device) against eager decodes (compaction between regions on the host's counts), FP32 with
    64-frame tiles and FP16/BF16 with 256-frame tiles: bits and iteration counts equal on two
    replays."""
    import torch
    oc, H = gallager()
    frames = 1024
    dec = gpu_decoder(H, alg, ITERS, sf=0.75, fmt=fmt, words=words)
    eager = gpu_decoder(H, alg, ITERS, sf=0.75, fmt=fmt, words=words)
    x = torch.zeros(frames * oc.N, dtype=torch.float32, device="cuda")
    g = F.BatchGraph([dec], [x])
    for rep in range(2):
        llr = staggered_llr(oc, frames, 90 + rep)
        x.copy_(torch.from_numpy(llr.astype(np.float32).ravel()))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        bits, it = eager.decode(x, iterations=True)
        torch.cuda.synchronize()
        n = graph_compactions(it.cpu().numpy(), dec.layout()[1], ITERS)
        print(f"alg={alg} fmt={fmt} words={words} replay {rep}: in-place compactions {n}", flush=True)
        assert n >= 1
        assert np.array_equal(g.msg[0].cpu().numpy(), bits.cpu().numpy()), rep
        assert np.array_equal(g.iterations[0].cpu().numpy(), it.cpu().numpy()), rep


def test_graph_with_fused_syndrome():
    """A hipGraph over a decoder whose syndrome is folded into the check kernel (that form does
    not compact): on two replays, bits and iteration counts equal the oracle's and the eager
    decode's."""
    import torch
    oc, H = gallager()
    frames = 512
    dec = gpu_decoder(H, F.Approximate, ITERS, words=1, fused=True)
    eager = gpu_decoder(H, F.Approximate, ITERS, words=1, fused=True)
    x = torch.zeros(frames * oc.N, dtype=torch.float32, device="cuda")
    g = F.BatchGraph([dec], [x])
    for rep in range(2):
        llr = staggered_llr(oc, frames, 140 + rep)
        ref = oracle_run(oc, llr, F.Approximate, ITERS, want=("hard", "iters"), threads=16)
        x.copy_(torch.from_numpy(llr.astype(np.float32).ravel()))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        got = g.msg[0].cpu().numpy().reshape(frames, oc.K)
        assert np.array_equal(got, ref["hard"]), rep
        assert np.array_equal(g.iterations[0].cpu().numpy(), ref["iters"]), rep
        bits, it = eager.decode(x, iterations=True)
        torch.cuda.synchronize()
        assert np.array_equal(bits.cpu().numpy().reshape(frames, oc.K), got), rep
        assert np.array_equal(it.cpu().numpy(), ref["iters"]), rep


def test_short_frame_compaction_matches_the_oracle():
    """DVB-S2 n=16200 r1/2, 512 frames mixing converging and non-converging Eb/N0 over 8 tiles
    of 64: soDecode outputs bit-identical to the oracle through the compactions."""
    from tests.test_gpu_decoder import dvbs2
    oc, H = dvbs2(16200, (1, 2))
    rng = np.random.default_rng(95)
    parts = [dvbs2_llr(n, e, seed=96 + i, n=16200) for i, (n, e) in enumerate(((200, 3.0), (150, 2.2), (100, 1.8), (62, -2.0)))]
    llr = np.concatenate(parts)
    llr = llr[rng.permutation(len(llr))].astype(np.float32).astype(np.float64)
    frames = llr.shape[0]
    ref = oracle_run(oc, llr, F.Approximate, ITERS, threads=16)
    n = compactions(ref["iters"], 64, ITERS)
    print(f"short frames: compactions {n}", flush=True)
    assert n >= 1
    dec = gpu_decoder(H, F.Approximate, ITERS, words=1)
    assert_bit_exact(dec.soDecode(llr), ref, frames, oc.K, oc.N, oc.E)
