"""Continuous-batching server: requests of mixed codes and sizes, from host or device memory,
submitted while others run, each decoded exactly as it is alone (frames share no state), with
one or several steps per round; plus admission by memory and the round-steps bounds."""
import threading
import time

import numpy as np
import pytest

import fecl_amd as F
from oracle import oracle as O
from tests.golden.vectors import dvbs2_llr

pytestmark = pytest.mark.gpu

CODES = [(64800, (1, 2), 1.6), (64800, (3, 4), 3.3), (16200, (1, 2), 2.2)]


def decoders(fmt=F.MsgFormat.F16, iters=50):
    out = []
    for n, rate, _ in CODES:
        H = F.DvbS2.matrix(n, rate)
        out.append(F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(iters).msgFormat(fmt)))
    return out


@pytest.mark.parametrize("inputs,steps", [("pageable", 1), ("pinned", 4), ("device", 4), ("mixed", 1),
                                          ("mixed", 4), ("mixed", 7)])
def test_server_results_same_alone_and_in_the_mix(inputs, steps):
    """inputs: where the requests' LLRs live: pageable host memory (staged by the server),
    page-locked host memory (copied to the device directly), device memory (read in place by
    the admission kernels), or a random mix of the three within each round.  steps: steps per
    round between the host's admissions and retirements."""
    import torch
    decs = decoders()
    rng = np.random.default_rng(5 + steps)
    pool = {i: dvbs2_llr(200, CODES[i][2], seed=50 + i, rate=CODES[i][1], n=CODES[i][0]).astype(np.float32)
            for i in range(len(CODES))}
    pinned = {i: torch.from_numpy(pool[i]).pin_memory() for i in pool}
    dev = {i: torch.from_numpy(pool[i]).cuda() for i in pool}
    torch.cuda.synchronize()
    reqs = []
    for _ in range(40):
        i = int(rng.integers(0, len(CODES)))
        f = int(rng.integers(1, 40))
        o = int(rng.integers(0, 200 - f))
        kind = inputs if inputs != "mixed" else ("pageable", "pinned", "device")[int(rng.integers(0, 3))]
        src = {"pageable": pool[i][o:o + f], "pinned": pinned[i].numpy()[o:o + f], "device": dev[i][o:o + f]}[kind]
        reqs.append((i, pool[i][o:o + f], src))
    # small pools so that requests wait for slots and join mid-flight
    tw = [d.layout()[1] for d in decs]
    sv = F.Server(decs, slots=[tw[0], tw[1], 2 * tw[2]])
    sv.setRoundSteps(steps)
    results = [None] * len(reqs)

    def submitter():
        for k, (i, _, src) in enumerate(reqs):
            results[k] = sv.submit(i, src)
            time.sleep(0.002)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    got = [sv.wait(t, timeout=120.0) for t in results]
    st = sv.stats()
    sv.close()
    assert all(g is not None for g in got), "server did not finish"
    assert st["frames_done"] == sum(x.shape[0] for _, x, _ in reqs)
    for (i, x, _), (msg, it, lat) in zip(reqs, got):
        alone, ait = decs[i].decode(x.astype(np.float64), iterations=True)
        assert np.array_equal(msg, alone) and np.array_equal(it, ait)
        assert lat > 0
    # and against the oracle for one request per code
    for i in range(len(CODES)):
        k = next(k for k, (j, _, _) in enumerate(reqs) if j == i)
        x = reqs[k][1]
        n, rate, _ = CODES[i]
        ref = O.decode(O.dvbs2(n, *rate), x.astype(np.float64), O.APPROXIMATE, 50, prec=32, msg_fmt=O.MSG_F16,
                       want=("hard", "iters"))
        assert np.array_equal(got[k][0].reshape(ref["hard"].shape), ref["hard"])
        assert np.array_equal(got[k][1], ref["iters"])


@pytest.mark.parametrize("fmt,words,tile", [(F.MsgFormat.F16, 1, 128), (F.MsgFormat.F32, 2, 128),
                                            (F.MsgFormat.F32, 4, 256)])
def test_server_pinned_layouts_over_several_tiles(fmt, words, tile):
    """Pools in a pinned lane layout (128- or 256-frame tiles), four tiles each, fed small
    requests from device memory over time: rounds admit a few frames into partly filled tiles
    (the fullest tile with room for the round first).  Every request equals its frames decoded
    alone."""
    import torch
    decs = decoders(fmt=fmt)
    for d in decs:
        d.setLayout(words)
    tw = [d.layout()[1] for d in decs]
    assert all(t == tile for t in tw)
    rng = np.random.default_rng(17 + words)
    pool = {i: dvbs2_llr(300, CODES[i][2], seed=70 + i, rate=CODES[i][1], n=CODES[i][0]).astype(np.float32)
            for i in range(len(CODES))}
    dev = {i: torch.from_numpy(pool[i]).cuda() for i in pool}
    torch.cuda.synchronize()
    sv = F.Server(decs, slots=[4 * t for t in tw])
    reqs, tickets = [], []
    for _ in range(60):
        i = int(rng.integers(0, len(CODES)))
        f = int(rng.integers(1, 24))
        o = int(rng.integers(0, 300 - f))
        reqs.append((i, pool[i][o:o + f]))
        tickets.append(sv.submit(i, dev[i][o:o + f]))
        time.sleep(0.001)
    got = [sv.wait(t, timeout=120.0) for t in tickets]
    sv.close()
    assert all(g is not None for g in got), "server did not finish"
    for (i, x), (msg, it, _) in zip(reqs, got):
        alone, ait = decs[i].decode(x.astype(np.float64), iterations=True)
        assert np.array_equal(msg, alone) and np.array_equal(it, ait)


def test_server_linear_and_exact_pools():
    """A server whose pools decode with Linear (FP32 messages) and Exact (BF16 messages), scaling
    factor 0.75: requests from device memory over time, each equal to its frames decoded alone."""
    import torch
    specs = [(0, F.Linear, F.MsgFormat.F32), (1, F.Exact, F.MsgFormat.BF16)]
    decs = []
    for ci, alg, fmt in specs:
        n, rate, _ = CODES[ci]
        decs.append(F.Ldpc(F.DvbS2.matrix(n, rate), F.DecoderOptions().algorithm(alg).iterations(30)
                           .scalingFactor(0.75).msgFormat(fmt)))
    rng = np.random.default_rng(23)
    pool = {k: dvbs2_llr(120, CODES[ci][2] - 0.4, seed=90 + k, rate=CODES[ci][1], n=CODES[ci][0]).astype(np.float32)
            for k, (ci, _, _) in enumerate(specs)}
    dev = {k: torch.from_numpy(pool[k]).cuda() for k in pool}
    torch.cuda.synchronize()
    sv = F.Server(decs, slots=[2 * d.layout()[1] for d in decs])
    reqs, tickets = [], []
    for _ in range(30):
        k = int(rng.integers(0, len(specs)))
        f = int(rng.integers(1, 30))
        o = int(rng.integers(0, 120 - f))
        reqs.append((k, pool[k][o:o + f]))
        tickets.append(sv.submit(k, dev[k][o:o + f]))
        time.sleep(0.001)
    got = [sv.wait(t, timeout=120.0) for t in tickets]
    sv.close()
    assert all(g is not None for g in got), "server did not finish"
    for (k, x), (msg, it, _) in zip(reqs, got):
        alone, ait = decs[k].decode(x.astype(np.float64), iterations=True)
        assert np.array_equal(msg, alone) and np.array_equal(it, ait)


def test_server_admission_into_a_finished_lane_group():
    """A tile's row loads skip the 4-lane groups that hold no live frame, so an admission must
    mark its slot's group live.  One 64-frame FP32 tile (one frame per lane, groups of 4
    slots): request A (4 frames at 4 dB) finishes within a few steps while request B (4 frames
    at -1 dB, which never converge, 1000 iterations) keeps decoding; request C is then admitted
    into free slots of groups that hold no live frame.  C finishes while B still decodes, and
    every request equals its frames decoded alone."""
    d = F.Ldpc(F.DvbS2.matrix(64800, (1, 2)), F.DecoderOptions().algorithm(F.Approximate).iterations(1000))
    d.setLayout(1)
    assert d.layout()[1] == 64
    a = dvbs2_llr(4, 4.0, seed=201).astype(np.float32)
    b = dvbs2_llr(4, -1.0, seed=202).astype(np.float32)
    c = dvbs2_llr(4, 1.8, seed=203).astype(np.float32)
    sv = F.Server([d], slots=[64])
    ta, tb = sv.submit(0, a), sv.submit(0, b)
    ga = sv.wait(ta, timeout=60.0)
    tc = sv.submit(0, c)
    gc = sv.wait(tc, timeout=60.0)
    done_with_c = sv.stats()["frames_done"]
    gb = sv.wait(tb, timeout=120.0)
    sv.close()
    assert ga is not None and gb is not None and gc is not None, "server did not finish"
    assert done_with_c == 8, "B finished before C: the case this test sets up did not arise"
    for x, (msg, it, _) in ((a, ga), (b, gb), (c, gc)):
        alone, ait = d.decode(x.astype(np.float64), iterations=True)
        assert np.array_equal(msg, alone) and np.array_equal(it, ait)
    assert int(gb[1].min()) == 1000 and int(ga[1].max()) < 10


def test_server_admission_by_memory_budget():
    decs = decoders()
    budget = 256 << 20
    sv = F.Server(decs, memory_bytes=budget)
    total = 0
    for i, d in enumerate(decs):
        slots = sv.slots(i)
        assert slots >= d.layout()[1] and slots % d.layout()[1] == 0
        total += slots
    sv.close()
    assert total > 0


def test_server_round_steps_bounds():
    decs = decoders()
    sv = F.Server(decs, slots=[d.layout()[1] for d in decs])
    for bad in (0, 65, -1):
        with pytest.raises(ValueError):
            sv.setRoundSteps(bad)
    sv.setRoundSteps(64)
    sv.setRoundSteps(1)
    sv.close()
