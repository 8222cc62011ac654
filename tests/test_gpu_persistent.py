"""The persistent multi-code decode (one kernel launch walking a work queue of check, variable,
syndrome and compaction items for every code; fecl_amd/csrc/ldpc_persist.hip) against the
per-step launches: captured in a graph, it gives the same hard decisions and iteration counts as
the per-step graph of the same batch, as eager decodes of each code, and as the oracle, on every
replay, with compaction on and off."""
import numpy as np
import pytest

import fecl_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RATES = {(1, 4): 1.3, (1, 2): 1.9, (3, 5): 2.6, (2, 3): 2.5, (3, 4): 2.8, (5, 6): 3.4, (9, 10): 4.3}


def frames_at(rate, ebn0s, seed, n=64800):
    code = O.dvbs2(n, *rate)
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (len(ebn0s), code.K), dtype=np.uint8)
    cw = O.encode(code, info)
    sigma2 = 1.0 / (2.0 * code.K / code.N * 10.0 ** (np.asarray(ebn0s, np.float64) / 10.0))
    y = (1.0 - 2.0 * cw) + rng.normal(0.0, 1.0, cw.shape) * np.sqrt(sigma2)[:, None]
    return (-2.0 * y / sigma2[:, None]).astype(np.float32)


def decoders(fmt, rates, persistent=True, compaction=True, iters=50):
    out = []
    for r in rates:
        d = F.Ldpc(F.DvbS2.matrix(64800, r), F.DecoderOptions().algorithm(F.Approximate).iterations(iters)
                   .msgFormat(fmt))
        d.setGraphPersistent(persistent)
        d.setCompaction(compaction)
        out.append(d)
    return out


def timed(fn, reps=3):
    import torch
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = 1e30
    for _ in range(reps):
        torch.cuda.synchronize()
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


@pytest.mark.parametrize("compaction", [True, False])
def test_persistent_seven_rates_equals_per_step_and_eager(compaction):
    import torch
    rng = np.random.default_rng(4000)
    rates = list(RATES)
    counts = [int(c) for c in rng.integers(100, 700, len(rates))]
    fmt = F.MsgFormat.F16
    pd = decoders(fmt, rates, True, compaction)
    sd = decoders(fmt, rates, False, compaction)
    bufs = [torch.zeros(c * 64800, dtype=torch.float32, device="cuda") for c in counts]
    gp = F.BatchGraph(pd, bufs)
    gs = F.BatchGraph(sd, bufs)
    assert gp.persistent and not gs.persistent
    info_bits = sum(c * d.msgSize() for c, d in zip(counts, pd))
    for rep in range(2):
        for i, (rate, c) in enumerate(zip(rates, counts)):
            ebn0 = RATES[rate] + rng.uniform(-0.3, 0.3, c)
            bufs[i].copy_(torch.from_numpy(frames_at(rate, ebn0, seed=4100 + 100 * rep + 1000 * i).ravel()))
        torch.cuda.synchronize()
        print(f"persistent graph, replay {rep}, compaction {compaction}, frames per rate {counts}", flush=True)
        gp.launch()
        torch.cuda.synchronize()
        gp.check()
        gs.launch()
        torch.cuda.synchronize()
        for i, (rate, c) in enumerate(zip(rates, counts)):
            assert torch.equal(gp.msg[i], gs.msg[i]), (rep, rate)
            assert torch.equal(gp.iterations[i], gs.iterations[i]), (rep, rate)
            bits, it = pd[i].decode(bufs[i], iterations=True)
            torch.cuda.synchronize()
            assert torch.equal(gp.msg[i], bits) and torch.equal(gp.iterations[i], it), (rep, rate)
    # the oracle on a sample of every code (last replay)
    for i, rate in enumerate(rates):
        x = bufs[i].view(counts[i], 64800)[:6].cpu().numpy().astype(np.float64)
        ref = O.decode(O.dvbs2(64800, *rate), x, O.APPROXIMATE, 50, prec=32, msg_fmt=O.MSG_F16,
                       want=("hard", "iters"), threads=16)
        assert np.array_equal(gp.msg[i].view(counts[i], -1)[:6].cpu().numpy(), ref["hard"]), rate
        assert np.array_equal(gp.iterations[i][:6].cpu().numpy(), ref["iters"]), rate
    tp = timed(gp.launch)
    ts = timed(gs.launch)

    def eager():
        for i in range(len(rates)):
            pd[i].decode(bufs[i])
    te = timed(eager)
    print(f"seven-rate batch ({sum(counts)} frames, compaction {compaction}): persistent {tp:.2f} ms "
          f"({info_bits / tp / 1e6:.3f} Gb/s), per-step graph {ts:.2f} ms ({info_bits / ts / 1e6:.3f} Gb/s), "
          f"eager {te:.2f} ms ({info_bits / te / 1e6:.3f} Gb/s)", flush=True)


@pytest.mark.parametrize("fmt,ofmt", [(F.MsgFormat.F32, O.MSG_F32), (F.MsgFormat.BF16, O.MSG_BF16)])
def test_persistent_single_code_matches_the_oracle(fmt, ofmt):
    """One code (r1/2) in the persistent form, 600 frames around the waterfall: every frame's
    bits and iteration count equal the oracle's (same storage rounding)."""
    import torch
    rate = (1, 2)
    d = decoders(fmt, [rate])
    llr = frames_at(rate, np.linspace(1.5, 2.1, 600), seed=4200)
    x = torch.from_numpy(llr.ravel()).cuda()
    g = F.BatchGraph(d, [x])
    assert g.persistent
    g.launch()
    torch.cuda.synchronize()
    g.check()
    ref = O.decode(O.dvbs2(64800, *rate), llr.astype(np.float64), O.APPROXIMATE, 50, prec=32, msg_fmt=ofmt,
                   want=("hard", "iters"), threads=16)
    assert np.array_equal(g.msg[0].view(600, -1).cpu().numpy(), ref["hard"])
    assert np.array_equal(g.iterations[0].cpu().numpy(), ref["iters"])
