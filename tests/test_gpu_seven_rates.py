"""The batching and serving forms on the seven DVB-S2 normal rates of the uneven workload
{1/4, 1/2, 3/5, 2/3, 3/4, 5/6, 9/10}, whose rows reach 4 to 30 edges and whose information
bits have degree 3 to 13 (so every check-kernel form the layouts use runs).  Per-frame Eb/N0
within +-0.3 dB of each rate's operating point, min-sum, FP16 messages, 50 iterations.

  * One hipGraph over all seven codes with uneven frame counts (8-64 per rate), replayed with
    new inputs: every branch equals the oracle (same FP16 storage rounding) on every replay.
  * The continuous-batching server fed Poisson-spaced requests of 1-48 frames over the seven
    codes, pools sized by a device-memory budget: every request equals the same frames decoded
    alone, and the oracle."""
import threading
import time

import numpy as np
import pytest

import fecl_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# rate: Eb/N0 (dB) of the operating point (first 0.1 dB step with min-sum FER <= 1e-2, + 0.3)
RATES = {(1, 4): 1.3, (1, 2): 1.9, (3, 5): 2.6, (2, 3): 2.5, (3, 4): 2.8, (5, 6): 3.4, (9, 10): 4.3}


def frames_at(rate, ebn0s, seed):
    """One frame per Eb/N0 value: seeded Bernoulli information bits, the oracle's encoder,
    BPSK over AWGN, FeCl-convention FP32 LLRs (as tests.golden.vectors.awgn_llr, per frame)."""
    code = O.dvbs2(64800, *rate)
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (len(ebn0s), code.K), dtype=np.uint8)
    cw = O.encode(code, info)
    sigma2 = 1.0 / (2.0 * code.K / code.N * 10.0 ** (np.asarray(ebn0s, np.float64) / 10.0))
    y = (1.0 - 2.0 * cw) + rng.normal(0.0, 1.0, cw.shape) * np.sqrt(sigma2)[:, None]
    return (-2.0 * y / sigma2[:, None]).astype(np.float32)


def decoders():
    return [F.Ldpc(F.DvbS2.matrix(64800, r), F.DecoderOptions().algorithm(F.Approximate).iterations(50)
                   .msgFormat(F.MsgFormat.F16)) for r in RATES]


def oracle_f16(rate, llr):
    return O.decode(O.dvbs2(64800, *rate), llr.astype(np.float64), O.APPROXIMATE, 50, prec=32, msg_fmt=O.MSG_F16,
                    want=("hard", "iters"), threads=16)


def test_seven_rate_graph_replays_match_the_oracle():
    import torch
    rng = np.random.default_rng(3000)
    rates = list(RATES)
    counts = [int(c) for c in rng.integers(8, 65, len(rates))]
    decs = decoders()
    bufs = [torch.zeros(c * 64800, dtype=torch.float32, device="cuda") for c in counts]
    g = F.BatchGraph(decs, bufs)
    for rep in range(2):
        refs = []
        for i, (rate, c) in enumerate(zip(rates, counts)):
            ebn0 = RATES[rate] + rng.uniform(-0.3, 0.3, c)
            llr = frames_at(rate, ebn0, seed=3100 + 100 * rep + 1000 * i)
            bufs[i].copy_(torch.from_numpy(llr.ravel()))
            refs.append(oracle_f16(rate, llr))
        torch.cuda.synchronize()
        print(f"seven-rate graph replay {rep}: frames per rate {counts}", flush=True)
        g.launch()
        torch.cuda.synchronize()
        for i, (rate, c) in enumerate(zip(rates, counts)):
            K = decs[i].msgSize()
            assert np.array_equal(g.msg[i].cpu().numpy().reshape(c, K), refs[i]["hard"]), (rep, rate)
            assert np.array_equal(g.iterations[i].cpu().numpy(), refs[i]["iters"]), (rep, rate)


def test_seven_rate_server_poisson_requests():
    rng = np.random.default_rng(3200)
    rates = list(RATES)
    decs = decoders()
    pools = {}
    for i, rate in enumerate(rates):
        ebn0 = RATES[rate] + rng.uniform(-0.3, 0.3, 96)
        pools[i] = frames_at(rate, ebn0, seed=3300 + 1000 * i)
    reqs = []
    for _ in range(60):
        i = int(rng.integers(0, len(rates)))
        f = int(rng.integers(1, 49))
        o = int(rng.integers(0, 96 - f))
        reqs.append((i, pools[i][o:o + f]))
    # a budget of 2.5 GB over seven pools: a few tiles each, so requests queue and join mid-flight
    sv = F.Server(decs, memory_bytes=2500 << 20)
    for i, d in enumerate(decs):
        assert sv.slots(i) >= d.layout()[1]
    tickets = [None] * len(reqs)
    gaps = rng.exponential(1.0 / 4000.0, len(reqs))   # ~4000 requests/s offered

    def submitter():
        for k, (i, x) in enumerate(reqs):
            tickets[k] = sv.submit(i, x)
            time.sleep(float(gaps[k]))

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    got = [sv.wait(t, timeout=120.0) for t in tickets]
    st = sv.stats()
    sv.close()
    assert all(x is not None for x in got), "server did not finish"
    assert st["frames_done"] == sum(x.shape[0] for _, x in reqs)
    for (i, x), (msg, it, lat) in zip(reqs, got):
        alone, ait = decs[i].decode(x.astype(np.float64), iterations=True)
        assert np.array_equal(msg, alone) and np.array_equal(it, ait), rates[i]
        assert lat > 0
    for i, rate in enumerate(rates):
        ks = [k for k, (j, _) in enumerate(reqs) if j == i]
        if not ks:
            continue
        x = np.concatenate([reqs[k][1] for k in ks])
        ref = oracle_f16(rate, x)
        msg = np.concatenate([got[k][0] for k in ks]).reshape(ref["hard"].shape)
        it = np.concatenate([got[k][1] for k in ks])
        assert np.array_equal(msg, ref["hard"]) and np.array_equal(it, ref["iters"]), rate
