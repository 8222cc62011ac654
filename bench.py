"""Benchmark: DVB-S2 n=64800 rate-1/2 flooding BP decode on one MI355X.

Workload (one "step" = one decode of one batch already resident in HBM):
    DVB-S2 normal frame, rate 1/2 (N=64800, K=32400, E=226799); batch of 1024 frames;
    info bits i.i.d. Bernoulli(0.5) from a seed, encoded by the library's encoder, BPSK over AWGN
    at Eb/N0 in {0.6, 0.8, 1.0, 1.2, 1.4} dB (steps cycle through the five batches), FeCl-convention
    FP32 LLR = -2y/sigma^2; min-sum (DecoderAlgorithm::Approximate), 50 max iterations with syndrome
    early termination, FP32 messages.
value = decoded info bits / wall time over the timed steps (Gb/s).

Also reported: per-batch latency p50/p99, the achieved HBM rate of the dominant kernel against
the MI355X peak (algorithmic bytes per launch / mean launch time from HIP events on the launch
stream), and the CPU baseline (stock FeCl built from the reference sources, or the oracle's C
restatement when that build is absent) timed on a bounded sample of the same frames.
"""
import argparse
import gc
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
EBN0 = (0.6, 0.8, 1.0, 1.2, 1.4)
METRIC = "decoded info Gb/s at FER parity vs FP64 (DVB-S2 r1/2); p50/p99 latency"

# HBM bytes moved per algorithmic byte by the dominant kernels, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of this workload (all frames live), each counter scaled by what
# it reads for copy kernels of known size (FETCH x2.000, WRITE x1.000).  Keyed by (msg format,
# lane words); a layout without an entry reports traffic null.
PMC_TRAFFIC_RATIO = {("f32", 4): {"vn_kernel": 1.0031, "cn_kernel": 1.0096},
                     ("f32", 2): {"vn_kernel": 1.0109, "cn_kernel": 1.0093},
                     ("f32", 1): {"vn_kernel": 0.9977, "cn_kernel": 1.0091},
                     ("f16", 2): {"vn_kernel": 1.0132, "cn_kernel": 1.0192},
                     ("f16", 1): {"vn_kernel": 1.0268, "cn_kernel": 1.0185},
                     ("bf16", 2): {"vn_kernel": 1.0132, "cn_kernel": 1.0192},
                     ("bf16", 1): {"vn_kernel": 1.0268, "cn_kernel": 1.0185}}

# Extra Eb/N0 points (dB) of the precision study, across the waterfall of min-sum on DVB-S2 r1/2.
AGREEMENT_EBN0 = (0.8, 0.9, 1.0, 1.1, 1.2)   # Linear / Exact waterfall of r1/2
PRECISION_WATERFALL = (1.45, 1.5, 1.55, 1.6, 1.65, 1.7)


def make_batches(F, dec, H, frames, seed, device, points=EBN0):
    """Seeded info bits, library encoder, BPSK/AWGN on the GPU -> FP32 LLR batches, one per
    Eb/N0 point."""
    import numpy as np
    import torch
    K, N = dec.msgSize(), dec.paritySize()
    rate = K / N
    batches = []
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    for e in points:
        info = rng.integers(0, 2, (frames, K), dtype=np.uint8)
        cw = dec.encode(info).reshape(frames, N)
        sigma2 = 1.0 / (2.0 * rate * 10.0 ** (e / 10.0))
        s = 1.0 - 2.0 * torch.from_numpy(cw).to(device=device, dtype=torch.float64)
        y = s + torch.randn(s.shape, generator=gen, device=device, dtype=torch.float64) * (sigma2 ** 0.5)
        llr = (-2.0 * y / sigma2).to(torch.float32).contiguous()
        batches.append({"ebn0": e, "llr": llr, "info": torch.from_numpy(info).to(device)})
    return batches


# DVB-S2 normal rates of the uneven-batching and serving workloads, with each rate's ~1e-3 FER
# Eb/N0 (dB) for this decoder (min-sum, FP16 messages, 50 max iterations).  Measured on MI355X
# with `bench.py --config calibrate` (8192 frames per 0.1 dB step).  FeCl's output rule (the
# hard decision after one more check update than the syndrome test saw) leaves a floor of a few
# 1e-3 below each waterfall, so the point is the first step with FER <= 1e-2 plus 0.3 dB: the
# whole +-0.3 dB window then lies on the floor, at FER ~1e-3..5e-3.
RATE_POINTS = {(1, 4): 1.3, (1, 2): 1.9, (3, 5): 2.6, (2, 3): 2.5, (3, 4): 2.8, (5, 6): 3.4, (9, 10): 4.3}


def run_calibrate(args, device):
    """FER of min-sum with FP16 messages (50 max iterations) against Eb/N0 in 0.1 dB steps for the
    7 DVB-S2 normal rates; all-zero codeword (BP is codeword-symmetric on this channel)."""
    import numpy as np
    import torch
    import fecl_amd as F
    frames = 8192
    gen = torch.Generator(device=device)
    gen.manual_seed(args.seed)
    start = {(1, 4): 0.4, (1, 2): 0.8, (3, 5): 1.3, (2, 3): 1.6, (3, 4): 2.1, (5, 6): 2.8, (9, 10): 3.6}
    out = {}
    for rate, e0 in start.items():
        dec = F.Ldpc(F.DvbS2.matrix(64800, rate), F.DecoderOptions().algorithm(F.Approximate)
                     .iterations(args.iterations).msgFormat(F.MsgFormat.F16))
        K, N = dec.msgSize(), dec.paritySize()
        pts = []
        for i in range(16):
            e = round(e0 + 0.1 * i, 2)
            sigma2 = 1.0 / (2 * (K / N) * 10 ** (e / 10))
            y = 1.0 + torch.randn(frames * N, generator=gen, device=device) * sigma2 ** 0.5
            bits, it = dec.decode((-2.0 * y / sigma2).float(), iterations=True)
            fer = float((bits.view(frames, K).sum(1) > 0).float().mean())
            pts.append({"ebn0": e, "fer": fer, "mean_iterations": round(float(it.float().mean()), 2)})
            print(f"[calibrate] {rate} {pts[-1]}", file=sys.stderr, flush=True)
            if fer == 0.0 and len(pts) > 2 and pts[-2]["fer"] == 0.0:
                break
        out[f"{rate[0]}/{rate[1]}"] = pts
    return {"workload": f"DVB-S2 n=64800, all-zero codeword, {frames} frames per point, min-sum, f16 messages, "
                        f"{args.iterations} max iterations", "fer": out}


def codeword_pool(dec, frames, rng):
    """Encoded frames of i.i.d. Bernoulli(0.5) info bits: (info [frames][K], codeword [frames][N])."""
    import numpy as np
    K, N = dec.msgSize(), dec.paritySize()
    info = rng.integers(0, 2, (frames, K), dtype=np.uint8)
    return info, dec.encode(info).reshape(frames, N)


def awgn_llr_gpu(cw_dev, ebn0_dev, rate, gen):
    """FeCl-convention FP32 LLRs of BPSK codewords over AWGN, one Eb/N0 per frame (GPU)."""
    import torch
    sigma2 = (1.0 / (2.0 * rate * torch.pow(10.0, ebn0_dev / 10.0))).unsqueeze(1)
    s = 1.0 - 2.0 * cw_dev.to(torch.float32)
    y = s + torch.randn(s.shape, generator=gen, device=cw_dev.device) * torch.sqrt(sigma2)
    return (-2.0 * y / sigma2).to(torch.float32).contiguous()


def run_uneven(args):
    """Uneven batching: 7 normal rates, per-rate frame counts U[64,1024], per-frame Eb/N0 within
    +-0.3 dB of the rate's ~1e-3 FER point; min-sum, FP16 messages; one hipGraph launch."""
    import numpy as np
    import torch
    import fecl_amd as F
    device = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(args.seed)
    decs, llrs, infos, counts = [], [], [], []
    for rate, e0 in RATE_POINTS.items():
        n = int(rng.integers(64, 1025))
        dec = F.Ldpc(F.DvbS2.matrix(64800, rate),
                     F.DecoderOptions().algorithm(F.Approximate).iterations(args.iterations)
                     .msgFormat(F.MsgFormat.F16))
        dec.autotune(n, probe_iterations=3)
        info, cw = codeword_pool(dec, min(n, 256), rng)
        idx = np.arange(n) % info.shape[0]
        ebn0 = torch.from_numpy(e0 + rng.uniform(-0.3, 0.3, n)).to(device)
        llrs.append(awgn_llr_gpu(torch.from_numpy(cw[idx]).to(device), ebn0, dec.msgSize() / dec.paritySize(), gen))
        infos.append(torch.from_numpy(info[idx]).to(device))
        decs.append(dec)
        counts.append(n)
    g = F.BatchGraph(decs, llrs)
    for _ in range(args.warmup):
        g.launch()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        g.launch()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    total = time.perf_counter() - t0
    # the same batches decoded by one eager decode() call per rate, for comparison
    for _ in range(args.warmup):
        for d, x in zip(decs, llrs):
            d.decode(x)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eager_out = []
    for _ in range(args.steps):
        eager_out = [d.decode(x, iterations=True) for d, x in zip(decs, llrs)]
    torch.cuda.synchronize()
    eager = time.perf_counter() - t1
    if not all(torch.equal(m, e[0]) and torch.equal(it, e[1]) for m, it, e in zip(g.msg, g.iterations, eager_out)):
        raise RuntimeError("the graph's results differ from eager decodes")
    # the persistent form of the graph (one kernel for all codes; one lane layout, so the
    # decoders keep the automatic one), checked against the per-step graph's outputs
    pdecs = []
    for rate in RATE_POINTS:
        d = F.Ldpc(F.DvbS2.matrix(64800, rate),
                   F.DecoderOptions().algorithm(F.Approximate).iterations(args.iterations).msgFormat(F.MsgFormat.F16))
        d.setGraphPersistent(True)
        pdecs.append(d)
    gp = F.BatchGraph(pdecs, llrs)
    persistent = None
    if gp.persistent:
        for _ in range(args.warmup):
            gp.launch()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.steps):
            gp.launch()
        torch.cuda.synchronize()
        persistent = time.perf_counter() - t2
        gp.check()
        same = all(torch.equal(a, b) and torch.equal(c, e)
                   for a, b, c, e in zip(gp.msg, g.msg, gp.iterations, g.iterations))
        if not same:
            raise RuntimeError("persistent graph differs from the per-step graph")
    info_bits = sum(c * d.msgSize() for c, d in zip(counts, decs)) * args.steps
    per_rate = {}
    for (rate, _), d, c, m, it, inf in zip(RATE_POINTS.items(), decs, counts, g.msg, g.iterations, infos):
        errs = int(torch.any(m.view(c, d.msgSize()) != inf, dim=1).sum())
        per_rate[f"{rate[0]}/{rate[1]}"] = {"frames": c, "fer": round(errs / c, 5),
                                            "mean_iterations": round(float(it.float().mean()), 2),
                                            "lane_words": d.layout()[0]}
    return {"metric": "decoded info Gb/s, uneven multi-rate batch in one hipGraph launch",
            "value": round(info_bits / total / 1e9, 4), "unit": "Gb/s", "frames_per_launch": sum(counts),
            "ms_per_launch": round(total / args.steps * 1e3, 3),
            "eager_per_rate_calls_Gb_per_s": round(info_bits / eager / 1e9, 4),
            "persistent_graph_Gb_per_s": round(info_bits / persistent / 1e9, 4) if persistent else None,
            "identical_results": "per-step graph, eager calls and persistent kernel: same bits and iteration counts",
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                           "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
            "dtype": "f16", "per_rate": per_rate}


class ServingSetup:
    """Seven DVB-S2 normal-rate decoders (min-sum, FP16 messages) and per-rate pools of FP32 LLR
    frames at Eb/N0 within +-0.3 dB of each rate's ~1e-3 FER point, in device memory (as a
    receiver's demapper on the same GPU leaves them) and in page-locked host memory."""

    POOL = 768

    def __init__(self, F, args, device):
        import numpy as np
        import torch
        rng = np.random.default_rng(args.seed + 1)
        gen = torch.Generator(device=device)
        gen.manual_seed(args.seed + 1)
        self.rates = list(RATE_POINTS)
        self.decs, self.pools, self.dev, self.infos, self.pinned = [], [], [], [], []
        for rate in self.rates:
            dec = F.Ldpc(F.DvbS2.matrix(64800, rate),
                         F.DecoderOptions().algorithm(F.Approximate).iterations(args.iterations)
                         .msgFormat(F.MsgFormat.F16))
            info, cw = codeword_pool(dec, 256, rng)
            idx = np.arange(self.POOL) % 256
            ebn0 = torch.from_numpy(RATE_POINTS[rate] + rng.uniform(-0.3, 0.3, self.POOL)).to(device)
            llr = awgn_llr_gpu(torch.from_numpy(cw[idx]).to(device), ebn0, dec.msgSize() / dec.paritySize(), gen)
            self.dev.append(llr)
            host = llr.cpu().pin_memory()
            self.pinned.append(host)
            self.pools.append(host.numpy())
            self.infos.append(info[idx])
            self.decs.append(dec)
        torch.cuda.synchronize()

    def eager_frames_per_s(self, reps=3):
        """Each rate's pool decoded by one eager decode() call from device memory: frames per
        second over the seven calls (the serving mix has equal frames per rate).  The bits and
        iteration counts of that decode are kept: every served request must equal its slice."""
        import torch
        self.eager_bits, self.eager_iters = [], []
        for d, x in zip(self.decs, self.dev):
            bits, it = d.decode(x, iterations=True)
            self.eager_bits.append(bits.view(self.POOL, d.msgSize()).cpu().numpy())
            self.eager_iters.append(it.cpu().numpy())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for d, x in zip(self.decs, self.dev):
                d.decode(x)
        torch.cuda.synchronize()
        return reps * self.POOL * len(self.decs) / (time.perf_counter() - t0)


def run_serving(F, setup, args, duration_s, offered_fps, seed, device_inputs=True):
    """Serving: Poisson-arriving requests of U[1,256] frames, rate uniform over the 7 rates;
    continuous batching in the library's server.  Arrivals are generated by this thread at
    their scheduled times for duration_s; every result is then collected and checked."""
    import numpy as np
    rng = np.random.default_rng(seed)
    slots = [args.serve_slots] * len(setup.rates)
    sv = F.Server(setup.decs, slots=slots)
    sv.setRoundSteps(args.serve_round_steps)
    lam = offered_fps / 128.5     # requests per second (mean request 128.5 frames)
    tickets = []
    t_start = time.perf_counter()
    t_next = t_start
    while True:
        now = time.perf_counter()
        if now - t_start >= duration_s:
            break
        if now < t_next:
            time.sleep(min(t_next - now, 0.0005))
            continue
        t_next += rng.exponential(1.0 / lam)
        ri = int(rng.integers(0, len(setup.rates)))
        f = int(rng.integers(1, 257))
        o = int(rng.integers(0, setup.POOL - f + 1))
        src = setup.dev[ri][o:o + f] if device_inputs else setup.pools[ri][o:o + f]
        tickets.append((ri, o, f, sv.submit(ri, src)))
    t_sub_end = time.perf_counter()
    offered_frames = sum(f for _, _, f, _ in tickets)
    lats, frames, bits, results = [], 0, 0, []
    for ri, o, f, t in tickets:
        r = sv.wait(t, timeout=600.0)
        if r is None:
            raise RuntimeError("serving request timed out")
        msg, it, lat = r
        lats.append(lat)
        results.append((ri, o, f, msg, it))
        frames += f
        bits += f * setup.decs[ri].msgSize()
    t_end = time.perf_counter()
    # checks, outside the timed region: FER against the sent bits, and every request against
    # the same frames decoded alone (one eager batch of the rate's pool)
    errs, differ = 0, 0
    for ri, o, f, msg, it in results:
        K = setup.decs[ri].msgSize()
        m = msg.reshape(f, K)
        errs += int(np.any(m != setup.infos[ri][o:o + f], axis=1).sum())
        if not (np.array_equal(m, setup.eager_bits[ri][o:o + f]) and
                np.array_equal(np.asarray(it), setup.eager_iters[ri][o:o + f])):
            differ += 1
    st = sv.stats()
    sv.close()
    wall = t_end - t_start
    return {"offered_frames_per_s": offered_fps, "inputs": "device memory" if device_inputs else "pinned host memory",
            "arrival_frames_per_s": round(offered_frames / (t_sub_end - t_start), 1),
            "arrival_seconds": round(t_sub_end - t_start, 2), "requests": len(tickets), "frames": frames,
            "throughput_frames_per_s": round(frames / wall, 1), "info_Gb_per_s": round(bits / wall / 1e9, 4),
            "latency_ms": {"p50": round(float(np.percentile(lats, 50)) * 1e3, 2),
                           "p99": round(float(np.percentile(lats, 99)) * 1e3, 2), "unit": "per request"},
            "fer": round(errs / max(1, frames), 5), "requests_differing_from_eager_decode": differ,
            "server_steps": st["steps"], "slots_per_rate": slots[0],
            "steps_per_round": args.serve_round_steps}


def serving_report(F, args, device, duration_s):
    """The offered load (default 200k frames/s) with requests' LLRs in device memory and in
    page-locked host memory; then 50, 70 and 90 % of the throughput the device-memory run
    sustained, for the latency of an unsaturated server; and the eager rate on the same mix."""
    setup = ServingSetup(F, args, device)
    eager = setup.eager_frames_per_s()
    sat = run_serving(F, setup, args, duration_s, args.offered_fps, args.seed + 2)
    print(f"[bench] serving at offered load: {sat}", file=sys.stderr, flush=True)
    host = run_serving(F, setup, args, duration_s, args.offered_fps, args.seed + 5, device_inputs=False)
    print(f"[bench] serving from host memory: {host}", file=sys.stderr, flush=True)
    loads = {}
    for i, frac in enumerate((0.5, 0.7, 0.9)):
        load = round(frac * sat["throughput_frames_per_s"], -2)
        loads[f"{int(frac * 100)}pct"] = run_serving(F, setup, args, duration_s, load, args.seed + 3 + 10 * i)
        print(f"[bench] serving at {frac:.0%} of sustained: {loads[f'{int(frac * 100)}pct']}", file=sys.stderr,
              flush=True)
    return {"workload": "Poisson arrivals, requests of U[1,256] frames, rate uniform over DVB-S2 normal "
                        "{1/4,1/2,3/5,2/3,3/4,5/6,9/10}, Eb/N0 within +-0.3 dB of each rate's ~1e-3 FER point, "
                        f"min-sum, f16 messages, {args.iterations} max iterations",
            "eager_same_mix_frames_per_s": round(eager, 1),
            "sustained_vs_eager": round(sat["throughput_frames_per_s"] / eager, 4),
            "at_offered": sat, "at_offered_host_inputs": host,
            "latency_at_fraction_of_sustained": loads}


def run_precision(args, device):
    """Precision form on DVB-S2 n=64800 r1/2, min-sum, 50 max iterations: FP32, FP16 and BF16
    message storage on the GPU against FeCl's FP64 decoder (the oracle's bit-identical
    restatement, host cores) on the same FP64 channel LLRs.  Per Eb/N0: FER of each form and
    frames whose hard decisions differ from FP64; the Eb/N0 gap of each form's FER curve to the
    FP64 curve at FER 1e-2 (log-linear interpolation); a-posteriori LLR error after 1 and 5
    iterations (before the decoder's dynamics separate the forms), max and mean, absolute and
    relative to max(1, |FP64|)."""
    import numpy as np
    import torch
    import fecl_amd as F
    from oracle import oracle as O
    threads = host_threads()[0]
    oc = O.dvbs2(64800, 1, 2)
    H = F.DvbS2.matrix(64800, 0.5)
    fmts = {"f32": F.MsgFormat.F32, "f16": F.MsgFormat.F16, "bf16": F.MsgFormat.BF16}
    decs = {k: F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(args.iterations).msgFormat(v))
            for k, v in fmts.items()}
    K, N = decs["f32"].msgSize(), decs["f32"].paritySize()
    rng = np.random.default_rng(args.seed + 7)
    points = list(EBN0) + list(PRECISION_WATERFALL)
    frames = args.precision_frames
    rows = []
    for e in points:
        info = rng.integers(0, 2, (frames, K), dtype=np.uint8)
        cw = decs["f32"].encode(info).reshape(frames, N)
        sigma2 = 1.0 / (2.0 * 0.5 * 10.0 ** (e / 10.0))
        llr = -2.0 * ((1.0 - 2.0 * cw) + rng.normal(0.0, np.sqrt(sigma2), cw.shape)) / sigma2
        t0 = time.perf_counter()
        ref = O.decode(oc, llr, O.APPROXIMATE, args.iterations, prec=64, want=("hard", "iters"), threads=threads)
        cpu_s = time.perf_counter() - t0
        ok64 = np.all(ref["hard"] == info, axis=1)
        row = {"ebn0": e, "frames": frames, "fer_fp64": float(1 - ok64.mean()), "fp64_seconds": round(cpu_s, 2)}
        d_llr = torch.from_numpy(llr.astype(np.float32)).to(device)
        for k, dec in decs.items():
            hard, it = dec.decode(d_llr, iterations=True)
            hard = hard.view(frames, K).cpu().numpy()
            row[f"fer_{k}"] = float(np.mean(np.any(hard != info, axis=1)))
            row[f"frames_differing_from_fp64_{k}"] = int(np.sum(np.any(hard != ref["hard"], axis=1)))
        rows.append(row)
        print(f"[precision] {row}", file=sys.stderr, flush=True)
        del d_llr, llr, cw

    x64 = fer_crossing(rows, "fer_fp64")
    gaps = {}
    for k in fmts:
        xk = fer_crossing(rows, f"fer_{k}")
        gaps[k] = None if (x64 is None or xk is None) else round(xk - x64, 4)
    # a-posteriori LLR error after a few iterations
    err = {}
    info = rng.integers(0, 2, (64, K), dtype=np.uint8)
    cw = decs["f32"].encode(info).reshape(64, N)
    sigma2 = 1.0 / 10.0 ** (1.0 / 10.0)
    llr = -2.0 * ((1.0 - 2.0 * cw) + rng.normal(0.0, np.sqrt(sigma2), cw.shape)) / sigma2
    for iters in (1, 5):
        ref = O.decode(oc, llr, O.APPROXIMATE, iters, prec=64, want=("msg",), threads=threads)["msg"]
        for k, v in fmts.items():
            d = F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(iters).msgFormat(v))
            app = np.asarray(d.soDecode(llr.astype(np.float32).astype(np.float64), outputs=("msg",))["msg"]).reshape(ref.shape)
            a = np.abs(app - ref)
            err[f"{k}_iter{iters}"] = {"max_abs": float(a.max()), "mean_abs": float(a.mean()),
                                       "max_rel": float((a / np.maximum(1.0, np.abs(ref))).max())}
    return {"workload": f"DVB-S2 n=64800 r1/2, min-sum, {args.iterations} max iterations, {frames} frames per "
                        "Eb/N0, FP64 channel LLRs (GPU forms take them rounded to FP32)",
            "points": rows, "fer_1e-2_gap_db_vs_fp64": gaps, "fp64_fer_1e-2_ebn0": x64,
            "app_llr_error_vs_fp64": err}


def fer_crossing(rows, col, target=1e-2):
    """Eb/N0 where the FER column crosses `target`, by log-linear interpolation (None if it
    does not cross between the points)."""
    import numpy as np
    pts = sorted((r["ebn0"], r[col]) for r in rows)
    for (x0, y0), (x1, y1) in zip(pts, pts[1:]):
        if y0 >= target > y1:
            if y1 <= 0:
                return x1
            t = (np.log(target) - np.log(y0)) / (np.log(y1) - np.log(y0))
            return x0 + t * (x1 - x0)
    return None


def run_agreement(args, device):
    """Linear and Exact FP32 on the GPU against FeCl's FP64 decoder of the same algorithm (the
    oracle's bit-identical restatement, host cores) on the same FP64 channel LLRs, DVB-S2
    n=64800 r1/2, 50 max iterations.  Per Eb/N0: both FERs, a two-proportion z statistic for
    their difference and the frames whose hard decisions differ; the gap of the FER curves at
    1e-2; the a-posteriori LLR error after 1 and 5 iterations (max relative to max(1, |FP64|))."""
    import numpy as np
    import torch
    import fecl_amd as F
    from oracle import oracle as O
    threads = host_threads()[0]
    oc = O.dvbs2(64800, 1, 2)
    H = F.DvbS2.matrix(64800, 0.5)
    frames = args.agreement_frames
    rng = np.random.default_rng(args.seed + 11)
    out = {"workload": f"DVB-S2 n=64800 r1/2, {args.iterations} max iterations, {frames} frames per Eb/N0, "
                       "FP64 channel LLRs (the GPU takes them rounded to FP32)"}
    for name, alg, oalg in (("linear", F.Linear, O.LINEAR), ("exact", F.Exact, O.EXACT)):
        dec = F.Ldpc(H, F.DecoderOptions().algorithm(alg).iterations(args.iterations))
        K, N = dec.msgSize(), dec.paritySize()
        rows = []
        for e in AGREEMENT_EBN0:
            info = rng.integers(0, 2, (frames, K), dtype=np.uint8)
            cw = dec.encode(info).reshape(frames, N)
            sigma2 = 1.0 / (2.0 * 0.5 * 10.0 ** (e / 10.0))
            llr = -2.0 * ((1.0 - 2.0 * cw) + rng.normal(0.0, np.sqrt(sigma2), cw.shape)) / sigma2
            t0 = time.perf_counter()
            ref = O.decode(oc, llr, oalg, args.iterations, prec=64, want=("hard",), threads=threads)["hard"]
            cpu_s = time.perf_counter() - t0
            hard = dec.decode(torch.from_numpy(llr.astype(np.float32)).to(device)).view(frames, K).cpu().numpy()
            f64 = float(np.mean(np.any(ref != info, axis=1)))
            f32 = float(np.mean(np.any(hard != info, axis=1)))
            p = (f64 + f32) / 2
            z = 0.0 if p in (0.0, 1.0) else (f32 - f64) / np.sqrt(2 * p * (1 - p) / frames)
            row = {"ebn0": e, "fer_fp64": f64, "fer_gpu_fp32": f32, "z": round(float(z), 3),
                   "frames_differing": int(np.sum(np.any(hard != ref, axis=1))), "fp64_seconds": round(cpu_s, 2)}
            rows.append(row)
            print(f"[agreement] {name} {row}", file=sys.stderr, flush=True)
        x64, x32 = fer_crossing(rows, "fer_fp64"), fer_crossing(rows, "fer_gpu_fp32")
        err = {}
        info = rng.integers(0, 2, (32, K), dtype=np.uint8)
        cw = dec.encode(info).reshape(32, N)
        sigma2 = 1.0 / 10.0 ** (1.0 / 10.0)
        llr = -2.0 * ((1.0 - 2.0 * cw) + rng.normal(0.0, np.sqrt(sigma2), cw.shape)) / sigma2
        llr = llr.astype(np.float32).astype(np.float64)
        for iters in (1, 5):
            ref = O.decode(oc, llr, oalg, iters, prec=64, want=("msg",), threads=threads)["msg"]
            d = F.Ldpc(H, F.DecoderOptions().algorithm(alg).iterations(iters))
            app = np.asarray(d.soDecode(llr, outputs=("msg",))["msg"]).reshape(ref.shape)
            fin = np.isfinite(ref)
            a = np.abs(app - ref)[fin]
            err[f"iter{iters}"] = {"max_rel": float((a / np.maximum(1.0, np.abs(ref[fin]))).max()),
                                   "mean_abs": float(a.mean())}
        out[name] = {"points": rows, "fer_1e-2_ebn0_fp64": x64, "fer_1e-2_ebn0_gpu": x32,
                     "gap_db": None if x64 is None or x32 is None else round(x32 - x64, 4),
                     "app_llr_error_vs_fp64": err}
    return out


def run_plain(args, device):
    """The reference's own CPU-runnable case: DVB-S2 n=64800 r1/2, 16 frames at Eb/N0 = 1.0 dB,
    FP64 channel LLRs, Exact, 50 max iterations with syndrome early termination.  The GPU takes
    the LLRs through FeCl's decode hook (FP64 in host memory, bits out, PCIe included) and, for
    scale, from device memory; stock FeCl decodes the same LLRs on the host cores.  Agreement:
    hard decisions and iteration counts against FeCl's FP64 decoder (the oracle's bit-identical
    restatement), APP error against the exact box-plus value."""
    import numpy as np
    import torch
    import fecl_amd as F
    from oracle import oracle as O
    frames, iters, ebn0, reps = 16, 50, 1.0, 20
    threads = host_threads()[0]
    H = F.DvbS2.matrix(64800, 0.5)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Exact).iterations(iters))
    K, N = dec.msgSize(), dec.paritySize()
    rng = np.random.default_rng(args.seed + 21)
    info = rng.integers(0, 2, (frames, K), dtype=np.uint8)
    cw = dec.encode(info).reshape(frames, N)
    sigma2 = 1.0 / (2.0 * (K / N) * 10.0 ** (ebn0 / 10.0))
    llr = -2.0 * ((1.0 - 2.0 * cw) + rng.normal(0.0, np.sqrt(sigma2), cw.shape)) / sigma2
    bits = dec.decode(llr)
    t0 = time.perf_counter()
    for _ in range(reps):
        bits = dec.decode(llr)
    host_s = (time.perf_counter() - t0) / reps
    x = torch.from_numpy(llr.astype(np.float32)).to(device)
    dec.decode(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dbits, dit = dec.decode(x, iterations=True)
    torch.cuda.synchronize()
    dev_s = (time.perf_counter() - t0) / reps
    out = {"metric": "decoded info Gb/s, DVB-S2 r1/2 Exact, 16 frames of FP64 LLRs (the reference's CPU case)",
           "workload": f"DVB-S2 n=64800 r1/2, {frames} frames, Eb/N0 {ebn0} dB, Exact, {iters} max iterations, "
                       "FP64 LLRs from host memory",
           "gpu_host_path_Gb_per_s": round(frames * K / host_s / 1e9, 4), "gpu_host_path_ms": round(host_s * 1e3, 3),
           "gpu_device_resident_Gb_per_s": round(frames * K / dev_s / 1e9, 4),
           "fer_gpu": float(np.mean(np.any(np.asarray(bits).reshape(frames, K) != info, axis=1)))}
    cpu = None
    exe = os.path.join(ROOT, "oracle", "_ref", "fecl_bench")
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "llr.bin")
            with open(path, "wb") as f:
                np.array(llr.shape, np.int64).tofile(f)
                llr.astype(np.float64).tofile(f)
            try:
                r = subprocess.run([exe, "dvbs2:64800:0.5", path, "0", str(iters), str(threads)],
                                   capture_output=True, text=True, timeout=600, check=True)
                cpu = {"seconds": json.loads(r.stdout.strip().splitlines()[-1])["seconds"], "kind": "reference"}
            except Exception as e:
                print(f"[bench] stock FeCl unavailable ({e}); using the C restatement", file=sys.stderr, flush=True)
    code = O.dvbs2(64800, 1, 2)
    t0 = time.perf_counter()
    ref = O.decode(code, llr, O.EXACT, iters, prec=64, want=("hard", "iters", "msg"), threads=threads)
    if cpu is None:
        cpu = {"seconds": time.perf_counter() - t0, "kind": "port"}
    out["cpu_baseline"] = {"value": round(frames * K / cpu["seconds"] / 1e9, 6), "unit": "Gb/s", "cores": threads,
                           "kind": cpu["kind"], "seconds": round(cpu["seconds"], 3), "cpu": _cpu_model()}
    exact = O.decode(code, llr.astype(np.float32).astype(np.float64), O.EXACT_STABLE, iters, prec=64,
                     want=("msg",), threads=threads)["msg"]
    app = np.asarray(dec.soDecode(llr, outputs=("msg",))["msg"]).reshape(frames, K)
    out["agreement"] = {
        "hard_identical_to_fecl_fp64": bool(np.array_equal(np.asarray(bits).reshape(frames, K), ref["hard"])),
        "iterations_identical_to_fecl_fp64": bool(np.array_equal(dit.cpu().numpy(), ref["iters"])),
        "iterations": [int(v) for v in ref["iters"]],
        "app_max_rel_error_vs_exact_box_plus": float((np.abs(app - exact) / np.maximum(1.0, np.abs(exact))).max())}
    return out


# Best stream-probe setting per mode on MI355X (`bench.py --config copy` sweeps them all):
# (order/hints flags, workgroups per CU)
CEILING_SETTINGS = {"read": (3, 2), "write": (2, 8), "update": (1, 2), "copy": (1, 2), "scatter_update": (0, 16)}


def stream_ceiling(device, gib=2, reps=5, settings=None):
    """Achievable HBM bandwidth on this GPU: hand-written streaming kernels
    (fecl_amd/libfecl_stream.so, 16-byte words) copying, updating in place, reading and writing
    gib GiB, swept over grid-stride and per-block-chunk order, with and without non-temporal
    hints, at 2 to 16 workgroups of 256 threads per CU (or only `settings[mode]`); best of
    `reps` per setting, timed with HIP events on the launch stream.  GB/s counts the bytes read
    plus written.  The full sweep adds scattered 1 KiB rows updated in place over 8 GiB.
    Returns the best setting of each mode and the whole sweep."""
    import ctypes
    import torch
    lib = ctypes.CDLL(os.path.join(ROOT, "fecl_amd", "libfecl_stream.so"))
    lib.fecl_stream_probe.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p]
    cus = torch.cuda.get_device_properties(device).multi_processor_count
    nbytes = gib << 30
    a = torch.ones(nbytes // 4, dtype=torch.float32, device=device)
    b = torch.empty_like(a)
    st = torch.cuda.current_stream(device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sweep, best = {}, {}
    # "update" is the copy with source and destination the same buffer: every word is read and
    # written back in place, as the decoder's check and variable passes update their rows
    for mode, name, moved in ((0, "copy", 2 * nbytes), (0, "update", 2 * nbytes), (1, "read", nbytes),
                              (2, "write", nbytes)):
        dst = a if name == "update" else b
        grid = [settings[name]] if settings else [(f, p) for f in range(4) for p in (2, 4, 8, 16)]
        for flags, per_cu in grid:
            t = 1e30
            for _ in range(reps + 1):
                e0.record(st)
                rc = lib.fecl_stream_probe(mode, a.data_ptr(), dst.data_ptr(), nbytes, flags, per_cu * cus,
                                           st.cuda_stream)
                e1.record(st)
                e1.synchronize()
                if rc:
                    raise RuntimeError(f"stream probe failed ({rc})")
                t = min(t, e0.elapsed_time(e1))
            gbs = round(moved / t / 1e6, 1)
            key = f"{name}/{'chunk' if flags & 2 else 'stride'}/{'nt' if flags & 1 else 'plain'}/{per_cu}"
            sweep[key] = gbs
            if gbs > best.get(name, (0.0, ""))[0]:
                best[name] = (gbs, key)
    del a, b
    torch.cuda.empty_cache()
    scatter = settings.get("scatter_update") if settings else None
    if settings is None or scatter:
        # scattered 1 KiB rows updated in place over 8 GiB (32x the Infinity Cache), the access
        # pattern of the variable pass; with `settings`, one setting and one timed launch
        sb = 8 << 30
        c = torch.zeros(sb // 4, dtype=torch.float32, device=device)
        for flags, per_cu in [scatter] if scatter else [(f, p) for f in (0, 1) for p in (2, 4, 8, 16)]:
            waves = per_cu * cus * 4
            moved = 2 * 8192 * max(1, (sb // 1024) // (8 * waves)) * waves
            t = 1e30
            for _ in range((1 if scatter else reps) + 1):
                e0.record(st)
                rc = lib.fecl_stream_probe(3, None, c.data_ptr(), sb, flags, per_cu * cus, st.cuda_stream)
                e1.record(st)
                e1.synchronize()
                if rc:
                    raise RuntimeError(f"stream probe failed ({rc})")
                t = min(t, e0.elapsed_time(e1))
            gbs = round(moved / t / 1e6, 1)
            key = f"scatter_update/rows1KiB/{'nt' if flags & 1 else 'plain'}/{per_cu}"
            sweep[key] = gbs
            if gbs > best.get("scatter_update", (0.0, ""))[0]:
                best["scatter_update"] = (gbs, key)
        del c
        torch.cuda.empty_cache()
    return {"bytes": nbytes, "best_GBs": {k: v[0] for k, v in best.items()},
            "best_setting": {k: v[1] for k, v in best.items()},
            "sweep_GBs": sweep, "key": "mode/order/hints/workgroups per CU"}


def copy_reference(device, gib=1, reps=5):
    """A plain device-to-device copy of gib GiB (the PMC byte-counter calibration kernel)."""
    import torch
    n = gib << 28
    a = torch.ones(n, dtype=torch.float32, device=device)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "reps": reps}


def time_form(F, H, alg, fmt, iterations, batches, frames, K, steps, device):
    """Gb/s, FER and kernel GB/s of one algorithm / message format on the bench batches
    (autotuned layout, one warm-up decode, `steps` timed decodes; then the same decodes again
    with per-kernel timing events, which slow the decode, for the kernel GB/s)."""
    import torch
    from fecl_amd._lib import lib, check
    import ctypes
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(alg).iterations(iterations).msgFormat(fmt))
    dec.autotune(frames, probe_iterations=3)
    stream = torch.cuda.current_stream(device).cuda_stream
    dec.decode(batches[0]["llr"], stream=stream)
    torch.cuda.synchronize()
    errs, its = 0, 0
    t0 = time.perf_counter()
    outs = []
    for i in range(steps):
        b = batches[i % len(batches)]
        outs.append((b, dec.decode(b["llr"], iterations=True, stream=stream)))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    check(lib().fecl_ldpc_profile(dec._h, 1))
    for i in range(steps):
        dec.decode(batches[i % len(batches)]["llr"], stream=stream)
    torch.cuda.synchronize()
    ms = (ctypes.c_double * 3)()
    launches = (ctypes.c_int64 * 3)()
    check(lib().fecl_ldpc_profile_read(dec._h, ms, launches))
    check(lib().fecl_ldpc_profile(dec._h, 0))
    for b, (hard, it) in outs:
        errs += int(torch.any(hard.view(frames, K) != b["info"], dim=1).sum())
        its += int(it.to(torch.int64).sum())
    E, N = dec.stateSize(), dec.paritySize()
    w = 4 if fmt == F.MsgFormat.F32 else 2
    vn_updates = its - steps * frames
    return {"Gb_per_s": round(steps * frames * K / total / 1e9, 4), "fer": round(errs / (steps * frames), 5),
            "lane_words": dec.layout()[0], "steps": steps,
            "cn_GBs": round(its * 2 * E * w / (ms[0] / 1e3) / 1e9, 1) if ms[0] > 0 else None,
            "vn_GBs": round(vn_updates * (2 * E * w + 4 * N) / (ms[1] / 1e3) / 1e9, 1) if ms[1] > 0 else None}


def host_threads():
    """Threads for the host legs: the CPUs this process may run on (sched_getaffinity), capped
    by the cgroup CPU quota when one is set and at 16 (the CPU share of one GPU on the test machines)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    n = aff if quota is None else min(aff, max(1, int(quota)))
    return min(n, 16), {"affinity_cpus": aff, "cgroup_cpu_quota": quota, "cap": 16}


def run_waterfall(F, H, device, threads, iterations, seed, frames=2048, points=(1.55, 1.6)):
    """FER on the r1/2 waterfall of the GPU's FP32, FP16 and BF16 message forms (min-sum) and of
    FeCl's FP64 decoder (the oracle's bit-identical restatement, host cores) on the same FP64
    channel LLRs (the GPU takes them rounded to FP32).  The gap of each form to FP64 in dB at
    each point is its log10-FER excess over FP64 divided by the FP64 curve's log10-FER slope
    between the points (None where a FER is 0)."""
    import numpy as np
    import torch
    from oracle import oracle as O
    oc = O.dvbs2(64800, 1, 2)
    fmts = {"f32": F.MsgFormat.F32, "f16": F.MsgFormat.F16, "bf16": F.MsgFormat.BF16}
    decs = {k: F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(iterations).msgFormat(v))
            for k, v in fmts.items()}
    K, N = decs["f32"].msgSize(), decs["f32"].paritySize()
    rng = np.random.default_rng(seed + 13)
    rows = []
    for e in points:
        info = rng.integers(0, 2, (frames, K), dtype=np.uint8)
        cw = decs["f32"].encode(info).reshape(frames, N)
        sigma2 = 1.0 / (2.0 * 0.5 * 10.0 ** (e / 10.0))
        llr = -2.0 * ((1.0 - 2.0 * cw) + rng.normal(0.0, np.sqrt(sigma2), cw.shape)) / sigma2
        t0 = time.perf_counter()
        ref = O.decode(oc, llr, O.APPROXIMATE, iterations, prec=64, want=("hard",), threads=threads)["hard"]
        row = {"ebn0": e, "frames": frames, "fer_fp64": float(np.mean(np.any(ref != info, axis=1))),
               "fp64_seconds": round(time.perf_counter() - t0, 2)}
        d_llr = torch.from_numpy(llr.astype(np.float32)).to(device)
        for k, dec in decs.items():
            hard = dec.decode(d_llr).view(frames, K).cpu().numpy()
            row[f"fer_{k}"] = float(np.mean(np.any(hard != info, axis=1)))
            row[f"frames_differing_from_fp64_{k}"] = int(np.sum(np.any(hard != ref, axis=1)))
        rows.append(row)
        print(f"[bench] waterfall {row}", file=sys.stderr, flush=True)
        del d_llr, llr, cw, ref
    gaps = {k: [None] * len(rows) for k in fmts}
    f0, f1 = rows[0]["fer_fp64"], rows[-1]["fer_fp64"]
    slope = None
    if f0 > f1 > 0:
        slope = (np.log10(f1) - np.log10(f0)) / (points[-1] - points[0])
        for k in fmts:
            for i, r in enumerate(rows):
                if r[f"fer_{k}"] > 0 and r["fer_fp64"] > 0:
                    gaps[k][i] = round(float((np.log10(r[f"fer_{k}"]) - np.log10(r["fer_fp64"])) / -slope), 4)
    return {"workload": f"DVB-S2 n=64800 r1/2, min-sum, {iterations} max iterations, {frames} frames per Eb/N0, "
                        "FP64 channel LLRs", "host_threads": threads, "points": rows,
            "fp64_log10_fer_slope_per_db": None if slope is None else round(float(slope), 3),
            "gap_db_vs_fp64": gaps}


def fp64_parity(llr_host, info_host, iters, threads, gpu_hard):
    """FER parity on the sample: FeCl's FP64 decoder (the oracle's bit-identical restatement)
    against the GPU's hard decisions on the same frames."""
    import numpy as np
    from oracle import oracle as O
    code = O.dvbs2(64800, 1, 2)
    ref = O.decode(code, llr_host.astype(np.float64), O.APPROXIMATE, iters, prec=64,
                   want=("hard",), threads=threads)
    ok64 = np.all(ref["hard"] == info_host, axis=1)
    same = np.all(gpu_hard == ref["hard"], axis=1)
    return {"frames": int(len(ok64)), "fer_fp64": float(1.0 - ok64.mean()),
            "fer_gpu": float(np.mean(np.any(gpu_hard != info_host, axis=1))),
            "frames_identical_to_fp64": float(same.mean()),
            "fp64_decoded_frames_identical": float(same[ok64].mean()) if ok64.any() else None}


def cpu_baseline(llr_host, info_host, iters, threads):
    """Stock FeCl (oracle/_ref/fecl_bench, built from the reference sources) on the host cores;
    falls back to the oracle's C restatement of the same decoder."""
    import numpy as np
    frames = llr_host.shape[0]
    exe = os.path.join(ROOT, "oracle", "_ref", "fecl_bench")
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "llr.bin")
            with open(path, "wb") as f:
                np.array(llr_host.shape, np.int64).tofile(f)
                llr_host.astype(np.float64).tofile(f)
            try:
                out = subprocess.run([exe, "dvbs2:64800:0.5", path, "2", str(iters), str(threads)],
                                     capture_output=True, text=True, timeout=600, check=True)
                r = json.loads(out.stdout.strip().splitlines()[-1])
                return {"seconds": r["seconds"], "kind": "reference", "fer": None}
            except Exception as e:  # report and fall back to the restatement
                print(f"[bench] stock FeCl baseline unavailable ({e}); using the C restatement",
                      file=sys.stderr, flush=True)
    from oracle import oracle as O
    code = O.dvbs2(64800, 1, 2)
    t0 = time.perf_counter()
    ref = O.decode(code, llr_host.astype(np.float64), O.APPROXIMATE, iters, prec=64,
                   want=("hard",), threads=threads)
    dt = time.perf_counter() - t0
    fer = float(np.mean(np.any(ref["hard"] != info_host, axis=1)))
    return {"seconds": dt, "kind": "port", "fer": fer}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=1024)
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--msg", choices=["f32", "f16", "bf16"], default="f32")
    ap.add_argument("--alg", choices=["minsum", "linear", "exact"], default="minsum")
    ap.add_argument("--cpu-frames", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--words", type=int, default=-1,
                    help="lane layout: 1, 2, 4 words per lane per edge; 0 automatic; -1 autotune")
    ap.add_argument("--config", choices=["batch", "uneven", "serving", "precision", "agreement", "calibrate", "copy",
                                         "plain"],
                    default="batch")
    ap.add_argument("--precision-frames", type=int, default=4096)
    ap.add_argument("--agreement-frames", type=int, default=1024)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-slots", type=int, default=2048)
    ap.add_argument("--serve-round-steps", type=int, default=1, help="server steps per round")
    ap.add_argument("--offered-fps", type=float, default=200000.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-uneven", action="store_true", help="skip the seven-rate uneven batch leg")
    ap.add_argument("--no-waterfall", action="store_true", help="skip the FP16/BF16 vs FP64 waterfall FER leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the reference's 16-frame Exact CPU case")
    ap.add_argument("--no-forms", action="store_true", help="skip the Linear / FP16 / BF16 lines")
    ap.add_argument("--ebn0", type=str, default="",
                    help="comma-separated Eb/N0 (dB) of the batches (default: the headline points)")
    ap.add_argument("--compaction", choices=["on", "off"], default="on",
                    help="compaction of live frames into fewer tiles (results are the same)")
    ap.add_argument("--pass-frames", type=int, default=0, help="frames per decode pass (0 = library default)")
    ap.add_argument("--cache-policy", type=int, default=-1,
                    help="message-row cache hints: -1 automatic, else bit 0 check / bit 1 variable non-temporal")
    ap.add_argument("--vpw", type=int, default=0, help="variables per wavefront (0 automatic; with --words >= 0)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the decoder's frames are independent, so N GPUs would "
              "decode disjoint frame slices with no exchange; only --gpus 1 is run here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import fecl_amd as F
    from fecl_amd._lib import lib, check
    import ctypes

    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    if args.config == "uneven":
        print(json.dumps(run_uneven(args)), flush=True)
        return
    if args.config == "serving":
        print(json.dumps(serving_report(F, args, device, args.serve_seconds)), flush=True)
        return
    if args.config == "calibrate":
        print(json.dumps(run_calibrate(args, device)), flush=True)
        return
    if args.config == "precision":
        print(json.dumps(run_precision(args, device)), flush=True)
        return
    if args.config == "agreement":
        print(json.dumps(run_agreement(args, device)), flush=True)
        return
    if args.config == "copy":
        out = copy_reference(device)
        out["streaming_kernels"] = stream_ceiling(device)
        print(json.dumps(out), flush=True)
        return
    if args.config == "plain":
        print(json.dumps(run_plain(args, device)), flush=True)
        return
    alg = {"minsum": F.Approximate, "linear": F.Linear, "exact": F.Exact}[args.alg]
    fmt = {"f32": F.MsgFormat.F32, "f16": F.MsgFormat.F16, "bf16": F.MsgFormat.BF16}[args.msg]
    H = F.DvbS2.matrix(64800, 0.5)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(alg).iterations(args.iterations).msgFormat(fmt))
    K, N, E = dec.msgSize(), dec.paritySize(), dec.stateSize()
    tune = None
    if args.words >= 0:
        dec.setLayout(args.words)
        dec.setVarsPerWave(args.vpw)
    else:
        best, times = dec.autotune(args.frames, probe_iterations=5)
        tune = {"picked_words": best, "picked_vars_per_wave": dec.varsPerWave(),
                "probe_ms": {f"w{k[0]}v{k[1]}": round(v, 3) for k, v in times.items()}}
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)
    dec.setPassFrames(args.pass_frames)
    dec.setCachePolicy(args.cache_policy)
    dec.setCompaction(args.compaction == "on")
    words, frames_per_tile = dec.layout()
    points = tuple(float(x) for x in args.ebn0.split(",")) if args.ebn0 else EBN0
    print(f"[bench] building {len(points)} batches of {args.frames} frames", file=sys.stderr, flush=True)
    batches = make_batches(F, dec, H, args.frames, args.seed, device, points)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(device)
    out = {}

    def step(i):
        b = batches[i % len(batches)]
        hard, it = dec.decode(b["llr"], iterations=True, stream=stream.cuda_stream)
        out[i % len(batches)] = (hard, it)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)

    check(lib().fecl_ldpc_profile(dec._h, 1))
    lat = []
    kept = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        step(i)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        kept.append((i % len(batches), out[i % len(batches)][1].clone()))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    # accounting outside the timed region: iteration counts of every timed step, FER of the
    # last decode of each batch
    iter_sum = sum(int(it.to(torch.int64).sum()) for _, it in kept)
    vn_sum = sum(int((it.to(torch.int64) - 1).sum()) for _, it in kept)
    errs, nfr = 0, 0
    for bi, (hard, _) in out.items():
        errs += int(torch.any(hard.view(args.frames, K) != batches[bi]["info"], dim=1).sum())
        nfr += args.frames
    ms = (ctypes.c_double * 3)()
    launches = (ctypes.c_int64 * 3)()
    check(lib().fecl_ldpc_profile_read(dec._h, ms, launches))
    check(lib().fecl_ldpc_profile(dec._h, 0))

    info_bits = args.steps * args.frames * K
    value = info_bits / total / 1e9
    w = 4 if args.msg == "f32" else 2
    # algorithmic bytes: check node reads and writes every edge message once per frame update;
    # variable node does the same plus the FP32 channel value of every variable.
    cn_bytes = iter_sum * 2 * E * w
    vn_bytes = vn_sum * (2 * E * w + 4 * N)
    cn_ms, vn_ms, sc_ms = ms[0], ms[1], ms[2]
    cn_gbs = cn_bytes / (cn_ms / 1e3) / 1e9 if cn_ms > 0 else 0.0
    vn_gbs = vn_bytes / (vn_ms / 1e3) / 1e9 if vn_ms > 0 else 0.0
    if vn_ms >= cn_ms:
        dom, dom_gbs, dom_launch_ms, dom_bytes_per = "vn_kernel", vn_gbs, vn_ms / max(1, launches[1]), vn_bytes / max(1, launches[1])
    else:
        dom, dom_gbs, dom_launch_ms, dom_bytes_per = "cn_kernel", cn_gbs, cn_ms / max(1, launches[0]), cn_bytes / max(1, launches[0])
    ratio = PMC_TRAFFIC_RATIO.get((args.msg, words), {}).get(dom)
    traffic = int(ratio * dom_bytes_per) if ratio else None
    b_iter = 4 * E * w + 4 * N
    whole_gbs = iter_sum * b_iter / total / 1e9

    result = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gb/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.msg,
        "data": "synthetic: seeded Bernoulli info bits, DVB-S2 encoder, BPSK/AWGN, FeCl-convention FP32 LLR",
        "config": {"workload": f"DVB-S2 n=64800 r1/2, {args.frames} frames per batch, "
                               f"{'min-sum' if args.alg == 'minsum' else args.alg}, {args.msg} messages, "
                               f"{args.iterations} max iterations with syndrome early exit, "
                               f"Eb/N0 {list(points)} dB",
                   "frames_per_step": args.frames, "iterations": args.iterations, "algorithm": args.alg,
                   "msg_format": args.msg, "code": "dvbs2 64800 1/2",
                   "lane_words": words, "frames_per_tile": frames_per_tile, "vars_per_wave": dec.varsPerWave(),
                   "compaction": args.compaction,
                   "autotune": tune,
                   "pass_frames": args.pass_frames, "cache_policy": args.cache_policy},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3), "unit": "per batch decode"},
        "fer": round(errs / max(1, nfr), 5),
        "mean_iterations": round(iter_sum / max(1, args.steps * args.frames), 3),
        "achieved_algorithmic_GBs_whole_step": round(whole_gbs, 1),
        "kernels_ms": {"cn": round(cn_ms, 3), "vn": round(vn_ms, 3), "syndrome_control": round(sc_ms, 3)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launch_ms": round(dom_launch_ms, 4), "algorithmic_bytes_per_launch": int(dom_bytes_per),
                     "cn_GBs": round(cn_gbs, 1), "vn_GBs": round(vn_gbs, 1)},
    }

    # the achievable ceiling for the dominant kernel's read/write mix: the best hand-written
    # streaming read and write rates on this GPU, combined in the kernel's proportions
    ceil = stream_ceiling(device, gib=1, reps=3, settings=CEILING_SETTINGS)["best_GBs"]
    rd = E * w + (4 * N if dom == "vn_kernel" else 0)
    wr = E * w
    mix = (rd + wr) / (rd / ceil["read"] + wr / ceil["write"])
    # the probe with the dominant kernel's own access pattern: the variable pass updates 1 KiB
    # rows scattered over the edge array in place, the check pass contiguous rows
    pattern = ceil["scatter_update"] if dom == "vn_kernel" else ceil["update"]
    result["roofline"]["measured_ceiling"] = {
        "read_GBs": ceil["read"], "write_GBs": ceil["write"], "update_in_place_GBs": ceil["update"],
        "copy_GBs": ceil["copy"], "scatter_update_1KiB_rows_GBs": ceil["scatter_update"],
        "mix_bound_GBs": round(mix, 1), "frac_of_mix_bound": round(dom_gbs / mix, 4),
        "pattern_probe_GBs": pattern, "frac_of_pattern_probe": round(dom_gbs / pattern, 4),
        "note": "streaming kernels of fecl_amd/libfecl_stream.so; mix bound = the kernel's bytes read and "
                "written at the best read-only and write-only rates; pattern probe = an in-place update "
                "with the kernel's access pattern (scattered 1 KiB rows over 8 GiB for the variable "
                "kernel, a contiguous stream for the check kernel), which the decoder can exceed where "
                "its rows stay in the 256 MB Infinity Cache between launches"}

    if not args.no_forms:
        # the same batches through the other forms of the path (one timed decode of each batch)
        forms = {}
        for name, a2, f2 in (("linear_f32", F.Linear, F.MsgFormat.F32), ("exact_f32", F.Exact, F.MsgFormat.F32),
                             ("minsum_f16", F.Approximate, F.MsgFormat.F16),
                             ("minsum_bf16", F.Approximate, F.MsgFormat.BF16)):
            if a2 == alg and f2 == fmt:
                continue
            forms[name] = time_form(F, H, a2, f2, args.iterations, batches, args.frames, K, len(batches), device)
            print(f"[bench] {name}: {forms[name]}", file=sys.stderr, flush=True)
        result["other_forms"] = forms
        # the precision form on its own 4096-frame batches (same channel points)
        p_batches = make_batches(F, dec, H, 4096, args.seed + 1, device, points)
        torch.cuda.synchronize()
        result["precision_4096_frames"] = {
            name: time_form(F, H, F.Approximate, f2, args.iterations, p_batches, 4096, K, len(p_batches), device)
            for name, f2 in (("minsum_f16", F.MsgFormat.F16), ("minsum_bf16", F.MsgFormat.BF16))}
        print(f"[bench] 4096-frame precision form: {result['precision_4096_frames']}", file=sys.stderr, flush=True)
        del p_batches
        torch.cuda.empty_cache()
        # FeCl's hooks hand over host FP64 buffers: the same decode from host memory, transfers
        # included (pageable numpy arrays, as a FeCl caller holds them)
        llr_h = batches[0]["llr"].cpu().numpy().astype(np.float64)
        dec.decode(llr_h)
        t0 = time.perf_counter()
        for _ in range(2):
            dec.decode(llr_h)
        th = (time.perf_counter() - t0) / 2
        result["host_buffer_path"] = {"Gb_per_s": round(args.frames * K / th / 1e9, 4), "ms_per_batch": round(th * 1e3, 2),
                                      "note": "FP64 LLRs in pageable host memory in, bits out, PCIe included"}
        print(f"[bench] host buffer path: {result['host_buffer_path']}", file=sys.stderr, flush=True)
        # soDecode from host memory as an iterative receiver calls it: FP64 LLRs in, FP64
        # a-posteriori LLRs and the state (one value per edge, 1.8 MB a frame) out
        nso = min(512, args.frames)
        llr_s = np.ascontiguousarray(llr_h[:nso])
        dec.soDecode(llr_s, outputs=("msg", "state"))
        t0 = time.perf_counter()
        for _ in range(2):
            dec.soDecode(llr_s, outputs=("msg", "state"))
        ts = (time.perf_counter() - t0) / 2
        result["host_buffer_path_sodecode"] = {
            "Gb_per_s": round(nso * K / ts / 1e9, 4), "ms_per_batch": round(ts * 1e3, 2), "frames": nso,
            "note": "FP64 LLRs in, FP64 a-posteriori LLRs and state out, pageable host memory, PCIe included"}
        print(f"[bench] host soDecode path: {result['host_buffer_path_sodecode']}", file=sys.stderr, flush=True)

    if not args.no_cpu:
        threads, cores_info = host_threads()
        cpu_batch = batches[min(2, len(batches) - 1)]  # Eb/N0 = 1.0 dB with the headline points
        nf = min(args.cpu_frames, args.frames)
        llr_host = cpu_batch["llr"][:nf].cpu().numpy().astype(np.float64)
        info_host = cpu_batch["info"][:nf].cpu().numpy()
        print(f"[bench] CPU baseline: {nf} frames on {threads} threads", file=sys.stderr, flush=True)
        cb = cpu_baseline(llr_host, info_host, args.iterations, threads)
        if args.alg == "minsum":
            # FP64 FeCl decoder vs this run's last decode of the 1.0 and 1.4 dB batches
            par = {}
            for bi in (2, 4):
                if bi not in out:
                    continue
                npar = min(256, args.frames)
                b = batches[bi]
                par[f"{b['ebn0']} dB"] = fp64_parity(
                    b["llr"][:npar].cpu().numpy().astype(np.float64), b["info"][:npar].cpu().numpy(),
                    args.iterations, threads, out[bi][0].view(args.frames, K)[:npar].cpu().numpy())
            result["fer_parity_sample"] = par
        result["cpu_baseline"] = {"value": round(nf * K / cb["seconds"] / 1e9, 6), "unit": "Gb/s",
                                  "cores": threads, "kind": cb["kind"],
                                  "sample": f"{nf} frames of the Eb/N0={cpu_batch['ebn0']} dB batch, FP64 min-sum, "
                                            f"{args.iterations} max iterations, {cb['seconds']:.2f} s wall",
                                  "cpu": _cpu_model(), "host_cpus": cores_info}
        if not args.no_waterfall:
            result["waterfall_fer_vs_fp64"] = run_waterfall(F, H, device, threads, args.iterations, args.seed)
        if not args.no_plain:
            # the reference's own CPU case (16 Exact frames of FP64 LLRs) beside stock FeCl
            print("[bench] the reference's CPU case", file=sys.stderr, flush=True)
            result["reference_cpu_case"] = run_plain(args, device)
    if not args.no_serving or not args.no_uneven:
        # the uneven batch and the serving pools start from a device heap without the batch
        # bench's buffers (decoder workspaces, input batches, torch's cached blocks)
        del dec, batches, out, kept
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if not args.no_serving:
        print(f"[bench] serving loop: {args.serve_seconds} s of arrivals", file=sys.stderr, flush=True)
        result["serving"] = serving_report(F, args, device, args.serve_seconds)
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if not args.no_uneven:
        print("[bench] uneven seven-rate batch", file=sys.stderr, flush=True)
        ua = argparse.Namespace(**vars(args))
        ua.steps, ua.warmup = 3, 1
        result["uneven"] = run_uneven(ua)
        print(f"[bench] uneven: {result['uneven']}", file=sys.stderr, flush=True)
    print(json.dumps(result), flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
