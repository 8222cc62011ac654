"""fecl_ldpc_decode_device_noalloc (the decode whose workspace is the library's fixed device
array) through the C ABI, against fecl_ldpc_decode_device on the same handle and inputs:
identical hard decisions and iteration counts on two calls in a row, the second with fewer
frames (the array keeps the first call's contents), on codes of several tiles and formats;
sizes beyond the array's capacity are refused with an error, and the call also runs inside a
captured hipGraph."""
import numpy as np
import pytest

import fecl_amd as F
from fecl_amd._lib import FECL_INVALID_ARGUMENT, lib
from oracle import oracle as O
from tests.golden.vectors import dvbs2_llr

pytestmark = pytest.mark.gpu


def noalloc():
    return lib().fecl_ldpc_decode_device_noalloc


def run(fn, dec, x, frames, K, stream=None):
    import torch
    bits = torch.full((frames * K,), 7, dtype=torch.uint8, device="cuda")
    iters = torch.full((frames,), -1, dtype=torch.int32, device="cuda")
    st = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    rc = fn(dec._h, x.data_ptr(), frames, bits.data_ptr(), iters.data_ptr(), st)
    return rc, bits, iters


def reference(dec, x, frames, K):
    import torch
    bits = torch.full((frames * K,), 7, dtype=torch.uint8, device="cuda")
    iters = torch.full((frames,), -1, dtype=torch.int32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    rc = lib().fecl_ldpc_decode_device(dec._h, x.data_ptr(), frames, bits.data_ptr(), iters.data_ptr(), st)
    assert rc == 0, lib().fecl_last_error()
    return bits, iters


@pytest.mark.parametrize("n,rate,alg,fmt,frames,ebn0", [
    (16200, (1, 2), F.Approximate, F.MsgFormat.F32, (150, 40), 2.0),   # 3 tiles, then 1
    (64800, (1, 2), F.Linear, F.MsgFormat.F32, (60, 17), 1.0),         # one 64-frame tile
    (64800, (3, 4), F.Approximate, F.MsgFormat.F16, (128, 5), 3.0),    # one 128-frame tile
])
def test_noalloc_equals_device_decode(n, rate, alg, fmt, frames, ebn0):
    import torch
    oc = O.dvbs2(n, *rate)
    H = F.CheckMatrix(oc.row_ptr, oc.col, oc.N)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(alg).iterations(30).msgFormat(fmt))
    fn = noalloc()
    for call, f in enumerate(frames):
        llr = dvbs2_llr(f, ebn0 + 0.3 * call, seed=500 + call, rate=rate, n=n).astype(np.float32)
        x = torch.from_numpy(llr.ravel()).cuda()
        print(f"noalloc n={n} rate={rate} alg={int(alg)} fmt={int(fmt)} call={call} frames={f}", flush=True)
        rc, bits, iters = run(fn, dec, x, f, oc.K)
        assert rc == 0, lib().fecl_last_error()
        rbits, riters = reference(dec, x, f, oc.K)
        torch.cuda.synchronize()
        assert torch.equal(bits, rbits), (call, f)
        assert torch.equal(iters, riters), (call, f)
        assert int(iters.min()) >= 1


def test_noalloc_refuses_batches_beyond_its_array():
    import torch
    oc = O.dvbs2(64800, 1, 2)
    H = F.CheckMatrix(oc.row_ptr, oc.col, oc.N)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(10))
    x = torch.zeros(200 * oc.N, dtype=torch.float32, device="cuda")
    rc, _, _ = run(noalloc(), dec, x, 200, oc.K)
    assert rc == FECL_INVALID_ARGUMENT
    assert "workspace" in lib().fecl_last_error().decode()
    rc, bits, iters = run(noalloc(), dec, x, 0, oc.K)
    assert rc == 0


def test_noalloc_inside_a_captured_graph():
    """Recorded into a hipGraph (no allocation may happen during capture) and replayed twice with
    new inputs: each replay equals the eager device decode."""
    import torch
    oc = O.dvbs2(16200, 1, 2)
    H = F.CheckMatrix(oc.row_ptr, oc.col, oc.N)
    dec = F.Ldpc(H, F.DecoderOptions().algorithm(F.Approximate).iterations(25))
    frames = 90
    fn = noalloc()
    x = torch.zeros(frames * oc.N, dtype=torch.float32, device="cuda")
    bits = torch.zeros(frames * oc.K, dtype=torch.uint8, device="cuda")
    iters = torch.zeros(frames, dtype=torch.int32, device="cuda")
    # first call outside capture: the array's address is looked up once per device
    run(fn, dec, x, frames, oc.K)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        rc = fn(dec._h, x.data_ptr(), frames, bits.data_ptr(), iters.data_ptr(), s.cuda_stream)
    assert rc == 0, lib().fecl_last_error()
    for rep in range(2):
        llr = dvbs2_llr(frames, 2.0 + 0.2 * rep, seed=700 + rep, rate=(1, 2), n=16200).astype(np.float32)
        x.copy_(torch.from_numpy(llr.ravel()))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        rbits, riters = reference(dec, x, frames, oc.K)
        torch.cuda.synchronize()
        assert torch.equal(bits, rbits) and torch.equal(iters, riters), rep
