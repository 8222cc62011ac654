"""Pins the oracle where the golden vectors leave room: closed forms of the check node on
single-row codes (brute force over random and special inputs), the 16-bit quantizers against
numpy / torch conversions, and FeCl's own invariants (test/Cpp/operations.h)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def single_row(d):
    """One check of degree d: with iterations = 1 the state output is CN(L) itself."""
    return O.Code(np.array([0, d], np.int32), np.arange(d, dtype=np.int32), d, d - 1)


def cn(x, alg, prec=64, sf=1.0):
    code = single_row(len(x))
    out = O.decode(code, np.asarray(x, np.float64)[None, :], alg, 1, sf=O.scaling_table(code, sf),
                   prec=prec, want=("state",))
    return out["state"][0]


SPECIAL = [0.0, -0.0, 1e-30, -1e-30, 0.5, -0.5, 3.0, -3.0, 20.0, -20.0]


@pytest.mark.parametrize("d", [2, 3, 4, 7, 11])
def test_minsum_check_is_sign_times_min_of_others(d):
    rng = np.random.default_rng(d)
    for trial in range(200):
        x = rng.normal(0, 3, d)
        if trial % 3 == 0:
            x[rng.integers(0, d, 2)] = rng.choice(SPECIAL, 2)
        if trial % 5 == 0:
            x[1] = x[0]  # ties
        y = cn(x, O.APPROXIMATE)
        for j in range(d):
            others = np.delete(x, j)
            mag = np.min(np.abs(others))
            # FeCl convention (L = ln p1/p0): box-sum of two has sign bit 1 ^ s_a ^ s_b, so a
            # fold over k operands has sign bit (k-1) ^ xor(s_i)
            sbit = (len(others) - 1) % 2 ^ (int(np.sum(np.signbit(others))) % 2)
            assert abs(y[j]) == mag
            assert bool(np.signbit(y[j])) == bool(sbit), (x, j, y)


@pytest.mark.parametrize("d", [3, 5, 7])
def test_exact_check_matches_high_precision_boxplus(d):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    rng = np.random.default_rng(100 + d)
    for _ in range(60):
        x = rng.normal(0, 4, d)
        y = cn(x, O.EXACT)
        for j in range(d):
            p = mp.mpf(1)
            for i in range(d):
                if i != j:
                    p *= mp.tanh(-mp.mpf(x[i]) / 2)
            ref = float(-2 * mp.atanh(p))
            assert abs(y[j] - ref) <= 1e-9 * max(1.0, abs(ref)), (x, j, y[j], ref)


@pytest.mark.parametrize("d", [3, 5, 7])
def test_exact_stable_check_matches_high_precision_boxplus(d):
    """The LLR-domain evaluation of BoxSum (the exact value FeCl's tanh form approximates)."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    rng = np.random.default_rng(200 + d)
    for _ in range(60):
        x = rng.normal(0, 8, d)
        y = cn(x, O.EXACT_STABLE)
        for j in range(d):
            p = mp.mpf(1)
            for i in range(d):
                if i != j:
                    p *= mp.tanh(-mp.mpf(x[i]) / 2)
            ref = float(-2 * mp.atanh(p))
            assert abs(y[j] - ref) <= 1e-12 * max(1.0, abs(ref)), (x, j, y[j], ref)


def test_linear_pairwise_is_table_form():
    """Degree 3 row: out_0 = x_2 # x_1 with # = LinearBoxSum; f is the 8-knot interpolant of
    log(1+e^-x) at x = i/2, zero from x = 3.5 (src/detail/LinearTable.h:57-81)."""
    y = [math.log(1 + math.exp(-i / 2)) for i in range(8)]

    def f(z):
        z *= 2
        if z >= 7:
            return 0.0
        i = int(z)
        return (y[i + 1] - y[i]) * (z - i) + y[i]

    for zz in np.linspace(0, 5, 101):
        assert abs(f(zz) - math.log1p(math.exp(-zz))) < 0.02 or zz >= 3.5
    rng = np.random.default_rng(5)
    for _ in range(300):
        a, b = rng.normal(0, 2, 2)
        out = cn([0.7, b, a], O.LINEAR)
        m = min(abs(a), abs(b))
        ref = (m if np.signbit(a) != np.signbit(b) else -m) - f(abs(a + b)) + f(abs(a - b))
        assert out[0] == ref


def test_scaling_factor_applies_per_iteration_row():
    x = np.array([1.0, -2.0, 3.0, -4.0])
    assert np.array_equal(cn(x, O.APPROXIMATE, sf=0.5), 0.5 * cn(x, O.APPROXIMATE))


def test_quantize_f16_matches_numpy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.normal(0, 10, 20000), rng.normal(0, 1e-5, 5000), rng.normal(0, 3e4, 5000),
                         [65504, 65519.99, 65520, 1e9, -1e9, 0.0, -0.0, 2 ** -24, 2 ** -25, 3 * 2 ** -26]])
    xs = xs.astype(np.float32)
    for x in xs:
        ref = np.float32(np.float16(np.clip(x, -65504, 65504)))
        got = np.float32(O.quantize(float(x), O.MSG_F16))
        assert ref.view(np.uint32) == got.view(np.uint32), (x, ref, got)


def test_quantize_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.normal(0, 10, 20000), rng.normal(0, 1e-38, 2000), [3.4e38, -3.4e38, 0.0, -0.0]])
    xs = xs.astype(np.float32)
    bmax = np.float32(3.38953138925153547590e38)
    ref = torch.from_numpy(np.clip(xs, -bmax, bmax)).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([O.quantize(float(x), O.MSG_BF16) for x in xs], np.float32)
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))


@pytest.fixture(scope="module")
def short_code():
    return O.dvbs2(16200, 1, 2)


def _llr(code, ebn0, frames, seed):
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (frames, code.K), dtype=np.uint8)
    cw = O.encode(code, info)
    sigma2 = 1 / (2 * code.K / code.N * 10 ** (ebn0 / 10))
    y = 1 - 2.0 * cw + rng.normal(0, math.sqrt(sigma2), cw.shape)
    return -2 * y / sigma2, info, cw


@pytest.mark.parametrize("alg", [O.EXACT, O.LINEAR, O.APPROXIMATE])
def test_fecl_invariants(short_code, alg):
    c = short_code
    llr, info, cw = _llr(c, 3.0, 3, 1)
    full = O.decode(c, llr, alg, 30)
    # recovery at high SNR; converged hard decisions are codewords
    assert np.array_equal(full["hard"], info)
    # systOut == parity-extrinsic prefix (test/Cpp/Ldpc.cpp:32-49)
    assert np.array_equal(full["syst"], full["parity"][:, :c.K])
    # soDecode a-posteriori signs == decode hard bits here
    assert np.array_equal((full["msg"] >= 0).astype(np.uint8), full["hard"])
    # extrinsic = sum of final c2v per variable, ascending edges
    ext = np.zeros((3, c.N))
    for e in range(c.E):
        ext[:, c.col[e]] += full["state"][:, e]
    assert np.array_equal(ext, full["parity"])
    # zero a-priori syst / state change nothing (test_soDecode_0systIn / 0stateIn)
    z = O.decode(c, llr, alg, 30, syst_in=np.zeros((3, c.K)))
    assert np.array_equal(z["msg"], full["msg"])


@pytest.mark.parametrize("alg", [O.EXACT, O.LINEAR, O.APPROXIMATE])
def test_two_phase_resume(short_code, alg):
    """test_soDecode_2phases (test/Cpp/operations.h:344-383) at -5 dB (nothing converges)."""
    c = short_code
    llr, _, _ = _llr(c, -5.0, 2, 2)
    a = O.decode(c, llr, alg, 10)
    b = O.decode(c, llr, alg, 10, state_in=a["state"])
    one = O.decode(c, llr, alg, 20)
    assert np.array_equal(b["msg"], one["msg"]) and np.array_equal(b["state"], one["state"])


def test_iteration_count_semantics(short_code):
    """Counts are check-node updates: t_break + 2, or the budget; I = 1 gives 1."""
    c = short_code
    llr, _, _ = _llr(c, 3.0, 2, 3)
    assert O.decode(c, llr, O.APPROXIMATE, 1)["iters"].tolist() == [1, 1]
    it = O.decode(c, llr, O.APPROXIMATE, 50)["iters"]
    assert np.all(it >= 2) and np.all(it < 50)
    hard_llr, _, _ = _llr(c, -3.0, 2, 4)
    assert O.decode(c, hard_llr, O.APPROXIMATE, 7)["iters"].tolist() == [7, 7]


def test_rejects_bad_arguments():
    c = single_row(3)
    with pytest.raises(ValueError):
        O.decode(c, np.zeros((1, 3)), O.APPROXIMATE, 0)
    bad = O.Code(np.array([0, 1], np.int32), np.array([0], np.int32), 2, 1)
    with pytest.raises(ValueError):
        O.decode(bad, np.zeros((1, 2)), O.APPROXIMATE, 5)


def _fnv_sides(sides):
    h = 0xcbf29ce484222325
    for s in sides:
        h = ((h ^ int(s)) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_linear_cutoff_side_hash():
    """The Linear diagnostic (lookup-side hash): on a degree-3 row with one iteration, the check
    update makes six table lookups in FeCl's order, at |x0 + x1|, |x0 - x1| (forward), then
    |x0 + x2|, |x0 - x2| and |x2 + x1|, |x2 - x1| (backward); each lookup records whether
    2|a +- b| >= 7.  The hash equals FNV-1a over those six bits computed here, in both
    precisions; an input that is not FP32-representable and sits just below the cutoff in FP64
    (x1 = 2.5 - 2^-40, so x0 + x1 = 3.5 - 2^-40) rounds onto it in FP32 (2.5), which flips the
    first lookup's side and so the hash."""
    code = single_row(3)

    def side_hash(x, prec):
        out = O.decode(code, np.asarray(x, np.float64)[None, :], O.LINEAR, 1, prec=prec, want=("lin_side",))
        return int(out["lin_side"][0])

    def expected(x0, x1, x2):
        pairs = [(x0, x1), (x0, x2), (x2, x1)]
        sides = [2.0 * abs(s) >= 7.0 for a, b in pairs for s in (a + b, a - b)]
        return _fnv_sides(sides)

    for x in ([1.0, 2.5, -0.5], [1.0, 2.0, 4.0], [-3.0, 0.25, 6.5], [0.1, 0.2, 0.3]):
        assert side_hash(x, 64) == side_hash(x, 32) == expected(*x)
    x = [1.0, 2.5 - 2.0 ** -40, -0.5]
    assert side_hash(x, 64) == expected(*x)
    assert side_hash(x, 32) == expected(1.0, 2.5, -0.5) != side_hash(x, 64)
    # other algorithms make no lookups: the hash stays at FNV's offset basis
    out = O.decode(code, np.asarray([1.0, 2.0, 3.0])[None, :], O.APPROXIMATE, 1, want=("lin_side",))
    assert int(out["lin_side"][0]) == 0xcbf29ce484222325
