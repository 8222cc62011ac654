"""Seeded input vectors shared by the golden-fixture generator and the tests.

Channel: BPSK with bit 0 -> +1, y = s + N(0, sigma^2), sigma^2 = 1 / (2 R 10^(EbN0/10)), and
FeCl-convention LLR L = ln p(1|y)/p(0|y) = -2 y / sigma^2 (src/Codec.h:204-205).
"""
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def awgn_llr(cw, rate, ebn0, rng):
    sigma2 = 1.0 / (2.0 * rate * 10.0 ** (ebn0 / 10.0))
    y = (1.0 - 2.0 * cw.astype(np.float64)) + rng.normal(0.0, np.sqrt(sigma2), cw.shape)
    return -2.0 * y / sigma2


def gallager_llr(n, k, frames, ebn0, seed):
    """All-zero codeword (a codeword of every linear code) through the channel."""
    rng = np.random.default_rng(seed)
    return awgn_llr(np.zeros((frames, n), np.uint8), k / n, ebn0, rng)


def dvbs2_llr(frames, ebn0, seed, rate=(1, 2), n=64800, return_info=False):
    from oracle import oracle as O
    code = O.dvbs2(n, *rate)
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (frames, code.K), dtype=np.uint8)
    cw = O.encode(code, info)
    llr = awgn_llr(cw, code.K / code.N, ebn0, rng)
    return (llr, info) if return_info else llr
