"""Pins of the oracle's random streams (numba xoroshiro128+, numba/cuda/random.py).

No reference test checks random values, so the generator is pinned by what
the mathematics fixes:
 * SplitMix64 against its published outputs (seed 0 and seed 1234567);
 * the state transition T (taken from the oracle) has full period 2^128 - 1,
   which a wrong shift/rotation constant breaks;
 * the oracle's jump equals T^(2^64) computed by repeated squaring;
 * output = s0 + s1 and the float32 conversion of numba.
numba's own seeding choice (SplitMix64 output in both words, stream i =
stream i-1 jumped) is restated from numba 0.53.1's source and is otherwise
unpinned.
"""
import numpy as np
import pytest

import pyoracle as orc

MASK = (1 << 64) - 1


def state_to_bits(s0, s1):
    v = np.zeros(128, dtype=np.uint8)
    for k in range(64):
        v[k] = (s0 >> k) & 1
        v[64 + k] = (s1 >> k) & 1
    return v


def bits_to_state(v):
    s0 = sum(int(v[k]) << k for k in range(64))
    s1 = sum(int(v[64 + k]) << k for k in range(64))
    return s0, s1


def basis(k):
    return (1 << k, 0) if k < 64 else (0, 1 << (k - 64))


def matrix_of(fn):
    """128x128 GF(2) matrix whose column k is fn(e_k)."""
    M = np.zeros((128, 128), dtype=np.uint8)
    for k in range(128):
        s0, s1 = fn(basis(k))
        M[:, k] = state_to_bits(s0, s1)
    return M


def gf2_mul(A, B):
    return (A.astype(np.int64) @ B.astype(np.int64) % 2).astype(np.uint8)


def gf2_pow(M, e):
    R = np.eye(128, dtype=np.uint8)
    P = M.copy()
    while e:
        if e & 1:
            R = gf2_mul(R, P)
        P = gf2_mul(P, P)
        e >>= 1
    return R


def is_prime(n):
    """Deterministic Miller-Rabin for n < 3.3e24."""
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    for p in small:
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def transition(st):
    _, _, new = orc.rng_draws(np.array(st, dtype=np.uint64), 1)
    return int(new[0]), int(new[1])


@pytest.fixture(scope="module")
def T():
    return matrix_of(transition)


def test_splitmix64_published_outputs():
    # first outputs of SplitMix64 (Steele, Lea, Flood 2014) for seeds 0 and 1234567
    assert orc.splitmix64(0) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(1234567) == 6457827717110365317


def test_full_period(T):
    I = np.eye(128, dtype=np.uint8)
    order = (1 << 128) - 1
    assert np.array_equal(gf2_pow(T, order), I)
    # prime factors of 2^128 - 1
    # (2^32-1)(2^32+1)(2^64+1)
    primes = [3, 5, 17, 257, 65537, 641, 6700417, 274177, 67280421310721]
    prod = 1
    for p in primes:
        assert is_prime(p)
        prod *= p
    assert prod == order
    for p in primes:
        assert not np.array_equal(gf2_pow(T, order // p), I), p


def test_jump_is_2_pow_64_steps(T):
    J = matrix_of(lambda st: tuple(int(v) for v in orc.rng_jump(np.array(st, dtype=np.uint64))))
    assert np.array_equal(J, gf2_pow(T, 1 << 64))


def test_output_and_float32_conversion():
    st = np.array([0x0123456789ABCDEF, 0xFEDCBA9876543210], dtype=np.uint64)
    raw, f, _ = orc.rng_draws(st, 1)
    assert int(raw[0]) == (0x0123456789ABCDEF + 0xFEDCBA9876543210) & MASK
    raw, f, _ = orc.rng_draws(st, 1000)
    expect = ((raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53).astype(np.float32)
    assert np.array_equal(f, expect)
    assert f.min() >= 0 and f.max() <= 1


def test_stream_seeding_is_jump_chain():
    s = orc.rng_states(1, 8)
    z = orc.splitmix64(1)
    assert int(s[0, 0]) == z and int(s[0, 1]) == z
    for i in range(1, 8):
        assert np.array_equal(orc.rng_jump(s[i - 1]), s[i])


def test_uniformity_of_draws():
    s = orc.rng_states(7, 1)[0]
    _, f, _ = orc.rng_draws(s, 200000)
    counts, _ = np.histogram(f, bins=20, range=(0, 1))
    chi2 = ((counts - 10000.0) ** 2 / 10000.0).sum()
    assert chi2 < 60  # 19 dof
