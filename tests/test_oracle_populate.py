"""Pins of the oracle's Gaussian populate (the counter-based stand-in for
populate_6d, populate.cc:128-145, which draws with GSL's ranlxd2):

- Philox-4x32-10 against the known-answer vectors published with Random123
  (Salmon et al., SC'11; kat_vectors, philox4x32_10) and against an
  independent restatement in Python integers;
- the basic-operation log and sin/cos(2 pi u) against 200-bit arithmetic
  (mpmath), within 1 and 2 ulp;
- whole bunches: Box-Muller assembled from those pieces (bit-equal), and the
  sample covariance against L L^T."""
import math

import numpy as np
import pytest

import oracle as O

mpmath = pytest.importorskip("mpmath")

KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def philox_py(ctr, key):
    """Philox-4x32-10 in Python integers (Salmon et al., SC'11, Fig. 2)."""
    M = 0xFFFFFFFF
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0 = 0xD2511F53 * c0
        p1 = 0xCD9E8D57 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & M, p1 & M, ((p0 >> 32) ^ c3 ^ k1) & M, p0 & M
        k0 = (k0 + 0x9E3779B9) & M
        k1 = (k1 + 0xBB67AE85) & M
    return (c0, c1, c2, c3)


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(O.philox4x32_10(ctr, key)) == want
    assert philox_py(ctr, key) == want


def test_philox_matches_restatement():
    rng = np.random.default_rng(5)
    for _ in range(300):
        ctr = tuple(int(x) for x in rng.integers(0, 2 ** 32, 4))
        key = tuple(int(x) for x in rng.integers(0, 2 ** 32, 2))
        assert tuple(O.philox4x32_10(ctr, key)) == philox_py(ctr, key)


def u53(a, b):
    return ((((a << 21) | (b >> 11))) + 0.5) * 2.0 ** -53


def sample_uniforms(n, seed=9):
    rng = np.random.default_rng(seed)
    us = [u53(int(a), int(b)) for a, b in rng.integers(0, 2 ** 32, (n, 2))]
    # the ends of the range, both sides of sqrt(1/2) and of the quadrants
    us += [u53(0, 0), u53(0, 2048), u53(0xffffffff, 0xfffff000), u53(0xffffffff, 0xffffe000)]
    us += [math.sqrt(0.5) * (1 + d) for d in (-2e-16, 0.0, 2e-16)]
    us += [0.25 + d for d in (-2 ** -54, 2 ** -54)] + [0.5 - 2 ** -54, 0.75 + 2 ** -53]
    return us


def ulps(got, exact):
    e = float(exact)
    return float(abs(mpmath.mpf(got) - exact)) / math.ulp(abs(e)) if e != 0 else abs(got)


def test_det_log_within_one_ulp():
    mpmath.mp.prec = 200
    worst = 0.0
    for u in sample_uniforms(20000):
        worst = max(worst, ulps(O.det_log(u), mpmath.log(mpmath.mpf(u))))
    assert O.det_log(1.0) == 0.0
    assert worst <= 1.0, worst


def test_det_sincos2pi_within_two_ulp():
    mpmath.mp.prec = 200
    ws = wc = 0.0
    for u in sample_uniforms(20000):
        s, c = O.det_sincos2pi(u)
        a = 2 * mpmath.pi * mpmath.mpf(u)
        ws = max(ws, ulps(s, mpmath.sin(a)))
        wc = max(wc, ulps(c, mpmath.cos(a)))
    # exact values at the quadrant boundaries (u = 1 is reachable: the
    # top uniform rounds up to 1.0)
    assert O.det_sincos2pi(0.25) == (1.0, 0.0) and O.det_sincos2pi(1.0) == (0.0, 1.0)
    assert O.det_sincos2pi(0.5)[1] == -1.0 and O.det_sincos2pi(0.75)[0] == -1.0
    assert ws <= 2.0 and wc <= 2.0, (ws, wc)


def identity_chol():
    return np.array([1.0 if r == c else 0.0 for r in range(6) for c in range(r + 1)])


def test_populate_is_box_muller_of_philox():
    seed = (0x1234567 << 32) | 0x89abcdef
    n, id0 = 300, (1 << 33) + 17  # ids above 2^32 exercise the counter's high word
    parts, masks = O.populate_gaussian(n, np.zeros(6), identity_chol(), seed, id0=id0)
    assert np.all(masks == 1)
    assert np.array_equal(parts[6], id0 + np.arange(n, dtype=np.float64))
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for i in range(n):
        pid = id0 + i
        u = []
        for call in range(3):
            r = philox_py((pid & 0xFFFFFFFF, pid >> 32, call, 0), key)
            u += [u53(r[0], r[1]), u53(r[2], r[3])]
        for k in range(3):
            rad = math.sqrt(-2.0 * O.det_log(u[2 * k]))
            s, c = O.det_sincos2pi(u[2 * k + 1])
            assert parts[2 * k, i] == rad * c and parts[2 * k + 1, i] == rad * s


def test_populate_covariance():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((6, 6)) * np.array([1e-3, 1e-4, 2e-3, 1e-4, 1e-2, 1e-4])[:, None]
    cov = A @ A.T
    L = np.linalg.cholesky(cov)
    chol = np.array([L[r, c] for r in range(6) for c in range(r + 1)])
    mean = np.array([1e-4, 0, -2e-4, 0, 3e-3, 1e-5])
    n = 400000
    parts, _ = O.populate_gaussian(n, mean, chol, seed=77)
    x = parts[:6]
    sd = np.sqrt(np.diag(cov))
    assert np.all(np.abs(x.mean(axis=1) - mean) < 5 * sd / math.sqrt(n))
    got = np.cov(x, bias=True)
    assert np.all(np.abs(got - cov) < 0.015 * np.outer(sd, sd))
    # standard-normal tails of the underlying draws (Box-Muller, not a
    # uniform or a truncated generator): kurtosis 3, |g| > 4 about 6.3e-5
    g = np.linalg.solve(L, x - mean[:, None])
    k4 = np.mean(g ** 4, axis=1)
    assert np.all(np.abs(k4 - 3) < 0.05)
    assert 20 < np.sum(np.abs(g) > 4) < 300
