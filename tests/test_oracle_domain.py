"""Bunch extent and domain (core_diagnostics.cc:139-153,422-462;
space_charge_3d_open_hockney.cc:94-109,493-540; rectangular_grid_domain.h:
20-36): the fixed summation tree restated independently in numpy, closeness
to the correctly rounded sum (math.fsum), exact closed forms on symmetric
fixtures, the negligible-extent error, and the derived scalars."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import BETA, GAMMA, PREF

MP = 0.938272046
E = 1.602176565e-19
C = 299792458.0


def tree_numpy(parts, masks):
    """Independent restatement of the device tree: chunks of 4096, lane t
    sums particle 4096c + 256j + t for j = 0..15 from +0.0, then pairwise
    over lanes (s[t] += s[t+w]), chunk partials combined the same way."""
    n = masks.shape[0]
    vals = np.stack([parts[0], parts[2], parts[4], parts[0] * parts[0], parts[2] * parts[2],
                     parts[4] * parts[4]])[:, :n]
    vals = np.where(masks.astype(bool)[None, :], vals, 0.0)

    def lanes_pairwise(s):  # s: (6, 256)
        s = s.copy()
        w = 1
        while w < 256:
            s[:, 0:256:2 * w] = s[:, 0:256:2 * w] + s[:, w:256:2 * w]
            w *= 2
        return s[:, 0]

    nch = (n + 4095) // 4096
    P = np.zeros((nch, 6))
    for c in range(nch):
        s = np.zeros((6, 256))
        for j in range(16):
            idx = c * 4096 + j * 256 + np.arange(256)
            ok = idx < n
            add = np.zeros((6, 256))
            add[:, ok] = vals[:, idx[ok]]
            s = s + add
        P[c] = lanes_pairwise(s)
    s = np.zeros((6, 256))
    for j in range(0, nch, 256):
        blk = np.zeros((6, 256))
        m = min(256, nch - j)
        blk[:, :m] = P[j:j + m].T
        s = s + blk
    return lanes_pairwise(s)


def test_tree_sums_restated_and_fsum():
    rng = np.random.default_rng(3)
    n = 9000  # 3 chunks, last partial
    parts = np.zeros((7, n + 5))
    parts[0, :n] = rng.standard_normal(n) * 1e-3 + 2e-4
    parts[2, :n] = rng.standard_normal(n) * 2e-3
    parts[4, :n] = rng.standard_normal(n) * 1e-2 - 0.3
    masks = (rng.random(n) > 0.05).astype(np.uint8)
    got = O.tree_sums(np.ascontiguousarray(parts), masks)
    ref = tree_numpy(parts, masks)
    assert np.array_equal(got, ref)
    sel = masks.astype(bool)
    for q, v in enumerate((parts[0, :n], parts[2, :n], parts[4, :n])):
        exact = math.fsum(v[sel])
        assert abs(got[q] - exact) <= 1e-13 * math.fsum(np.abs(v[sel]))
        exact2 = math.fsum((v * v)[sel])
        assert abs(got[q + 3] - exact2) <= 1e-14 * exact2


def make(parts, masks, total=None, real=1e10):
    n = masks.shape[0]
    return O.make_bunch(np.ascontiguousarray(parts), masks, total or n, real, 1, MP, GAMMA, BETA,
                        PREF)


def test_symmetric_fixture_closed_form():
    # +-a pairs: sums exactly 0, squares exactly 2 a^2 per pair, mean 0, std a
    a = (1.5e-3, 2.5e-3, 0.125)  # exactly representable squares not needed:
    n = 64
    parts = np.zeros((7, n))
    for k, ax in ((0, a[0]), (2, a[1]), (4, a[2])):
        parts[k, 0::2] = ax
        parts[k, 1::2] = -ax
    masks = np.ones(n, np.uint8)
    opts = O.make_options((32, 32, 64))
    rc, d = O.compute_domain(opts, make(parts, masks), 1e-9)
    assert rc == 0
    assert d["mean"] == [0.0, 0.0, 0.0]
    for q, ax in enumerate(a):
        assert d["stddev"][q] == math.sqrt((n * (ax * ax)) / n)
        assert d["size"][q] == 8.0 * d["stddev"][q]
        assert d["left"][q] == -d["size"][q] / 2.0
        assert d["h"][q] == d["size"][q] / (1.0 * (32, 32, 64)[q])


def test_smallest_non_tiny_and_errors():
    n = 10
    parts = np.zeros((7, n))
    parts[0, :] = np.linspace(-1, 1, n) * 1e-3  # only x has extent
    masks = np.ones(n, np.uint8)
    rc, d = O.compute_domain(O.make_options((32, 32, 64)), make(parts, masks), 1e-9)
    assert rc == 0
    sx = d["stddev"][0]
    # y and z have std 0 < tiny: their size falls back to max(others) = sx
    assert d["size"] == [8 * sx, 8 * sx, 8 * sx]
    parts[0, :] = 0.0
    rc, _ = O.compute_domain(O.make_options((32, 32, 64)), make(parts, masks), 1e-9)
    assert rc == -3
    assert "neglible extent" in O.last_error()
    # grid_entire_period: z offset 0, z size = z_period
    parts[0, :] = np.linspace(-1, 1, n) * 1e-3
    o = O.make_options((32, 32, 64), grid_entire_period=True, z_period=2.5)
    rc, d = O.compute_domain(o, make(parts, masks), 1e-9)
    assert d["offset"][2] == 0.0 and d["size"][2] == 2.5


def test_derived_scalars_closed_form():
    rng = np.random.default_rng(5)
    n = 1000
    parts = np.zeros((7, n))
    for k in (0, 2, 4):
        parts[k] = rng.standard_normal(n) * 1e-3
    masks = np.ones(n, np.uint8)
    dt = 3.3e-10
    rc, d = O.compute_domain(O.make_options((32, 32, 64)), make(parts, masks, real=5e10), dt)
    h = d["h"]
    eps0 = 1.0 / (C * C * (4 * math.pi * 1e-7))
    w0 = (5e10 / n) * 1 * E / (h[0] * h[1] * h[2])
    assert d["w0"] == pytest.approx(w0, rel=1e-15)
    fn = h[0] * h[1] * h[2] / (4 * math.pi * eps0) / (h[2] * h[2]) / (64 * 64 * 128)
    assert d["fn_norm"] == pytest.approx(fn, rel=1e-14)
    factor = (C / (1e9 * E)) * E * dt * fn / (PREF * GAMMA * GAMMA * BETA)
    assert d["factor"] == pytest.approx(factor, rel=1e-14)
    # pointlike drops the 1/hz^2 (space_charge_3d_open_hockney.cc:726-729)
    rc, d2 = O.compute_domain(O.make_options((32, 32, 64), green_fn="pointlike"),
                              make(parts, masks, real=5e10), dt)
    assert d2["fn_norm"] == pytest.approx(fn * h[2] * h[2], rel=1e-14)
