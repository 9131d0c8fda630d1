"""Pins of the oracle's kick branches and of its moment diagnostics.

Kick (space_charge_3d_kernels.h:91-215, factor at
space_charge_3d_open_hockney.cc:752-770): probes in a spherical Gaussian
charge, whose field is known in closed form,
    E_r(r) = Q/(4 pi eps0) [erf(r/(sqrt2 s))/r^2 - sqrt(2/pi) e^{-r^2/2s^2}/(s r)]
(the potential Q erf(r/(sqrt2 s))/(4 pi eps0 r) of
docs/devel/solvers/3d_open_hockney.py:10-22, differentiated).  A probe on
each axis must receive only its own kick: dp/p on the cdt axis (energy kick
through the energy-momentum relation, negative behind the centre), y' on the
y axis, x' on the x axis, and nothing else.  The transverse scale of the kick
is pinned independently by the rod of charge (below, on the y axis; the x
axis case is in test_oracle_deposit_field.py), after
test_space_charge_3d_open_hockney_mpi.cc:99-107.

Moments (core_diagnostics.cc:202-229,318-362): np.mean and the centred
np.cov over the active particles, an exact Hadamard-pattern fixture, and an
independent numpy restatement of the fixed summation tree (bit-equal)."""
import math

import numpy as np
import pytest
from scipy.special import erf

import oracle as O

MP = 0.938272046
E_CHARGE = 1.602176565e-19
C_LIGHT = 299792458.0
EPS0 = 1.0 / (C_LIGHT ** 2 * 4 * math.pi * 1e-7)


def sphere_field(r, sigma, Q):
    """Radial field of a spherical Gaussian charge Q of rms sigma."""
    a = math.sqrt(2.0) * sigma
    return Q / (4 * math.pi * EPS0) * (erf(r / a) / r ** 2 -
                                       math.sqrt(2 / math.pi) * math.exp(-r * r / (2 * sigma ** 2))
                                       / (sigma * r))


@pytest.fixture(scope="module")
def sphere_kick():
    """One kick of a Gaussian sphere (sigma equal in x, y and cdt, the
    solver's three axes) on a fixed +-6 sigma domain of 64^3 nodes, with
    probes at 2.5 and 3 sigma on each axis (their own charge is 1 part in
    4e5 of the bunch)."""
    gamma = 2.0
    beta = math.sqrt(1 - 1 / gamma ** 2)
    pref = gamma * beta * MP
    sigma, n, real = 1e-3, 400000, 1e10
    radii = (2.5, 3.0)
    probes = []
    for r0 in radii:
        for col in (4, 0, 2):  # cdt, x, y
            probes.append((r0, col))
    npr = len(probes)
    rng = np.random.default_rng(7)
    p = np.zeros((7, n + npr))
    for j in (0, 2, 4):
        p[j, npr:] = sigma * rng.standard_normal(n)
    for k, (r0, col) in enumerate(probes):
        p[col, k] = r0 * sigma
    p[6] = np.arange(n + npr)
    before = p.copy()
    b = O.make_bunch(p, np.ones(n + npr, np.uint8), n + npr, real, 1, MP, gamma, beta, pref)
    size = 12 * sigma
    o = O.make_options((64, 64, 64), fixed=((0, 0, 0), (size, size, size)))
    dt = 0.1 / (beta * C_LIGHT)
    rc, _ = O.apply(o, b, dt, nthreads=4)
    assert rc == 0
    return dict(p=p, before=before, probes=probes, gamma=gamma, beta=beta, pref=pref,
                sigma=sigma, Q=real * E_CHARGE, dt=dt)


def transverse_kick(s, r0):
    """x' (or y') impulse of a probe at r0 sigma on that axis: the field
    times q dt, in GeV/c (c / (1e9 e)), over p_ref, reduced by gamma^2
    (electric minus magnetic force) and by beta (the density is deposited
    per unit cdt, beta times lower than per unit length).  The rod case pins
    this scale against 2 N r_p / (L beta^2 gamma^3) D / r."""
    Er = sphere_field(r0 * s["sigma"], s["sigma"], s["Q"])
    return (C_LIGHT / (1e9 * E_CHARGE)) * E_CHARGE * s["dt"] * Er / (
        s["pref"] * s["gamma"] ** 2 * s["beta"])


@pytest.mark.parametrize("r0", (2.5, 3.0))
def test_kick_dpop_on_cdt_axis(sphere_kick, r0):
    s = sphere_kick
    k = s["probes"].index((r0, 4))
    kick = transverse_kick(s, r0)
    # the same impulse taken as an energy change -p_ref * kick (the field
    # points away from the centre, so a particle behind it, cdt > 0, loses
    # energy), turned into momentum through E^2 = p^2 + m^2
    p_i = s["pref"] * (1 + s["before"][5, k])
    e_i = math.sqrt(p_i ** 2 + MP ** 2)
    e_f = e_i - s["pref"] * kick
    want = (math.sqrt(e_f ** 2 - MP ** 2) - p_i) / s["pref"]
    got = s["p"][5, k] - s["before"][5, k]
    assert want < 0 and got < 0
    assert got == pytest.approx(want, rel=0.01)
    # linearised: dp = dE / beta_particle (independent of the exact relation)
    assert got == pytest.approx(-kick * e_i / p_i, rel=0.011)
    # no transverse kick on the axis (noise of the sampled bunch only)
    assert abs(s["p"][1, k]) < 0.02 * kick and abs(s["p"][3, k]) < 0.02 * kick


@pytest.mark.parametrize("r0", (2.5, 3.0))
@pytest.mark.parametrize("col,other", ((0, 3), (2, 1)))
def test_kick_transverse_on_axis(sphere_kick, r0, col, other):
    """x' for the probe on the x axis and y' for the one on the y axis,
    none of the other transverse component and no dp/p."""
    s = sphere_kick
    k = s["probes"].index((r0, col))
    kick = transverse_kick(s, r0)
    own = col + 1
    assert s["p"][own, k] == pytest.approx(kick, rel=0.01)
    assert abs(s["p"][other, k]) < 0.02 * kick
    kz = s["probes"].index((r0, 4))
    dpop_axis = abs(s["p"][5, kz] - s["before"][5, kz])
    assert abs(s["p"][5, k] - s["before"][5, k]) < 0.02 * dpop_axis


def test_rod_of_charge_y_axis():
    """Rod fixture (rod_bunch.h:4-95) with the probe on the y axis at y0:
    y' = 2 N r_p / (L beta^2 gamma^3) * D / y0 within 3 % and x' ~ 0
    (test_space_charge_3d_open_hockney_mpi.cc:99-107 places it on x)."""
    for gamma in (61.0 / 60.0, 61.0 / 11.0):
        beta = math.sqrt(1 - 1 / gamma ** 2)
        pref = gamma * beta * MP
        nl = 4001
        nrod = 1 + 8 * nl
        L, rr, y0 = 0.1, 1e-6, 80e-6
        p = np.zeros((7, nrod))
        zs = -L / 2 + np.arange(nl) * (L / (nl - 1))
        r2o2 = math.sqrt(2) / 2
        ring = [(rr, 0), (rr * r2o2, rr * r2o2), (0, rr), (-rr * r2o2, rr * r2o2), (-rr, 0),
                (-rr * r2o2, -rr * r2o2), (0, -rr), (rr * r2o2, -rr * r2o2)]
        for k, (xx, yy) in enumerate(ring):
            p[0, 1 + k::8] = xx
            p[2, 1 + k::8] = yy
            p[4, 1 + k::8] = zs / beta
        p[2, 0] = y0
        p[6] = np.arange(nrod)
        N = 5e9
        b = O.make_bunch(p, np.ones(nrod, np.uint8), nrod, N, 1, MP, gamma, beta, pref)
        o = O.make_options((64, 64, 32), fixed=((0, 0, 0), (4 * y0, 4 * y0, L / beta)))
        step = 0.1
        rc, _ = O.apply(o, b, step / (beta * C_LIGHT), nthreads=4)
        assert rc == 0
        rp = E_CHARGE / (4 * math.pi * EPS0 * MP * 1e9)
        expect = 2 * N * rp / (L * beta ** 2 * gamma ** 3) * step / y0
        assert p[3, 0] == pytest.approx(expect, rel=0.03), (gamma, p[3, 0], expect)
        assert abs(p[1, 0]) < 1e-3 * expect
        assert abs(p[5, 0]) < 1e-3 * expect


# ---------------------------------------------------------------------------
# moments
# ---------------------------------------------------------------------------


def tree_numpy_rows(vals, masks):
    """Independent restatement of the fixed summation tree on K rows: chunks
    of 4096, lane t sums item 4096c + 256j + t for j = 0..15 from +0.0, then
    pairwise over lanes (s[t] += s[t+w]); chunk partials are summed per lane
    (chunk c to lane c mod 256, in chunk order) and combined pairwise."""
    K = vals.shape[0]
    n = masks.shape[0]
    vals = np.where(masks.astype(bool)[None, :], vals[:, :n], 0.0)

    def lanes_pairwise(s):
        s = s.copy()
        w = 1
        while w < 256:
            s[:, 0:256:2 * w] = s[:, 0:256:2 * w] + s[:, w:256:2 * w]
            w *= 2
        return s[:, 0]

    nch = (n + 4095) // 4096
    P = np.zeros((nch, K))
    for c in range(nch):
        s = np.zeros((K, 256))
        for j in range(16):
            idx = c * 4096 + j * 256 + np.arange(256)
            ok = idx < n
            add = np.zeros((K, 256))
            add[:, ok] = vals[:, idx[ok]]
            s = s + add
        P[c] = lanes_pairwise(s)
    s = np.zeros((K, 256))
    for j in range(0, nch, 256):
        blk = np.zeros((K, 256))
        m = min(256, nch - j)
        blk[:, :m] = P[j:j + m].T
        s = s + blk
    return lanes_pairwise(s)


def correlated_bunch(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((6, 6)) * np.array([1e-3, 1e-4, 2e-3, 1e-4, 1e-2, 1e-4])[:, None]
    p = np.zeros((7, n + 7))
    p[:6, :n] = A @ rng.standard_normal((6, n))
    p[:6, :n] += np.array([3e-4, -2e-5, 1e-3, 5e-6, -0.2, 1e-4])[:, None]
    p[6] = np.arange(n + 7)
    masks = (rng.random(n) > 0.07).astype(np.uint8)
    return p, masks


def test_moments_match_numpy_mean_and_cov():
    n = 9000  # 3 tree chunks, the last one partial
    p, masks = correlated_bunch(n, 11)
    sel = masks.astype(bool)
    na = int(sel.sum())
    mean, mom2 = O.moments(p, masks, total_num=na)
    x = p[:6, :n][:, sel]
    want_mean = x.mean(axis=1)
    want_cov = np.cov(x, bias=True)
    scale = np.abs(x).max(axis=1)
    assert np.all(np.abs(mean - want_mean) <= 1e-13 * scale)
    sd = np.sqrt(np.diag(want_cov))
    assert np.all(np.abs(mom2 - want_cov) <= 1e-12 * np.outer(sd, sd))
    assert np.array_equal(mom2, mom2.T)
    # divided by total_num, not by the active count
    mean2, _ = O.moments(p, masks, total_num=2 * na)
    assert np.array_equal(mean2, mean / 2)


def test_moments_bit_equal_to_restated_tree():
    n = 9000
    p, masks = correlated_bunch(n, 12)
    total = float(masks.sum())
    mean, mom2 = O.moments(p, masks, total_num=total)
    s1 = tree_numpy_rows(p[:6], masks)
    want_mean = s1 / total
    assert np.array_equal(mean, want_mean)
    rows, pairs = [], []
    for j in range(6):
        for k in range(j + 1):
            rows.append((p[j] - want_mean[j]) * (p[k] - want_mean[k]))
            pairs.append((j, k))
    s2 = tree_numpy_rows(np.stack(rows), masks)
    for q, (j, k) in enumerate(pairs):
        assert mom2[j, k] == s2[q] / total and mom2[k, j] == s2[q] / total


def test_moments_exact_hadamard_fixture():
    """Columns 1..6 of an 8x8 Hadamard matrix scaled by powers of two, plus
    dyadic means: each column sums to 0 and distinct columns are orthogonal,
    so mean and mom2 = diag(a^2) are exact."""
    H = np.array([[1]])
    for _ in range(3):
        H = np.block([[H, H], [H, -H]])
    a = 2.0 ** np.array([-10, -13, -9, -14, -6, -12])
    mu = 2.0 ** np.array([-4, -20, 3, -22, -1, -16])
    reps = 600  # 4800 particles: two tree chunks
    p = np.zeros((7, 8 * reps))
    for j in range(6):
        p[j] = np.tile(mu[j] + a[j] * H[:, j + 1], reps)
    p[6] = np.arange(8 * reps)
    masks = np.ones(8 * reps, np.uint8)
    mean, mom2 = O.moments(p, masks)
    assert np.array_equal(mean, mu)
    assert np.array_equal(mom2, np.diag(a ** 2))
    # a negated copy of an active particle masked out changes nothing
    p2 = np.concatenate([p, -p[:, :1]], axis=1)
    m2 = np.concatenate([masks, np.zeros(1, np.uint8)])
    mean2, mom22 = O.moments(p2, m2, total_num=8 * reps)
    assert np.array_equal(mean2, mu) and np.array_equal(mom22, np.diag(a ** 2))
