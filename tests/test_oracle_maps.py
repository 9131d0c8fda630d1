"""libFF maps restated in the oracle (drift_unit, ff_algorithm.h:45-79;
thin_quadrupole_unit :928-936; yoshida6 :1835-2008): closed forms, the
thick-lens linear limit, symplecticity."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import BETA, GAMMA, PREF

MP = 0.938272046


def run(parts, spec):
    n = parts.shape[1]
    p = np.ascontiguousarray(parts.copy())
    b = O.make_bunch(p, np.ones(n, np.uint8), n, 1e10, 1, MP, GAMMA, BETA, PREF)
    O.apply_maps(b, [(k, L, k1, 4, O.reference_cdt(k, L, k1, 4, PREF, MP)) for k, L, k1 in spec])
    return p


def test_drift_closed_form():
    n = 5
    p = np.zeros((7, n))
    p[0] = np.linspace(-1e-3, 1e-3, n)
    p[1] = np.linspace(-2e-4, 2e-4, n)
    p[5] = np.linspace(-1e-3, 1e-3, n)
    L = 0.7
    q = run(p, [(0, L, 0.0)])
    for i in range(n):
        dp = 1 + p[5, i]
        pz = math.sqrt(dp * dp - p[1, i] ** 2)
        assert q[0, i] == pytest.approx(p[0, i] + L * p[1, i] / pz, rel=1e-15, abs=1e-18)
        # cdt - reference cdt = path length / beta - L / beta_ref
        P = dp * PREF
        beta = P / math.sqrt(P * P + MP * MP)
        path = L * dp / pz
        assert q[4, i] == pytest.approx(path / beta - L / BETA, rel=1e-9, abs=1e-15)
    assert O.reference_cdt(0, L, 0.0, 4, PREF, MP) == pytest.approx(L / BETA, rel=1e-15)


def test_quadrupole_linear_limit_and_symplectic():
    L, k1 = 0.2, 0.6
    w = math.sqrt(k1)
    Mx = np.array([[math.cos(w * L), math.sin(w * L) / w], [-w * math.sin(w * L), math.cos(w * L)]])
    My = np.array([[math.cosh(w * L), math.sinh(w * L) / w], [w * math.sinh(w * L), math.cosh(w * L)]])
    eps = 1e-7
    base = np.zeros((7, 1))
    out = {}
    for col in (0, 1, 2, 3):
        p = np.zeros((7, 2))
        p[col, 0] = eps
        p[col, 1] = -eps
        q = run(p, [(1, L, k1)])
        out[col] = (q[:4, 0] - q[:4, 1]) / (2 * eps)
    J = np.array([out[c] for c in (0, 1, 2, 3)]).T
    np.testing.assert_allclose(J[:2, :2], Mx, rtol=0, atol=1e-8)
    np.testing.assert_allclose(J[2:, 2:], My, rtol=0, atol=1e-8)
    assert abs(np.linalg.det(J) - 1) < 1e-8
    # defocusing sign flips the planes
    q = run(np.array([[1e-6], [0], [0], [0], [0], [0], [0]]), [(1, L, -k1)])
    assert q[0, 0] == pytest.approx(My[0, 0] * 1e-6, rel=1e-6)
