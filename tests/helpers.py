"""Shared helpers of the parity tests: seeded bunches and oracle runs."""
import numpy as np

import oracle as O
import synergia2_amd as sc

KE = 0.8  # GeV, proton
GAMMA = (sc.PROTON_MASS + KE) / sc.PROTON_MASS
BETA, PREF = sc.kinematics(sc.PROTON_MASS, GAMMA)


def gaussian_parts(n, sig, seed, mean=(0, 0, 0, 0, 0, 0)):
    """(7, n) float64 host particles, uncorrelated Gaussian, id = slot."""
    rng = np.random.default_rng(seed)
    p = np.zeros((7, n))
    for j in range(6):
        p[j] = mean[j] + sig[j] * rng.standard_normal(n)
    p[6] = np.arange(n)
    return p


C1_SIGMA = (1e-3, 1e-4, 1e-3, 1e-4, 0.010 / BETA, 1e-4)


def oracle_run(opts_kw, parts, masks, real_num, time_step, total_num=None, want=(), nthreads=4,
               gamma=GAMMA):
    beta, pref = sc.kinematics(sc.PROTON_MASS, gamma)
    p = np.ascontiguousarray(parts.copy())
    m = np.ascontiguousarray(masks.copy())
    b = O.make_bunch(p, m, total_num if total_num is not None else parts.shape[1], real_num, 1,
                     sc.PROTON_MASS, gamma, beta, pref)
    o = O.make_options(**opts_kw)
    rc, out = O.apply(o, b, time_step, nthreads=nthreads, want=want)
    out["parts"] = p
    out["masks"] = m
    return rc, out


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def by_id(parts):
    """Return parts with columns reordered by particle id (column 6)."""
    order = np.argsort(parts[6], kind="stable")
    return parts[:, order]
