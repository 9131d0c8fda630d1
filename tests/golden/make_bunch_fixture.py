"""Writes tests/golden/bunch_1k.npz: a 1e3-particle seeded bunch and the
oracle's outputs of one space-charge apply on it (the CPU restatement of
Space_charge_3d_open_hockney::apply_bunch, space_charge_3d_open_hockney.cc:
422-453): domain record, cell keys, stable sort permutation, rho, phi, Ex, Ey,
Ez (octants, zyx) and the bunch after the kick.

Bunch: 6D Gaussian from the counter-based populate (Philox-4x32-10 +
Box-Muller, seed 1000), sigma (1 mm, 1e-4, 1.5 mm, 1e-4, 10 mm / beta,
1e-4), mean offsets in x and cdt, 3 % of the slots masked out, 4 slots far
just outside the +-4 sigma domain (without widening it much); proton at 0.8 GeV kinetic energy, population 5e10;
grid 8x8x16, linear Green's function, adaptive domain (n_sigma 8); kick of
0.1 m (dt = 0.1 / (beta c)).

    python tests/golden/make_bunch_fixture.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

MP = 0.938272046
KE = 0.8
SHAPE = (8, 8, 16)
N = 1000


def kinematics():
    gamma = (MP + KE) / MP
    beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
    return gamma, beta, gamma * beta * MP


def inputs():
    gamma, beta, p = kinematics()
    sig = (1e-3, 1e-4, 1.5e-3, 1e-4, 0.010 / beta, 1e-4)
    chol = np.zeros(21)
    for r in range(6):
        chol[r * (r + 1) // 2 + r] = sig[r]
    mean = np.array([2e-4, 0.0, 0.0, 0.0, -3e-3, 0.0])
    parts, masks = O.populate_gaussian(N, mean, chol, seed=1000)
    masks[5::33] = 0
    parts[0, [7, 300]] = 2e-4 + 5.5e-3            # outside the grid in x
    parts[4, [11, 901]] = -3e-3 - 5.5 * sig[4]    # outside in cdt
    return parts, masks


def main():
    gamma, beta, p = kinematics()
    parts, masks = inputs()
    out = {"parts_in": parts.copy(), "masks": masks.copy(), "shape": np.array(SHAPE),
           "real_num": np.array(5e10), "gamma": np.array(gamma),
           "time_step": np.array(0.1 / (beta * 299792458.0))}
    po = np.ascontiguousarray(parts.copy())
    mo = np.ascontiguousarray(masks.copy())
    b = O.make_bunch(po, mo, N, 5e10, 1, MP, gamma, beta, p)
    rc, res = O.apply(O.make_options(SHAPE), b, float(out["time_step"]), nthreads=1,
                      want=("keys", "perm", "rho", "phi", "ex", "ey", "ez"))
    assert rc == 0, O.last_error()
    for k in ("keys", "perm", "rho", "phi", "ex", "ey", "ez"):
        out[k] = res[k]
    for k, v in res["domain"].items():
        out["domain_" + k] = np.array(v, np.float64)
    out["parts_out"] = po
    np.savez_compressed(os.path.join(HERE, "bunch_1k.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
