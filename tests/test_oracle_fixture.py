"""The committed bunch fixture (tests/golden/bunch_1k.npz, written by
tests/golden/make_bunch_fixture.py from the oracle): its inputs come from the
counter-based populate bit for bit, and the oracle still produces every
recorded output -- domain record, keys and permutation exactly, rho, phi, E
and the kicked bunch exactly on one thread and within 1e-14 on four (the
per-thread density grids are summed in another order)."""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys_path = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(os.path.join(sys_path, "bunch_1k.npz")))


def test_fixture_inputs_from_populate(fx):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_bunch_fixture", os.path.join(sys_path, "make_bunch_fixture.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    parts, masks = mod.inputs()
    assert np.array_equal(parts, fx["parts_in"]) and np.array_equal(masks, fx["masks"])


@pytest.mark.parametrize("nthreads,tol", [(1, 0.0), (4, 1e-14)])
def test_oracle_reproduces_fixture(fx, nthreads, tol):
    import math
    MP = 0.938272046
    gamma = float(fx["gamma"])
    beta = math.sqrt(1 - 1 / gamma ** 2)
    n = fx["masks"].shape[0]
    po = np.ascontiguousarray(fx["parts_in"].copy())
    mo = np.ascontiguousarray(fx["masks"].copy())
    b = O.make_bunch(po, mo, n, float(fx["real_num"]), 1, MP, gamma, beta, gamma * beta * MP)
    rc, out = O.apply(O.make_options(tuple(int(x) for x in fx["shape"])), b,
                      float(fx["time_step"]), nthreads=nthreads,
                      want=("keys", "perm", "rho", "phi", "ex", "ey", "ez"))
    assert rc == 0
    for k, v in out["domain"].items():
        assert np.array_equal(np.array(v, np.float64), fx["domain_" + k]), k
    assert np.array_equal(out["keys"], fx["keys"]) and np.array_equal(out["perm"], fx["perm"])
    for k in ("rho", "phi", "ex", "ey", "ez"):
        err = np.abs(out[k] - fx[k]).max() / np.abs(fx[k]).max()
        assert err <= tol, (k, err)
    err = np.abs(po - fx["parts_out"]).max(axis=1) / np.maximum(np.abs(fx["parts_out"]).max(axis=1), 1e-300)
    assert np.all(err <= tol), err
    # the fixture exercises what it should: masked and out-of-grid particles
    assert (fx["masks"] == 0).sum() >= 20
    assert np.count_nonzero(fx["keys"] == fx["keys"].max()) >= 1
