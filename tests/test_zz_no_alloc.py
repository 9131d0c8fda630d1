"""sc3d_moments_noalloc (the allocation-free variant of sc3d_moments, its
partials in the library's fixed device array) against sc3d_moments through
the C ABI: bit-identical means and second moments on two calls in a row, the
second on a smaller bunch (the array keeps the first call's partials beyond
the second's chunks, which the second must not read), and the size limit
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import synergia2_amd as sc
from helpers import C1_SIGMA, GAMMA

pytestmark = pytest.mark.gpu


def noalloc_moments(b):
    mean = np.zeros(6)
    mom2 = np.zeros(36)
    d = b.desc()
    rc = sc.load().sc3d_moments_noalloc(C.byref(d), mean.ctypes.data_as(C.POINTER(C.c_double)),
                                        mom2.ctypes.data_as(C.POINTER(C.c_double)), None)
    return rc, mean, mom2.reshape(6, 6)


def test_moments_noalloc_matches_moments(gpu):
    op = sc.SpaceCharge3dOpenHockney(sc.default_options(32, 32, 64))
    for n, seed in ((1_000_003, 5), (70_001, 6)):
        print(f"case noalloc n={n}", flush=True)
        b = sc.Bunch(n, 1e10, gamma=GAMMA)
        b.populate_gaussian(np.zeros(6), sc.diag_chol(C1_SIGMA), seed=seed)
        b.masks[::7] = 0  # some inactive slots
        mean_ref, mom_ref = op.moments(b)
        rc, mean, mom = noalloc_moments(b)
        assert rc == 0, sc.load().sc3d_last_error()
        assert np.array_equal(mean.view(np.int64), mean_ref.view(np.int64))
        assert np.array_equal(mom.view(np.int64), mom_ref.view(np.int64))
        # and the moments are those of the active particles
        p, m = b.host()
        a = p[:6, m == 1]
        assert np.allclose(mean, a.sum(axis=1) / n, rtol=1e-12, atol=0)


def test_moments_noalloc_size_limit(gpu):
    b = sc.Bunch(16, 1e10, gamma=GAMMA)
    d = b.desc()
    d.n_active = (1 << 26) + 1
    d.stride = d.n_active
    mean = np.zeros(6)
    mom2 = np.zeros(36)
    rc = sc.load().sc3d_moments_noalloc(C.byref(d), mean.ctypes.data_as(C.POINTER(C.c_double)),
                                        mom2.ctypes.data_as(C.POINTER(C.c_double)), None)
    assert rc == sc.SC3D_ERR_SIZE
    assert b"2^26" in sc.load().sc3d_last_error()
