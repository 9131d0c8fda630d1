"""Green's functions on the doubled grid (space_charge_3d_open_hockney.cc:
136-348): the oracle against an independent numpy evaluation of the closed
form with the mirror mapping, evenness, the far field, and the reference's
float-reciprocal index decomposition quirk."""
import math

import numpy as np
import pytest

import oracle as O


def linear_closed_form(x, y, z, hx, hy, hz):
    rr_h = hx * hx + hy * hy
    r1 = math.sqrt(hx * hx + hy * hy + hz * hz)
    G000 = (2.0 / rr_h) * (hz * r1 + rr_h * math.log((hz + r1) / math.sqrt(rr_h)) - hz * hz)
    rr = x * x + y * y
    s0 = math.sqrt(rr + z * z)
    sn = math.sqrt(rr + (z - hz) ** 2)
    sp = math.sqrt(rr + (z + hz) ** 2)
    G = 2 * s0 - sn - sp
    eps = 1e-12 * hz
    if abs(z) < eps:
        if abs(x) + abs(y) < 2 * eps:
            G += hz * G000
        else:
            G += 2 * hz * math.log((math.sqrt(hz * hz + rr) + hz) / math.sqrt(rr))
    elif abs(z - hz) < eps:
        G += (hz + z) * math.log((sp + z + hz) / (s0 + z))
    elif z > hz:
        G += (hz - z) * math.log((s0 + z) / (sn + z - hz)) + (hz + z) * math.log((sp + z + hz) / (s0 + z))
    return G


@pytest.mark.parametrize("green", ["linear", "pointlike"])
def test_green_closed_form_and_mirrors(green):
    g = (8, 8, 16)
    h = (1.3e-4, 2.1e-4, 7.0e-4)
    g2 = O.green_fn2(green, g, h)
    D = [2 * g[0], 2 * g[1], 2 * g[2]]
    G000p = 2.8 / min(h)
    for Z in range(D[2]):
        for Y in range(D[1]):
            for X in range(D[0]):
                x = min(X, D[0] - X) * h[0]
                y = min(Y, D[1] - Y) * h[1]
                z = min(Z, D[2] - Z) * h[2]
                if green == "linear":
                    ref = linear_closed_form(x, y, z, *h)
                else:
                    ref = G000p if X == Y == Z == 0 else 1.0 / math.sqrt(x * x + y * y + z * z)
                assert g2[Z, Y, X] == pytest.approx(ref, rel=1e-14, abs=0), (X, Y, Z)
    assert np.all(g2[:, :, D[0]:] == 0.0)  # padding untouched


def test_linear_far_field():
    # G_linear is hz * (1/r integrated against the unit-area triangle of
    # half-width hz), so G/hz^2 -> 1/r far from the source; hence the
    # 1/(hz*hz) of get_normalization_force (space_charge_3d_open_hockney.cc:726-729)
    h = (1e-3, 1e-3, 1e-3)
    for (x, y, z) in ((20e-3, 0, 0), (0, 15e-3, 6e-3), (5e-3, 5e-3, 30e-3)):
        r = math.sqrt(x * x + y * y + z * z)
        assert abs(linear_closed_form(x, y, z, *h) / h[2] ** 2 * r - 1) < 2e-3


def reference_decomposition_exact(g0, g1, nz):
    gx, gy = g0 + 1, g1 + 1
    igxgy, igx = 1.0 / (gx * gy), 1.0 / gx
    for i in range(gx * gy * nz):
        iz = int(i * igxgy)
        iy = int((i - iz * gx * gy) * igx)
        if (iz, iy) != (i // (gx * gy), (i % (gx * gy)) // gx):
            return False
    return True


def test_reference_index_quirk():
    """alg_g2_* decomposes the flat index with int(i * (1.0/(gx*gy))).  At
    every configured shape this equals integer division, so the device
    path's integer indexing gives the same g2; at some other shapes (e.g.
    16x32x16) the reference mis-assigns a few points, which the oracle keeps
    and the device path does not."""
    for shape in ((32, 32, 64), (64, 64, 128), (128, 128, 256)):
        assert reference_decomposition_exact(shape[0], shape[1], 2 * shape[2])
        assert reference_decomposition_exact(shape[0], shape[1], shape[2] + 1)
    assert not reference_decomposition_exact(16, 32, 32)


def closed_form(green, x, y, z, h):
    if green == "linear":
        return linear_closed_form(x, y, z, *h)
    return 2.8 / min(h) if x == y == z == 0 else 1.0 / math.sqrt(x * x + y * y + z * z)


@pytest.mark.parametrize("green", ["linear", "pointlike"])
@pytest.mark.parametrize("shape", [(16, 32, 16), (32, 16, 8)])
def test_green_patches_reproduce_reference_g2(green, shape):
    """At the shapes where alg_g2_*'s float-reciprocal decomposition
    misplaces indices, the library's patch list (host side of the device
    path) names exactly the points where the oracle's g2 -- every index
    replayed in the reference's serial order -- differs from the symmetric
    closed form, and the value the oracle holds there: the closed form at
    the misplaced writer's coordinates, or 0 where nothing is written."""
    import synergia2_amd as sc
    gfn = sc.GREEN_LINEAR if green == "linear" else sc.GREEN_POINTLIKE
    patches = sc.green_patches(gfn, shape)
    assert 0 < len(patches) <= 512
    h = (1.3e-4, 2.1e-4, 7.0e-4)
    g2 = O.green_fn2(green, shape, h)
    D = [2 * shape[0], 2 * shape[1], 2 * shape[2]]
    targets = {}
    for t0, t1, t2, s0, s1, s2 in patches:
        targets[(t0, t1, t2)] = None if s0 < 0 else (s0, s1, s2)
    assert len(targets) == len(patches)
    for Z in range(D[2]):
        for Y in range(D[1]):
            for X in range(D[0]):
                if (Z, Y, X) in targets:
                    src = targets[(Z, Y, X)]
                    if src is None:
                        assert g2[Z, Y, X] == 0.0
                    else:
                        iz, iy, ix = src
                        z = (D[2] - iz if (green == "linear" and iz > shape[2]) else iz) * h[2]
                        ref = closed_form(green, ix * h[0], iy * h[1], z, h)
                        assert g2[Z, Y, X] == pytest.approx(ref, rel=1e-14, abs=0)
                        sym = closed_form(green, min(X, D[0] - X) * h[0],
                                          min(Y, D[1] - Y) * h[1], min(Z, D[2] - Z) * h[2], h)
                        assert abs(g2[Z, Y, X] / sym - 1) > 1e-6  # a real difference
                else:
                    sym = closed_form(green, min(X, D[0] - X) * h[0], min(Y, D[1] - Y) * h[1],
                                      min(Z, D[2] - Z) * h[2], h)
                    assert g2[Z, Y, X] == pytest.approx(sym, rel=1e-14, abs=0), (Z, Y, X)


def test_green_patches_none_at_configured_shapes():
    import synergia2_amd as sc
    for shape in ((32, 32, 64), (64, 64, 128), (128, 128, 256), (8, 8, 16), (64, 64, 64)):
        for gfn in (sc.GREEN_LINEAR, sc.GREEN_POINTLIKE):
            assert len(sc.green_patches(gfn, shape)) == 0
