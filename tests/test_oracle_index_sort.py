"""Cell index (collective/utils.h:9-24) and the build-side stable cell sort,
pinned by integer closed forms on hand-built coordinates and numpy's stable
argsort."""
import numpy as np

import oracle as O
from helpers import BETA, GAMMA, PREF


def test_cell_index_closed_forms():
    left, h = -0.5, 0.25
    ih = 1.0 / h
    # cell centres map to idx = k - 1/2 ... s = (pos-left)/h - 0.5
    for k in range(-3, 8):
        pos = left + (k + 0.5) * h          # s = k exactly (h is a power of 2)
        idx, off = O.cell_index(pos, left, ih)
        assert (idx, off) == (k, 0.0)
        pos = left + (k + 0.75) * h          # s = k + 0.25
        idx, off = O.cell_index(pos, left, ih)
        assert (idx, off) == (k, 0.25)
    # far outside and NaN: outside every grid, defined result
    assert O.cell_index(1e300, 0.0, 1.0)[0] < -10 ** 8
    assert O.cell_index(float("nan"), 0.0, 1.0)[0] < -10 ** 8
    # negative side floors toward -inf
    assert O.cell_index(left, left, ih) == (-1, 0.5)


def test_keys_and_stable_sort():
    g = (4, 4, 4)
    h = 0.25
    pts = [  # (x, y, z) -> expected key; domain left = -0.5 in each axis
        ((-0.5 + 0.5 * h, -0.5 + 0.5 * h, -0.5 + 0.5 * h), 0),
        ((-0.5 + 1.7 * h, -0.5 + 2.9 * h, -0.5 + 1.2 * h), (0 * 4 + 2) * 4 + 1),
        ((-0.5 + 3.6 * h, 0.0, 0.0), 64),      # ix = 3 = g-1 -> not in [0, g-2]
        ((-0.5 + 0.2 * h, 0.0, 0.0), 64),      # ix = -1
        ((-0.5 + 2.5 * h, -0.5 + 1.5 * h, -0.5 + 2.5 * h), (2 * 4 + 1) * 4 + 2),
        ((-0.5 + 1.7 * h, -0.5 + 2.9 * h, -0.5 + 1.2 * h), 64),  # masked out
    ]
    n = len(pts)
    parts = np.zeros((7, n))
    for i, (p, _) in enumerate(pts):
        parts[0, i], parts[2, i], parts[4, i] = p
    masks = np.ones(n, np.uint8)
    masks[5] = 0
    b = O.make_bunch(parts, masks, n, 1e10, 1, 0.938272046, GAMMA, BETA, PREF)
    dom = O.Domain()
    for a in range(3):
        dom.left[a] = -0.5
        dom.h[a] = h
    keys = np.zeros(n, np.int32)
    import ctypes as C
    O.lib().orc_cell_keys(C.byref(b), C.byref(dom), np.array(g, np.int32).ctypes.data_as(
        C.POINTER(C.c_int32)), keys.ctypes.data_as(C.POINTER(C.c_int32)))
    assert list(keys) == [k for _, k in pts]
    rng = np.random.default_rng(0)
    k = rng.integers(0, 50, 10000).astype(np.int32)
    perm = np.zeros(10000, np.int32)
    O.lib().orc_stable_sort(k.ctypes.data_as(C.POINTER(C.c_int32)), 10000,
                            perm.ctypes.data_as(C.POINTER(C.c_int32)))
    assert np.array_equal(perm, np.argsort(k, kind="stable"))


def _keys(g, cells, h=0.25, left=-0.5):
    """orc_cell_keys of one particle at the centre of each cell (ix, iy, iz)."""
    import ctypes as C
    n = len(cells)
    parts = np.zeros((7, n))
    for i, c in enumerate(cells):
        for a in range(3):
            parts[2 * a, i] = left + (c[a] + 1.0) * h  # s = c + 1/2
    masks = np.ones(n, np.uint8)
    b = O.make_bunch(parts, masks, n, 1e10, 1, 0.938272046, GAMMA, BETA, PREF)
    dom = O.Domain()
    for a in range(3):
        dom.left[a] = left
        dom.h[a] = h
    keys = np.zeros(n, np.int32)
    O.lib().orc_cell_keys(C.byref(b), C.byref(dom), np.array(g, np.int32).ctypes.data_as(
        C.POINTER(C.c_int32)), keys.ctypes.data_as(C.POINTER(C.c_int32)))
    return keys


def test_brick_keys():
    """Brick order: 4 cells per axis where 4 divides the axis (x and z of
    8 x 6 x 12), 1 elsewhere, bricks on a serpentine path; keys worked out
    by hand, then the bijection onto [0, G), the bricks as contiguous key
    ranges and consecutive bricks as face neighbours."""
    g = (8, 6, 12)
    # (ix, iy, iz) -> brick (r*2 + bx'), r = bz*6 + by' (by' = 5 - by for odd bz,
    # bx' = 1 - bx for odd r), cell (lz*4 + lx); key brick*16 + cell
    hand = {(0, 0, 0): 0,
            (5, 2, 7): (9 * 2 + 0) * 16 + 3 * 4 + 1,     # bz 1: by' 3, r 9 odd: bx' 0
            (3, 4, 0): (4 * 2 + 0) * 16 + 3,             # r 4
            (4, 0, 1): (0 * 2 + 1) * 16 + 1 * 4,         # r 0, bx 1, lz 1
            (1, 1, 2): (1 * 2 + 1) * 16 + 2 * 4 + 1,     # r 1 odd: bx' 1
            (6, 4, 10): (16 * 2 + 1) * 16 + 2 * 4 + 2}   # bz 2: r 16
    cells = list(hand)
    assert list(_keys(g, cells)) == [hand[c] for c in cells]
    G = g[0] * g[1] * g[2]
    # every cell whose lower corner is in [0, g-2]^3 has a distinct key < G;
    # the last layer of each axis gets G
    cells = [(x, y, z) for z in range(g[2]) for y in range(g[1]) for x in range(g[0])]
    k = _keys(g, cells)
    inner = np.array([all(c[a] <= g[a] - 2 for a in range(3)) for c in cells])
    assert np.all(k[~inner] == G)
    assert len(set(k[inner].tolist())) == inner.sum() and k[inner].max() < G
    # the 4 x 1 x 4 cells of one brick hold 16 consecutive keys
    for bz in range(g[2] // 4 - 1):
        for by in range(g[1] - 1):
            kb = _keys(g, [(x, by, 4 * bz + z) for z in range(4) for x in range(4)])
            assert sorted(kb.tolist()) == list(range(kb.min(), kb.min() + 16))
    # serpentine: the bricks of 8 x 8 x 12 in key order step to a face neighbour
    g = (8, 8, 12)
    bricks = [(bx, by, bz) for bz in range(3) for by in range(2) for bx in range(2)]
    rank = _keys(g, [(4 * bx, 4 * by, 4 * bz) for bx, by, bz in bricks]) // 64
    path = [bricks[i] for i in np.argsort(rank)]
    assert sorted(rank.tolist()) == list(range(12))
    for a, b in zip(path, path[1:]):
        assert sum(abs(u - v) for u, v in zip(a, b)) == 1, (a, b)
    # no axis that 4 divides: bricks of one cell, the serpentine path over
    # the cells of 5 x 6 x 7 (r = z*6 + y', y' = 5 - y for odd z; x' = 4 - x
    # for odd r; key r*5 + x')
    hand = {(0, 0, 0): 0, (1, 1, 0): 1 * 5 + 3, (2, 3, 1): 8 * 5 + 2, (3, 4, 3): 19 * 5 + 1}
    assert list(_keys((5, 6, 7), list(hand))) == list(hand.values())
