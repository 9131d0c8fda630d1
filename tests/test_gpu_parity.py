"""Parity of the HIP space-charge step (through the C ABI) against the CPU
restatement of the reference operator (oracle/), on identical seeded bunches.

Bars: domain record, cell keys and the stable sort
permutation bit-exact; rho, phi, E rel-L2 <= 1e-12; post-kick coordinates
rel <= 1e-12 (matched by particle id, since the device path reorders the
bunch by cell).
"""
import numpy as np
import pytest

import oracle as O
import synergia2_amd as sc
from helpers import BETA, C1_SIGMA, GAMMA, by_id, gaussian_parts, oracle_run, rel_l2

pytestmark = pytest.mark.gpu

DOMAIN_FIELDS = ("sums", "mean", "stddev", "size", "offset", "left", "h", "w0", "fn_norm",
                 "factor")


def gpu_run(opts, parts, masks, real_num, time_step, steps=1, total_num=None):
    n = parts.shape[1]
    b = sc.Bunch(n, real_num, gamma=GAMMA, capacity=n + 17)
    if total_num is not None:
        b.total_num = float(total_num)
    b.set_host(parts, masks)
    op = sc.SpaceCharge3dOpenHockney(opts)
    for _ in range(steps):
        op.apply_bunch(b, time_step)
    return op, b


def make_opts(shape, green=sc.GREEN_LINEAR, sort_period=1, fixed=None, precision=sc.FP64):
    o = sc.default_options(*shape)
    o.green_fn = green
    o.sort_period = sort_period
    o.precision = precision
    if fixed is not None:
        sc.set_fixed_domain(o, *fixed)
    return o


def oracle_kw(shape, green="linear", fixed=None, reorder=False):
    return dict(shape=shape, green_fn=green, fixed=fixed, reorder=reorder)


def check_domain(dg, do):
    for f in DOMAIN_FIELDS:
        a = np.array(dg[f], np.float64)
        b = np.array(do[f], np.float64)
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), (f, a, b)


def check_kicked(parts_gpu, masks_gpu, parts_orc, parts0, tol=1e-12, dk_tol=1e-10):
    pg = by_id(parts_gpu)
    po = parts_orc[:, np.argsort(parts_orc[6], kind="stable")]
    p0 = parts0[:, np.argsort(parts0[6], kind="stable")]
    assert np.array_equal(pg[6], po[6])
    for c in (0, 2, 4, 6):  # positions and ids are untouched
        assert np.array_equal(pg[c], po[c]), c
    for c in (1, 3, 5):
        scale = np.abs(po[c]).max()
        assert np.abs(pg[c] - po[c]).max() <= tol * scale, (c, np.abs(pg[c] - po[c]).max(), scale)
        dk_g, dk_o = pg[c] - p0[c], po[c] - p0[c]
        assert rel_l2(dk_g, dk_o) <= dk_tol, (c, rel_l2(dk_g, dk_o))


def test_step_parity_c1_geometry(gpu):
    """One kick on a C1-shaped bunch (32x32x64, linear G, adaptive domain)."""
    n = 50000
    parts = gaussian_parts(n, C1_SIGMA, seed=0)
    masks = np.ones(n, np.uint8)
    dt = 0.1 / (BETA * sc.C_LIGHT)
    print(f"case c1 n={n}", flush=True)
    rc, out = oracle_run(oracle_kw((32, 32, 64)), parts, masks, 5e10, dt,
                         want=("keys", "perm", "rho", "phi", "ex", "ey", "ez"))
    assert rc == 0
    op, b = gpu_run(make_opts((32, 32, 64)), parts, masks, 5e10, dt)
    dom = op.debug(sc.DBG_DOMAIN)
    check_domain(dom, out["domain"])
    assert dom["error"] == 0 and dom["ghat_dirty"] == 1
    keys = op.debug(sc.DBG_KEYS, n)
    perm = op.debug(sc.DBG_PERM, n)
    assert np.array_equal(keys, out["keys"])
    assert np.array_equal(perm, out["perm"])
    assert rel_l2(op.debug(sc.DBG_RHO), out["rho"]) <= 1e-13
    assert rel_l2(op.debug(sc.DBG_PHI), out["phi"]) <= 1e-12
    for tap, name in ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez")):
        assert rel_l2(op.debug(tap), out[name]) <= 1e-12, name
    pg, mg = b.host()
    # the bunch is now in cell order
    assert np.array_equal(pg[6].astype(np.int64), out["perm"].astype(np.int64))
    check_kicked(pg, mg, out["parts"], parts)


def test_green_spectrum_matches_full_fft(gpu):
    """Real Green's spectrum octant vs the full FFT of the reference's g2.
    (Shapes where the reference's float-reciprocal index decomposition of
    alg_g2_* is exact; see test_reference_index_quirk in test_oracle_green.)"""
    shape = (16, 64, 32)
    n = 20000
    parts = gaussian_parts(n, (2e-3, 1e-4, 1e-3, 1e-4, 5e-3, 1e-4), seed=3)
    masks = np.ones(n, np.uint8)
    for green, gname in ((sc.GREEN_LINEAR, "linear"), (sc.GREEN_POINTLIKE, "pointlike")):
        print(f"case green {gname}", flush=True)
        rc, out = oracle_run(oracle_kw(shape, gname), parts, masks, 1e10, 1e-9,
                             want=("ghat_full",))
        assert rc == 0
        op, _ = gpu_run(make_opts(shape, green), parts, masks, 1e10, 1e-9)
        gx, gy, gz = shape
        D = [2 * gx, 2 * gy, 2 * gz]
        hx = gx + 1
        full = out["ghat_full"].view(np.complex128).reshape(D[2], D[1], hx)
        ref = full[: gz + 1, : gy + 1, :].real
        got = op.debug(sc.DBG_GHAT).reshape(gz + 1, gy + 1, hx)
        assert rel_l2(got, ref) <= 1e-13, gname
        assert np.abs(full.imag).max() <= 1e-12 * np.abs(full.real).max()
        # precision form: the same FP64 values, transformed in FP32 (measured
        # 9e-8 for both functions; a wrong twiddle or index is an O(1) error)
        op32, _ = gpu_run(make_opts(shape, green, precision=sc.FP32_GRID), parts, masks, 1e10,
                          1e-9)
        got32 = op32.debug(sc.DBG_GHAT).reshape(gz + 1, gy + 1, hx)
        err32 = rel_l2(got32, ref)
        print(f"  FP32 spectrum rel-L2 {err32:.2e}", flush=True)
        assert err32 <= 1e-6, gname


@pytest.mark.parametrize("case", ["pointlike_fixed_masked", "anisotropic_outliers"])
def test_step_parity_variants(gpu, case):
    rng = np.random.default_rng(7)
    if case == "pointlike_fixed_masked":
        shape, green = (32, 8, 64), "pointlike"
        n = 12345
        parts = gaussian_parts(n, (1e-3, 1e-4, 2e-3, 1e-4, 1e-2, 1e-4), seed=11)
        masks = (rng.random(n) > 0.1).astype(np.uint8)
        fixed = ((1e-4, -2e-4, 3e-4), (8e-3, 1.6e-2, 8e-2))
    else:
        shape, green = (64, 16, 32), "linear"
        n = 40961
        parts = gaussian_parts(n, (3e-3, 1e-4, 0.5e-3, 1e-4, 2e-2, 1e-4), seed=12,
                               mean=(1e-3, 0, -2e-3, 0, 0.5, 0))
        masks = np.ones(n, np.uint8)
        far = rng.choice(n, 200, replace=False)
        parts[0, far[:100]] += 1.0       # far outside the grid
        parts[4, far[100:]] = np.nan     # invalid coordinate -> outside
        masks[far[150:]] = 0
        fixed = None
    dt = 2e-9
    print(f"case {case}", flush=True)
    rc, out = oracle_run(oracle_kw(shape, green, fixed), parts, masks, 2e10, dt,
                         want=("keys", "perm", "rho", "phi", "ex", "ey", "ez"))
    assert rc == 0
    gid = sc.GREEN_LINEAR if green == "linear" else sc.GREEN_POINTLIKE
    op, b = gpu_run(make_opts(shape, gid, fixed=fixed), parts, masks, 2e10, dt)
    check_domain(op.debug(sc.DBG_DOMAIN), out["domain"])
    assert np.array_equal(op.debug(sc.DBG_KEYS, n), out["keys"])
    assert np.array_equal(op.debug(sc.DBG_PERM, n), out["perm"])
    assert rel_l2(op.debug(sc.DBG_RHO), out["rho"]) <= 1e-13
    assert rel_l2(op.debug(sc.DBG_PHI), out["phi"]) <= 1e-12
    for tap, name in ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez")):
        assert rel_l2(op.debug(tap), out[name]) <= 1e-12, name
    pg, mg = b.host()
    po = out["parts"]
    # NaN coordinates stay NaN in both; compare with NaN-aware equality
    g_sorted = by_id(pg)
    o_sorted = po[:, np.argsort(po[6], kind="stable")]
    for c in (0, 2, 4):
        assert np.array_equal(g_sorted[c], o_sorted[c], equal_nan=True)
    for c in (1, 3, 5):
        scale = np.nanmax(np.abs(o_sorted[c]))
        assert np.nanmax(np.abs(g_sorted[c] - o_sorted[c])) <= 1e-12 * scale
    # masks travel with their particles
    assert np.array_equal(mg[np.argsort(pg[6], kind="stable")], masks)


@pytest.mark.parametrize("precision", [sc.FP64, sc.FP32_GRID])
@pytest.mark.parametrize("dense", [True, False])
def test_deposit_lane_widths(gpu, dense, precision):
    """rho for both lane widths of the deposit (4 particles per lane on a
    bunch with >= 6 particles per cell, else 2), walking the bunch right after
    a reorder (sort_period 1) and walking the sort permutation (sort_period 0,
    never reordered), with masked and off-grid particles and a count that is
    no multiple of the lane width."""
    shape = (16, 16, 32)  # G = 8192
    n = 60001 if dense else 20001
    parts = gaussian_parts(n, C1_SIGMA, seed=31 + int(dense))
    rng = np.random.default_rng(8)
    masks = (rng.random(n) > 0.05).astype(np.uint8)
    parts[0, rng.choice(n, 50, replace=False)] += 1.0  # off the grid
    fixed = ((0.0, 0.0, 0.0), (8e-3, 8e-3, 0.1))
    dt = 0.1 / (BETA * sc.C_LIGHT)
    print(f"case deposit dense={dense} precision={precision}", flush=True)
    rc, out = oracle_run(oracle_kw(shape, fixed=fixed), parts, masks, 5e10, dt, want=("rho",))
    assert rc == 0
    tol = 1e-13 if precision == sc.FP64 else 2e-7
    for sp in (1, 0):
        op, _ = gpu_run(make_opts(shape, sort_period=sp, fixed=fixed, precision=precision), parts,
                        masks, 5e10, dt)
        err = rel_l2(op.debug(sc.DBG_RHO), out["rho"])
        print("sort_period", sp, "rho rel-L2", err)
        assert err <= tol


@pytest.mark.parametrize("precision", [sc.FP64, sc.FP32_GRID])
def test_deposit_windows_out_of_cell_order(gpu, precision):
    """rho when the bunch storage is no longer in cell order (the deposit's
    windows are sorted by cell in LDS): one apply reorders the bunch, the
    particles are then moved by several cells (and some masked out), and the
    next apply (no reorder) must give the oracle's rho of the moved bunch."""
    shape = (32, 32, 64)
    n = 200003
    parts = gaussian_parts(n, C1_SIGMA, seed=41)
    masks = np.ones(n, np.uint8)
    fixed = ((0.0, 0.0, 0.0), (8e-3, 8e-3, 0.1))
    dt = 0.1 / (BETA * sc.C_LIGHT)
    op, b = gpu_run(make_opts(shape, sort_period=8, fixed=fixed, precision=precision), parts,
                    masks, 5e10, dt)
    moved, mm = b.host()  # now in cell order
    rng = np.random.default_rng(12)
    h = (8e-3 / 32, 8e-3 / 32, 0.1 / 64)
    for col, hc in ((0, h[0]), (2, h[1]), (4, h[2])):
        moved[col] += rng.normal(0.0, 3.0 * hc, n)
    mm[rng.random(n) < 0.03] = 0
    b.set_host(moved, mm)
    op.apply_bunch(b, dt)
    print(f"case windows precision={precision}", flush=True)
    rc, out = oracle_run(oracle_kw(shape, fixed=fixed), moved, mm, 5e10, dt, want=("rho",))
    assert rc == 0
    err = rel_l2(op.debug(sc.DBG_RHO), out["rho"])
    print("rho rel-L2", err)
    # FP32: every cell is now split over many windows, so each node adds
    # many once-rounded FP32 partial sums (2e-7 measured; the sorted bunch of
    # test_fp32_grid_field_accuracy stays within 2e-7)
    assert err <= (1e-13 if precision == sc.FP64 else 5e-7)


@pytest.mark.parametrize("precision", [sc.FP64, sc.FP32_GRID])
def test_deposit_crowded_cells(gpu, precision):
    """rho when a few cells hold thousands of particles (one cell fills
    whole wavefronts and windows): after a reorder (compact windows, LDS node
    tiles), through the sort permutation (never reordered: lane runs and the
    segmented scan carry partial sums across rows, wavefronts and blocks),
    and with the storage then shuffled (every window scattered: wavefront
    node caches, crowded nodes summed in FP64 in LDS)."""
    shape = (16, 16, 32)
    fixed = ((0.0, 0.0, 0.0), (8e-3, 8e-3, 0.1))
    left = np.array([-4e-3, -4e-3, -0.05])
    h = np.array([8e-3 / 16, 8e-3 / 16, 0.1 / 32])
    rng = np.random.default_rng(21)
    crowd = (((5, 6, 10), 9000), ((6, 6, 10), 6000), ((5, 7, 11), 300))
    n = 40000
    parts = gaussian_parts(n, C1_SIGMA, seed=22)
    at = 0
    for cell, count in crowd:
        for k, col in enumerate((0, 2, 4)):
            u = rng.uniform(0.01, 0.99, count)
            parts[col, at:at + count] = left[k] + (cell[k] + 0.5 + u) * h[k]
        at += count
    masks = (rng.random(n) > 0.02).astype(np.uint8)
    dt = 0.1 / (BETA * sc.C_LIGHT)
    print(f"case crowded cells precision={precision}", flush=True)
    # FP32: a crowded cell's total arrives as once-rounded FP32 partial sums
    # (per window or wavefront, FP64 inside it; per run of the permutation
    # path), added in an order that varies from run to run.  A dropped or
    # doubled partial sum would be an error above 1e-3.
    tol = 1e-13 if precision == sc.FP64 else 1e-6
    rc, out = oracle_run(oracle_kw(shape, fixed=fixed), parts, masks, 5e10, dt, want=("rho",))
    assert rc == 0
    for sp in (1, 0):
        op, _ = gpu_run(make_opts(shape, sort_period=sp, fixed=fixed, precision=precision), parts,
                        masks, 5e10, dt)
        err = rel_l2(op.debug(sc.DBG_RHO), out["rho"])
        print("sort_period", sp, "rho rel-L2", err, flush=True)
        assert err <= tol, sp
    # reorder once, then shuffle the storage: the next apply (no reorder)
    # sorts every window in LDS
    op, b = gpu_run(make_opts(shape, sort_period=8, fixed=fixed, precision=precision), parts,
                    masks, 5e10, dt)
    p1, m1 = b.host()
    order = rng.permutation(n)
    p1, m1 = p1[:, order], m1[order]
    b.set_host(p1, m1)
    op.apply_bunch(b, dt)
    rc, out = oracle_run(oracle_kw(shape, fixed=fixed), p1, m1, 5e10, dt, want=("rho",))
    assert rc == 0
    err = rel_l2(op.debug(sc.DBG_RHO), out["rho"])
    print("shuffled rho rel-L2", err, flush=True)
    assert err <= tol


@pytest.mark.parametrize("precision", [sc.FP64, sc.FP32_GRID])
def test_kick_field_tile_and_direct_gathers(gpu, precision):
    """Post-kick momenta against the oracle for both ways the map-free kick
    reads the field: from the LDS tile of a window's node box (storage in
    cell order after a reorder; crowded cells fill whole windows) and from
    global memory (storage never reordered, or shuffled after a reorder: the
    windows' boxes exceed the tile), with masked and off-grid particles and a
    count that is no multiple of the window."""
    shape = (16, 16, 32)
    fixed = ((0.0, 0.0, 0.0), (8e-3, 8e-3, 0.1))
    left = np.array([-4e-3, -4e-3, -0.05])
    h = np.array([8e-3 / 16, 8e-3 / 16, 0.1 / 32])
    rng = np.random.default_rng(51)
    n = 50001
    parts = gaussian_parts(n, C1_SIGMA, seed=52)
    for k, col in enumerate((0, 2, 4)):  # 5000 particles in one cell
        parts[col, :5000] = left[k] + (np.array((7, 8, 15))[k] + 0.5 +
                                       rng.uniform(0.01, 0.99, 5000)) * h[k]
    masks = (rng.random(n) > 0.03).astype(np.uint8)
    parts[0, rng.choice(n, 40, replace=False)] += 1.0  # off the grid
    dt = 0.1 / (BETA * sc.C_LIGHT)
    # FP32 grid: the field within ~2e-7 of the FP64 one, so the kicks too (a
    # wrong node in a gather is an O(1) error in the kick of its particles)
    tol, dk_tol = (1e-12, 1e-10) if precision == sc.FP64 else (1e-6, 1e-5)
    print(f"case kick tiles precision={precision}", flush=True)
    rc, out = oracle_run(oracle_kw(shape, fixed=fixed), parts, masks, 5e10, dt)
    assert rc == 0
    for sp in (1, 0):
        op, b = gpu_run(make_opts(shape, sort_period=sp, fixed=fixed, precision=precision), parts,
                        masks, 5e10, dt)
        pg, mg = b.host()
        print("sort_period", sp, flush=True)
        check_kicked(pg, mg, out["parts"], parts, tol=tol, dk_tol=dk_tol)
    # reorder once, shuffle the storage, kick again
    op, b = gpu_run(make_opts(shape, sort_period=8, fixed=fixed, precision=precision), parts,
                    masks, 5e10, dt)
    p1, m1 = b.host()
    order = rng.permutation(n)
    p1, m1 = p1[:, order], m1[order]
    b.set_host(p1, m1)
    op.apply_bunch(b, dt)
    rc, out = oracle_run(oracle_kw(shape, fixed=fixed), p1, m1, 5e10, dt)
    assert rc == 0
    pg, mg = b.host()
    check_kicked(pg, mg, out["parts"], p1, tol=tol, dk_tol=dk_tol)


def test_timing_modes(gpu):
    """Phase timing and whole-apply timing: each read back by its own entry
    point (the other refuses), a mode change refused while events are
    unread, and every timed apply counted."""
    n = 20000
    shape = (16, 16, 32)
    parts = gaussian_parts(n, C1_SIGMA, seed=61)
    masks = np.ones(n, np.uint8)
    b = sc.Bunch(n, 5e10, gamma=GAMMA)
    b.set_host(parts, masks)
    op = sc.SpaceCharge3dOpenHockney(make_opts(shape))
    op.set_timing(sc.TIMING_APPLY)
    for _ in range(3):
        op.apply_bunch(b, 1e-9)
    with pytest.raises(sc.Sc3dError):
        op.set_timing(sc.TIMING_PHASES)  # three applies' events unread
    with pytest.raises(sc.Sc3dError):
        op.timing()
    ms, napplies = op.apply_timing(reset=True)
    assert napplies == 3 and ms > 0.0
    op.set_timing(sc.TIMING_PHASES)
    for _ in range(2):
        op.apply_bunch(b, 1e-9)
    with pytest.raises(sc.Sc3dError):
        op.apply_timing()
    phases, napplies = op.timing(reset=True)
    assert napplies == 2 and all(v >= 0.0 for v in phases.values())
    assert phases["sc3d_kick"] > 0.0
    op.set_timing(sc.TIMING_OFF)
    op.apply_bunch(b, 1e-9)
    assert op.timing()[1] == 0


@pytest.mark.parametrize("precision", [sc.FP64, sc.FP32_GRID])
def test_green_side_stream(gpu, precision):
    """The Green's spectrum on its own stream beside the sort and the deposit
    (forced by the SC3D_TEST_GREEN_FORK hook: the automatic choice forks only
    when the caller enqueues ahead of the device) against the same applies
    with every pass on one stream (phase timing), over applies whose domain
    widens by 5 % each time (the spectrum recomputed each time; a spectrum
    of the previous domain would change the kicks by percents) and sort
    periods of 2.  Positions, ids and masks identical; the kicks agree to
    the rounding of the deposit's atomic sums, whose order varies from run
    to run."""
    n = 40000
    shape = (16, 32, 64)
    parts = gaussian_parts(n, C1_SIGMA, seed=71)
    masks = np.ones(n, np.uint8)
    masks[::29] = 0
    out = []
    import os
    for mode, fork in ((sc.TIMING_OFF, "1"), (sc.TIMING_PHASES, "0")):
        print(f"case green side stream precision={precision} mode={mode}", flush=True)
        os.environ["SC3D_TEST_GREEN_FORK"] = fork
        try:
            op = sc.SpaceCharge3dOpenHockney(make_opts(shape, sort_period=2, precision=precision))
            op.set_timing(mode)
            b = sc.Bunch(n, 5e10, gamma=GAMMA)
            got = []
            cur = parts.copy()
            for step in range(3):
                cur[0] *= 1.0 + 0.05 * step  # a wider domain each apply
                b.set_host(cur, masks)
                op.apply_bunch(b, 2e-9)
                p, m = b.host()
                got.append((cur.copy(), p.copy(), m.copy()))
        finally:
            del os.environ["SC3D_TEST_GREEN_FORK"]
        if mode == sc.TIMING_PHASES:
            assert op.timing(reset=True)[1] == 3
        out.append(got)
    dk_tol = 1e-10 if precision == sc.FP64 else 1e-5
    for step, ((p0, pa, ma), (_, pb, mb)) in enumerate(zip(*out)):
        assert np.array_equal(ma, mb)
        for c in (0, 2, 4, 6):
            assert np.array_equal(pa[c], pb[c]), (step, c)
        a, z, s0 = by_id(pa), by_id(pb), p0[:, np.argsort(p0[6], kind="stable")]
        for c in (1, 3, 5):
            err = rel_l2(a[c] - s0[c], z[c] - s0[c])
            print(f"  step {step} column {c}: kick rel-L2 {err:.2e}", flush=True)
            assert err <= dk_tol, (step, c, err)


def test_green_fork_automatic(gpu):
    """The automatic choice on the workload it is for: 2^20 + 4096 particles
    on 64 x 64 x 64 (2^18 cells) tracked through FODO steps with the maps
    between kicks and no wait between the steps, so the Green's spectrum
    forks beside the particle front whenever the previous apply is still
    running; against the same steps with the fork switched off
    (SC3D_TEST_GREEN_FORK=0).  Ids and masks identical; every coordinate
    within 1e-10 of its largest value (the runs differ by the order of the
    deposit's atomic sums only)."""
    import os
    from synergia2_amd import stepper as S
    n = (1 << 20) + 4096
    shape = (64, 64, 64)
    parts = gaussian_parts(n, C1_SIGMA, seed=72)
    masks = np.ones(n, np.uint8)
    out = []
    for fork in (None, "0"):
        print(f"case green fork automatic, hook {fork}", flush=True)
        if fork is not None:
            os.environ["SC3D_TEST_GREEN_FORK"] = fork
        try:
            op = sc.SpaceCharge3dOpenHockney(make_opts(shape, sort_period=24))
            b = sc.Bunch(n, 5e10, gamma=GAMMA)
            b.set_host(parts, masks)
            S.Propagator(S.fodo_channel(2), 8, b, op).run(6, fuse=False)
            op.sync()
            p, m = b.host()
        finally:
            os.environ.pop("SC3D_TEST_GREEN_FORK", None)
        out.append((by_id(p), m[np.argsort(p[6], kind="stable")]))
    (pa, ma), (pb, mb) = out
    assert np.array_equal(ma, mb)
    assert np.array_equal(pa[6], pb[6])
    for c in range(6):
        scale = np.abs(pb[c]).max()
        err = np.abs(pa[c] - pb[c]).max() / scale
        print(f"  column {c}: max rel {err:.2e}", flush=True)
        assert err <= 1e-10, (c, err)


def test_multi_step_parity(gpu):
    """Three kicks in a row (sort every 2nd apply) vs three oracle kicks."""
    n = 30000
    shape = (32, 32, 32)
    parts = gaussian_parts(n, C1_SIGMA, seed=5)
    masks = np.ones(n, np.uint8)
    dt = 5e-9
    op = sc.SpaceCharge3dOpenHockney(make_opts(shape, sort_period=2))
    b = sc.Bunch(n, 3e11, gamma=GAMMA)
    b.set_host(parts, masks)
    cur, curm = parts.copy(), masks.copy()
    for step in range(3):
        print(f"case multi step={step}", flush=True)
        # the oracle reorders its storage on the same applies as the device,
        # so the order-dependent domain sums stay bit-identical
        rc, out = oracle_run(oracle_kw(shape, reorder=(step % 2 == 0)), cur, curm, 3e11, dt,
                             want=("phi",))
        assert rc == 0
        op.apply_bunch(b, dt)
        check_domain(op.debug(sc.DBG_DOMAIN), out["domain"])
        assert rel_l2(op.debug(sc.DBG_PHI), out["phi"]) <= 1e-12
        pg, _ = b.host()
        assert np.array_equal(pg[6], out["parts"][6])  # same storage order
        check_kicked(pg, None, out["parts"], cur, tol=1e-12)
        cur, curm = out["parts"], out["masks"]


def test_green_cache_fixed_domain(gpu):
    """With a fixed domain h never changes: the spectrum is computed once."""
    n = 10000
    parts = gaussian_parts(n, C1_SIGMA, seed=9)
    masks = np.ones(n, np.uint8)
    fixed = ((0, 0, 0), (8e-3, 8e-3, 0.1))
    op = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64), fixed=fixed))
    b = sc.Bunch(n, 1e10, gamma=GAMMA)
    b.set_host(parts, masks)
    op.apply_bunch(b, 1e-9)
    assert op.debug(sc.DBG_DOMAIN)["ghat_dirty"] == 1
    phi1 = op.debug(sc.DBG_PHI)
    g1 = op.debug(sc.DBG_GHAT)
    op.apply_bunch(b, 1e-9)
    assert op.debug(sc.DBG_DOMAIN)["ghat_dirty"] == 0
    assert np.array_equal(op.debug(sc.DBG_GHAT), g1)
    # positions unchanged by the kick -> same rho, same phi up to atomics order
    assert rel_l2(op.debug(sc.DBG_PHI), phi1) <= 1e-14
    # precision form: its FP32 spectrum is cached the same way, and the field
    # of the second apply matches the first (same positions)
    op32 = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64), fixed=fixed,
                                                 precision=sc.FP32_GRID))
    b32 = sc.Bunch(n, 1e10, gamma=GAMMA)
    b32.set_host(parts, masks)
    op32.apply_bunch(b32, 1e-9)
    assert op32.debug(sc.DBG_DOMAIN)["ghat_dirty"] == 1
    g32 = op32.debug(sc.DBG_GHAT)
    e32 = op32.debug(sc.DBG_EX)
    assert rel_l2(g32, g1) <= 1e-6
    op32.apply_bunch(b32, 1e-9)
    assert op32.debug(sc.DBG_DOMAIN)["ghat_dirty"] == 0
    assert np.array_equal(op32.debug(sc.DBG_GHAT), g32)
    assert rel_l2(op32.debug(sc.DBG_EX), e32) <= 1e-6


def test_errors_and_empty(gpu):
    # all particles at one point: negligible extent (space_charge_3d_open_hockney.cc:510-516)
    n = 1000
    parts = np.zeros((7, n))
    parts[6] = np.arange(n)
    b = sc.Bunch(n, 1e10, gamma=GAMMA)
    b.set_host(parts, np.ones(n, np.uint8))
    op = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64)))
    with pytest.raises(sc.Sc3dError) as ei:
        op.apply_bunch(b, 1e-9)
    assert ei.value.code == sc.SC3D_ERR_DOMAIN
    assert "neglible extent" in str(ei.value)
    assert np.array_equal(b.host()[0], parts)
    # periodic_z throws (:594-598, :619-623)
    o = make_opts((32, 32, 64))
    o.periodic_z = 1
    op2 = sc.SpaceCharge3dOpenHockney(o)
    with pytest.raises(sc.Sc3dError) as ei:
        op2.apply_bunch(b, 1e-9)
    assert ei.value.code == sc.SC3D_ERR_PERIODIC
    assert "periodic_z not yet implemented" in str(ei.value)
    # empty bunch with a fixed domain: nothing happens, no failure
    e = sc.Bunch(0, 1e10, gamma=GAMMA, capacity=8)
    op3 = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64), fixed=((0, 0, 0), (1, 1, 1))))
    op3.apply_bunch(e, 1e-9)
    assert np.all(op3.debug(sc.DBG_RHO) == 0)
    # unsupported shapes: an axis below 4 or above 512 (doubled length 1024)
    for shape in ((2, 32, 64), (32, 600, 8)):
        with pytest.raises(sc.Sc3dError) as ei:
            sc.SpaceCharge3dOpenHockney(sc.default_options(*shape))
        assert ei.value.code == sc.SC3D_ERR_ARG and "[4, 512]" in str(ei.value)


def test_populate_bit_exact_vs_oracle(gpu):
    """k_populate (Philox-4x32-10, Box-Muller from basic operations, Cholesky
    factor) equals the oracle's restatement bit for bit, with a correlated
    factor, a mean, a 64-bit seed and ids above 2^32."""
    n = 300000
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6)) * np.array([1e-3, 2e-4, 3e-3, 4e-4, 1e-2, 1e-4])[:, None]
    L = np.linalg.cholesky(A @ A.T)
    chol = np.array([L[r, c] for r in range(6) for c in range(r + 1)])
    mean = np.array([1e-4, -2e-5, 3e-4, 0.0, -1e-3, 2e-6])
    seed, id0 = (0xfedcba98 << 32) | 0x76543210, (1 << 32) - 1000
    b = sc.Bunch(n, 1e10, gamma=GAMMA)
    b.populate_gaussian(mean, chol, seed=seed, id0=id0)
    got, m = b.host()
    want, _ = O.populate_gaussian(n, mean, chol, seed, id0=id0)
    assert np.all(m == 1)
    assert np.array_equal(got, want)
    # deterministic: a second fill of another bunch is the same
    b2 = sc.Bunch(n, 1e10, gamma=GAMMA)
    b2.populate_gaussian(mean, chol, seed=seed, id0=id0)
    assert np.array_equal(b2.host()[0], got)


def test_fp32_grid_field_accuracy(gpu):
    """FP32 deposit/FFT/field with FP64 particles: field rel-L2 vs FP64 oracle."""
    n = 100000
    shape = (32, 32, 64)
    parts = gaussian_parts(n, C1_SIGMA, seed=21)
    masks = np.ones(n, np.uint8)
    dt = 0.1 / (BETA * sc.C_LIGHT)
    rc, out = oracle_run(oracle_kw(shape), parts, masks, 5e10, dt,
                         want=("rho", "ex", "ey", "ez"))
    op, b = gpu_run(make_opts(shape, precision=sc.FP32_GRID), parts, masks, 5e10, dt)
    errs = [rel_l2(op.debug(t), out[nm]) for t, nm in
            ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez"))]
    rho = op.debug(sc.DBG_RHO)
    print("fp32 field rel-L2", errs, "rho", rel_l2(rho, out["rho"]))
    # the field is transformed straight from the spectrum: rounding relative
    # to the field, not to the potential (about 2e-7 at this size)
    assert max(errs) <= 1e-6
    with pytest.raises(RuntimeError):
        op.debug(sc.DBG_PHI)  # no potential in the precision form
    # compensated per-wavefront sums: the FP32 density is within a few FP32
    # roundings of the FP64 one
    assert rel_l2(rho, out["rho"]) <= 2e-7
    # the unit-peak scale is the peak of the unit-weight density
    dom = op.debug(sc.DBG_DOMAIN)
    peak = np.array([dom["rho_peak"]], np.int32).view(np.float32)[0]
    assert peak > 0
    assert abs(float(peak) * dom["w0"] / rho.max() - 1) <= 1e-12


def test_maps_parity(gpu):
    """Fused drift/quadrupole program vs the oracle's libFF restatement:
    bit-exact (same IEEE operations, sqrt and division correctly rounded)."""
    from synergia2_amd import stepper as S
    n = 50000
    b = sc.Bunch(n, 1e11, gamma=GAMMA)
    chol = S.matched_chol(S.fodo_channel(1), 1e-6, 0.01 / BETA, 1e-3)
    b.populate_gaussian(np.zeros(6), chol, seed=3)
    p0, m0 = b.host()
    spec = [(sc.MAP_DRIFT, 0.37, 0.0), (sc.MAP_QUADRUPOLE, 0.2, 0.6),
            (sc.MAP_DRIFT, 0.5, 0.0), (sc.MAP_QUADRUPOLE, 0.13, -0.6),
            (sc.MAP_QUADRUPOLE, 0.07, -0.6)]
    slices = sc.make_slices(spec, b.p, b.mass)
    for i, (kd, L, k1) in enumerate(spec):
        assert slices[0][i].ref_cdt == O.reference_cdt(kd, L, k1, 4, b.p, b.mass)
    sc.apply_maps(b, slices)
    pg, _ = b.host()
    po = np.ascontiguousarray(p0.copy())
    ob = O.make_bunch(po, np.ascontiguousarray(m0.copy()), n, 1e11, 1, b.mass, b.gamma, b.beta, b.p)
    O.apply_maps(ob, [(kd, L, k1, 4, slices[0][i].ref_cdt) for i, (kd, L, k1) in enumerate(spec)])
    for c in range(7):
        assert np.array_equal(pg[c], po[c]), c


def test_fused_kick_maps_equal_unfused(gpu):
    """Kick with the post-kick map program fused into the kick kernel gives
    the same bits as the kick followed by a separate map pass."""
    from synergia2_amd import stepper as S
    n = 40000
    chol = S.matched_chol(S.fodo_channel(1), 1e-6, 0.01 / BETA, 1e-4)
    prog_spec = [(sc.MAP_DRIFT, 0.05, 0.0), (sc.MAP_QUADRUPOLE, 0.06, 0.6),
                 (sc.MAP_DRIFT, 0.03, 0.0)]
    outs = []
    for fused in (True, False):
        b = sc.Bunch(n, 1e11, gamma=GAMMA)
        b.populate_gaussian(np.zeros(6), chol, seed=8)
        op = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64)))
        prog = sc.make_slices(prog_spec, b.p, b.mass)
        if fused:
            op.set_post_kick_maps(prog)
            op.apply_bunch(b, 1e-9)
        else:
            op.apply_bunch(b, 1e-9)
            sc.apply_maps(b, prog)
            torch_sync()
        outs.append(b.host()[0])
    # rho atomics order may differ between the two runs (last-bit level);
    # dp/p has an absolute floor of ~2e-16 from the energy-kick difference
    # sqrt(Ef^2-m^2) - sqrt(Ei^2-m^2) (space_charge_3d_kernels.h:170-177)
    a, c = outs
    for col in range(7):
        scale = max(np.abs(c[col]).max(), 1e-300)
        assert np.abs(a[col] - c[col]).max() <= 1e-12 * scale, col


def torch_sync():
    import torch
    torch.cuda.synchronize()


def test_propagation_parity(gpu):
    """Six split-operator steps through the FODO channel (fused maps, sort
    every second kick) vs the oracle running the same sequence."""
    from synergia2_amd import stepper as S
    n = 30000
    shape = (32, 32, 64)
    lattice = S.fodo_channel(4)
    chol = S.matched_chol(S.fodo_channel(1), 1e-6, 0.01 / BETA, 1e-4)
    b = sc.Bunch(n, 1e11, gamma=GAMMA)
    b.populate_gaussian(np.zeros(6), chol, seed=9)
    p0, m0 = b.host()
    op = sc.SpaceCharge3dOpenHockney(make_opts(shape, sort_period=2))
    prop = S.Propagator(lattice, 40, b, op)
    prop.run(6, sync=True)
    pg, _ = b.host()
    # oracle: identical step sequence
    po = np.ascontiguousarray(p0.copy())
    mo = np.ascontiguousarray(m0.copy())
    ob = O.make_bunch(po, mo, n, 1e11, 1, b.mass, b.gamma, b.beta, b.p)
    for k in range(6):
        first, second, length = prop.steps[k]
        O.apply_maps(ob, [(kd, L, k1, 4, O.reference_cdt(kd, L, k1, 4, b.p, b.mass))
                          for kd, L, k1 in first])
        o = O.make_options(shape, reorder=(k % 2 == 0))
        rc, _ = O.apply(o, ob, length / (b.beta * sc.C_LIGHT), nthreads=4)
        assert rc == 0
        O.apply_maps(ob, [(kd, L, k1, 4, O.reference_cdt(kd, L, k1, 4, b.p, b.mass))
                          for kd, L, k1 in second])
    g = by_id(pg)
    r = po[:, np.argsort(po[6], kind="stable")]
    assert np.array_equal(g[6], r[6])
    for col in range(6):
        scale = np.abs(r[col]).max()
        err = np.abs(g[col] - r[col]).max() / scale
        assert err <= 1e-12, (col, err)


def test_moments_bit_exact(gpu):
    """Device moments (mean, 6x6 second moments) equal the oracle's fixed-tree
    restatement bit for bit."""
    from synergia2_amd import stepper as S
    n = 70001
    b = sc.Bunch(n, 1e10, gamma=GAMMA)
    chol = S.matched_chol(S.fodo_channel(1), 1e-6, 0.01 / BETA, 1e-4)
    b.populate_gaussian(np.array([1e-4, 0, -2e-4, 1e-5, 0.3, 1e-5]), chol, seed=12)
    b.masks[5:50:3] = 0
    op = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64)))
    mean, mom2 = op.moments(b)
    p, m = b.host()
    om, o2 = O.moments(p, m, b.total_num)
    assert np.array_equal(mean, om)
    assert np.array_equal(mom2, o2)
    ex, ey, ez = sc.emittances(mom2)
    assert abs(ex / 1e-6 - 1) < 0.02 and abs(ey / 1e-6 - 1) < 0.02


def test_fp32_grid_emittance_tracks_fp64(gpu):
    """Precision form over 100 FODO steps: rms emittances within 1e-4
    relative of the FP64 form on the same seeded bunch (the FP64 form is
    itself pinned to the oracle by the parity tests above)."""
    from synergia2_amd import stepper as S
    n = 200000
    chol = S.matched_chol(S.fodo_channel(1), 1e-6, 0.01 / BETA, 1e-4)
    em = {}
    for prec in (sc.FP64, sc.FP32_GRID):
        b = sc.Bunch(n, 1e11 * n / 4e6, gamma=GAMMA)
        b.populate_gaussian(np.zeros(6), chol, seed=1)
        op = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64), precision=prec))
        prop = S.Propagator(S.fodo_channel(4), 40, b, op)
        e0 = sc.emittances(op.moments(b)[1])
        prop.run(100, sync=True)
        em[prec] = (e0, sc.emittances(op.moments(b)[1]))
    print("emittance FP64", em[sc.FP64], "FP32-grid", em[sc.FP32_GRID])
    for k in range(3):
        growth64 = em[sc.FP64][1][k] / em[sc.FP64][0][k]
        assert abs(em[sc.FP32_GRID][1][k] / em[sc.FP64][1][k] - 1) <= 1e-4, k
        assert growth64 > 0.9  # sanity: emittances stay finite/physical


def test_tail_fraction_exact(gpu):
    """Device tail counts equal numpy's on the same particles (integers)."""
    from synergia2_amd import stepper as S
    n = 300000
    chol = S.matched_chol_sigma(S.fodo_channel(1), 2e-3, 1.5e-3, 0.02 / BETA, 1e-4)
    parts = S.halo_mix_parts(n, chol, seed=3)
    masks = np.ones(n, np.uint8)
    masks[::97] = 0
    b = sc.Bunch(n, 1e11, gamma=GAMMA)
    b.set_host(parts, masks)
    op = sc.SpaceCharge3dOpenHockney(make_opts((32, 32, 64)))
    mean, mom2 = op.moments(b)
    width = np.sqrt(np.diag(mom2))
    for k in (2.0, 3.0, 4.0):
        got = op.tail_fraction(b, mean, width, k)
        on = masks.astype(bool)
        want = [np.count_nonzero(np.abs(parts[j][on] - mean[j]) > k * width[j]) / n
                for j in range(6)]
        assert np.array_equal(got, np.array(want)), (k, got, want)
    assert got[0] > 0  # the Gaussian core reaches beyond 4 sigma in x


def test_convolution_vs_rocfft(gpu):
    """The pruned Stockham passes with the real-octant Green's spectrum vs
    an independent library on the full doubled grid: torch.fft (hipFFT /
    rocFFT) computes phi = IDFT(DFT(rho padded) * Ghat) in FP64 on the
    device from this apply's rho and Green's spectrum; rel-L2 <= 1e-12 at
    64x64x128 (doubled 128x128x256)."""
    import torch
    gx, gy, gz = 64, 64, 128
    n = 400000
    parts = gaussian_parts(n, C1_SIGMA, seed=17)
    op, b = gpu_run(make_opts((gx, gy, gz)), parts, np.ones(n, np.uint8), 5e10,
                    0.1 / (BETA * sc.C_LIGHT))
    rho = torch.from_numpy(op.debug(sc.DBG_RHO).reshape(gz, gy, gx)).cuda()
    ghat = torch.from_numpy(op.debug(sc.DBG_GHAT).reshape(gz + 1, gy + 1, gx + 1)).cuda()
    phi = op.debug(sc.DBG_PHI)
    Dx, Dy, Dz = 2 * gx, 2 * gy, 2 * gz
    kz = torch.arange(Dz, device="cuda")
    ky = torch.arange(Dy, device="cuda")
    mz = torch.minimum(kz, Dz - kz)
    my = torch.minimum(ky, Dy - ky)
    gfull = ghat[mz][:, my]  # (Dz, Dy, gx+1): real, even in kz and ky
    pad = torch.zeros((Dz, Dy, Dx), dtype=torch.float64, device="cuda")
    pad[:gz, :gy, :gx] = rho
    spec = torch.fft.rfftn(pad) * gfull
    ref = torch.fft.irfftn(spec, s=(Dz, Dy, Dx)) * (Dx * Dy * Dz)  # unnormalised inverse
    ref = ref[:gz, :gy, :gx].cpu().numpy().ravel()
    err = rel_l2(phi, ref)
    print("phi vs rocFFT rel-L2", err)
    assert err <= 1e-12


def test_apply_after_failed_growth(gpu, monkeypatch):
    """A failed growth of the particle buffers (forced on the host with the
    library's SC3D_TEST_MAX_ALLOC_BYTES hook, before any device work) leaves
    no stale buffer behind: a later, smaller apply on the same context
    reallocates and matches the oracle."""
    shape = (32, 32, 64)
    dt = 0.1 / (BETA * sc.C_LIGHT)
    op = sc.SpaceCharge3dOpenHockney(make_opts(shape))
    first = gaussian_parts(60000, C1_SIGMA, seed=71)
    b1 = sc.Bunch(60000, 5e10, gamma=GAMMA)
    b1.set_host(first, np.ones(60000, np.uint8))
    op.apply_bunch(b1, dt)
    big = sc.Bunch(400000, 5e10, gamma=GAMMA)
    big.set_host(gaussian_parts(400000, C1_SIGMA, seed=72), np.ones(400000, np.uint8))
    monkeypatch.setenv("SC3D_TEST_MAX_ALLOC_BYTES", str(8 << 20))
    with pytest.raises(sc.Sc3dError):
        op.apply_bunch(big, dt)
    monkeypatch.delenv("SC3D_TEST_MAX_ALLOC_BYTES")
    with pytest.raises(RuntimeError):
        op.debug(sc.DBG_PERM, 60000)  # the old sort is gone with its buffers
    n = 30000
    parts = gaussian_parts(n, C1_SIGMA, seed=73)
    masks = np.ones(n, np.uint8)
    b = sc.Bunch(n, 5e10, gamma=GAMMA)
    b.set_host(parts, masks)
    op.apply_bunch(b, dt)
    rc, out = oracle_run(oracle_kw(shape), parts, masks, 5e10, dt, want=("perm", "rho"))
    assert rc == 0
    assert np.array_equal(op.debug(sc.DBG_PERM, n), out["perm"])
    assert rel_l2(op.debug(sc.DBG_RHO), out["rho"]) <= 1e-13
    pg, mg = b.host()
    check_kicked(pg, mg, out["parts"], parts)


def test_fixture_bunch_1k(gpu):
    """The HIP path against the committed fixture (tests/golden/bunch_1k.npz,
    the oracle's outputs on a seeded 1e3-particle bunch with masked and
    out-of-grid particles, 8x8x16): domain record, keys and permutation
    exactly, rho <= 1e-13, phi and E <= 1e-12, kicked bunch <= 1e-12 by id."""
    import os
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                              "bunch_1k.npz"))
    parts, masks = fx["parts_in"], fx["masks"]
    n = masks.shape[0]
    shape = tuple(int(x) for x in fx["shape"])
    op, b = gpu_run(make_opts(shape), parts, masks, float(fx["real_num"]), float(fx["time_step"]))
    dom = op.debug(sc.DBG_DOMAIN)
    for f in DOMAIN_FIELDS:
        assert np.array_equal(np.array(dom[f], np.float64), fx["domain_" + f]), f
    assert np.array_equal(op.debug(sc.DBG_KEYS, n), fx["keys"])
    assert np.array_equal(op.debug(sc.DBG_PERM, n), fx["perm"])
    assert rel_l2(op.debug(sc.DBG_RHO), fx["rho"]) <= 1e-13
    assert rel_l2(op.debug(sc.DBG_PHI), fx["phi"]) <= 1e-12
    for tap, name in ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez")):
        assert rel_l2(op.debug(tap), fx[name]) <= 1e-12, name
    pg, mg = b.host()
    check_kicked(pg, mg, fx["parts_out"], parts)


@pytest.mark.parametrize("green,gname", [(sc.GREEN_LINEAR, "linear"),
                                         (sc.GREEN_POINTLIKE, "pointlike")])
@pytest.mark.parametrize("shape", [(16, 32, 16), (32, 16, 8)])
def test_reference_g2_index_quirk_shapes(gpu, green, gname, shape):
    """At the shapes where alg_g2_*'s float-reciprocal index decomposition
    misplaces points (space_charge_3d_open_hockney.cc:170-176,272-278), the
    oracle replays it (serial write order) and the device adds the patch
    points' convolution: phi and E within 1e-12 of the oracle in the FP64
    form, E within 2e-6 in the precision form (bar 1e-5), kicks within
    1e-12."""
    n = 30000
    parts = gaussian_parts(n, (2e-3, 1e-4, 1e-3, 1e-4, 5e-3, 1e-4), seed=11)
    masks = np.ones(n, np.uint8)
    dt = 0.1 / (BETA * sc.C_LIGHT)
    print(f"case quirk {shape} {gname}", flush=True)
    rc, out = oracle_run(oracle_kw(shape, gname), parts, masks, 1e10, dt,
                         want=("rho", "phi", "ex", "ey", "ez"))
    assert rc == 0
    op, b = gpu_run(make_opts(shape, green), parts, masks, 1e10, dt)
    assert rel_l2(op.debug(sc.DBG_RHO), out["rho"]) <= 1e-13
    assert rel_l2(op.debug(sc.DBG_PHI), out["phi"]) <= 1e-12
    for tap, name in ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez")):
        assert rel_l2(op.debug(tap), out[name]) <= 1e-12, name
    pg, mg = b.host()
    check_kicked(pg, mg, out["parts"], parts)
    op32, _ = gpu_run(make_opts(shape, green, precision=sc.FP32_GRID), parts, masks, 1e10, dt)
    errs = [rel_l2(op32.debug(t), out[nm]) for t, nm in
            ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez"))]
    print(f"  precision form field rel-L2 {errs}", flush=True)
    assert max(errs) <= 2e-6


@pytest.mark.parametrize("shape", [(24, 24, 48), (48, 48, 96), (48, 64, 128), (12, 20, 37),
                                   (15, 16, 300), (320, 8, 8)])
def test_any_shape_vs_oracle(gpu, shape):
    """Shapes outside the tuned power-of-two transforms -- factors 3 and 5
    (24x24x48, 48x48x96), a mix of generic x and tuned y, z passes
    (48x64x128), a prime factor (Dz = 74 = 2 * 37), an odd gx with a 600-long
    z axis (15x16x300) and gx above 256 (Dx = 640) -- against the oracle, whose
    non-power-of-two transforms are direct DFTs pinned by numpy golden
    vectors.  At gx = 48 the reference's g2 index decomposition misplaces
    points in most rows of every plane (29376 / 63744 points), which the
    device applies as a correction spectrum.  Green's spectrum <= 1e-13 of
    the full FFT of the reference's g2 (where it is symmetric),
    rho <= 1e-13, phi and E <= 1e-12, kicks <= 1e-12; precision form field
    <= 2e-6 (bar 1e-5)."""
    n = 30000
    parts = gaussian_parts(n, (2e-3, 1e-4, 1.5e-3, 1e-4, 5e-3, 1e-4), seed=12)
    masks = np.ones(n, np.uint8)
    dt = 0.1 / (BETA * sc.C_LIGHT)
    print(f"case any shape {shape}", flush=True)
    rc, out = oracle_run(oracle_kw(shape), parts, masks, 1e10, dt,
                         want=("rho", "phi", "ex", "ey", "ez", "ghat_full"))
    assert rc == 0, O.last_error()
    op, b = gpu_run(make_opts(shape), parts, masks, 1e10, dt)
    gx, gy, gz = shape
    if len(sc.green_patches(sc.GREEN_LINEAR, shape)) == 0:
        # (at 48x48x96 and 48x64x128 the reference's g2 has misplaced points,
        # which the device adds as a separate correction spectrum)
        full = out["ghat_full"].view(np.complex128).reshape(2 * gz, 2 * gy, gx + 1)
        got = op.debug(sc.DBG_GHAT).reshape(gz + 1, gy + 1, gx + 1)
        assert rel_l2(got, full[: gz + 1, : gy + 1, :].real) <= 1e-13
    assert rel_l2(op.debug(sc.DBG_RHO), out["rho"]) <= 1e-13
    assert rel_l2(op.debug(sc.DBG_PHI), out["phi"]) <= 1e-12
    for tap, name in ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez")):
        assert rel_l2(op.debug(tap), out[name]) <= 1e-12, name
    pg, mg = b.host()
    check_kicked(pg, mg, out["parts"], parts)
    op32, _ = gpu_run(make_opts(shape, precision=sc.FP32_GRID), parts, masks, 1e10, dt)
    errs = [rel_l2(op32.debug(t), out[nm]) for t, nm in
            ((sc.DBG_EX, "ex"), (sc.DBG_EY, "ey"), (sc.DBG_EZ, "ez"))]
    print(f"  precision form field rel-L2 {errs}", flush=True)
    assert max(errs) <= 2e-6


def test_adjust_moments_bit_exact_vs_oracle(gpu):
    """adjust_moments (populate.cc:30-61) on the device equals the oracle's
    bit for bit (same fixed-tree moments, same Cholesky / inverse order, same
    per-particle sums), and the adjusted bunch has the requested mean and
    second moments; populate_6d = unit Gaussians + adjust_moments; the
    reference's symmetric-check error."""
    n = 200003
    rng = np.random.default_rng(8)
    A = rng.standard_normal((6, 6)) * np.array([1e-3, 2e-4, 3e-3, 4e-4, 1e-2, 1e-4])[:, None]
    cov = A @ A.T
    means = np.array([1e-4, -2e-5, 3e-4, 1e-6, -1e-3, 2e-6])
    b = sc.Bunch(n, 1e10, gamma=GAMMA, capacity=n + 11)
    b.populate_gaussian(np.zeros(6), sc.diag_chol([1.0] * 6), seed=9)
    b.masks[::13] = 0  # lost particles: the total number counts the active ones
    b.total_num = float(n - len(range(0, n, 13)))
    p0, m0 = b.host()
    b.adjust_moments(means, cov)
    pg, _ = b.host()
    po = np.ascontiguousarray(p0.copy())
    assert O.adjust_moments(po, np.ascontiguousarray(m0), means, cov, total_num=b.total_num) == 0
    assert np.array_equal(pg.view(np.int64), po.view(np.int64))
    op = sc.SpaceCharge3dOpenHockney(sc.default_options(32, 32, 64))
    mean, mom2 = op.moments(b)
    assert np.allclose(mean, means, rtol=1e-12, atol=1e-16)
    assert np.allclose(mom2, cov, rtol=1e-11, atol=1e-13 * np.abs(cov).max())
    b2 = sc.Bunch(n, 1e10, gamma=GAMMA)
    b2.populate_6d(means, cov, seed=9)
    mean2, mom22 = op.moments(b2)
    assert np.allclose(mean2, means, rtol=1e-12, atol=1e-16)
    assert np.allclose(mom22, cov, rtol=1e-11, atol=1e-13 * np.abs(cov).max())
    bad = cov.copy()
    bad[0, 1] *= 1 + 1e-10
    with pytest.raises(sc.Sc3dError) as ei:
        b2.adjust_moments(means, bad)
    assert "adjust_moments: covariance matrix must be symmetric" in str(ei.value)
