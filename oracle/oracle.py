"""ctypes wrapper over liboracle_sc3d.so, the CPU restatement of the reference
3D open-Hockney step.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by the cpu_baseline leg of
bench.py as the checker / CPU baseline.  The product path never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_sc3d.so")


def build():
    """Compile the oracle with its Makefile (g++, OpenMP, no FMA contraction)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle_sc3d.so"], check=True)


class Options(C.Structure):
    _fields_ = [
        ("shape", C.c_int32 * 3),
        ("size", C.c_double * 3),
        ("offset", C.c_double * 3),
        ("green_fn", C.c_int32),
        ("periodic_z", C.c_int32),
        ("grid_entire_period", C.c_int32),
        ("domain_fixed", C.c_int32),
        ("z_period", C.c_double),
        ("n_sigma", C.c_double),
        ("kick_scale", C.c_double),
        ("reorder", C.c_int32),
    ]


class Bunch(C.Structure):
    _fields_ = [
        ("parts", C.POINTER(C.c_double)),
        ("stride", C.c_int64),
        ("n_active", C.c_int64),
        ("masks", C.POINTER(C.c_uint8)),
        ("total_num", C.c_double),
        ("real_num", C.c_double),
        ("charge", C.c_int32),
        ("mass", C.c_double),
        ("ref_gamma", C.c_double),
        ("ref_beta", C.c_double),
        ("ref_p", C.c_double),
    ]


class Domain(C.Structure):
    _fields_ = [
        ("sums", C.c_double * 6),
        ("mean", C.c_double * 3),
        ("stddev", C.c_double * 3),
        ("size", C.c_double * 3),
        ("offset", C.c_double * 3),
        ("left", C.c_double * 3),
        ("h", C.c_double * 3),
        ("w0", C.c_double),
        ("fn_norm", C.c_double),
        ("factor", C.c_double),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else float(v)
        return out


class Slice(C.Structure):
    _fields_ = [("kind", C.c_int32), ("yoshida_steps", C.c_int32), ("length", C.c_double),
                ("k1", C.c_double), ("k1s", C.c_double), ("ref_cdt", C.c_double)]


class Debug(C.Structure):
    _fields_ = [
        ("domain", C.POINTER(Domain)),
        ("keys", C.POINTER(C.c_int32)),
        ("perm", C.POINTER(C.c_int32)),
        ("rho", C.POINTER(C.c_double)),
        ("phi", C.POINTER(C.c_double)),
        ("ex", C.POINTER(C.c_double)),
        ("ey", C.POINTER(C.c_double)),
        ("ez", C.POINTER(C.c_double)),
        ("rho2_full", C.POINTER(C.c_double)),
        ("g2_full", C.POINTER(C.c_double)),
        ("ghat_full", C.POINTER(C.c_double)),
        ("seconds", C.c_double * 6),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_apply.argtypes = [C.POINTER(Options), C.POINTER(Bunch), C.c_double, C.c_int, C.POINTER(Debug)]
        L.orc_apply.restype = C.c_int
        L.orc_last_error.restype = C.c_char_p
        L.orc_tree_sums.argtypes = [dp, C.c_int64, C.c_int64, C.POINTER(C.c_uint8), dp]
        L.orc_compute_domain.argtypes = [C.POINTER(Options), C.POINTER(Bunch), C.c_double, C.POINTER(Domain)]
        L.orc_compute_domain.restype = C.c_int
        L.orc_cell_index.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_cell_index.restype = C.c_int32
        L.orc_cell_keys.argtypes = [C.POINTER(Bunch), C.POINTER(Domain), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_stable_sort.argtypes = [C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32)]
        L.orc_fft1d.argtypes = [dp, C.c_int, C.c_int]
        L.orc_fft3d_r2c.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_fft3d_c2r.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_green_fn2.argtypes = [C.c_int, C.POINTER(C.c_int32), dp, dp]
        L.orc_deposit.argtypes = [C.POINTER(Bunch), C.POINTER(Domain), C.POINTER(C.c_int32), dp, C.c_int]
        L.orc_force.argtypes = [dp, C.POINTER(C.c_int32), dp, dp, dp, dp]
        L.orc_kick.argtypes = [C.POINTER(Bunch), dp, dp, dp, C.POINTER(C.c_int32), C.POINTER(Domain)]
        L.orc_reference_cdt.argtypes = [C.POINTER(Slice), C.c_double, C.c_double]
        L.orc_reference_cdt.restype = C.c_double
        L.orc_apply_maps.argtypes = [C.POINTER(Bunch), C.POINTER(Slice), C.c_int]
        L.orc_moments.argtypes = [dp, C.c_int64, C.c_int64, C.POINTER(C.c_uint8), C.c_double, dp, dp]
        u32p = C.POINTER(C.c_uint32)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_det_log.argtypes = [C.c_double]
        L.orc_det_log.restype = C.c_double
        L.orc_det_sincos2pi.argtypes = [C.c_double, dp, dp]
        L.orc_populate_gaussian.argtypes = [dp, C.c_int64, C.c_int64, C.POINTER(C.c_uint8), dp, dp,
                                            C.c_uint64, C.c_int64]
        L.orc_adjust_matrix.argtypes = [dp, dp, dp]
        L.orc_adjust_moments.argtypes = [dp, C.c_int64, C.c_int64, C.POINTER(C.c_uint8),
                                         C.c_double, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_options(shape, green_fn="linear", n_sigma=8.0, kick_scale=1.0, fixed=None,
                 periodic_z=False, grid_entire_period=False, z_period=0.0, reorder=False):
    o = Options()
    o.reorder = int(reorder)
    o.shape[:] = list(shape)
    o.green_fn = 1 if green_fn == "linear" else 0
    o.periodic_z = int(periodic_z)
    o.grid_entire_period = int(grid_entire_period)
    o.z_period = z_period
    o.n_sigma = n_sigma
    o.kick_scale = kick_scale
    if fixed is not None:
        offset, size = fixed
        o.offset[:] = list(offset)
        o.size[:] = list(size)
        o.domain_fixed = 1
    return o


def make_bunch(parts, masks, total_num, real_num, charge, mass, gamma, beta, p):
    """parts: float64 array of shape (7, capacity) (column j contiguous)."""
    assert parts.dtype == np.float64 and parts.flags.c_contiguous and parts.shape[0] == 7
    assert masks.dtype == np.uint8
    b = Bunch()
    b.parts = _dp(parts)
    b.stride = parts.shape[1]
    b.n_active = masks.shape[0]
    b.masks = masks.ctypes.data_as(C.POINTER(C.c_uint8))
    b.total_num = float(total_num)
    b.real_num = float(real_num)
    b.charge = int(charge)
    b.mass, b.ref_gamma, b.ref_beta, b.ref_p = float(mass), float(gamma), float(beta), float(p)
    b._keep = (parts, masks)
    return b


def apply(opts, bunch, time_step, nthreads=0, want=()):
    """Run one operator application in place on the bunch's host particles.

    want: names of debug outputs to return (domain, keys, perm, rho, phi, ex,
    ey, ez, rho2_full, g2_full, ghat_full).  Returns (rc, dict)."""
    g = list(opts.shape)
    G = g[0] * g[1] * g[2]
    D = [2 * g[0], 2 * g[1], 2 * g[2]]
    px = 2 * (D[0] // 2 + 1)
    tot = px * D[1] * D[2]
    n = bunch.n_active
    dbg = Debug()
    out = {}
    dom = Domain()
    dbg.domain = C.pointer(dom)
    for name, size, dt in (("keys", n, np.int32), ("perm", n, np.int32), ("rho", G, np.float64),
                           ("phi", G, np.float64), ("ex", G, np.float64), ("ey", G, np.float64),
                           ("ez", G, np.float64), ("rho2_full", tot, np.float64),
                           ("g2_full", tot, np.float64), ("ghat_full", tot, np.float64)):
        if name in want:
            arr = np.zeros(size, dt)
            out[name] = arr
            setattr(dbg, name, _ip(arr) if dt == np.int32 else _dp(arr))
    rc = lib().orc_apply(C.byref(opts), C.byref(bunch), float(time_step), int(nthreads), C.byref(dbg))
    out["domain"] = dom.as_dict()
    out["seconds"] = list(dbg.seconds)
    if rc != 0:
        out["error"] = lib().orc_last_error().decode()
    return rc, out


def last_error():
    return lib().orc_last_error().decode()


def tree_sums(parts, masks):
    out = np.zeros(6)
    lib().orc_tree_sums(_dp(parts), parts.shape[1], masks.shape[0],
                        masks.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(out))
    return out


def compute_domain(opts, bunch, time_step):
    d = Domain()
    rc = lib().orc_compute_domain(C.byref(opts), C.byref(bunch), float(time_step), C.byref(d))
    return rc, d.as_dict()


def cell_index(pos, left, inv_h):
    off = C.c_double(0.0)
    idx = lib().orc_cell_index(float(pos), float(left), float(inv_h), C.byref(off))
    return idx, off.value


def fft1d(x, sign):
    a = np.ascontiguousarray(np.asarray(x, np.complex128)).copy()
    lib().orc_fft1d(a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), a.shape[0], int(sign))
    return a


def fft3d_r2c(real_padded, nx, ny, nz, nthreads=0):
    """real_padded: float64 (nz, ny, 2*(nx//2+1)). Returns complex (nz, ny, nx//2+1)."""
    inp = np.ascontiguousarray(real_padded, np.float64)
    out = np.zeros(inp.size, np.float64)
    lib().orc_fft3d_r2c(nx, ny, nz, _dp(inp), _dp(out), int(nthreads) or os.cpu_count())
    return out.view(np.complex128).reshape(nz, ny, nx // 2 + 1)


def fft3d_c2r(cplx, nx, ny, nz, nthreads=0):
    inp = np.ascontiguousarray(cplx, np.complex128).view(np.float64).ravel()
    out = np.zeros(inp.size, np.float64)
    lib().orc_fft3d_c2r(nx, ny, nz, _dp(inp), _dp(out), int(nthreads) or os.cpu_count())
    return out.reshape(nz, ny, 2 * (nx // 2 + 1))


def green_fn2(green_fn, g, h):
    g32 = np.array(g, np.int32)
    D = [2 * x for x in g]
    px = 2 * (D[0] // 2 + 1)
    out = np.zeros(px * D[1] * D[2])
    lib().orc_green_fn2(1 if green_fn == "linear" else 0, _ip(g32), _dp(np.array(h, np.float64)), _dp(out))
    return out.reshape(D[2], D[1], px)


def force(phi2_padded, g, h):
    G = g[0] * g[1] * g[2]
    ex, ey, ez = np.zeros(G), np.zeros(G), np.zeros(G)
    phi = np.ascontiguousarray(phi2_padded, np.float64)
    lib().orc_force(_dp(phi), _ip(np.array(g, np.int32)), _dp(np.array(h, np.float64)), _dp(ex), _dp(ey), _dp(ez))
    return ex, ey, ez


def make_slices(spec):
    """spec: list of (kind, length, k1, yoshida_steps, ref_cdt)."""
    arr = (Slice * len(spec))()
    for i, (kind, length, k1, steps, ref_cdt) in enumerate(spec):
        arr[i].kind = kind
        arr[i].length = length
        arr[i].k1 = k1
        arr[i].k1s = 0.0
        arr[i].yoshida_steps = steps
        arr[i].ref_cdt = ref_cdt
    return arr


def reference_cdt(kind, length, k1, steps, ref_p, mass):
    s = make_slices([(kind, length, k1, steps, 0.0)])
    return lib().orc_reference_cdt(s, float(ref_p), float(mass))


def apply_maps(bunch, spec):
    s = make_slices(spec)
    lib().orc_apply_maps(C.byref(bunch), s, len(spec))


def moments(parts, masks, total_num=None):
    """(mean[6], mom2[6,6]) with the fixed tree."""
    parts = np.ascontiguousarray(parts)
    n = masks.shape[0]
    mean = np.zeros(6)
    mom2 = np.zeros(36)
    lib().orc_moments(_dp(parts), parts.shape[1], n, masks.ctypes.data_as(C.POINTER(C.c_uint8)),
                      float(total_num if total_num is not None else n), _dp(mean), _dp(mom2))
    return mean, mom2.reshape(6, 6)


def adjust_matrix(cov, mom2):
    """(rc, A): A = G H^-1, G and H the lower Cholesky factors of cov and
    mom2 (adjust_moments_host, populate_host.cc:8-44); rc -1 not symmetric,
    -2 not positive definite."""
    A = np.zeros(36)
    rc = lib().orc_adjust_matrix(_dp(np.ascontiguousarray(cov, np.float64).ravel()),
                                 _dp(np.ascontiguousarray(mom2, np.float64).ravel()), _dp(A))
    return rc, A.reshape(6, 6)


def adjust_moments(parts, masks, means, cov, total_num=None):
    """adjust_moments (populate.cc:30-61) in place on parts (7, capacity):
    the first len(masks) rows are mapped so that the bunch's mean and second
    moments become means and cov.  Returns rc (0, -1, -2)."""
    n = masks.shape[0]
    return lib().orc_adjust_moments(_dp(parts), parts.shape[1], n,
                                     masks.ctypes.data_as(C.POINTER(C.c_uint8)),
                                     float(total_num if total_num is not None else n),
                                     _dp(np.ascontiguousarray(means, np.float64)),
                                     _dp(np.ascontiguousarray(cov, np.float64).ravel()))


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, out)
    return [int(x) for x in out]


def det_log(u):
    return lib().orc_det_log(float(u))


def det_sincos2pi(u):
    s, c = C.c_double(), C.c_double()
    lib().orc_det_sincos2pi(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def populate_gaussian(n, mean, chol, seed, id0=0, capacity=None):
    """(parts (7, capacity), masks (n,)) of the counter-based Gaussian
    populate: chol is the packed lower-triangular factor (21 entries)."""
    cap = capacity or n
    parts = np.zeros((7, cap))
    masks = np.zeros(n, np.uint8)
    lib().orc_populate_gaussian(_dp(parts), cap, n, masks.ctypes.data_as(C.POINTER(C.c_uint8)),
                                _dp(np.ascontiguousarray(mean, np.float64)),
                                _dp(np.ascontiguousarray(chol, np.float64)), C.c_uint64(seed),
                                int(id0))
    return parts, masks
