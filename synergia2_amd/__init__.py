"""synergia2_amd -- MI355X-native 3D open-boundary Hockney space-charge step.

Python binding of the C ABI in include/sc3d_capi.h (libsc3d_hip.so) over
ctypes.  Device memory comes from torch (ROCm) tensors; the library only sees
raw device pointers.  There is no CPU fallback: if the HIP library is not
built, load() raises.

Mirrors the reference operator interface:
  Space_charge_3d_open_hockney_options  (implemented_collective_options.h:67-126)
  Space_charge_3d_open_hockney          (space_charge_3d_open_hockney.h:16-73)
  Bunch columns x, xp, y, yp, cdt, dpop, id (bunch.h:66-74)
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsc3d_hip.so")

# physical_constants.h:17-24 (PDG 2012), :94-106
PROTON_MASS = 0.938272046
E_CHARGE = 1.602176565e-19
C_LIGHT = 299792458.0

# bunch.h:66-74
X, XP, Y, YP, CDT, DPOP, ID = range(7)

SC3D_OK = 0
SC3D_ERR_ARG = -1
SC3D_ERR_HIP = -2
SC3D_ERR_DOMAIN = -3
SC3D_ERR_PERIODIC = -4
SC3D_ERR_SIZE = -5

GREEN_POINTLIKE = 0
GREEN_LINEAR = 1
FP64 = 0
FP32_GRID = 1

DBG_DOMAIN, DBG_KEYS, DBG_PERM, DBG_RHO, DBG_PHI, DBG_EX, DBG_EY, DBG_EZ, DBG_GHAT = range(9)


class Sc3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("shape", C.c_int32 * 3),
        ("size", C.c_double * 3),
        ("offset", C.c_double * 3),
        ("green_fn", C.c_int32),
        ("periodic_z", C.c_int32),
        ("grid_entire_period", C.c_int32),
        ("domain_fixed", C.c_int32),
        ("comm_group_size", C.c_int32),
        ("z_period", C.c_double),
        ("n_sigma", C.c_double),
        ("kick_scale", C.c_double),
        ("precision", C.c_int32),
        ("sort_period", C.c_int32),
    ]


class BunchDesc(C.Structure):
    _fields_ = [
        ("parts", C.c_void_p),
        ("stride", C.c_int64),
        ("n_active", C.c_int64),
        ("masks", C.c_void_p),
        ("total_num", C.c_double),
        ("real_num", C.c_double),
        ("charge", C.c_int32),
        ("mass", C.c_double),
        ("ref_gamma", C.c_double),
        ("ref_beta", C.c_double),
        ("ref_p", C.c_double),
    ]


class DomainRecord(C.Structure):
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
        ("inv_h", C.c_double * 3),
        ("half_inv_h", C.c_double * 3),
        ("ghat_h", C.c_double * 3),
        ("error", C.c_int32),
        ("ghat_dirty", C.c_int32),
        ("ghat_valid", C.c_int32),
        ("rho_peak", C.c_int32),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class MapSlice(C.Structure):
    _fields_ = [("kind", C.c_int32), ("yoshida_steps", C.c_int32), ("length", C.c_double),
                ("k1", C.c_double), ("k1s", C.c_double), ("ref_cdt", C.c_double)]


MAP_DRIFT = 0
MAP_QUADRUPOLE = 1
MAX_SLICES = 16
PHASES = ("sc3d_domain", "sc3d_sort", "sc3d_local_rho", "sc3d_green_fn2", "sc3d_local_f",
          "sc3d_force", "sc3d_kick")
TIMING_OFF, TIMING_PHASES, TIMING_APPLY = 0, 1, 2

EXPORTS = ("sc3d_options_default", "sc3d_create", "sc3d_apply", "sc3d_sync", "sc3d_destroy",
           "sc3d_last_error", "sc3d_debug_get", "sc3d_populate_gaussian", "sc3d_reference_cdt",
           "sc3d_apply_maps", "sc3d_set_timing", "sc3d_get_timing", "sc3d_get_apply_timing",
           "sc3d_set_post_kick_maps",
           "sc3d_moments", "sc3d_tail_fraction", "sc3d_green_patches",
           "sc3d_moments_noalloc", "sc3d_adjust_moments",
           # served form (include/sc3d_serve.h)
           "sc3d_server_options_default", "sc3d_server_create", "sc3d_server_submit",
           "sc3d_server_iterate", "sc3d_server_job", "sc3d_server_job_particles",
           "sc3d_server_get_stats", "sc3d_server_destroy", "sc3d_job_footprint")

_lib = None


def load(path=None):
    """Load libsc3d_hip.so (raises if it is not built; there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is not built; run __graft_entry__.build() or "
                           "python -m synergia2_amd.build")
    L = C.CDLL(p)
    L.sc3d_options_default.argtypes = [C.POINTER(Options), C.c_int32, C.c_int32, C.c_int32]
    L.sc3d_options_default.restype = None
    L.sc3d_create.argtypes = [C.POINTER(Options), C.POINTER(C.c_void_p)]
    L.sc3d_apply.argtypes = [C.c_void_p, C.POINTER(BunchDesc), C.c_double, C.c_void_p]
    L.sc3d_sync.argtypes = [C.c_void_p]
    L.sc3d_destroy.argtypes = [C.c_void_p]
    L.sc3d_last_error.restype = C.c_char_p
    L.sc3d_debug_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
    L.sc3d_populate_gaussian.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.c_uint64, C.c_int64, C.c_void_p]
    L.sc3d_reference_cdt.argtypes = [C.POINTER(MapSlice), C.c_double, C.c_double]
    L.sc3d_reference_cdt.restype = C.c_double
    L.sc3d_apply_maps.argtypes = [C.POINTER(BunchDesc), C.POINTER(MapSlice), C.c_int, C.c_void_p]
    L.sc3d_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.sc3d_set_post_kick_maps.argtypes = [C.c_void_p, C.POINTER(MapSlice), C.c_int]
    L.sc3d_moments.argtypes = [C.c_void_p, C.POINTER(BunchDesc), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.c_void_p]
    L.sc3d_get_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                  C.c_int]
    L.sc3d_get_apply_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_int64), C.c_int]
    L.sc3d_tail_fraction.argtypes = [C.c_void_p, C.POINTER(BunchDesc), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.c_double,
                                     C.POINTER(C.c_double), C.c_void_p]
    L.sc3d_green_patches.argtypes = [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.c_int32]
    L.sc3d_green_patches.restype = C.c_int
    L.sc3d_moments_noalloc.argtypes = [C.POINTER(BunchDesc), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.c_void_p]
    L.sc3d_moments_noalloc.restype = C.c_int
    L.sc3d_adjust_moments.argtypes = [C.POINTER(BunchDesc), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_void_p]
    L.sc3d_adjust_moments.restype = C.c_int
    for name in ("sc3d_create", "sc3d_apply", "sc3d_sync", "sc3d_destroy", "sc3d_debug_get",
                 "sc3d_populate_gaussian", "sc3d_apply_maps", "sc3d_set_timing",
                 "sc3d_get_timing", "sc3d_get_apply_timing", "sc3d_set_post_kick_maps",
                 "sc3d_moments",
                 "sc3d_tail_fraction"):
        getattr(L, name).restype = C.c_int
    if path is None:
        _lib = L
    return L


def _check(rc):
    if rc != SC3D_OK:
        raise Sc3dError(rc, load().sc3d_last_error().decode())


def default_options(gx=32, gy=32, gz=64):
    """Space_charge_3d_open_hockney_options(gridx, gridy, gridz) defaults."""
    o = Options()
    load().sc3d_options_default(C.byref(o), gx, gy, gz)
    return o


def set_fixed_domain(opts, offset, size):
    """Space_charge_3d_open_hockney_options::set_fixed_domain
    (implemented_collective_options.h:97-109); its grid_entire_period branch
    is overwritten by the assignments that follow it, so it has no effect."""
    opts.offset[:] = list(offset)
    opts.size[:] = list(size)
    opts.domain_fixed = 1


def green_patches(green_fn, shape):
    """The points where the reference's g2 index decomposition leaves another
    value than the symmetric one at this shape (host only): an (n, 6) int32
    array of target (z, y, x) and the writer's (iz, iy, ix), -1 for none."""
    L = load()
    shp = (C.c_int32 * 3)(*shape)
    n = L.sc3d_green_patches(int(green_fn), shp, None, 0)
    if n < 0:
        _check(n)
    out = np.zeros((max(n, 1), 6), np.int32)
    L.sc3d_green_patches(int(green_fn), shp, out.ctypes.data_as(C.POINTER(C.c_int32)), n)
    return out[:n]


def kinematics(mass, gamma):
    """Four_momentum::update_from_gamma (four_momentum.cc:9-17)."""
    beta = np.sqrt(1.0 - 1.0 / (gamma * gamma))
    return float(beta), float(gamma * beta * mass)


class Bunch:
    """Device bunch: a (7, capacity) float64 torch tensor (column j contiguous,
    the reference's LayoutLeft particle view) and a uint8 mask per slot."""

    def __init__(self, n, real_num, mass=PROTON_MASS, charge=1, gamma=None,
                 kinetic_energy=None, capacity=None, device="cuda"):
        import torch
        if gamma is None:
            gamma = (mass + kinetic_energy) / mass
        self.mass = float(mass)
        self.charge = int(charge)
        self.gamma = float(gamma)
        self.beta, self.p = kinematics(self.mass, self.gamma)
        self.n = int(n)
        self.capacity = int(capacity or n)
        self.total_num = float(n)
        self.real_num = float(real_num)
        self.parts = torch.zeros((7, max(self.capacity, 1)), dtype=torch.float64, device=device)
        self.masks = torch.ones(max(self.capacity, 1), dtype=torch.uint8, device=device)

    def desc(self):
        d = BunchDesc()
        d.parts = self.parts.data_ptr()
        d.stride = self.parts.shape[1]
        d.n_active = self.n
        d.masks = self.masks.data_ptr()
        d.total_num = self.total_num
        d.real_num = self.real_num
        d.charge = self.charge
        d.mass = self.mass
        d.ref_gamma = self.gamma
        d.ref_beta = self.beta
        d.ref_p = self.p
        return d

    def set_host(self, parts_np, masks_np=None):
        import torch
        n = parts_np.shape[1]
        self.parts[:, :n].copy_(torch.from_numpy(np.ascontiguousarray(parts_np)))
        if masks_np is not None:
            self.masks[:n].copy_(torch.from_numpy(np.ascontiguousarray(masks_np)))

    def host(self):
        return self.parts[:, :self.n].cpu().numpy().copy(), self.masks[:self.n].cpu().numpy().copy()

    def adjust_moments(self, means, cov, stream=None):
        """adjust_moments (populate.cc:30-61): make the bunch's mean and second
        moments exactly means (6) and cov (6x6)."""
        means = np.ascontiguousarray(means, np.float64)
        cov = np.ascontiguousarray(cov, np.float64).reshape(36)
        d = self.desc()
        _check(load().sc3d_adjust_moments(C.byref(d), means.ctypes.data_as(C.POINTER(C.c_double)),
                                          cov.ctypes.data_as(C.POINTER(C.c_double)),
                                          _stream_ptr(stream)))

    def populate_6d(self, means, cov, seed, stream=None):
        """populate_6d (populate.cc:128-145): unit 6D Gaussians (the
        counter-based generator in place of GSL's), then adjust_moments to
        means and cov."""
        self.populate_gaussian(np.zeros(6), diag_chol([1.0] * 6), seed, stream=stream)
        self.adjust_moments(means, cov, stream)

    def populate_gaussian(self, mean, chol, seed, id0=0, stream=None):
        mean = np.ascontiguousarray(mean, np.float64)
        chol = np.ascontiguousarray(chol, np.float64)
        assert mean.shape == (6,) and chol.shape == (21,)
        _check(load().sc3d_populate_gaussian(
            self.parts.data_ptr(), self.parts.shape[1], self.n, self.masks.data_ptr(),
            mean.ctypes.data_as(C.POINTER(C.c_double)), chol.ctypes.data_as(C.POINTER(C.c_double)),
            int(seed), int(id0), _stream_ptr(stream)))


def _stream_ptr(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


def diag_chol(sigmas):
    """Packed lower-triangular factor of an uncorrelated 6D Gaussian."""
    L = np.zeros(21)
    for r in range(6):
        L[r * (r + 1) // 2 + r] = sigmas[r]
    return L


class SpaceCharge3dOpenHockney:
    """Space_charge_3d_open_hockney over the C ABI: one context (device
    workspaces + cached Green's spectrum) per operator instance."""

    def __init__(self, options):
        self.options = options
        self._ctx = C.c_void_p()
        _check(load().sc3d_create(C.byref(options), C.byref(self._ctx)))
        g = list(options.shape)
        self.G = g[0] * g[1] * g[2]
        self.ghat_elems = (g[2] + 1) * (g[1] + 1) * (g[0] + 1)

    def apply_bunch(self, bunch, time_step, stream=None, sync=True):
        d = bunch.desc()
        _check(load().sc3d_apply(self._ctx, C.byref(d), float(time_step), _stream_ptr(stream)))
        if sync:
            self.sync()

    def sync(self):
        _check(load().sc3d_sync(self._ctx))

    def debug(self, what, n=None):
        if what == DBG_DOMAIN:
            rec = DomainRecord()
            _check(load().sc3d_debug_get(self._ctx, what, C.byref(rec), C.sizeof(rec)))
            return rec.as_dict()
        if what in (DBG_KEYS, DBG_PERM):
            out = np.zeros(n, np.int32)
        elif what == DBG_GHAT:
            out = np.zeros(self.ghat_elems)
        else:
            out = np.zeros(self.G)
        _check(load().sc3d_debug_get(self._ctx, what, out.ctypes.data, out.nbytes))
        return out

    def set_post_kick_maps(self, slices):
        """Fuse a map program (from make_slices, or None) into the kick."""
        if slices is None:
            _check(load().sc3d_set_post_kick_maps(self._ctx, None, 0))
            return
        arr, n = slices
        _check(load().sc3d_set_post_kick_maps(self._ctx, arr, n))

    def moments(self, bunch, stream=None):
        """(mean[6], mom2[6,6]) of the bunch, computed on the device."""
        mean = np.zeros(6)
        mom2 = np.zeros(36)
        d = bunch.desc()
        _check(load().sc3d_moments(self._ctx, C.byref(d), mean.ctypes.data_as(C.POINTER(C.c_double)),
                                   mom2.ctypes.data_as(C.POINTER(C.c_double)), _stream_ptr(stream)))
        return mean, mom2.reshape(6, 6)

    def tail_fraction(self, bunch, center, width, nsigma=4.0, stream=None):
        """Per-plane fraction of active particles with |p - center| >
        nsigma * width (device counts, exact)."""
        c = np.ascontiguousarray(center, np.float64)
        w = np.ascontiguousarray(width, np.float64)
        out = np.zeros(6)
        d = bunch.desc()
        _check(load().sc3d_tail_fraction(self._ctx, C.byref(d),
                                         c.ctypes.data_as(C.POINTER(C.c_double)),
                                         w.ctypes.data_as(C.POINTER(C.c_double)), float(nsigma),
                                         out.ctypes.data_as(C.POINTER(C.c_double)),
                                         _stream_ptr(stream)))
        return out

    def set_timing(self, mode=TIMING_PHASES):
        """TIMING_OFF, TIMING_PHASES (timing()) or TIMING_APPLY (apply_timing());
        True / False for phases / off."""
        _check(load().sc3d_set_timing(self._ctx, int(mode)))

    def timing(self, reset=False):
        """(dict phase -> total ms, number of timed applies)."""
        ms = (C.c_double * len(PHASES))()
        n = C.c_int64()
        _check(load().sc3d_get_timing(self._ctx, ms, C.byref(n), int(reset)))
        return dict(zip(PHASES, list(ms))), n.value

    def apply_timing(self, reset=False):
        """(total ms of the whole applies timed, their number)."""
        ms = C.c_double()
        n = C.c_int64()
        _check(load().sc3d_get_apply_timing(self._ctx, C.byref(ms), C.byref(n), int(reset)))
        return ms.value, n.value

    def close(self):
        if self._ctx:
            load().sc3d_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_slices(spec, ref_p, mass):
    """spec: iterable of (kind, length, k1[, yoshida_steps]) -> MapSlice array
    with the design reference cdt of every slice filled in."""
    spec = list(spec)
    arr = (MapSlice * max(len(spec), 1))()
    for i, item in enumerate(spec):
        kind, length, k1 = item[:3]
        steps = item[3] if len(item) > 3 else 4
        arr[i].kind = kind
        arr[i].length = float(length)
        arr[i].k1 = float(k1)
        arr[i].k1s = 0.0
        arr[i].yoshida_steps = int(steps)
        arr[i].ref_cdt = 0.0
        arr[i].ref_cdt = load().sc3d_reference_cdt(C.byref(arr[i]), float(ref_p), float(mass))
    return arr, len(spec)


def apply_maps(bunch, slices, stream=None):
    """Apply a map program (from make_slices) to the bunch on the device."""
    arr, n = slices
    if n == 0:
        return
    d = bunch.desc()
    _check(load().sc3d_apply_maps(C.byref(d), arr, n, _stream_ptr(stream)))


def emittances(mom2):
    """rms emittances (x, y, z) = sqrt(det) of the 2x2 blocks of the second
    moments (calculate_emittances, diagnostics_full2_host.cc:20-40), with
    small negatives (> -1e-15) clamped to 0 as the reference does."""
    out = []
    for k in (0, 2, 4):
        det = mom2[k, k] * mom2[k + 1, k + 1] - mom2[k + 1, k] * mom2[k, k + 1]
        if -1e-15 < det < 0:
            det = 0.0
        out.append(float(np.sqrt(det)))
    return out
