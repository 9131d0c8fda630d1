"""Python binding of the served form (include/sc3d_serve.h) and the job
builder for a FODO channel.

A job is one bunch tracked through the split-operator steps of a lattice
(Propagator::propagate / do_step, simulation/propagator.cc:36-118,157-283)
with the space-charge kick of each step; the server keeps many such jobs
resident on the device and advances all of them one step per iteration.
"""
import ctypes as C
import math

import numpy as np

import synergia2_amd as sc
from synergia2_amd import stepper as S


class StepMaps(C.Structure):
    _fields_ = [("n", C.c_int32), ("s", sc.MapSlice * sc.MAX_SLICES)]


class JobSpec(C.Structure):
    _fields_ = [("n_particles", C.c_int64), ("shape", C.c_int32 * 3), ("steps", C.c_int32),
                ("seed", C.c_uint64), ("mean", C.c_double * 6), ("chol", C.c_double * 21),
                ("real_num", C.c_double), ("charge", C.c_int32), ("mass", C.c_double),
                ("ref_gamma", C.c_double), ("ref_beta", C.c_double), ("ref_p", C.c_double),
                ("turn_steps", C.c_int32), ("kick_dt", C.POINTER(C.c_double)),
                ("first", C.POINTER(StepMaps)), ("between", C.POINTER(StepMaps)),
                ("last", C.POINTER(StepMaps))]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_resident", C.c_int32),
                ("precision", C.c_int32), ("sort_period", C.c_int32),
                ("use_graphs", C.c_int32), ("batch_grids", C.c_int32),
                ("admit_interval", C.c_int32), ("branches", C.c_int32),
                ("keep_particles", C.c_int32)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("state", C.c_int32), ("steps_done", C.c_int32),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("emit0", C.c_double * 3), ("emit", C.c_double * 3), ("mean", C.c_double * 6),
                ("mom2", C.c_double * 36)]


class ServerStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("graph_builds", C.c_int64),
                ("jobs_submitted", C.c_int64), ("jobs_done", C.c_int64),
                ("jobs_resident", C.c_int64), ("jobs_waiting", C.c_int64),
                ("bytes_resident", C.c_int64), ("hbm_budget", C.c_int64),
                ("particle_steps", C.c_double), ("busy_s", C.c_double),
                ("iter_ms_p50", C.c_double), ("iter_ms_p99", C.c_double),
                ("iter_ms_max", C.c_double)]


JOB_WAITING, JOB_RESIDENT, JOB_DONE, JOB_FAILED = 0, 1, 2, 3

_bound = False


def _bind():
    global _bound
    L = sc.load()
    if _bound:
        return L
    P = C.POINTER
    L.sc3d_server_options_default.argtypes = [P(ServerOptions)]
    L.sc3d_server_options_default.restype = None
    L.sc3d_server_create.argtypes = [P(ServerOptions), P(C.c_void_p)]
    L.sc3d_server_submit.argtypes = [C.c_void_p, P(JobSpec), P(C.c_int64)]
    L.sc3d_server_iterate.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_int32)]
    L.sc3d_server_job.argtypes = [C.c_void_p, C.c_int64, P(JobResult)]
    L.sc3d_server_job_particles.argtypes = [C.c_void_p, C.c_int64, P(C.c_double),
                                            P(C.c_uint8)]
    L.sc3d_server_get_stats.argtypes = [C.c_void_p, P(ServerStats)]
    L.sc3d_server_destroy.argtypes = [C.c_void_p]
    L.sc3d_job_footprint.argtypes = [P(JobSpec), C.c_int32]
    L.sc3d_job_footprint.restype = C.c_int64
    for name in ("sc3d_server_create", "sc3d_server_submit", "sc3d_server_iterate",
                 "sc3d_server_job", "sc3d_server_job_particles", "sc3d_server_get_stats",
                 "sc3d_server_destroy"):
        getattr(L, name).restype = C.c_int
    _bound = True
    return L


def _maps_array(progs):
    arr = (StepMaps * len(progs))()
    for k, (slices, n) in enumerate(progs):
        arr[k].n = n
        for q in range(n):
            arr[k].s[q] = slices[q]
    return arr


class Job:
    """A FODO-channel tracking job: a 6D Gaussian bunch matched to the cell
    (rms sizes sigma_x, sigma_y at the cell entrance), sigma_z longitudinal
    (cdt = z / beta), taking `steps` split-operator steps with
    kicks_per_cell kicks per cell.  Holds the ctypes spec and its arrays."""

    def __init__(self, n, shape, steps, sigma_x, sigma_y, sigma_z, seed, sigma_dpop=1e-4,
                 real_num=1e11, kinetic_energy=0.8, ncells=4, kicks_per_cell=10,
                 ld=0.5, lq=0.2, k1=0.6):
        mass = sc.PROTON_MASS
        gamma = (mass + kinetic_energy) / mass
        beta, p = sc.kinematics(mass, gamma)
        lattice = S.fodo_channel(ncells, ld, lq, k1)
        cell = S.fodo_channel(1, ld, lq, k1)
        chol = S.matched_chol_sigma(cell, sigma_x, sigma_y, sigma_z / beta, sigma_dpop)
        nsteps = ncells * kicks_per_cell
        steps_tab = S.split_operator_steps(lattice, nsteps)
        first = [sc.make_slices(st[0], p, mass) for st in steps_tab]
        last = [sc.make_slices(st[1], p, mass) for st in steps_tab]
        between = []
        for i, st in enumerate(steps_tab):
            nxt = steps_tab[(i + 1) % len(steps_tab)][0]
            between.append(sc.make_slices(list(st[1]) + list(nxt), p, mass))
        self._dt = (C.c_double * nsteps)(*[st[2] / (beta * sc.C_LIGHT) for st in steps_tab])
        self._first = _maps_array(first)
        self._between = _maps_array(between)
        self._last = _maps_array(last)
        s = JobSpec()
        s.n_particles = int(n)
        for a in range(3):
            s.shape[a] = int(shape[a])
        s.steps = int(steps)
        s.seed = int(seed)
        for q in range(21):
            s.chol[q] = float(chol[q])
        s.real_num = float(real_num)
        s.charge = 1
        s.mass, s.ref_gamma, s.ref_beta, s.ref_p = mass, gamma, beta, p
        s.turn_steps = nsteps
        s.kick_dt = self._dt
        s.first = self._first
        s.between = self._between
        s.last = self._last
        self.spec = s
        self.n, self.steps, self.shape = int(n), int(steps), tuple(shape)
        self.gamma, self.beta, self.p, self.mass = gamma, beta, p, mass
        self.real_num = real_num
        self.chol = chol
        self.seed = seed
        self.lattice = lattice
        self.nsteps = nsteps

    def footprint(self, precision=sc.FP32_GRID):
        return int(_bind().sc3d_job_footprint(C.byref(self.spec), precision))


def options(precision=sc.FP32_GRID, hbm_budget=0, max_resident=0, use_graphs=True,
            batch_grids=True, admit_interval=8, branches=8, keep_particles=False,
            sort_period=24):
    o = ServerOptions()
    _bind().sc3d_server_options_default(C.byref(o))
    o.precision = precision
    o.hbm_budget = int(hbm_budget)
    o.max_resident = int(max_resident)
    o.use_graphs = int(bool(use_graphs))
    o.batch_grids = int(bool(batch_grids))
    o.admit_interval = int(admit_interval)
    o.branches = int(branches)
    o.keep_particles = int(bool(keep_particles))
    o.sort_period = int(sort_period)
    return o


class Server:
    def __init__(self, opts=None):
        L = _bind()
        self._s = C.c_void_p()
        sc._check(L.sc3d_server_create(C.byref(opts or options()), C.byref(self._s)))
        self._jobs = {}

    def submit(self, job):
        L = _bind()
        jid = C.c_int64()
        sc._check(L.sc3d_server_submit(self._s, C.byref(job.spec), C.byref(jid)))
        self._jobs[jid.value] = job
        return jid.value

    def iterate(self):
        r, w = C.c_int32(), C.c_int32()
        sc._check(_bind().sc3d_server_iterate(self._s, C.byref(r), C.byref(w)))
        return r.value, w.value

    def result(self, jid):
        out = JobResult()
        sc._check(_bind().sc3d_server_job(self._s, jid, C.byref(out)))
        return out

    def particles(self, jid):
        job = self._jobs[jid]
        parts = np.zeros((7, job.n))
        masks = np.zeros(job.n, np.uint8)
        sc._check(_bind().sc3d_server_job_particles(
            self._s, jid, parts.ctypes.data_as(C.POINTER(C.c_double)),
            masks.ctypes.data_as(C.POINTER(C.c_uint8))))
        return parts, masks

    def stats(self):
        out = ServerStats()
        sc._check(_bind().sc3d_server_get_stats(self._s, C.byref(out)))
        return {name: getattr(out, name) for name, _ in out._fields_}

    def close(self):
        if self._s:
            _bind().sc3d_server_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_stream(count=256, rate=20.0, seed=4, n_range=(1e4, 4e6),
               shapes=((32, 32, 64), (64, 64, 128), (128, 128, 256)), steps_range=(20, 200),
               sigma_t=(0.5e-3, 3e-3), sigma_z=(5e-3, 30e-3)):
    """The served workload: Poisson arrivals (mean `rate` jobs/s) of `count`
    FODO-channel jobs; particle count log-uniform in n_range, grid drawn from
    `shapes`, steps uniform in steps_range (inclusive), transverse sigmas
    uniform in sigma_t (each plane), sigma_z uniform in sigma_z; one numpy
    Generator seeded with `seed` draws everything in this order per job.
    Returns [(arrival_s, Job kwargs)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        shape = shapes[int(rng.integers(len(shapes)))]
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        sx = float(rng.uniform(*sigma_t))
        sy = float(rng.uniform(*sigma_t))
        sz = float(rng.uniform(*sigma_z))
        out.append((t, dict(n=n, shape=shape, steps=steps, sigma_x=sx, sigma_y=sy, sigma_z=sz,
                            seed=seed * 1_000_003 + i)))
    return out
