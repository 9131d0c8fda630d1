"""Served form (include/sc3d_serve.h) on the GPU, through the C ABI.

Bars:
  - a job served alone ends where the same bunch tracked step by step
    through the operator API ends (same kick, maps, reorder cadence and
    precision); the two differ only in the order of the deposit atomics:
    FP64 form, particle coordinates (matched by id) within 1e-10 relative;
    FP32-grid form within 1e-6 (FP32 atomics reorder at ~6e-8 per node);
  - the same job alone and in a mix of jobs of other sizes and grid shapes
    (some admitted while it runs) gives the same result: coordinates within
    1e-6, rms emittances within 1e-6 (the stated bar is 1e-4);
  - graphs on / off, grid batching on / off: same results within 1e-6;
  - admission never holds more than the memory budget and every job ends.
"""
import numpy as np
import pytest

import synergia2_amd as sc
from helpers import by_id, rel_l2
from synergia2_amd import serve as SV
from synergia2_amd import stepper as S

pytestmark = pytest.mark.gpu


def run_all(server, jobs, arrivals=None):
    """Submit jobs (at the given iteration numbers, else all at once) and
    iterate until every job is done.  Returns the job ids."""
    arrivals = arrivals or [0] * len(jobs)
    ids = [None] * len(jobs)
    it = 0
    while True:
        for k, job in enumerate(jobs):
            if ids[k] is None and arrivals[k] <= it:
                ids[k] = server.submit(job)
        res, wait = server.iterate()
        it += 1
        if all(i is not None for i in ids) and res == 0 and wait == 0:
            break
        assert it < 10000
    return ids


def coords(parts):
    return by_id(parts)


def close(a, b, tol):
    a, b = coords(a), coords(b)
    assert np.array_equal(a[6], b[6])
    errs = [np.abs(a[col] - b[col]).max() / np.abs(b[col]).max() for col in range(6)]
    print("max relative coordinate differences", errs)
    assert max(errs) <= tol, errs


@pytest.mark.parametrize("precision,tol,shape", [(sc.FP64, 1e-10, (32, 32, 64)),
                                                 (sc.FP32_GRID, 1e-6, (32, 32, 64)),
                                                 (sc.FP64, 1e-10, (24, 24, 48)),
                                                 (sc.FP32_GRID, 1e-6, (48, 48, 96))])
def test_served_job_matches_operator_api(gpu, precision, tol, shape):
    """One job in the server == the bunch tracked by the Propagator over the
    operator API with the same options (same sort_period); also on grids
    of other lengths (generic transform passes; at 48x48x96 with the
    reference's g2 correction spectrum)."""
    job = SV.Job(30000, shape, 23, 1.5e-3, 1.0e-3, 0.012, seed=11)
    srv = SV.Server(SV.options(precision=precision, keep_particles=True))
    jid = run_all(srv, [job])[0]
    res = srv.result(jid)
    assert res.state == SV.JOB_DONE and res.steps_done == 23
    got, _ = srv.particles(jid)
    srv.close()

    b = sc.Bunch(job.n, job.real_num, gamma=job.gamma)
    b.populate_gaussian(np.zeros(6), job.chol, seed=job.seed)
    opts = sc.default_options(*job.shape)
    opts.precision = precision
    op = sc.SpaceCharge3dOpenHockney(opts)
    e0 = sc.emittances(op.moments(b)[1])
    prop = S.Propagator(job.lattice, job.nsteps, b, op)
    prop.run(job.steps, sync=True)
    ref, _ = b.host()
    close(got, ref, tol)
    e1 = sc.emittances(op.moments(b)[1])
    for k in range(3):
        assert res.emit0[k] == e0[k]  # same populate, same fixed-tree moments
        assert abs(res.emit[k] / e1[k] - 1) <= tol


def test_served_alone_equals_in_mix(gpu):
    """The same job alone and inside a mix (other sizes and shapes, two of
    them arriving while it runs) ends at the same particles."""
    def target():
        return SV.Job(40000, (64, 64, 128), 31, 2.0e-3, 1.2e-3, 0.020, seed=5)
    alone = SV.Server(SV.options(keep_particles=True))
    ja = run_all(alone, [target()])[0]
    pa, _ = alone.particles(ja)
    ra = alone.result(ja)
    alone.close()

    others = [SV.Job(20000, (32, 32, 64), 17, 1e-3, 1e-3, 0.010, seed=21),
              SV.Job(60000, (64, 64, 128), 40, 3e-3, 2e-3, 0.030, seed=22),
              SV.Job(15000, (32, 32, 64), 9, 0.6e-3, 0.8e-3, 0.006, seed=23),
              SV.Job(80000, (128, 128, 256), 12, 2.5e-3, 2.5e-3, 0.025, seed=24)]
    mix = SV.Server(SV.options(keep_particles=True, admit_interval=4))
    ids = run_all(mix, [others[0], target()] + others[1:], arrivals=[0, 0, 3, 6, 11])
    pm, _ = mix.particles(ids[1])
    rm = mix.result(ids[1])
    st = mix.stats()
    mix.close()
    print("mix stats", st)
    assert st["jobs_done"] == 5 and st["graph_builds"] >= 2
    close(pm, pa, 1e-6)
    for k in range(3):
        assert abs(rm.emit[k] / ra.emit[k] - 1) <= 1e-6


@pytest.mark.parametrize("graphs,batch", [(False, True), (True, False)])
def test_served_execution_modes_agree(gpu, graphs, batch):
    """Eager launches or unbatched grid passes give the graph + batched
    results."""
    def jobs():
        return [SV.Job(25000, (32, 32, 64), 14, 1e-3, 1.5e-3, 0.010, seed=31),
                SV.Job(25000, (32, 32, 64), 19, 2e-3, 1.0e-3, 0.015, seed=32),
                SV.Job(50000, (64, 64, 128), 11, 1.5e-3, 1.5e-3, 0.020, seed=33)]
    base = SV.Server(SV.options(keep_particles=True))
    ib = run_all(base, jobs())
    ref = [base.particles(i)[0] for i in ib]
    base.close()
    other = SV.Server(SV.options(keep_particles=True, use_graphs=graphs, batch_grids=batch))
    io = run_all(other, jobs())
    for k, i in enumerate(io):
        close(other.particles(i)[0], ref[k], 1e-6)
    other.close()


def test_served_admission_respects_budget(gpu):
    """With a budget that holds two of these jobs, never more than two are
    resident and all five finish."""
    jobs = [SV.Job(50000, (64, 64, 128), 6 + k, 1.5e-3, 1.5e-3, 0.015, seed=40 + k)
            for k in range(5)]
    fp = max(j.footprint() for j in jobs)
    srv = SV.Server(SV.options(hbm_budget=int(2.5 * fp)))
    for j in jobs:
        srv.submit(j)
    peak = 0
    for _ in range(200):
        res, wait = srv.iterate()
        st = srv.stats()
        peak = max(peak, st["jobs_resident"])
        assert st["bytes_resident"] <= st["hbm_budget"]
        if res == 0 and wait == 0:
            break
    st = srv.stats()
    srv.close()
    assert st["jobs_done"] == 5
    assert 1 <= peak <= 2
    # a job larger than the whole budget is refused at submission
    small = SV.Server(SV.options(hbm_budget=fp // 2))
    with pytest.raises(sc.Sc3dError):
        small.submit(jobs[0])
    small.close()


def test_failed_admission_leaves_server_consistent(gpu, monkeypatch):
    """A job whose allocation fails at admission (forced on the host with the
    library's SC3D_TEST_MAX_ALLOC_BYTES hook: its 128x128x256 spectrum
    workspace exceeds the limit) is marked failed and dropped; the job
    admitted before it in the same iteration stays resident once (stepped
    once per iteration), the one behind it is admitted later, both end where
    they end alone, and no memory stays accounted."""
    def small(seed):
        return SV.Job(20000, (32, 32, 64), 13, 1e-3, 1.2e-3, 0.010, seed=seed)
    alone = SV.Server(SV.options(keep_particles=True))
    ia = run_all(alone, [small(61), small(62)])
    want = [alone.particles(i)[0] for i in ia]
    alone.close()

    big = SV.Job(30000, (128, 128, 256), 5, 2e-3, 2e-3, 0.02, seed=63)
    srv = SV.Server(SV.options(keep_particles=True))
    ids = [srv.submit(small(61)), srv.submit(big), srv.submit(small(62))]
    monkeypatch.setenv("SC3D_TEST_MAX_ALLOC_BYTES", str(32 << 20))
    with pytest.raises(sc.Sc3dError):
        srv.iterate()
    monkeypatch.delenv("SC3D_TEST_MAX_ALLOC_BYTES")
    assert srv.result(ids[0]).state == SV.JOB_RESIDENT
    assert srv.result(ids[1]).state == SV.JOB_FAILED
    assert srv.result(ids[2]).state == SV.JOB_WAITING
    st = srv.stats()
    assert st["jobs_resident"] == 1 and st["jobs_waiting"] == 1
    for _ in range(100):
        res, wait = srv.iterate()
        if res == 0 and wait == 0:
            break
    for k, jid in ((0, ids[0]), (1, ids[2])):
        r = srv.result(jid)
        assert r.state == SV.JOB_DONE and r.steps_done == 13
        close(srv.particles(jid)[0], want[k], 1e-6)
    assert srv.result(ids[1]).state == SV.JOB_FAILED
    st = srv.stats()
    srv.close()
    assert st["jobs_done"] == 2 and st["bytes_resident"] == 0
