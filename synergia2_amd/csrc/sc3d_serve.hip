// sc3d_serve.hip -- the served form (include/sc3d_serve.h): bunch-tracking
// jobs resident on the device, advanced one split-operator step per
// iteration (continuous batching), with
//   - admission by device footprint against a memory budget,
//   - the per-job particle passes spread over a few streams (independent
//     branches), the grid passes batched over all same-shape jobs (one
//     launch per pass, blockIdx.y = job),
//   - the whole iteration captured as two hipGraphs, replayed while the
//     resident set is unchanged: one sorts and reorders every bunch by cell
//     (the first iteration after a rebuild, which brings newly admitted,
//     unsorted bunches, and every sort_period iterations after it), the other
//     does not; schedule entry, kick time step and "finished" are decided on
//     the device by each job's StepCtl, so the launch sequence is the same
//     every step; a finished job stays in the graphs as a no-op until the
//     next rebuild, which admissions trigger at most every admit_interval
//     iterations.
// Each job's step is the operator's apply (sc3d_ctx pieces, as sc3d_apply)
// between its scheduled maps, exactly as Propagator::do_step runs Step::apply
// (simulation/propagator.cc:157-283, step.cc:44-78) for one bunch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sc3d_serve.h"
#include "sc3d_ctx.h"

using namespace sc3d;

namespace {

#define HIP_TRY SC3D_HIP_TRY

struct Job {
    int64_t id = 0;
    sc3d_job_spec spec{};           // schedule pointers cleared after the copy
    std::vector<MapProgram> progs;  // [first | between | last] x turn_steps
    std::vector<double> dt;
    int64_t footprint = 0;
    int32_t state = SC3D_JOB_WAITING;
    bool zombie = false;            // done, still in the captured graph
    int32_t steps_done = 0;
    double t_submit = -1, t_admit = -1, t_done = -1;
    double emit0[3] = {0, 0, 0}, emit[3] = {0, 0, 0};
    double mean[6] = {0}, mom2[36] = {0};
    std::vector<double> final_parts;
    std::vector<uint8_t> final_masks;
    // device
    sc3d_ctx* ctx = nullptr;
    double* parts = nullptr;
    uint8_t* masks = nullptr;
    StepCtl* ctl = nullptr;
    MapProgram* d_progs = nullptr;
    double* d_dt = nullptr;
    sc3d_bunch b{};
    Sched sched(bool sort_now) const { return Sched{ctl, d_progs, d_dt, sort_now ? 1 : 0}; }
};

struct Group {
    int32_t shape[3];
    std::vector<Job*> jobs;
    std::vector<sc3d_ctx*> ctxs;
    GridJob* d_table = nullptr;
};

// calculate_emittances (diagnostics_full2_host.cc:20-40): sqrt(det) of the
// 2x2 blocks, small negatives (> -1e-15) clamped to 0
void emittances(const double* mom2, double out[3])
{
    for (int p = 0; p < 3; ++p) {
        const int a = 2 * p, c = a + 1;
        double det = mom2[a * 6 + a] * mom2[c * 6 + c] - mom2[c * 6 + a] * mom2[a * 6 + c];
        if (det < 0.0 && det > -1e-15) det = 0.0;
        out[p] = std::sqrt(det);
    }
}

int check_spec(const sc3d_job_spec* s)
{
    if (!s) return fail(SC3D_ERR_ARG, "sc3d_server_submit: null spec");
    if (s->n_particles <= 0 || s->steps <= 0 || s->turn_steps <= 0 || !s->kick_dt ||
        !s->first || !s->between || !s->last)
        return fail(SC3D_ERR_ARG, "sc3d_server_submit: bad job spec");
    if (s->n_particles >= (int64_t(1) << 31) - kSortTileItems)
        return fail(SC3D_ERR_SIZE, "sc3d_server_submit: more than 2^31 particles in one job");
    for (const sc3d_step_maps* tab : {s->first, s->between, s->last})
        for (int k = 0; k < s->turn_steps; ++k) {
            if (tab[k].n < 0 || tab[k].n > SC3D_MAX_SLICES)
                return fail(SC3D_ERR_ARG, "sc3d_server_submit: bad map program");
            for (int q = 0; q < tab[k].n; ++q) {
                const sc3d_map_slice& m = tab[k].s[q];
                if (m.kind != SC3D_MAP_DRIFT && m.kind != SC3D_MAP_QUADRUPOLE)
                    return fail(SC3D_ERR_ARG, "sc3d_server_submit: unknown slice kind");
                if (m.kind == SC3D_MAP_QUADRUPOLE && m.yoshida_steps < 1)
                    return fail(SC3D_ERR_ARG, "sc3d_server_submit: yoshida_steps < 1");
            }
        }
    return SC3D_OK;
}

sc3d_options job_options(const sc3d_job_spec& s, int precision, int sort_period)
{
    sc3d_options o;
    sc3d_options_default(&o, s.shape[0], s.shape[1], s.shape[2]);
    o.precision = precision;
    o.sort_period = sort_period;
    return o;
}

int64_t job_bytes(const sc3d_job_spec& s, int precision, int sort_period)
{
    const int64_t n = s.n_particles;
    return ctx_footprint(job_options(s, precision, sort_period), n) + n * (7 * 8 + 1) +
           int64_t(sizeof(StepCtl)) + 3 * s.turn_steps * int64_t(sizeof(MapProgram)) +
           s.turn_steps * 8;
}

}  // namespace

struct sc3d_server {
    sc3d_server_options opt{};
    int device = 0;
    hipStream_t st = nullptr;
    std::vector<hipStream_t> aux;
    hipEvent_t fork_ev[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> join_ev[2];
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::vector<Job*> waiting;   // first come, first admitted
    std::vector<Job*> resident;  // admitted, including finished ones still in the graph
    std::vector<Group> groups;
    int64_t next_id = 1;
    int64_t used = 0;
    bool dirty = true;
    // the iteration's graph, without and with the sort + reorder of every job
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int64_t iters_since_build = 0;
    int64_t iterations = 0, graph_builds = 0, jobs_done = 0, jobs_submitted = 0;
    double particle_steps = 0.0, busy_s = 0.0;
    std::vector<double> iter_ms;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    int active() const
    {
        int a = 0;
        for (Job* j : resident) a += j->zombie ? 0 : 1;
        return a;
    }
};

namespace {

void drop_job_buffers(Job* j);

void release_job(sc3d_server* s, Job* j)
{
    drop_job_buffers(j);
    s->used -= j->footprint;
}

void release_groups(sc3d_server* s)
{
    for (Group& g : s->groups)
        if (g.d_table) (void)hipFree(g.d_table);
    s->groups.clear();
}

void release_graph(sc3d_server* s)
{
    for (int k = 0; k < 2; ++k) {
        if (s->exec[k]) (void)hipGraphExecDestroy(s->exec[k]);
        if (s->graph[k]) (void)hipGraphDestroy(s->graph[k]);
        s->exec[k] = nullptr;
        s->graph[k] = nullptr;
    }
}

// Frees what an admission allocated for j (all of it, or the part that a
// failed admission got to) and forgets the pointers.
void drop_job_buffers(Job* j)
{
    if (j->ctx) sc3d_destroy(j->ctx);
    for (void* p : {(void*)j->parts, (void*)j->masks, (void*)j->ctl, (void*)j->d_progs,
                    (void*)j->d_dt})
        if (p) (void)hipFree(p);
    j->ctx = nullptr;
    j->parts = nullptr;
    j->masks = nullptr;
    j->ctl = nullptr;
    j->d_progs = nullptr;
    j->d_dt = nullptr;
}

int admit_alloc(sc3d_server* s, Job* j)
{
    const sc3d_job_spec& sp = j->spec;
    const int64_t n = sp.n_particles;
    const sc3d_options o = job_options(sp, s->opt.precision, s->opt.sort_period);
    if (int rc = sc3d_create(&o, &j->ctx)) return rc;
    HIP_TRY(hipMalloc(&j->parts, size_t(n) * 7 * sizeof(double)));
    HIP_TRY(hipMalloc(&j->masks, size_t(n)));
    HIP_TRY(hipMalloc(&j->ctl, sizeof(StepCtl)));
    HIP_TRY(hipMalloc(&j->d_progs, j->progs.size() * sizeof(MapProgram)));
    HIP_TRY(hipMalloc(&j->d_dt, j->dt.size() * sizeof(double)));
    StepCtl c{};
    c.turn_steps = sp.turn_steps;
    c.total_steps = sp.steps;
    c.first_prog = -1;
    c.after_prog = -1;
    HIP_TRY(hipMemcpy(j->ctl, &c, sizeof(c), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(j->d_progs, j->progs.data(), j->progs.size() * sizeof(MapProgram),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(j->d_dt, j->dt.data(), j->dt.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    if (int rc = sc3d_populate_gaussian(j->parts, n, n, j->masks, sp.mean, sp.chol, sp.seed, 0,
                                        s->st))
        return rc;
    sc3d_bunch& b = j->b;
    b.parts = j->parts;
    b.stride = n;
    b.n_active = n;
    b.masks = j->masks;
    b.total_num = double(n);
    b.real_num = sp.real_num;
    b.charge = sp.charge;
    b.mass = sp.mass;
    b.ref_gamma = sp.ref_gamma;
    b.ref_beta = sp.ref_beta;
    b.ref_p = sp.ref_p;
    if (int rc = prepare_apply(j->ctx, &b)) return rc;
    double mean[6], mom2[36];
    if (int rc = sc3d_moments(j->ctx, &b, mean, mom2, s->st)) return rc;
    emittances(mom2, j->emit0);
    return SC3D_OK;
}

// Allocates and fills one job's device state.  On failure nothing of it
// stays allocated and the job is marked failed (never retried).
int admit(sc3d_server* s, Job* j)
{
    if (int rc = admit_alloc(s, j)) {
        (void)hipStreamSynchronize(s->st);  // the populate may still be queued
        drop_job_buffers(j);
        j->state = SC3D_JOB_FAILED;
        j->t_done = s->now();
        return rc;
    }
    j->ctx->last_stream = s->st;
    j->state = SC3D_JOB_RESIDENT;
    j->t_admit = s->now();
    s->used += j->footprint;
    return SC3D_OK;
}

int finish(sc3d_server* s, Job* j)
{
    if (int rc = sc3d_moments(j->ctx, &j->b, j->mean, j->mom2, s->st)) return rc;
    emittances(j->mom2, j->emit);
    if (s->opt.keep_particles) {
        const int64_t n = j->spec.n_particles;
        j->final_parts.resize(size_t(n) * 7);
        j->final_masks.resize(size_t(n));
        HIP_TRY(hipMemcpy(j->final_parts.data(), j->parts, size_t(n) * 7 * sizeof(double),
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(j->final_masks.data(), j->masks, size_t(n), hipMemcpyDeviceToHost));
    }
    j->state = SC3D_JOB_DONE;
    j->zombie = true;
    j->t_done = s->now();
    s->jobs_done++;
    return SC3D_OK;
}

// One step of every resident job, enqueued on s->st (and the branch streams);
// sort_now: the jobs' bunches are also sorted by cell and reordered.
void enqueue_iteration(sc3d_server* s, bool sort_now)
{
    const int B = int(std::min<size_t>(s->aux.size(), std::max<size_t>(1, s->resident.size())));
    hipEvent_t ev[SC3D_NPHASES + 1];
    auto fork = [&](int phase) {
        (void)hipEventRecord(s->fork_ev[phase], s->st);
        for (int q = 0; q < B; ++q) (void)hipStreamWaitEvent(s->aux[q], s->fork_ev[phase], 0);
    };
    auto join = [&](int phase) {
        for (int q = 0; q < B; ++q) {
            (void)hipEventRecord(s->join_ev[phase][q], s->aux[q]);
            (void)hipStreamWaitEvent(s->st, s->join_ev[phase][q], 0);
        }
    };
    fork(0);
    for (size_t i = 0; i < s->resident.size(); ++i) {
        Job* j = s->resident[i];
        hipStream_t q = s->aux[i % B];
        const Sched sc = j->sched(sort_now);
        launch_step_begin(j->ctl, q);
        launch_maps_sched(j->parts, j->b.stride, j->masks, j->b.n_active, sc, q);
        enqueue_front(j->ctx, &j->b, j->dt[0], sc, q, ev);
    }
    join(0);
    for (Group& g : s->groups) {
        if (s->opt.batch_grids) {
            enqueue_grid(g.ctxs.data(), int(g.ctxs.size()), g.d_table, s->st, ev);
        } else {
            for (sc3d_ctx* c : g.ctxs) enqueue_grid(&c, 1, c->d_gj, s->st, ev);
        }
    }
    fork(1);
    for (size_t i = 0; i < s->resident.size(); ++i) {
        Job* j = s->resident[i];
        enqueue_kick(j->ctx, &j->b, j->sched(sort_now), s->aux[i % B]);
    }
    join(1);
}

int rebuild(sc3d_server* s)
{
    // finished jobs leave (everything before this point has completed)
    std::vector<Job*> keep;
    for (Job* j : s->resident) {
        if (j->zombie) {
            release_job(s, j);
            j->zombie = false;
        } else {
            keep.push_back(j);
        }
    }
    s->resident.swap(keep);
    release_groups(s);
    release_graph(s);
    for (Job* j : s->resident) {
        Group* g = nullptr;
        for (Group& x : s->groups)
            if (std::memcmp(x.shape, j->spec.shape, sizeof(x.shape)) == 0) g = &x;
        if (!g) {
            s->groups.emplace_back();
            g = &s->groups.back();
            std::memcpy(g->shape, j->spec.shape, sizeof(g->shape));
        }
        g->jobs.push_back(j);
        g->ctxs.push_back(j->ctx);
    }
    for (Group& g : s->groups) {
        std::vector<GridJob> tab;
        for (Job* j : g.jobs) tab.push_back(j->ctx->gj);
        HIP_TRY(hipMalloc(&g.d_table, tab.size() * sizeof(GridJob)));
        HIP_TRY(hipMemcpy(g.d_table, tab.data(), tab.size() * sizeof(GridJob),
                          hipMemcpyHostToDevice));
    }
    s->dirty = false;
    s->iters_since_build = 0;
    if (!s->opt.use_graphs || s->resident.empty()) return SC3D_OK;
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
        enqueue_iteration(s, k == 1);
        HIP_TRY(hipStreamEndCapture(s->st, &s->graph[k]));
        HIP_TRY(hipGraphInstantiate(&s->exec[k], s->graph[k], nullptr, nullptr, 0));
    }
    s->graph_builds++;
    return SC3D_OK;
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * double(v.size() - 1);
    const size_t lo = size_t(pos);
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (pos - double(lo)) * (v[hi] - v[lo]);
}

}  // namespace

extern "C" {

void sc3d_server_options_default(sc3d_server_options* o)
{
    std::memset(o, 0, sizeof(*o));
    o->precision = SC3D_PRECISION_FP32_GRID;
    o->sort_period = 24;
    o->use_graphs = 1;
    o->batch_grids = 1;
    o->admit_interval = 8;
    o->branches = 8;
}

int sc3d_server_create(const sc3d_server_options* opt, sc3d_server** out)
{
    if (!opt || !out) return fail(SC3D_ERR_ARG, "sc3d_server_create: null argument");
    if (opt->precision != SC3D_PRECISION_FP64 && opt->precision != SC3D_PRECISION_FP32_GRID)
        return fail(SC3D_ERR_ARG, "sc3d_server_create: unknown precision");
    if (opt->sort_period < 0 || opt->hbm_budget < 0 || opt->max_resident < 0 ||
        opt->admit_interval < 0 || opt->branches < 0)
        return fail(SC3D_ERR_ARG, "sc3d_server_create: bad option");
    *out = nullptr;
    auto s = std::make_unique<sc3d_server>();
    s->opt = *opt;
    if (s->opt.admit_interval == 0) s->opt.admit_interval = 8;
    if (s->opt.branches == 0) s->opt.branches = 8;
    HIP_TRY(hipGetDevice(&s->device));
    if (s->opt.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        s->opt.hbm_budget = int64_t(fr / 2);
    }
    HIP_TRY(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    for (int q = 0; q < s->opt.branches; ++q) {
        hipStream_t a;
        HIP_TRY(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
        s->aux.push_back(a);
    }
    for (int ph = 0; ph < 2; ++ph) {
        HIP_TRY(hipEventCreateWithFlags(&s->fork_ev[ph], hipEventDisableTiming));
        for (int q = 0; q < s->opt.branches; ++q) {
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            s->join_ev[ph].push_back(e);
        }
    }
    *out = s.release();
    return SC3D_OK;
}

int sc3d_server_submit(sc3d_server* s, const sc3d_job_spec* spec, int64_t* id)
{
    if (!s) return fail(SC3D_ERR_ARG, "sc3d_server_submit: null server");
    if (int rc = check_spec(spec)) return rc;
    auto j = std::make_unique<Job>();
    j->spec = *spec;
    const int T = spec->turn_steps;
    j->progs.resize(size_t(3) * T);
    for (int t = 0; t < 3; ++t) {
        const sc3d_step_maps* tab = t == 0 ? spec->first : t == 1 ? spec->between : spec->last;
        for (int k = 0; k < T; ++k) {
            MapProgram& p = j->progs[size_t(t) * T + k];
            p = MapProgram{};
            for (int q = 0; q < tab[k].n; ++q) p.s[q] = tab[k].s[q];
            p.n = tab[k].n;
            p.ref_p = spec->ref_p;
            p.mass = spec->mass;
        }
    }
    j->dt.assign(spec->kick_dt, spec->kick_dt + T);
    j->spec.kick_dt = nullptr;
    j->spec.first = j->spec.between = j->spec.last = nullptr;
    j->footprint = job_bytes(*spec, s->opt.precision, s->opt.sort_period);
    if (j->footprint > s->opt.hbm_budget)
        return fail(SC3D_ERR_SIZE, "sc3d_server_submit: job footprint exceeds the memory budget");
    j->id = s->next_id++;
    j->t_submit = s->now();
    if (id) *id = j->id;
    s->waiting.push_back(j.get());
    s->jobs[j->id] = std::move(j);
    s->jobs_submitted++;
    return SC3D_OK;
}

int sc3d_server_iterate(sc3d_server* s, int32_t* n_resident, int32_t* n_waiting)
{
    if (!s) return fail(SC3D_ERR_ARG, "sc3d_server_iterate: null server");
    HIP_TRY(hipSetDevice(s->device));
    // admission: always when nothing runs or without graphs; with graphs at
    // most every admit_interval iterations (each admission rebuilds the graph)
    const bool may_admit = !s->opt.use_graphs || s->active() == 0 ||
                           s->iters_since_build >= s->opt.admit_interval;
    if (may_admit && !s->waiting.empty()) {
        size_t taken = 0;
        for (Job* j : s->waiting) {
            int64_t zombie_bytes = 0;
            for (Job* r : s->resident) zombie_bytes += r->zombie ? r->footprint : 0;
            const bool slot = s->opt.max_resident == 0 || s->active() < s->opt.max_resident;
            if (!slot) break;
            if (s->used + j->footprint > s->opt.hbm_budget) {
                if (zombie_bytes == 0 || s->used - zombie_bytes + j->footprint > s->opt.hbm_budget)
                    break;
                if (int rc = rebuild(s)) return rc;  // frees the finished jobs
            }
            if (int rc = admit(s, j)) {
                // the jobs admitted before it stay resident; the failed one
                // (marked SC3D_JOB_FAILED, nothing of it allocated) leaves
                // the queue with them
                s->waiting.erase(s->waiting.begin(), s->waiting.begin() + taken + 1);
                return rc;
            }
            s->resident.push_back(j);
            s->dirty = true;
            ++taken;
        }
        s->waiting.erase(s->waiting.begin(), s->waiting.begin() + taken);
    }
    const int act = s->active();
    if (act == 0) {
        if (!s->resident.empty()) {
            if (int rc = rebuild(s)) return rc;  // frees the finished jobs
        }
        if (n_resident) *n_resident = 0;
        if (n_waiting) *n_waiting = int32_t(s->waiting.size());
        return SC3D_OK;
    }
    // many finished jobs still in the graph: rebuild without them
    if (int(s->resident.size()) - act > std::max(2, act / 4)) s->dirty = true;
    if (s->dirty)
        if (int rc = rebuild(s)) return rc;
    // the bunches are sorted and reordered on the first iteration after a
    // rebuild (newly admitted bunches come unsorted) and every sort_period
    // iterations after it
    const bool sort_now =
        s->opt.sort_period > 0 && s->iters_since_build % s->opt.sort_period == 0;
    const auto t_start = std::chrono::steady_clock::now();
    if (s->exec[sort_now ? 1 : 0]) {
        HIP_TRY(hipGraphLaunch(s->exec[sort_now ? 1 : 0], s->st));
    } else {
        enqueue_iteration(s, sort_now);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->st));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
            .count();
    s->iter_ms.push_back(ms);
    s->busy_s += ms * 1e-3;
    s->iterations++;
    s->iters_since_build++;
    for (Job* j : s->resident) {
        if (j->zombie) continue;
        j->steps_done++;
        s->particle_steps += double(j->spec.n_particles);
        if (j->steps_done == j->spec.steps)
            if (int rc = finish(s, j)) return rc;
    }
    // nothing left to run: the finished jobs (results already copied out)
    // release their memory now, not at the next call
    if (s->active() == 0 && !s->resident.empty())
        if (int rc = rebuild(s)) return rc;
    if (n_resident) *n_resident = int32_t(s->active());
    if (n_waiting) *n_waiting = int32_t(s->waiting.size());
    return SC3D_OK;
}

int sc3d_server_job(sc3d_server* s, int64_t id, sc3d_job_result* out)
{
    if (!s || !out) return fail(SC3D_ERR_ARG, "sc3d_server_job: null argument");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(SC3D_ERR_ARG, "sc3d_server_job: unknown job id");
    const Job& j = *it->second;
    std::memset(out, 0, sizeof(*out));
    out->id = j.id;
    out->state = j.state;
    out->steps_done = j.steps_done;
    out->t_submit = j.t_submit;
    out->t_admit = j.t_admit;
    out->t_done = j.t_done;
    std::memcpy(out->emit0, j.emit0, sizeof(out->emit0));
    std::memcpy(out->emit, j.emit, sizeof(out->emit));
    std::memcpy(out->mean, j.mean, sizeof(out->mean));
    std::memcpy(out->mom2, j.mom2, sizeof(out->mom2));
    return SC3D_OK;
}

int sc3d_server_job_particles(sc3d_server* s, int64_t id, double* parts, uint8_t* masks)
{
    if (!s || !parts || !masks) return fail(SC3D_ERR_ARG, "sc3d_server_job_particles: null");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || it->second->final_parts.empty())
        return fail(SC3D_ERR_ARG, "sc3d_server_job_particles: no particles kept for this job");
    const Job& j = *it->second;
    std::memcpy(parts, j.final_parts.data(), j.final_parts.size() * sizeof(double));
    std::memcpy(masks, j.final_masks.data(), j.final_masks.size());
    return SC3D_OK;
}

int sc3d_server_get_stats(sc3d_server* s, sc3d_server_stats* out)
{
    if (!s || !out) return fail(SC3D_ERR_ARG, "sc3d_server_get_stats: null argument");
    std::memset(out, 0, sizeof(*out));
    out->iterations = s->iterations;
    out->graph_builds = s->graph_builds;
    out->jobs_submitted = s->jobs_submitted;
    out->jobs_done = s->jobs_done;
    out->jobs_resident = s->active();
    out->jobs_waiting = int64_t(s->waiting.size());
    out->bytes_resident = s->used;
    out->hbm_budget = s->opt.hbm_budget;
    out->particle_steps = s->particle_steps;
    out->busy_s = s->busy_s;
    out->iter_ms_p50 = percentile(s->iter_ms, 0.50);
    out->iter_ms_p99 = percentile(s->iter_ms, 0.99);
    out->iter_ms_max = s->iter_ms.empty()
                           ? 0.0
                           : *std::max_element(s->iter_ms.begin(), s->iter_ms.end());
    return SC3D_OK;
}

int sc3d_server_destroy(sc3d_server* s)
{
    if (!s) return SC3D_OK;
    (void)hipSetDevice(s->device);
    if (s->st) (void)hipStreamSynchronize(s->st);
    release_graph(s);
    release_groups(s);
    for (Job* j : s->resident) release_job(s, j);
    for (hipStream_t a : s->aux) (void)hipStreamDestroy(a);
    for (int ph = 0; ph < 2; ++ph) {
        if (s->fork_ev[ph]) (void)hipEventDestroy(s->fork_ev[ph]);
        for (hipEvent_t e : s->join_ev[ph]) (void)hipEventDestroy(e);
    }
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
    return SC3D_OK;
}

int64_t sc3d_job_footprint(const sc3d_job_spec* spec, int32_t precision)
{
    if (!spec || spec->n_particles < 0 || spec->turn_steps < 0) return -1;
    return job_bytes(*spec, precision, 8);
}

}  // extern "C"
