// sc3d_capi.hip -- the C ABI of libsc3d_hip.so (declared in
// include/sc3d_capi.h): context = device workspaces + cached Green's
// spectrum of one operator instance, and the per-apply launch sequence of
// Space_charge_3d_open_hockney::apply_bunch
// (space_charge_3d_open_hockney.cc:422-453):
//   domain -> cell sort [+ reorder] -> deposit -> Green's spectrum (only when
//   h changed) -> x/y/z FFT passes with the convolution fused into the z pass
//   -> E field -> kick.
// Everything is enqueued on the caller's stream; no host synchronisation
// happens inside sc3d_apply (the domain is finished on the device).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sc3d_ctx.h"

using namespace sc3d;

namespace {
thread_local std::string g_err;

#define HIP_TRY SC3D_HIP_TRY

// Releases the particle-side buffers and forgets them: no pointer to freed
// memory and no stale capacity survive, so a failed reallocation leaves a
// context that simply reallocates on its next apply.
void drop_particle_buffers(sc3d_ctx* c)
{
    for (void** p : {(void**)&c->sb.keys0, (void**)&c->sb.keysA, (void**)&c->sb.keysB,
                     (void**)&c->sb.valsA, (void**)&c->sb.valsB, (void**)&c->sb.counts,
                     (void**)&c->sb.totals, (void**)&c->scratch, (void**)&c->smask}) {
        dfree(c, *p);
        *p = nullptr;
    }
    c->sb.perm = nullptr;
    c->pcap = 0;
    c->sorted_n = -1;
}

int ensure_particle_buffers(sc3d_ctx* c, int64_t n)
{
    if (n <= c->pcap) return 0;
    drop_particle_buffers(c);
    const int64_t cap = n + n / 8 + 4096;
    const int64_t tiles = (cap + kSortTileItems - 1) / kSortTileItems;
    int rc = dmalloc(c, &c->sb.keys0, cap);
    if (!rc) rc = dmalloc(c, &c->sb.keysA, cap);
    if (!rc) rc = dmalloc(c, &c->sb.keysB, cap);
    if (!rc) rc = dmalloc(c, &c->sb.valsA, cap);
    if (!rc) rc = dmalloc(c, &c->sb.valsB, cap);
    if (!rc) rc = dmalloc(c, &c->sb.counts, 2048 * tiles);
    if (!rc) rc = dmalloc(c, &c->sb.totals, 2048);
    if (!rc) rc = dmalloc(c, &c->scratch, 7 * cap);
    if (!rc) rc = dmalloc(c, &c->smask, cap);
    if (rc) {
        drop_particle_buffers(c);
        return rc;
    }
    c->pcap = cap;
    return 0;
}

// e^{-2 pi i k / L}, k < L, appended to t64 / t32 (long double, exact on the
// axes where L allows)
void append_twiddles(int L, std::vector<double2>& t64, std::vector<float2>& t32)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < L; ++k) {
        long double a = -2.0L * pi * k / L;
        long double cr = std::cos(a), si = std::sin(a);
        if (4 * k == L) { cr = 0; si = -1; }
        if (2 * k == L) { cr = -1; si = 0; }
        if (4 * k == 3 * L) { cr = 0; si = 1; }
        if (k == 0) { cr = 1; si = 0; }
        t64.push_back(make_double2(double(cr), double(si)));
        t32.push_back(make_float2(float(cr), float(si)));
    }
}

// The power-of-two tables (twiddle_offset), then the tables of the
// generic axes of g at g.tw_any.
void make_twiddles(const Geometry& g, std::vector<double2>& t64, std::vector<float2>& t32)
{
    t64.clear();
    t32.clear();
    for (int L = 1; L <= kMaxLen; L *= 2) append_twiddles(L, t64, t32);
    const int len[3] = {g.Dx, g.Dy, g.Dz};
    for (int a = 0; a < 3; ++a)
        if (!g.tuned[a]) append_twiddles(len[a], t64, t32);
}

// all_generic: every axis on the generic passes (the shapes whose g2
// correction is a spectrum, which only those passes apply)
Geometry geometry_of(const sc3d_options& o, bool all_generic = false)
{
    Geometry g;
    g.gx = o.shape[0]; g.gy = o.shape[1]; g.gz = o.shape[2];
    g.Dx = 2 * g.gx; g.Dy = 2 * g.gy; g.Dz = 2 * g.gz;
    g.Hx = g.gx + 1;
    g.G = int64_t(g.gx) * g.gy * g.gz;
    g.S_elems = int64_t(g.gz) * g.Dy * g.Hx;
    g.Ghat_elems = int64_t(g.gz + 1) * (g.gy + 1) * g.Hx;
    g.tuned[0] = !all_generic && tuned_row_length(g.gx);
    g.tuned[1] = !all_generic && tuned_col_length(g.Dy);
    g.tuned[2] = !all_generic && tuned_col_length(g.Dz);
    const int len[3] = {g.Dx, g.Dy, g.Dz};
    int64_t off = kTwiddleElems;
    for (int a = 0; a < 3; ++a) {
        g.plan[a] = AnyPlan{};
        g.tw_any[a] = -1;
        if (!g.tuned[a]) {
            g.plan[a] = any_plan(len[a]);
            g.tw_any[a] = off;
            off += len[a];
        }
    }
    g.tw_elems = off;
    return g;
}

}  // namespace

namespace sc3d {

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int64_t test_alloc_limit()
{
    const char* v = std::getenv("SC3D_TEST_MAX_ALLOC_BYTES");
    if (!v || !*v) return -1;
    return std::strtoll(v, nullptr, 10);
}

// Test hook: SC3D_TEST_GREEN_FORK=1 forks the Green's spectrum onto its
// stream on every apply (timing phases aside), =0 never; unset: the
// automatic choice of sc3d_apply.
static int test_green_fork()
{
    const char* v = std::getenv("SC3D_TEST_GREEN_FORK");
    if (!v || !*v) return -1;
    return std::atoi(v) != 0 ? 1 : 0;
}

void dfree(sc3d_ctx* c, void* p)
{
    if (!p) return;
    for (auto& q : c->owned)
        if (q == p) {
            (void)hipFree(q);
            q = nullptr;
        }
}

hipEvent_t mark(sc3d_ctx* c, hipStream_t st, bool edge)
{
    if (c->timing != SC3D_TIMING_PHASES && !(c->timing == SC3D_TIMING_APPLY && edge))
        return nullptr;
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    hipEvent_t e = c->ev_pool[c->ev_used++];
    (void)hipEventRecord(e, st);
    return e;
}

int64_t ctx_footprint(const sc3d_options& o, int64_t n)
{
    std::vector<GreenPatch> patches;
    const int np = green_patches(o.green_fn, o.shape, patches);
    const bool spectral = np > kMaxGreenPatches;
    const Geometry g = geometry_of(o, spectral);
    const bool fp32 = o.precision == SC3D_PRECISION_FP32_GRID;
    const int64_t rs = fp32 ? 4 : 8;
    int64_t b = sizeof(DomainRec) + sizeof(GridJob) + sizeof(StepCtl);
    b += int64_t(np) * sizeof(GreenPatch);
    if (spectral) b += int64_t(g.Dz) * g.Dy * (g.Dx + 2 * g.Hx) * rs;  // C + chat
    b += (g.G * rs + 15) / 16 * 16;              // rho
    b += g.S_elems * 2 * rs * (fp32 ? 2 : 1);    // S (+ S2)
    b += fp32 ? 0 : g.G * rs;                    // phi
    b += g.G * 4 * rs;                           // E
    b += g.Ghat_elems * (fp32 ? 4 : 8);          // Green's spectrum
    b += g.tw_elems * (16 + 8);                  // twiddles
    const int64_t cap = n + n / 8 + 4096;        // ensure_particle_buffers
    const int64_t tiles = (cap + kSortTileItems - 1) / kSortTileItems;
    b += cap * 4 * 5 + 2048 * tiles * 4 + 2048 * 4 + cap * 8 * 7 + cap;
    b += (((n + 4095) / 4096) * 6 + 64) * 8;     // domain partials
    return b;
}

int prepare_apply(sc3d_ctx* c, const sc3d_bunch* b)
{
    if (!c || !b) return fail(SC3D_ERR_ARG, "sc3d_apply: null argument");
    const int64_t n = b->n_active;
    if (n < 0 || b->stride < n) return fail(SC3D_ERR_ARG, "sc3d_apply: bad n_active/stride");
    if (n > 0 && (!b->parts || !b->masks))
        return fail(SC3D_ERR_ARG, "sc3d_apply: null particle or mask pointer");
    if (n >= (int64_t(1) << 31) - kSortTileItems)
        return fail(SC3D_ERR_SIZE, "sc3d_apply: more than 2^31 particles in one bunch");
    if (c->opt.periodic_z)
        return fail(SC3D_ERR_PERIODIC,
                    c->opt.green_fn == SC3D_GREEN_POINTLIKE
                        ? "Space_charge_3d_open_hockney::get_green_fn2_pointlike: periodic_z "
                          "not yet implemented"
                        : "Space_charge_3d_open_hockney::get_green_fn2_linear: periodic_z not "
                          "yet implemented");
    HIP_TRY(hipSetDevice(c->device));
    const int64_t nch = (n + 4095) / 4096;
    if (!c->opt.domain_fixed && nch * 6 > c->P_cap) {
        dfree(c, c->P);
        c->P = nullptr;
        c->P_cap = 0;
        if (int rc = dmalloc(c, &c->P, nch * 6 + 64)) return rc;
        c->P_cap = nch * 6 + 64;
    }
    if (n > 0)
        if (int rc = ensure_particle_buffers(c, n)) return rc;
    return SC3D_OK;
}

// Green's spectrum of nj contexts of one shape (and the spectral patch
// correction of the contexts that have one), on st.
static void enqueue_green(sc3d_ctx* const* ctxs, int nj, const GridJob* table, hipStream_t st)
{
    sc3d_ctx* c = ctxs[0];
    if (!c->fp32)
        launch_green<double>(c->g, c->opt.green_fn, table, nj, c->tw64, st);
    else
        launch_green<float>(c->g, c->opt.green_fn, table, nj, c->tw32, st);
    for (int j = 0; j < nj; ++j) launch_green_chat(ctxs[j], st);
}

int enqueue_front(sc3d_ctx* c, const sc3d_bunch* b, double time_step, const Sched& sc,
                  hipStream_t st, hipEvent_t* ev, hipStream_t green_st)
{
    const int64_t n = b->n_active;
    const Geometry& g = c->g;
    const sc3d_options& o = c->opt;
    DomainFinishArgs a{};
    a.fixed = o.domain_fixed;
    a.green_fn = o.green_fn;
    a.grid_entire_period = o.grid_entire_period;
    a.charge = b->charge;
    for (int k = 0; k < 3; ++k) {
        a.shape[k] = o.shape[k];
        a.size[k] = o.size[k];
        a.offset[k] = o.offset[k];
    }
    a.n_sigma = o.n_sigma;
    a.z_period = o.z_period;
    a.kick_scale = o.kick_scale;
    a.total_num = b->total_num;
    a.real_num = b->real_num;
    a.time_step = time_step;
    a.ref_p = b->ref_p;
    a.ref_gamma = b->ref_gamma;
    a.ref_beta = b->ref_beta;
    // The bunch storage is cell-sorted (sort + reorder) every sort_period
    // applies (served form: on the iterations the server picks, sc.sort_now);
    // the deposit then works on windows of the storage, sorted by cell in
    // LDS.  sort_period 0 never reorders: the sort then runs on every apply
    // and the deposit walks its permutation.
    const bool local_dep = o.sort_period > 0;
    const bool do_reorder =
        n > 0 && local_dep &&
        (sc.ctl != nullptr ? sc.sort_now != 0 : (c->applies % o.sort_period) == 0);
    const bool do_sort = n > 0 && (do_reorder || !local_dep);
    ev[0] = mark(c, st, true);
    launch_domain(b->parts, b->stride, b->masks, n, c->P, a, c->dom, sc, c->rho, c->rho_bytes,
                  st);
    ev[1] = mark(c, st);
    if (green_st) {
        (void)hipEventRecord(c->fork_ev, st);
        (void)hipStreamWaitEvent(green_st, c->fork_ev, 0);
        enqueue_green(&c, 1, c->d_gj, green_st);
        (void)hipEventRecord(c->join_ev, green_st);
    }
    const int32_t* dep_perm = nullptr;
    if (do_sort) {
        launch_sort(b->parts, b->stride, b->masks, n, g.gx, g.gy, g.gz, c->dom, c->sb, st);
        c->sorted_n = n;
        if (do_reorder)
            launch_reorder(b->parts, b->stride, b->masks, n, c->sb.perm, c->scratch, c->smask,
                           c->dom, st);
        if (!local_dep) dep_perm = c->sb.perm;
    }
    ev[2] = mark(c, st);
    // after a reorder the cell-sorted bunch is the gathered copy until the
    // kick writes it back
    c->reordering = do_reorder;
    const double* dp = do_reorder ? c->scratch : b->parts;
    const int64_t ds = do_reorder ? n : b->stride;
    const uint8_t* dm = do_reorder ? c->smask : b->masks;
    if (!c->fp32) {
        launch_deposit<double>(dp, ds, dm, dep_perm, local_dep, n, g.gx, g.gy, g.gz, c->dom,
                               static_cast<double*>(c->rho), st);
    } else {
        launch_deposit<float>(dp, ds, dm, dep_perm, local_dep, n, g.gx, g.gy, g.gz, c->dom,
                              static_cast<float*>(c->rho), st);
    }
    ev[3] = mark(c, st);
    return SC3D_OK;
}

void enqueue_grid(sc3d_ctx* const* ctxs, int nj, const GridJob* table, hipStream_t st,
                  hipEvent_t* ev, bool green_forked)
{
    sc3d_ctx* c = ctxs[0];
    const Geometry& g = c->g;
    if (green_forked)
        (void)hipStreamWaitEvent(st, c->join_ev, 0);
    else
        enqueue_green(ctxs, nj, table, st);
    ev[4] = mark(c, st);
    if (!c->fp32) {
        launch_solve<double>(g, table, nj, c->tw64, st);
        ev[5] = mark(c, st);
        for (int j = 0; j < nj; ++j) {
            sc3d_ctx* cj = ctxs[j];
            launch_green_patches(cj, st);
            launch_efield<double>(static_cast<double*>(cj->phi), g.gx, g.gy, g.gz, cj->dom,
                                  static_cast<double4*>(cj->E), st);
        }
        ev[6] = mark(c, st);
    } else {
        // the field comes straight out of the spectrum (no potential, no
        // separate difference pass): sc3d_force stays empty in this form
        launch_solve_field<float>(g, table, nj, c->tw32, st);
        for (int j = 0; j < nj; ++j) launch_green_patches(ctxs[j], st);
        ev[5] = mark(c, st);
        ev[6] = mark(c, st);
    }
}

void enqueue_kick(sc3d_ctx* c, const sc3d_bunch* b, const Sched& sc, hipStream_t st)
{
    const Geometry& g = c->g;
    MapProgram post = c->post;
    post.ref_p = b->ref_p;
    post.mass = b->mass;
    const double* src = c->reordering ? c->scratch : nullptr;
    const uint8_t* smasks = c->reordering ? c->smask : nullptr;
    c->reordering = false;
    if (!c->fp32)
        launch_kick<double>(b->parts, b->stride, b->masks, b->n_active, g.gx, g.gy, g.gz, c->dom,
                            static_cast<double4*>(c->E), b->ref_p, b->mass, post, sc, src,
                            smasks, st);
    else
        launch_kick<float>(b->parts, b->stride, b->masks, b->n_active, g.gx, g.gy, g.gz, c->dom,
                           static_cast<float4*>(c->E), b->ref_p, b->mass, post, sc, src, smasks,
                           st);
}

}  // namespace sc3d

extern "C" {

__attribute__((visibility("hidden"))) int sc3d_internal_fail(int code, const char* msg)
{
    return fail(code, msg);
}

const char* sc3d_last_error(void) { return g_err.c_str(); }

void sc3d_options_default(sc3d_options* o, int32_t gx, int32_t gy, int32_t gz)
{
    std::memset(o, 0, sizeof(*o));
    o->shape[0] = gx;
    o->shape[1] = gy;
    o->shape[2] = gz;
    o->green_fn = SC3D_GREEN_LINEAR;
    o->periodic_z = 0;
    o->z_period = 0.0;
    o->grid_entire_period = 0;
    o->n_sigma = 8.0;
    o->kick_scale = 1.0;
    o->domain_fixed = 0;
    o->comm_group_size = 4;
    o->precision = SC3D_PRECISION_FP64;
    o->sort_period = 24;
}

int sc3d_create(const sc3d_options* opt, sc3d_ctx** out)
{
    if (!opt || !out) return fail(SC3D_ERR_ARG, "sc3d_create: null argument");
    *out = nullptr;
    const int gx = opt->shape[0], gy = opt->shape[1], gz = opt->shape[2];
    if (gx < 4 || gy < 4 || gz < 4 || 2 * gx > kAnyMaxLen || 2 * gy > kAnyMaxLen ||
        2 * gz > kAnyMaxLen)
        return fail(SC3D_ERR_ARG, "sc3d_create: grid shape " + std::to_string(gx) + "x" +
                                      std::to_string(gy) + "x" + std::to_string(gz) +
                                      ": each axis must be in [4, 512] (doubled length <= 1024)");
    if (int64_t(gx) * gy * gz >= (int64_t(1) << 31))
        return fail(SC3D_ERR_ARG, "sc3d_create: more than 2^31 grid cells");
    if (opt->green_fn != SC3D_GREEN_LINEAR && opt->green_fn != SC3D_GREEN_POINTLIKE)
        return fail(SC3D_ERR_ARG, "sc3d_create: unknown green_fn");
    if (opt->precision != SC3D_PRECISION_FP64 && opt->precision != SC3D_PRECISION_FP32_GRID)
        return fail(SC3D_ERR_ARG, "sc3d_create: unknown precision");
    if (opt->sort_period < 0) return fail(SC3D_ERR_ARG, "sc3d_create: sort_period < 0");
    sc3d_ctx* c = new sc3d_ctx();
    c->opt = *opt;
    c->fp32 = opt->precision == SC3D_PRECISION_FP32_GRID;
    // the shapes where the reference's g2 index decomposition misplaces
    // points get its values there (sc3d_quirk.hip): few points in direct
    // space, many as a correction spectrum on the generic passes
    std::vector<GreenPatch> patches;
    c->npatch = green_patches(opt->green_fn, opt->shape, patches);
    c->spectral_patch = c->npatch > kMaxGreenPatches;
    c->g = geometry_of(*opt, c->spectral_patch);
    const Geometry& g = c->g;
    auto cleanup = [&](int rc) {
        sc3d_destroy(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess)
        return cleanup(fail(SC3D_ERR_HIP, "sc3d_create: no HIP device"));
    const int64_t rs = c->fp32 ? 4 : 8;
    int rc = 0;
    rc |= dmalloc(c, &c->dom, 1);
    char* p = nullptr;
    c->rho_bytes = (g.G * rs + 15) / 16 * 16;
    rc |= dmalloc(c, &p, c->rho_bytes);
    c->rho = p;
    rc |= dmalloc(c, &p, g.S_elems * 2 * rs);
    c->S = p;
    if (c->fp32) {
        rc |= dmalloc(c, &p, g.S_elems * 2 * rs);
        c->S2 = p;
    } else {
        rc |= dmalloc(c, &p, g.G * rs);
        c->phi = p;
    }
    rc |= dmalloc(c, &p, g.G * 4 * rs);
    c->E = p;
    if (c->fp32)
        rc |= dmalloc(c, &c->ghat32, g.Ghat_elems);
    else
        rc |= dmalloc(c, &c->ghat64, g.Ghat_elems);
    rc |= dmalloc(c, &c->tw64, g.tw_elems);
    rc |= dmalloc(c, &c->tw32, g.tw_elems);
    rc |= dmalloc(c, &c->d_gj, 1);
    if (c->spectral_patch) {
        rc |= dmalloc(c, &p, int64_t(g.Dz) * g.Dy * g.Dx * rs);
        c->Cbuf = p;
        rc |= dmalloc(c, &p, int64_t(g.Dz) * g.Dy * g.Hx * 2 * rs);
        c->chat = p;
    }
    if (rc) return cleanup(rc);
    if (c->spectral_patch &&
        hipMemset(c->Cbuf, 0, size_t(g.Dz) * g.Dy * g.Dx * rs) != hipSuccess)
        return cleanup(fail(SC3D_ERR_HIP, "sc3d_create: workspace initialisation failed"));
    if (c->npatch > 0) {
        if (int rp = dmalloc(c, &c->d_patch, c->npatch)) return cleanup(rp);
        if (hipMemcpy(c->d_patch, patches.data(), patches.size() * sizeof(GreenPatch),
                      hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(SC3D_ERR_HIP, "sc3d_create: workspace initialisation failed"));
    }
    c->gj.rho = c->rho;
    c->gj.S = c->S;
    c->gj.S2 = c->S2;
    c->gj.phi = c->phi;
    c->gj.E = c->E;
    c->gj.ghat64 = c->ghat64;
    c->gj.ghat = c->fp32 ? static_cast<void*>(c->ghat32) : static_cast<void*>(c->ghat64);
    c->gj.dom = c->dom;
    c->gj.chat = c->chat;
    std::vector<double2> t64;
    std::vector<float2> t32;
    make_twiddles(g, t64, t32);
    if (hipMemcpy(c->tw64, t64.data(), t64.size() * sizeof(double2), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->tw32, t32.data(), t32.size() * sizeof(float2), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->d_gj, &c->gj, sizeof(GridJob), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->dom, 0, sizeof(DomainRec)) != hipSuccess)
        return cleanup(fail(SC3D_ERR_HIP, "sc3d_create: workspace initialisation failed"));
    *out = c;
    return SC3D_OK;
}

int sc3d_destroy(sc3d_ctx* c)
{
    if (!c) return SC3D_OK;
    if (c->last_stream || c->applies) (void)hipStreamSynchronize(c->last_stream);
    if (c->green_st) {
        (void)hipStreamSynchronize(c->green_st);
        (void)hipStreamDestroy(c->green_st);
    }
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    if (c->done_ev) (void)hipEventDestroy(c->done_ev);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (void* q : c->owned)
        if (q) (void)hipFree(q);
    delete c;
    return SC3D_OK;
}

int sc3d_apply(sc3d_ctx* c, const sc3d_bunch* b, double time_step, void* stream)
{
    if (int rc = prepare_apply(c, b)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const Sched none{};
    hipEvent_t ev[SC3D_NPHASES + 1];
    // The Green's spectrum beside the sort and the deposit on grids of at
    // least 2^18 cells with at least 2^20 particles, when the caller enqueues
    // ahead of the device (the previous apply still running), and not when
    // timing phases (whose events bracket one phase at a time).  Otherwise
    // the cross-stream fork and join cost more than the overlap gives
    // (measured at 1e5 particles on 32x32x64, synchronised every step: 0.095
    // -> 0.150 ms per step with the fork; on the 64-job mix 2.08e9 -> 2.03e9
    // particle-steps/s when every apply of a grid of 2^18 cells forked).
    const bool eligible = c->timing != SC3D_TIMING_PHASES &&
                          int64_t(c->g.gx) * c->g.gy * c->g.gz >= (int64_t(1) << 18) &&
                          b->n_active >= (int64_t(1) << 20);
    bool ahead = false;
    if (eligible && c->done_rec) {
        ahead = hipEventQuery(c->done_ev) == hipErrorNotReady;
        (void)hipGetLastError();  // a pending event is not an error
    }
    const int forced = test_green_fork();
    const bool fork = forced >= 0 ? forced == 1 && c->timing != SC3D_TIMING_PHASES
                                  : eligible && ahead;
    if (fork && !c->green_st) {
        if (!c->fork_ev) HIP_TRY(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
        if (!c->join_ev) HIP_TRY(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
        HIP_TRY(hipStreamCreateWithFlags(&c->green_st, hipStreamNonBlocking));
    }
    enqueue_front(c, b, time_step, none, st, ev, fork ? c->green_st : nullptr);
    enqueue_grid(&c, 1, c->d_gj, st, ev, fork);
    enqueue_kick(c, b, none, st);
    ev[7] = mark(c, st, true);
    (void)ev;
    c->done_rec = false;
    if (eligible) {
        if (!c->done_ev) HIP_TRY(hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(c->done_ev, st));
        c->done_rec = true;
    }
    HIP_TRY(hipGetLastError());
    c->applies++;
    c->last_stream = st;
    return SC3D_OK;
}

int sc3d_sync(sc3d_ctx* c)
{
    if (!c) return fail(SC3D_ERR_ARG, "sc3d_sync: null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->last_stream));
    int32_t err = 0;
    HIP_TRY(hipMemcpy(&err, &c->dom->error, sizeof(err), hipMemcpyDeviceToHost));
    if (err == SC3D_ERR_DOMAIN)
        return fail(SC3D_ERR_DOMAIN,
                    "Space_charge_3d_open_hockney_eigen::update_domain: all three spatial "
                    "dimensions have neglible extent");
    if (err) return fail(err, "sc3d: device error");
    return SC3D_OK;
}

int sc3d_debug_get(sc3d_ctx* c, int what, void* dst, int64_t bytes)
{
    if (!c || !dst) return fail(SC3D_ERR_ARG, "sc3d_debug_get: null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->last_stream));
    const Geometry& g = c->g;
    auto need = [&](int64_t nb) -> bool { return bytes >= nb; };
    // FP32 grids are returned in the FP64 form's units: the density grid
    // holds unit weights (x w0), the field comes from the density scaled to
    // unit peak (x w0 * peak)
    auto fp32_scale = [&](bool peak_scaled, double& scale) -> int {
        DomainRec r;
        HIP_TRY(hipMemcpy(&r, c->dom, sizeof(r), hipMemcpyDeviceToHost));
        float peak;
        std::memcpy(&peak, &r.rho_peak, sizeof(peak));
        scale = peak_scaled && peak > 0.0f ? r.w0 * double(peak) : r.w0;
        return SC3D_OK;
    };
    auto grid_out = [&](const void* src) -> int {
        if (!need(g.G * 8)) return fail(SC3D_ERR_SIZE, "sc3d_debug_get: buffer too small");
        if (!c->fp32) {
            HIP_TRY(hipMemcpy(dst, src, g.G * 8, hipMemcpyDeviceToHost));
        } else {
            double scale = 1.0;
            if (int rc = fp32_scale(false, scale)) return rc;
            std::vector<float> tmp(g.G);
            HIP_TRY(hipMemcpy(tmp.data(), src, g.G * 4, hipMemcpyDeviceToHost));
            double* d = static_cast<double*>(dst);
            for (int64_t i = 0; i < g.G; ++i) d[i] = double(tmp[i]) * scale;
        }
        return SC3D_OK;
    };
    switch (what) {
    case SC3D_DBG_DOMAIN:
        if (!need(sizeof(DomainRec))) return fail(SC3D_ERR_SIZE, "sc3d_debug_get: too small");
        HIP_TRY(hipMemcpy(dst, c->dom, sizeof(DomainRec), hipMemcpyDeviceToHost));
        return SC3D_OK;
    case SC3D_DBG_KEYS:
    case SC3D_DBG_PERM: {
        if (c->sorted_n < 0) return fail(SC3D_ERR_ARG, "sc3d_debug_get: no sort has run");
        if (!need(c->sorted_n * 4)) return fail(SC3D_ERR_SIZE, "sc3d_debug_get: too small");
        const int32_t* src = what == SC3D_DBG_KEYS ? c->sb.keys0 : c->sb.perm;
        HIP_TRY(hipMemcpy(dst, src, c->sorted_n * 4, hipMemcpyDeviceToHost));
        return SC3D_OK;
    }
    case SC3D_DBG_RHO: return grid_out(c->rho);
    case SC3D_DBG_PHI:
        if (c->fp32)
            return fail(SC3D_ERR_ARG,
                        "sc3d_debug_get: the precision form transforms the field straight "
                        "from the spectrum and keeps no potential");
        return grid_out(c->phi);
    case SC3D_DBG_EX:
    case SC3D_DBG_EY:
    case SC3D_DBG_EZ: {
        if (!need(g.G * 8)) return fail(SC3D_ERR_SIZE, "sc3d_debug_get: buffer too small");
        int comp = what - SC3D_DBG_EX;
        double* d = static_cast<double*>(dst);
        if (!c->fp32) {
            std::vector<double> tmp(g.G * 4);
            HIP_TRY(hipMemcpy(tmp.data(), c->E, g.G * 32, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < g.G; ++i) d[i] = tmp[4 * i + comp];
        } else {
            double scale = 1.0;
            if (int rc = fp32_scale(true, scale)) return rc;
            std::vector<float> tmp(g.G * 4);
            HIP_TRY(hipMemcpy(tmp.data(), c->E, g.G * 16, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < g.G; ++i) d[i] = double(tmp[4 * i + comp]) * scale;
        }
        return SC3D_OK;
    }
    case SC3D_DBG_GHAT:
        if (!need(g.Ghat_elems * 8)) return fail(SC3D_ERR_SIZE, "sc3d_debug_get: too small");
        if (c->fp32) {  // the precision form keeps the spectrum's FP32 copy only
            std::vector<float> tmp(g.Ghat_elems);
            HIP_TRY(hipMemcpy(tmp.data(), c->ghat32, g.Ghat_elems * 4, hipMemcpyDeviceToHost));
            double* out = static_cast<double*>(dst);
            for (int64_t i = 0; i < g.Ghat_elems; ++i) out[i] = tmp[i];
        } else {
            HIP_TRY(hipMemcpy(dst, c->ghat64, g.Ghat_elems * 8, hipMemcpyDeviceToHost));
        }
        return SC3D_OK;
    default: return fail(SC3D_ERR_ARG, "sc3d_debug_get: unknown tap");
    }
}

int sc3d_set_post_kick_maps(sc3d_ctx* c, const sc3d_map_slice* slices, int nslices)
{
    if (!c || nslices < 0 || nslices > SC3D_MAX_SLICES || (nslices > 0 && !slices))
        return fail(SC3D_ERR_ARG, "sc3d_set_post_kick_maps: bad argument");
    for (int k = 0; k < nslices; ++k) {
        if (slices[k].kind != SC3D_MAP_DRIFT && slices[k].kind != SC3D_MAP_QUADRUPOLE)
            return fail(SC3D_ERR_ARG, "sc3d_set_post_kick_maps: unknown slice kind");
        if (slices[k].kind == SC3D_MAP_QUADRUPOLE && slices[k].yoshida_steps < 1)
            return fail(SC3D_ERR_ARG, "sc3d_set_post_kick_maps: yoshida_steps < 1");
        c->post.s[k] = slices[k];
    }
    c->post.n = nslices;
    return SC3D_OK;
}

int sc3d_set_timing(sc3d_ctx* c, int mode)
{
    if (!c) return fail(SC3D_ERR_ARG, "sc3d_set_timing: null context");
    if (mode < SC3D_TIMING_OFF || mode > SC3D_TIMING_APPLY)
        return fail(SC3D_ERR_ARG, "sc3d_set_timing: unknown mode");
    if (mode != c->timing && c->ev_used > 0)
        return fail(SC3D_ERR_ARG, "sc3d_set_timing: read the timing before changing the mode");
    c->timing = mode;
    return SC3D_OK;
}

int sc3d_get_timing(sc3d_ctx* c, double ms[SC3D_NPHASES], int64_t* napplies, int reset)
{
    if (!c || !ms) return fail(SC3D_ERR_ARG, "sc3d_get_timing: null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->last_stream));
    if (c->timing == SC3D_TIMING_APPLY)
        return fail(SC3D_ERR_ARG, "sc3d_get_timing: whole applies timed (sc3d_get_apply_timing)");
    const size_t per = SC3D_NPHASES + 1;
    for (size_t k = 0; k + per <= c->ev_used; k += per) {
        for (int p = 0; p < SC3D_NPHASES; ++p) {
            float t = 0.f;
            HIP_TRY(hipEventElapsedTime(&t, c->ev_pool[k + p], c->ev_pool[k + p + 1]));
            c->ph_ms[p] += t;
        }
        c->ph_applies++;
    }
    c->ev_used = 0;
    for (int p = 0; p < SC3D_NPHASES; ++p) ms[p] = c->ph_ms[p];
    if (napplies) *napplies = c->ph_applies;
    if (reset) {
        for (int p = 0; p < SC3D_NPHASES; ++p) c->ph_ms[p] = 0.0;
        c->ph_applies = 0;
    }
    return SC3D_OK;
}

int sc3d_get_apply_timing(sc3d_ctx* c, double* ms, int64_t* napplies, int reset)
{
    if (!c || !ms) return fail(SC3D_ERR_ARG, "sc3d_get_apply_timing: null argument");
    if (c->timing == SC3D_TIMING_PHASES)
        return fail(SC3D_ERR_ARG, "sc3d_get_apply_timing: phases timed (sc3d_get_timing)");
    HIP_TRY(hipSetDevice(c->device));
    if (c->ev_used > 0) HIP_TRY(hipStreamSynchronize(c->last_stream));
    for (size_t k = 0; k + 2 <= c->ev_used; k += 2) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, c->ev_pool[k], c->ev_pool[k + 1]));
        c->apply_ms += t;
        c->apply_n++;
    }
    c->ev_used = 0;
    *ms = c->apply_ms;
    if (napplies) *napplies = c->apply_n;
    if (reset) {
        c->apply_ms = 0.0;
        c->apply_n = 0;
    }
    return SC3D_OK;
}

int sc3d_moments(sc3d_ctx* c, const sc3d_bunch* b, double mean[6], double mom2[36],
                 void* stream)
{
    if (!c || !b || !mean || !mom2) return fail(SC3D_ERR_ARG, "sc3d_moments: null argument");
    const int64_t n = b->n_active;
    if (n < 0 || b->stride < n || (n > 0 && (!b->parts || !b->masks)))
        return fail(SC3D_ERR_ARG, "sc3d_moments: bad bunch");
    HIP_TRY(hipSetDevice(c->device));
    const int64_t need = ((n + 4095) / 4096) * 21 + 64;
    if (need > c->M_cap) {
        dfree(c, c->M);
        c->M = nullptr;
        c->M_cap = 0;
        if (int rc = dmalloc(c, &c->M, need + 27)) return rc;
        c->M_cap = need;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* out = c->M + c->M_cap;  // 27 result doubles after the partials
    launch_moments(b->parts, b->stride, b->masks, n, b->total_num, c->M, out, st);
    HIP_TRY(hipGetLastError());
    double h[27];
    HIP_TRY(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int j = 0; j < 6; ++j) mean[j] = h[j];
    int q = 6;
    for (int j = 0; j < 6; ++j)
        for (int k = 0; k <= j; ++k) {
            mom2[j * 6 + k] = h[q];
            mom2[k * 6 + j] = h[q];
            ++q;
        }
    return SC3D_OK;
}

int sc3d_tail_fraction(sc3d_ctx* c, const sc3d_bunch* b, const double center[6],
                       const double width[6], double nsigma, double frac[6], void* stream)
{
    if (!c || !b || !center || !width || !frac)
        return fail(SC3D_ERR_ARG, "sc3d_tail_fraction: null argument");
    const int64_t n = b->n_active;
    if (n < 0 || b->stride < n || (n > 0 && (!b->parts || !b->masks)))
        return fail(SC3D_ERR_ARG, "sc3d_tail_fraction: bad bunch");
    HIP_TRY(hipSetDevice(c->device));
    if (!c->tails)
        if (int rc = dmalloc(c, &c->tails, 8)) return rc;
    TailArgs a{};
    for (int j = 0; j < 6; ++j) {
        a.c[j] = center[j];
        a.lim[j] = nsigma * width[j];
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    launch_tail_counts(b->parts, b->stride, b->masks, n, a, c->tails, st);
    HIP_TRY(hipGetLastError());
    unsigned long long h[6];
    HIP_TRY(hipMemcpyAsync(h, c->tails, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int j = 0; j < 6; ++j) frac[j] = double(h[j]) / b->total_num;
    return SC3D_OK;
}

int sc3d_adjust_moments(const sc3d_bunch* b, const double means[6], const double cov[36],
                        void* stream)
{
    if (!b || !means || !cov) return fail(SC3D_ERR_ARG, "adjust_moments: null argument");
    const int64_t n = b->n_active;
    if (n < 1 || b->stride < n || !b->parts || !b->masks || !(b->total_num > 0.0))
        return fail(SC3D_ERR_ARG, "adjust_moments: bad bunch");
    // is_symmetric66 (populate.cc:12-27): floating_point_equal, tolerance 1e-14
    auto fp_equal = [](double a, double c) {
        return std::abs(a) < 1.0e-14 ? std::abs(a - c) < 1.0e-14
                                     : std::abs((a - c) / a) < 1.0e-14;
    };
    for (int i = 0; i < 6; ++i)
        for (int j = i + 1; j < 6; ++j)
            if (!fp_equal(cov[i * 6 + j], cov[j * 6 + i]))
                return fail(SC3D_ERR_ARG, "adjust_moments: covariance matrix must be symmetric");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the bunch's mean and second moments (sc3d_moments' kernels); the
    // partials live only for this call
    const int64_t parts_n = ((n + 4095) / 4096) * 21 + 64;
    double* P = nullptr;
    HIP_TRY(hipMalloc(&P, size_t(parts_n + 27) * sizeof(double)));
    launch_moments(b->parts, b->stride, b->masks, n, b->total_num, P, P + parts_n, st);
    double h[27];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, P + parts_n, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(P);
    if (e != hipSuccess)
        return fail(SC3D_ERR_HIP, std::string("adjust_moments: ") + hipGetErrorString(e));
    double mom2[36];
    int q = 6;
    for (int j = 0; j < 6; ++j)
        for (int k = 0; k <= j; ++k) {
            mom2[j * 6 + k] = mom2[k * 6 + j] = h[q];
            ++q;
        }
    // A = G H^-1 (Eigen LLT factors there): Cholesky-Banachiewicz, sums in
    // ascending k; H^-1 by forward substitution
    auto chol = [](const double* C, double* L) {
        for (int i = 0; i < 36; ++i) L[i] = 0.0;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = C[i * 6 + j];
                for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k];
                if (i == j) {
                    if (!(s > 0.0)) return false;
                    L[i * 6 + i] = std::sqrt(s);
                } else {
                    L[i * 6 + j] = s / L[j * 6 + j];
                }
            }
        return true;
    };
    double G[36], H[36], Hi[36];
    if (!chol(cov, G))
        return fail(SC3D_ERR_ARG, "adjust_moments: covariance matrix is not positive definite");
    if (!chol(mom2, H))
        return fail(SC3D_ERR_ARG,
                    "adjust_moments: the bunch's second moments are not positive definite");
    for (int i = 0; i < 36; ++i) Hi[i] = 0.0;
    for (int c = 0; c < 6; ++c) {
        Hi[c * 6 + c] = 1.0 / H[c * 6 + c];
        for (int i = c + 1; i < 6; ++i) {
            double s = 0.0;
            for (int k = c; k < i; ++k) s += H[i * 6 + k] * Hi[k * 6 + c];
            Hi[i * 6 + c] = -s / H[i * 6 + i];
        }
    }
    AdjustArgs a{};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += G[i * 6 + k] * Hi[k * 6 + j];
            a.A[i * 6 + j] = s;
        }
    for (int j = 0; j < 6; ++j) {
        a.mean[j] = h[j];
        a.means[j] = means[j];
    }
    launch_adjust(b->parts, b->stride, n, a, st);
    HIP_TRY(hipGetLastError());
    return SC3D_OK;
}

int sc3d_populate_gaussian(double* parts, int64_t stride, int64_t n, uint8_t* masks,
                           const double mean[6], const double chol[21], uint64_t seed,
                           int64_t id0, void* stream)
{
    if (n < 0 || stride < n || (n > 0 && (!parts || !masks)) || !mean || !chol)
        return fail(SC3D_ERR_ARG, "sc3d_populate_gaussian: bad argument");
    if (n == 0) return SC3D_OK;
    PopulateArgs a{};
    std::memcpy(a.mean, mean, sizeof(a.mean));
    std::memcpy(a.chol, chol, sizeof(a.chol));
    a.seed = seed;
    a.id0 = id0;
    launch_populate(parts, stride, n, masks, a, static_cast<hipStream_t>(stream));
    HIP_TRY(hipGetLastError());
    return SC3D_OK;
}

}  // extern "C"
