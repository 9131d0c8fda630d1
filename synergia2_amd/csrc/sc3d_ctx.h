// sc3d_ctx.h -- the operator context (device workspaces of one operator
// instance, Space_charge_3d_open_hockney's rho2/g2/phi2/En* and FFT plans,
// space_charge_3d_open_hockney.h:22-37) and the three pieces one apply is
// made of, shared by the C ABI (sc3d_capi.hip) and the served form
// (sc3d_serve.hip).  Internal to libsc3d_hip.so.
#pragma once

#include <string>
#include <vector>

#include "sc3d_particles.h"

struct sc3d_ctx {
    sc3d_options opt;
    sc3d::Geometry g;
    int device = 0;
    bool fp32 = false;
    sc3d::DomainRec* dom = nullptr;
    double* P = nullptr;
    int64_t P_cap = 0;
    void* rho = nullptr;
    int64_t rho_bytes = 0;  // the density grid, rounded up to 16 bytes
    void* S = nullptr;
    void* S2 = nullptr;   // precision form: second spectrum workspace
    void* phi = nullptr;  // FP64 form only
    void* E = nullptr;
    double* ghat64 = nullptr;
    float* ghat32 = nullptr;
    double2* tw64 = nullptr;
    float2* tw32 = nullptr;
    sc3d::GridJob gj{};               // this context's grid pointers
    sc3d::GridJob* d_gj = nullptr;    // the same, as a device table of one
    int64_t pcap = 0;
    sc3d::SortBuffers sb{};
    double* scratch = nullptr;
    uint8_t* smask = nullptr;
    // the apply being enqueued reorders the bunch: its deposit reads the
    // gathered copy in scratch / smask, and its kick writes it back
    bool reordering = false;
    // Green's spectrum stream of the plain form (sc3d_apply): the spectrum
    // depends only on the domain, so it runs beside the sort and the deposit
    // (fork after the domain, join before the convolution) when the caller
    // enqueues ahead of the device; created on first use, default priority.
    // done_ev marks the end of the last apply (recorded once done_rec).
    hipStream_t green_st = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, done_ev = nullptr;
    bool done_rec = false;
    int64_t applies = 0;
    int64_t sorted_n = -1;
    hipStream_t last_stream = nullptr;
    std::vector<void*> owned;
    // timing (SC3D_TIMING_*): SC3D_NPHASES+1 events per timed apply, or 2
    // (whole applies)
    int timing = SC3D_TIMING_OFF;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double ph_ms[SC3D_NPHASES] = {0};
    double apply_ms = 0.0;
    int64_t apply_n = 0;
    int64_t ph_applies = 0;
    sc3d::MapProgram post{};  // maps fused into the kick kernel (post.n == 0: none)
    double* M = nullptr;      // moment partials (diagnostics)
    int64_t M_cap = 0;
    unsigned long long* tails = nullptr;  // tail counts (diagnostics)
    int npatch = 0;                 // reference g2 patches of this shape
    sc3d::GreenPatch* d_patch = nullptr;
    bool spectral_patch = false;    // npatch > kMaxGreenPatches: correction as chat
    void* Cbuf = nullptr;           // R[Dz][Dy][Dx] g2 correction (spectral_patch)
    void* chat = nullptr;           // its spectrum, cplx<R>[Dz][Dy][Hx]
};

namespace sc3d {

// Sets this thread's last error message (sc3d_last_error) and returns code.
int fail(int code, const std::string& msg);

#define SC3D_HIP_TRY(expr)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::sc3d::fail(SC3D_ERR_HIP, std::string(#expr " failed: ") +            \
                                                  hipGetErrorString(e_));                  \
    } while (0)

// Test hook: with the environment variable SC3D_TEST_MAX_ALLOC_BYTES set to
// B, every workspace allocation of more than B bytes fails as a failed
// hipMalloc would (before any device call), so tests can drive the error
// paths; -1 when unset.
int64_t test_alloc_limit();

template <typename T>
int dmalloc(sc3d_ctx* c, T** p, int64_t n)
{
    void* q = nullptr;
    if (n <= 0) n = 1;
    const int64_t lim = test_alloc_limit();
    if (lim >= 0 && int64_t(n) * int64_t(sizeof(T)) > lim)
        return ::sc3d::fail(SC3D_ERR_HIP, "hipMalloc failed: out of memory "
                                          "(SC3D_TEST_MAX_ALLOC_BYTES)");
    SC3D_HIP_TRY(hipMalloc(&q, size_t(n) * sizeof(T)));
    c->owned.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

void dfree(sc3d_ctx* c, void* p);

// Records the next pooled timing event on st when timing phases, or when
// timing whole applies and `edge` (the apply's first and last mark); else
// nullptr.
hipEvent_t mark(sc3d_ctx* c, hipStream_t st, bool edge = false);

// Device bytes a context of these options holds once it has applied to a
// bunch of n particles (the admission footprint of the served form).
int64_t ctx_footprint(const sc3d_options& o, int64_t n);

// Argument checks of an apply and the particle-side buffers for its bunch.
int prepare_apply(sc3d_ctx* c, const sc3d_bunch* b);

// The particle front of an apply: domain, cell sort, reorder, density
// (events ev[0..3] when timing).  With sc.ctl (served form) the step
// control decides reorder, schedule and kick time step on the device.
// With green_st, the context's Green's spectrum is enqueued there right
// after the domain (fork on c->fork_ev, done at c->join_ev), beside the
// sort and the deposit on st.
int enqueue_front(sc3d_ctx* c, const sc3d_bunch* b, double time_step, const Sched& sc,
                  hipStream_t st, hipEvent_t* ev, hipStream_t green_st = nullptr);

// The grid middle for nj contexts of one shape and precision: Green's
// spectrum (where h changed), convolution and field, as batched launches
// over the device table `table` (events ev[4..6] when timing).
// green_forked: enqueue_front put the spectrum of the one context on its
// side stream; st waits for it instead.
void enqueue_grid(sc3d_ctx* const* ctxs, int nj, const GridJob* table, hipStream_t st,
                  hipEvent_t* ev, bool green_forked = false);

// The points where the reference's g2 differs from the symmetric one at
// this shape (sc3d_quirk.hip); returns their number.
int green_patches(int green_fn, const int shape[3], std::vector<GreenPatch>& out);

// Adds the patches' convolution with rho: to phi (FP64 form, before the
// field) or, differenced, to E (precision form, after the field passes);
// nothing for a spectral_patch context.
void launch_green_patches(const sc3d_ctx* c, hipStream_t st);

// spectral_patch contexts: the step's patch values into Cbuf and their
// spectrum into chat (where h changed), after the Green's spectrum.
void launch_green_chat(const sc3d_ctx* c, hipStream_t st);

// The kick (with the fused post-kick maps: c->post, or the schedule's).
void enqueue_kick(sc3d_ctx* c, const sc3d_bunch* b, const Sched& sc, hipStream_t st);

}  // namespace sc3d
