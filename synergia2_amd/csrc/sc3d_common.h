// sc3d_common.h -- shared device definitions of the MI355X space-charge step.
//
// The whole library is compiled with -ffp-contract=off: every a*b+c below is
// two IEEE-rounded operations, as in the reference's CPU build, which is what
// keeps cell indices and per-particle weights identical bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sc3d_capi.h"

namespace sc3d {

// physical_constants.h:17-24 (PDG 2012) and :94-106
constexpr double kPi = 3.141592653589793;
constexpr double kE = 1.602176565e-19;
constexpr double kC = 299792458.0;
constexpr double kMu0 = 4 * kPi * 1.0e-7;
constexpr double kEpsilon0 = 1.0 / (kC * kC * kMu0);

constexpr int kThreads = 256;

// Device-resident domain record (same layout as sc3d_domain_record).
struct DomainRec {
    double sums[6];
    double mean[3], stddev[3];
    double size[3], offset[3];
    double left[3], h[3];
    double w0, fn_norm, factor;
    double inv_h[3], half_inv_h[3];
    double ghat_h[3];
    int32_t error, ghat_dirty, ghat_valid;
    int32_t rho_peak;  // precision form: FP32 bits of the unit-weight density peak
};
static_assert(sizeof(DomainRec) == sizeof(sc3d_domain_record), "layout");

// collective/utils.h:9-24: s = (pos-left)*inv_h - 0.5, idx = floor(s),
// off = s - idx.  Positions whose scaled location is beyond +-1e9 cells (or
// NaN) return an index far outside any grid, keeping the int conversion
// defined; such particles are outside the grid either way.
__device__ __forceinline__ int cell_index(double pos, double left, double inv_h,
                                          double& off)
{
    double s = (pos - left) * inv_h - 0.5;
    if (!(s > -1.0e9 && s < 1.0e9)) {
        off = 0.0;
        return -1000000000;
    }
    int idx = static_cast<int>(floor(s));
    off = s - idx;
    return idx;
}

// Device-side control of the applies of one context in the served form
// (kernels take it as a nullable pointer; null = the host decides, as in
// sc3d_apply).  k_step_begin opens each apply: which scheduled map programs
// run before / after the kick and which kick time step applies, and whether
// the job has finished (then every kernel of the apply exits).  The launch
// sequence of an apply is then the same every step, so captured graphs
// serve every step (the server picks the one with or without the reorder).
struct StepCtl {
    int64_t applies;      // applies opened so far
    int32_t turn_steps;   // schedule length (steps per turn)
    int32_t total_steps;  // steps of the job
    int32_t cur;          // schedule index of this apply
    int32_t first_prog;   // program table index before this kick, -1: none
    int32_t after_prog;   // program table index fused after this kick
    int32_t done;         // all steps taken: this apply does nothing
};

// Domain-record error value of an apply of a finished job (served form):
// every kernel of that apply exits.  Not a C-ABI error code.
constexpr int32_t kDoneMark = 1;

template <typename R> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };

template <typename T>
__device__ __forceinline__ T cmul(T a, T b)
{
    T r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <typename T>
__device__ __forceinline__ T cmul_conj(T a, T b)  // a * conj(b)
{
    T r;
    r.x = a.x * b.x + a.y * b.y;
    r.y = a.y * b.x - a.x * b.y;
    return r;
}

}  // namespace sc3d
