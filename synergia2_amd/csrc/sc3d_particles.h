// sc3d_particles.h -- host launch interface of the particle-side kernels.
// Internal to libsc3d_hip.so.
#pragma once

#include "sc3d_internal.h"
#include "sc3d_maps.h"

namespace sc3d {

template <typename R> struct vec4;
template <> struct vec4<double> { using type = double4; };
template <> struct vec4<float> { using type = float4; };

struct DomainFinishArgs {
    int32_t fixed, green_fn, grid_entire_period, charge;
    int32_t shape[3];
    double size[3], offset[3];
    double n_sigma, z_period, kick_scale;
    double total_num, real_num;
    double time_step, ref_p, ref_gamma, ref_beta;
};

// Served form: the device-side step control of a context and its schedule
// (all null for host-decided applies).
struct Sched {
    const StepCtl* ctl;
    const MapProgram* progs;  // [first | between | last] x turn_steps
    const double* dt;         // kick time step per schedule entry
    int32_t sort_now;         // served form: this iteration sorts and reorders the bunch
};

struct PopulateArgs {
    double mean[6];
    double chol[21];
    uint64_t seed;
    int64_t id0;
};

struct SortBuffers {
    int32_t* keys0;   // unsorted keys (kept for inspection)
    int32_t* keysA;
    int32_t* keysB;
    int32_t* valsA;
    int32_t* valsB;
    int32_t* counts;  // 256 * tiles
    int32_t* totals;  // 256
    int32_t* sorted_keys;  // result views (point into A/B)
    int32_t* perm;
};

constexpr int kSortTileItems = 4096;

void launch_domain(const double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                   double* P, const DomainFinishArgs& a, DomainRec* dom, const Sched& sc,
                   void* grid, int64_t grid_bytes, hipStream_t st);
void launch_step_begin(StepCtl* ctl, hipStream_t st);
// scheduled maps before the kick (the first step's first half-step)
void launch_maps_sched(double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                       const Sched& sc, hipStream_t st);
int sort_passes(int64_t G);
void launch_sort(const double* parts, int64_t stride, const uint8_t* masks, int64_t n, int gx,
                 int gy, int gz, const DomainRec* dom, SortBuffers& b, hipStream_t st);
void launch_reorder(double* parts, int64_t stride, uint8_t* masks, int64_t n, const int32_t* perm,
                    double* scratch, uint8_t* smask, const DomainRec* dom, hipStream_t st);
// Deposit: local (k_deposit_tile): windows of the bunch storage through an
// LDS node tile of each window's cell box (the storage is cell-sorted at
// every reorder); else (k_deposit): through the permutation `perm` of this
// step's sort, compensated run sums in FP32.  FP32: unit weights; then
// k_rho_peak records the density peak in the domain record.
template <typename R>
void launch_deposit(const double* parts, int64_t stride, const uint8_t* masks,
                    const int32_t* perm, bool local, int64_t n, int gx, int gy, int gz,
                    DomainRec* dom, R* rho, hipStream_t st);
template <>
void launch_deposit<float>(const double* parts, int64_t stride, const uint8_t* masks,
                           const int32_t* perm, bool local, int64_t n, int gx, int gy, int gz,
                           DomainRec* dom, float* rho, hipStream_t st);
template <typename R>
void launch_efield(const R* phi, int gx, int gy, int gz, const DomainRec* dom,
                   typename vec4<R>::type* E, hipStream_t st);
// Kick (+ maps): src / smasks non-null on the apply that reordered the bunch:
// the particles are read from the reorder's gathered copy (7 columns of
// stride n) and written whole to the storage (the reorder's copy back).
template <typename R>
void launch_kick(double* parts, int64_t stride, uint8_t* masks, int64_t n, int gx, int gy, int gz,
                 const DomainRec* dom, const typename vec4<R>::type* E, double pref, double m,
                 const MapProgram& prog, const Sched& sc, const double* src,
                 const uint8_t* smasks, hipStream_t st);
void launch_moments(const double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                    double total_num, double* P, double* out, hipStream_t st);
struct TailArgs {
    double c[6];    // centre per plane
    double lim[6];  // |p - c| beyond this counts as tail
};
void launch_tail_counts(const double* parts, int64_t stride, const uint8_t* masks, int64_t n,
                        const TailArgs& a, unsigned long long* counts, hipStream_t st);
void launch_populate(double* parts, int64_t stride, int64_t n, uint8_t* masks,
                     const PopulateArgs& a, hipStream_t st);
// adjust_moments' particle map (populate_host.cc:30-43): every row i < n
// becomes (p - mean) A^T + means, sums in ascending k.
struct AdjustArgs {
    double A[36];
    double mean[6];   // the bunch's mean
    double means[6];  // the requested mean
};
void launch_adjust(double* parts, int64_t stride, int64_t n, const AdjustArgs& a, hipStream_t st);

}  // namespace sc3d
