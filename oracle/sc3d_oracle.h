/*
 * sc3d_oracle.h -- CPU restatement of Synergia's 3D open-boundary Hockney
 * space-charge step, used ONLY as test infrastructure: the parity checker of
 * the HIP library and the timed CPU baseline.  Nothing in the product links it.
 *
 * The reference cannot be built in this image (empty Kokkos/Eigen submodules,
 * no FFTW-MPI/MPI/GSL/HDF5/cereal), so this is a line-by-line restatement of
 *   src/synergia/collective/space_charge_3d_open_hockney.cc:94-109,136-372,493-784
 *   src/synergia/collective/space_charge_3d_kernels.h:91-307
 *   src/synergia/collective/deposit.cc:21-30,309-395,852-1031
 *   src/synergia/collective/utils.h:9-24
 *   src/synergia/collective/rectangular_grid_domain.h:20-36
 *   src/synergia/utils/distributed_fft3d.h:18-72, distributed_fft3d_fftw.cc:50-119
 *   src/synergia/bunch/core_diagnostics.cc:139-153,422-462
 *   src/synergia/foundation/physical_constants.h:17-24,94-106
 * with its own FP64 FFT in place of FFTW 3.3.10.
 *
 * Every function is pinned by a check in tests/ (golden numpy FFT vectors,
 * closed forms, brute-force convolution, charge conservation, rod-of-charge
 * physics) -- see the per-function comments in sc3d_oracle.cc.
 */
#ifndef SC3D_ORACLE_H
#define SC3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_options {
    int32_t shape[3];            /* x, y, z grid shape                      */
    double size[3], offset[3];   /* used when domain_fixed                   */
    int32_t green_fn;            /* 0 pointlike, 1 linear                    */
    int32_t periodic_z;
    int32_t grid_entire_period;
    int32_t domain_fixed;
    double z_period;
    double n_sigma;
    double kick_scale;
    int32_t reorder;             /* 1: after the domain, reorder the bunch by
                                    the stable cell sort, as the device
                                    library does when its sort_period is due
                                    (storage order only; no physics change)  */
} orc_options;

typedef struct orc_bunch {
    double* parts;               /* host, column j at parts + j*stride       */
    int64_t stride;
    int64_t n_active;
    uint8_t* masks;
    double total_num;
    double real_num;
    int32_t charge;
    double mass, ref_gamma, ref_beta, ref_p;
} orc_bunch;

/* Domain and derived scalars of one step, in the same order as the device
 * library's domain record. */
typedef struct orc_domain {
    double sums[6];              /* Sx Sy Scdt Sxx Syy Scdtcdt              */
    double mean[3], stddev[3];
    double size[3], offset[3];
    double left[3], h[3];
    double w0;                   /* deposit weight0                          */
    double fn_norm;              /* get_normalization_force                  */
    double factor;               /* kick factor                              */
} orc_domain;

/* Optional outputs of orc_apply; any pointer may be NULL. */
typedef struct orc_debug {
    orc_domain* domain;
    int32_t* keys;               /* n_active cell keys                       */
    int32_t* perm;               /* n_active stable sort permutation         */
    double* rho;                 /* G = gx*gy*gz octant of rho2              */
    double* phi;                 /* G octant of phi2 (unnormalised)          */
    double* ex; double* ey; double* ez;   /* G each                          */
    double* rho2_full;           /* Px*Dy*Dz padded doubled rho2             */
    double* g2_full;             /* Px*Dy*Dz padded doubled g2 (real space)  */
    double* ghat_full;           /* Px*Dy*Dz FFT of g2 (complex interleaved) */
    double seconds[6];           /* domain, rho, green, phi, force, kick     */
} orc_debug;

int orc_apply(const orc_options* opt, orc_bunch* bunch, double time_step,
              int nthreads, orc_debug* dbg);
const char* orc_last_error(void);

/* ---- pieces, exported for unit tests ---- */
void orc_tree_sums(const double* parts, int64_t stride, int64_t n,
                   const uint8_t* masks, double out[6]);
int orc_compute_domain(const orc_options* opt, const orc_bunch* b,
                       double time_step, orc_domain* d);
int32_t orc_cell_index(double pos, double left, double inv_h, double* off);
void orc_cell_keys(const orc_bunch* b, const orc_domain* d, const int32_t g[3],
                   int32_t* keys);
void orc_stable_sort(const int32_t* keys, int64_t n, int32_t* perm);
/* 1-D complex FFT, in place, interleaved re/im; sign -1 forward, +1 inverse,
 * unnormalised. */
void orc_fft1d(double* data, int n, int sign);
/* 3-D real-to-complex on dims (nz, ny, nx), x fastest; in: padded real rows
 * of 2*(nx/2+1); out: complex rows of nx/2+1 (may alias in). */
void orc_fft3d_r2c(int nx, int ny, int nz, const double* in, double* out,
                   int nthreads);
void orc_fft3d_c2r(int nx, int ny, int nz, const double* in, double* out,
                   int nthreads);
void orc_green_fn2(int green_fn, const int32_t g[3], const double h[3],
                   double* g2 /* Px*Dy*Dz */);
void orc_deposit(const orc_bunch* b, const orc_domain* d, const int32_t g[3],
                 double* rho2 /* Px*Dy*Dz */, int nthreads);
void orc_force(const double* phi2, const int32_t g[3], const double h[3],
               double* ex, double* ey, double* ez);
void orc_kick(orc_bunch* b, const double* ex, const double* ey,
              const double* ez, const int32_t g[3], const orc_domain* d);

/* Single-particle maps (libFF restatement): kind 0 drift (drift_unit,
 * ff_algorithm.h:45-79), kind 1 quadrupole (yoshida6 + thin quadrupole
 * kicks, ff_algorithm.h:928-936,1835-2008, ff_quadrupole.h:438-468). */
typedef struct orc_slice {
    int32_t kind, yoshida_steps;
    double length, k1, k1s, ref_cdt;
} orc_slice;
double orc_reference_cdt(const orc_slice* s, double ref_p, double mass);
/* Core_diagnostics::calculate_mean / calculate_mom2 (core_diagnostics.cc:
 * 202-229,318-362) with the fixed tree of orc_tree_sums. */
void orc_moments(const double* parts, int64_t stride, int64_t n, const uint8_t* masks,
                 double total_num, double mean[6], double mom2[36]);
void orc_apply_maps(orc_bunch* b, const orc_slice* s, int nslices);

/* Counter-based Gaussian populate (the build's generator; see the .cc). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_det_log(double u);
void orc_det_sincos2pi(double u, double* sn, double* cs);
int orc_adjust_matrix(const double cov[36], const double mom2[36], double A[36]);
int orc_adjust_moments(double* parts, int64_t stride, int64_t n, const uint8_t* masks,
                       double total_num, const double means[6], const double cov[36]);
void orc_populate_gaussian(double* parts, int64_t stride, int64_t n, uint8_t* masks,
                           const double mean[6], const double chol[21], uint64_t seed,
                           int64_t id0);

#ifdef __cplusplus
}
#endif
#endif
