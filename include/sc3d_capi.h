/*
 * sc3d_capi.h -- C ABI of libsc3d_hip.so, the MI355X (gfx950) implementation
 * of Synergia's 3D open-boundary Hockney space-charge step.
 *
 * No C++ or torch types cross this boundary: plain structs, device pointers
 * and sizes.  The C++ operator class synergia2_amd/host/
 * space_charge_3d_open_hockney.h is the only in-tree caller; every other
 * binding (ctypes in synergia2_amd/__init__.py) uses
 * exactly these entry points.
 *
 * Reference interfaces replaced (paths relative to the synergia2 tree):
 *   sc3d_options   <- Space_charge_3d_open_hockney_options,
 *                     src/synergia/simulation/implemented_collective_options.h:67-126
 *   sc3d_bunch     <- the Bunch accessors the operator reads,
 *                     src/synergia/bunch/bunch.h:267-289,317,458-472 and the
 *                     7-column LayoutLeft particle view, bunch_particles.h:24-43
 *   sc3d_create    <- Space_charge_3d_open_hockney::Space_charge_3d_open_hockney
 *                     + construct_workspaces,
 *                     src/synergia/collective/space_charge_3d_open_hockney.cc:376-397,455-491
 *   sc3d_apply     <- Space_charge_3d_open_hockney::apply_bunch,
 *                     space_charge_3d_open_hockney.cc:422-453 (called per bunch by
 *                     apply_impl, :399-420, behind Operator::apply, operator.h:62-65)
 *   sc3d_sync      <- the Kokkos::fence() that ends every stage of apply_bunch
 *                     (space_charge_3d_open_hockney.cc:613,638,650,653,671,675,749,783)
 *   sc3d_last_error<- the std::runtime_error messages thrown by the operator
 *                     (space_charge_3d_open_hockney.cc:512-516,594-598,619-623)
 *   sc3d_debug_get <- no single counterpart: copies out the workspaces the
 *                     reference keeps as members (rho2, phi2, enx/eny/enz,
 *                     space_charge_3d_open_hockney.h:22-37) for parity tests.
 *   sc3d_populate_gaussian <- populate_6d, src/synergia/bunch/populate.cc:128-145
 *                     (own counter-based generator, not GSL ranlxd2)
 *   sc3d_adjust_moments    <- adjust_moments, populate.cc:30-61
 */
#ifndef SC3D_CAPI_H
#define SC3D_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define SC3D_OK 0
#define SC3D_ERR_ARG (-1)         /* invalid argument / unsupported shape    */
#define SC3D_ERR_HIP (-2)         /* HIP runtime failure                      */
#define SC3D_ERR_DOMAIN (-3)      /* all three extents negligible            */
#define SC3D_ERR_PERIODIC (-4)    /* periodic_z not implemented (reference)  */
#define SC3D_ERR_SIZE (-5)        /* workspace/bunch size limit exceeded     */

#define SC3D_GREEN_POINTLIKE 0
#define SC3D_GREEN_LINEAR 1

#define SC3D_PRECISION_FP64 0     /* plain form: everything FP64             */
#define SC3D_PRECISION_FP32_GRID 1 /* deposit/FFT/field FP32, particles FP64 */

typedef struct sc3d_options {
    /* the reference option fields, same meaning and defaults */
    int32_t shape[3];            /* x, y, z; default 32, 32, 64              */
    double size[3];              /* fixed-domain size  (domain_fixed only)   */
    double offset[3];            /* fixed-domain centre (domain_fixed only)  */
    int32_t green_fn;            /* default SC3D_GREEN_LINEAR                */
    int32_t periodic_z;          /* default 0; 1 -> SC3D_ERR_PERIODIC        */
    int32_t grid_entire_period;  /* default 0                                */
    int32_t domain_fixed;        /* default 0                                */
    int32_t comm_group_size;     /* default 4; accepted, one GPU per bunch   */
    double z_period;             /* default 0                                */
    double n_sigma;              /* default 8                                */
    double kick_scale;           /* default 1                                */
    /* build extensions (no reference counterpart) */
    int32_t precision;           /* SC3D_PRECISION_*; default FP64           */
    int32_t sort_period;         /* every N applies the particles are sorted
                                    by cell and the bunch storage is reordered
                                    into that order (the deposit sorts windows
                                    of the storage on the device in between);
                                    0 = never reorder: sort on every apply and
                                    deposit through the permutation; default 24*/
} sc3d_options;

typedef struct sc3d_bunch {
    double* parts;               /* device; column j at parts + j*stride     */
    int64_t stride;              /* capacity (LayoutLeft leading dimension)  */
    int64_t n_active;            /* Bunch::size()                            */
    uint8_t* masks;              /* device, one byte per slot, 1 = active    */
    double total_num;            /* Bunch::get_total_num()                   */
    double real_num;             /* Bunch::get_real_num()                    */
    int32_t charge;              /* Bunch::get_particle_charge()             */
    double mass;                 /* Bunch::get_mass() [GeV/c^2]              */
    double ref_gamma, ref_beta, ref_p;  /* reference particle, p in GeV/c    */
} sc3d_bunch;

typedef struct sc3d_ctx sc3d_ctx;

/* Fills the reference defaults for a gx x gy x gz grid
 * (implemented_collective_options.h:82-95). */
void sc3d_options_default(sc3d_options* opt, int32_t gx, int32_t gy, int32_t gz);

/* Creates a context on the current HIP device: validates the options and
 * allocates the grid workspaces.  Any shape with every axis in [4, 512]
 * (doubled lengths up to 1024) and fewer than 2^31 cells; power-of-two axes
 * (x in [8, 256], doubled y and z in [8, 512]) run the tuned transform
 * passes, other lengths generic mixed-radix ones. */
int sc3d_create(const sc3d_options* opt, sc3d_ctx** out);

/* Enqueues one space-charge kick of length time_step [s] on `stream`
 * (hipStream_t, NULL = default stream).  Asynchronous: errors found on the
 * device (negligible extent) are reported by sc3d_sync.  When sort_period
 * says so, the first n_active slots of the bunch (all 7 columns and the
 * masks) are reordered in place into cell order; particle identity is
 * column 6. */
int sc3d_apply(sc3d_ctx* ctx, const sc3d_bunch* bunch, double time_step,
               void* stream);

/* Waits for the context's last stream and reports device-side errors. */
int sc3d_sync(sc3d_ctx* ctx);

int sc3d_destroy(sc3d_ctx* ctx);

/* Message of the last failure on this thread ("" if none). */
const char* sc3d_last_error(void);

/* Debug taps: synchronise and copy a workspace to host memory. */
#define SC3D_DBG_DOMAIN 0        /* sc3d_domain_record                       */
#define SC3D_DBG_KEYS 1          /* int32[n] cell keys of the last sort, in  */
                                 /* the bunch order before that sort (brick */
                                 /* order: bricks of 4 cells per axis where */
                                 /* 4 divides it, on a serpentine path, x   */
                                 /* fastest; cells of a brick z-major, x    */
                                 /* fastest; G for particles outside        */
                                 /* [0, g-2]^3 or masked out)               */
#define SC3D_DBG_PERM 2          /* int32[n] stable sort permutation         */
#define SC3D_DBG_RHO 3           /* double[G] charge density octant (zyx)    */
#define SC3D_DBG_PHI 4           /* double[G] potential octant, unnormalised */
                                 /* (FP64 form; the precision form forms no  */
                                 /* potential and returns SC3D_ERR_ARG)      */
#define SC3D_DBG_EX 5            /* double[G]                                */
#define SC3D_DBG_EY 6
#define SC3D_DBG_EZ 7
#define SC3D_DBG_GHAT 8          /* double[(gz+1)(gy+1)(gx+1)] real Green's  */
                                 /* spectrum, kz<=gz, ky<=gy, kx<=gx         */
int sc3d_debug_get(sc3d_ctx* ctx, int what, void* host_dst, int64_t bytes);

/* Device-side phase timing of sc3d_apply with HIP events recorded on the
 * apply stream.  Phase names follow the reference's simple_timer labels
 * (space_charge_3d_open_hockney.cc:496,545,625,644,757); each phase but
 * SC3D_PH_SOLVE and SC3D_PH_SORT is exactly one kernel:
 *   SC3D_PH_DOMAIN  sc3d_domain          (extent partial sums + finish)
 *   SC3D_PH_SORT    (no counterpart)     (keys, radix passes, reorder)
 *   SC3D_PH_DEPOSIT sc3d_local_rho       (the deposit kernel only)
 *   SC3D_PH_GREEN   sc3d_green_fn2_*     (spectrum passes, cheap when cached)
 *   SC3D_PH_SOLVE   sc3d_local_f         (x, y, z-with-convolution, y', x')
 *   SC3D_PH_FIELD   get_force            (E field kernel)
 *   SC3D_PH_KICK    sc3d_kick            (the kick kernel only)
 * sc3d_get_timing synchronises and returns milliseconds summed over the
 * applies timed since the last reset, and their count.
 * sc3d_set_timing modes: SC3D_TIMING_OFF; SC3D_TIMING_PHASES (events at
 * every phase boundary, read by sc3d_get_timing); SC3D_TIMING_APPLY (one
 * event pair around each whole apply, read by sc3d_get_apply_timing: the
 * phase events add a few microseconds of device time each, so the whole
 * apply is timed best without them).  No reference counterpart: the
 * reference times with simple_timer on the host. */
#define SC3D_TIMING_OFF 0
#define SC3D_TIMING_PHASES 1
#define SC3D_TIMING_APPLY 2
#define SC3D_PH_DOMAIN 0
#define SC3D_PH_SORT 1
#define SC3D_PH_DEPOSIT 2
#define SC3D_PH_GREEN 3
#define SC3D_PH_SOLVE 4
#define SC3D_PH_FIELD 5
#define SC3D_PH_KICK 6
#define SC3D_NPHASES 7
int sc3d_set_timing(sc3d_ctx* ctx, int enable);
int sc3d_get_timing(sc3d_ctx* ctx, double ms[SC3D_NPHASES], int64_t* napplies, int reset);
int sc3d_get_apply_timing(sc3d_ctx* ctx, double* ms, int64_t* napplies, int reset);

/* Domain record copied by SC3D_DBG_DOMAIN (all FP64). */
typedef struct sc3d_domain_record {
    double sums[6];              /* Sx Sy Scdt Sxx Syy Scdtcdt               */
    double mean[3], stddev[3];
    double size[3], offset[3];
    double left[3], h[3];
    double w0, fn_norm, factor;
    double inv_h[3], half_inv_h[3];
    double ghat_h[3];            /* cell size the cached spectrum is for     */
    int32_t error, ghat_dirty, ghat_valid;
    int32_t rho_peak;            /* precision form: bits of the FP32 density */
                                 /* peak (unit charge weights), 0 otherwise  */
} sc3d_domain_record;

/* Single-particle maps of a FODO channel, applied on the device as one
 * fused program (each particle read and written once for all slices).
 *   SC3D_MAP_DRIFT      <- FF_drift::apply, src/synergia/libFF/ff_drift.h:92-153
 *                          (drift_unit, libFF/ff_algorithm.h:45-79)
 *   SC3D_MAP_QUADRUPOLE <- FF_quadrupole::apply, libFF/ff_quadrupole.h:309-565
 *                          (yoshida6 with thin_quadrupole_unit kicks,
 *                          ff_algorithm.h:928-936,1835-2008); no offsets,
 *                          tilt or multipoles.
 * ref_cdt is the design reference particle's cdt over the slice, from
 * sc3d_reference_cdt (get_reference_cdt, ff_drift.h:68-87,
 * ff_quadrupole.h:262-300, for an on-axis reference). */
#define SC3D_MAP_DRIFT 0
#define SC3D_MAP_QUADRUPOLE 1
#define SC3D_MAX_SLICES 16
typedef struct sc3d_map_slice {
    int32_t kind;                /* SC3D_MAP_*                               */
    int32_t yoshida_steps;       /* quadrupole only; reference default 4     */
    double length;               /* slice length [m]                         */
    double k1, k1s;              /* quadrupole strengths [m^-2]              */
    double ref_cdt;              /* design reference cdt over the slice      */
} sc3d_map_slice;

double sc3d_reference_cdt(const sc3d_map_slice* slice, double ref_p, double mass);

/* Map program that every subsequent sc3d_apply of this context runs on each
 * active particle right after its kick, in the kick kernel itself (the
 * independent operators that follow the collective one in a split-operator
 * step, step.cc:44-78, fused into the same particle pass).  nslices = 0
 * removes it.  No counterpart in the reference operator. */
int sc3d_set_post_kick_maps(sc3d_ctx* ctx, const sc3d_map_slice* slices, int nslices);
int sc3d_apply_maps(const sc3d_bunch* bunch, const sc3d_map_slice* slices, int nslices,
                    void* stream);

/* Moment diagnostics of the bunch on the device: mean[6] and the
 * symmetric 6x6 second moments about the mean, both divided by total_num
 * (Core_diagnostics::calculate_mean / calculate_mom2,
 * src/synergia/bunch/core_diagnostics.cc:202-229,318-362), summed with the
 * same fixed tree as the bunch extent.  Synchronises `stream`.  rms
 * emittances follow as sqrt(det) of the 2x2 blocks (calculate_emittances,
 * diagnostics_full2_host.cc:20-40). */
int sc3d_moments(sc3d_ctx* ctx, const sc3d_bunch* bunch, double mean[6], double mom2[36],
                 void* stream);

/* Tail fractions: per plane j, the active particles with
 * |coordinate_j - center[j]| > nsigma * width[j], divided by total_num
 * (halo diagnostic of long tracking, e.g. beyond 4 sigma; no counterpart in
 * the reference's diagnostics).  Integer counts, so independent of the
 * reduction order.  Synchronises `stream`. */
int sc3d_tail_fraction(sc3d_ctx* ctx, const sc3d_bunch* bunch, const double center[6],
                       const double width[6], double nsigma, double frac[6], void* stream);

/* Fills slots [0, n) of a bunch: coordinates = mean + L * u with u a
 * standard normal 6-vector from Philox-4x32-10 keyed (seed, id) and
 * Box-Muller, L lower-triangular packed row-major (21 entries); column 6 =
 * id0 + slot; masks = 1.  Deterministic for a given (seed, id). */
int sc3d_populate_gaussian(double* parts, int64_t stride, int64_t n,
                           uint8_t* masks, const double mean[6],
                           const double chol[21], uint64_t seed, int64_t id0,
                           void* stream);

/* adjust_moments (src/synergia/bunch/populate.cc:30-61, the last step of
 * populate_6d :128-224): maps the first n_active rows so that the bunch's
 * mean and second moments (over its active particles, divided by total_num)
 * become exactly `means` and `cov` (6x6 row-major):  (p - mean) A^T + means
 * with A = G H^-1, G and H the lower Cholesky factors of cov and of the
 * bunch's second moments (adjust_moments_host, populate_host.cc:8-44).
 * Errors: the reference's "adjust_moments: covariance matrix must be
 * symmetric" (SC3D_ERR_ARG), a non-positive-definite matrix (SC3D_ERR_ARG).
 * Synchronises `stream` once (the moments go to the host for A). */
int sc3d_adjust_moments(const sc3d_bunch* bunch, const double means[6], const double cov[36],
                        void* stream);

/* sc3d_moments without a context and without allocating: no reference
 * counterpart; for callers that cannot call hipMalloc (code under stream
 * capture, tight serving loops, harnesses that forbid allocation).  The
 * chunk partials live in a fixed device array of the library (2.6 MiB per
 * GPU, held until the process ends), so at most 2^26 active particles
 * (larger: SC3D_ERR_SIZE, use sc3d_moments).  Results equal sc3d_moments'
 * bit for bit.  Calls that overlap on one GPU (other streams or threads)
 * race on that array: serialise them. */
int sc3d_moments_noalloc(const sc3d_bunch* bunch, double mean[6], double mom2[36],
                         void* stream);

/* Host only (no device call).  The points of the doubled-grid Green's
 * function where the reference's flat-index decomposition in alg_g2_linear /
 * alg_g2_pointlike (space_charge_3d_open_hockney.cc:170-176,272-278) leaves
 * another value than the symmetric one at this shape, replayed in its serial
 * write order; sc3d_apply adds their effect, so the field is the reference's
 * at every shape.  Writes min(count, cap) records of 6 int32 -- target
 * (z, y, x), then the (iz, iy, ix) whose value stands there, or -1 -1 -1 for
 * a point left at zero -- and returns the count (<0: error).  No reference
 * counterpart: it exposes that behaviour for tests and callers. */
int sc3d_green_patches(int32_t green_fn, const int32_t shape[3], int32_t* out,
                       int32_t cap);

#ifdef __cplusplus
}
#endif
#endif
