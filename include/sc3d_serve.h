/* sc3d_serve.h -- served form of the 3D open-Hockney space-charge step on
 * one MI355X: a request loop that keeps bunch-tracking jobs resident on the
 * device and advances all of them one split-operator step per iteration
 * (continuous batching).
 *
 * There is no counterpart in the reference: Synergia runs one
 * Bunch_simulator per process through Propagator::propagate
 * (simulation/propagator.cc:36-118) with the collective operator applied
 * once per step (Step::apply, simulation/step.cc:44-78).  Each served job is
 * that loop for one bunch -- maps of the first half-step, the space-charge
 * kick (Space_charge_3d_open_hockney::apply_bunch,
 * space_charge_3d_open_hockney.cc:422-453), maps of the second half-step --
 * with the schedule of a lattice given as map programs (see sc3d_map_slice
 * in sc3d_capi.h).
 *
 * Per iteration the server:
 *   - admits waiting jobs while their device footprint fits the memory
 *     budget (populating their bunches on the device);
 *   - runs one step of every resident job: the particle passes per job, the
 *     grid passes (Green's spectrum, convolution, field) as one launch per
 *     pass for all resident jobs of the same grid shape;
 *   - with graphs on, replays one captured hipGraph of the whole iteration
 *     while the resident set is unchanged (the step-dependent decisions --
 *     reorder, schedule entry, kick time step, finished -- are taken on the
 *     device);
 *   - retires finished jobs with their final moments.
 * One server per host thread; the calls are not thread-safe. */
#ifndef SC3D_SERVE_H
#define SC3D_SERVE_H

#include "sc3d_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One map program of a schedule entry. */
typedef struct sc3d_step_maps {
    int32_t n;                          /* slices, <= SC3D_MAX_SLICES */
    sc3d_map_slice s[SC3D_MAX_SLICES];
} sc3d_step_maps;

/* A bunch-tracking job.  The bunch is a 6D Gaussian (sc3d_populate_gaussian
 * with mean, chol, seed) of n_particles macroparticles; it takes `steps`
 * steps of the schedule, starting at entry 0 and wrapping every turn_steps
 * entries.  Schedule arrays are copied by sc3d_server_submit. */
typedef struct sc3d_job_spec {
    int64_t n_particles;
    int32_t shape[3];                   /* grid, as sc3d_options.shape        */
    int32_t steps;
    uint64_t seed;
    double mean[6];
    double chol[21];
    double real_num;
    int32_t charge;
    double mass, ref_gamma, ref_beta, ref_p;
    int32_t turn_steps;                 /* schedule entries                   */
    const double* kick_dt;              /* [turn_steps] kick time step [s]    */
    const sc3d_step_maps* first;        /* [turn_steps] before the kick of a  */
                                        /*   job's first step at entry k      */
    const sc3d_step_maps* between;      /* [turn_steps] after kick k and      */
                                        /*   before kick k+1                  */
    const sc3d_step_maps* last;         /* [turn_steps] after the kick of a   */
                                        /*   job's last step at entry k       */
} sc3d_job_spec;

typedef struct sc3d_server_options {
    int64_t hbm_budget;       /* bytes the resident jobs may hold (0: half of */
                              /* the device memory free at creation)          */
    int32_t max_resident;     /* 0: no limit                                  */
    int32_t precision;        /* SC3D_PRECISION_* of every job                */
    int32_t sort_period;      /* as sc3d_options.sort_period                  */
    int32_t use_graphs;       /* capture and replay each iteration            */
    int32_t batch_grids;      /* same-shape grid passes in one launch         */
    int32_t admit_interval;   /* with graphs: admit at most every this many   */
                              /* iterations while jobs are resident (0: 8)    */
    int32_t branches;         /* streams the per-job chains are spread over   */
                              /* (independent graph branches; 0: 8)           */
    int32_t keep_particles;   /* copy each finished job's particles to host   */
                              /* memory (sc3d_server_job_particles)           */
} sc3d_server_options;

#define SC3D_JOB_WAITING 0
#define SC3D_JOB_RESIDENT 1
#define SC3D_JOB_DONE 2
#define SC3D_JOB_FAILED 3             /* admission failed (allocation); not retried */

typedef struct sc3d_job_result {
    int64_t id;
    int32_t state;                      /* SC3D_JOB_*                          */
    int32_t steps_done;
    double t_submit, t_admit, t_done;   /* server clock [s], -1 if not yet     */
    double emit0[3], emit[3];           /* rms emittances x, y, z: admission,  */
                                        /* end (sqrt det of the 2x2 blocks)    */
    double mean[6], mom2[36];           /* moments at the end                  */
} sc3d_job_result;

typedef struct sc3d_server_stats {
    int64_t iterations, graph_builds;
    int64_t jobs_submitted, jobs_done, jobs_resident, jobs_waiting;
    int64_t bytes_resident, hbm_budget;
    double particle_steps;              /* sum of n_particles over steps taken */
    double busy_s;                      /* sum of iteration wall times         */
    double iter_ms_p50, iter_ms_p99, iter_ms_max;
} sc3d_server_stats;

typedef struct sc3d_server sc3d_server;

void sc3d_server_options_default(sc3d_server_options* o);
int sc3d_server_create(const sc3d_server_options* opt, sc3d_server** out);
/* Queues a job (server clock starts its latency now); SC3D_ERR_SIZE if its
 * footprint alone exceeds the budget. */
int sc3d_server_submit(sc3d_server* s, const sc3d_job_spec* spec, int64_t* id);
/* One iteration: admit, one step of every resident job (synchronous),
 * retire.  Returns the counts after it; does nothing when idle. */
int sc3d_server_iterate(sc3d_server* s, int32_t* n_resident, int32_t* n_waiting);
int sc3d_server_job(sc3d_server* s, int64_t id, sc3d_job_result* out);
/* Final particles of a finished job (7 columns of n_particles, column j at
 * parts + j*n_particles, and the masks), kept when keep_particles is set;
 * SC3D_ERR_ARG otherwise or before the job is done. */
int sc3d_server_job_particles(sc3d_server* s, int64_t id, double* parts, uint8_t* masks);
int sc3d_server_get_stats(sc3d_server* s, sc3d_server_stats* out);
int sc3d_server_destroy(sc3d_server* s);
/* Device bytes a job holds while resident (host only, no device needed). */
int64_t sc3d_job_footprint(const sc3d_job_spec* spec, int32_t precision);

#ifdef __cplusplus
}
#endif
#endif /* SC3D_SERVE_H */
