/*
 * mbir_tem.h -- C-ABI of the MI355X-native MBIR reconstruction engine.
 *
 * The boundary replaces the TBB CPU path behind OpenMBIR-TEM's reconstruction
 * engine interface.  Reference citations are relative to the reference's
 * Code/MBIRLib/ directory:
 *
 *   class HAADF_ReconstructionEngine : public AbstractFilter
 *     HAADF/HAADF_ReconstructionEngine.h:67-99   (properties TomoInputs, Sinogram,
 *     Geometry, AdvParams, ForwardModel; void execute(); getErrorCondition())
 *   inner seam updateVoxels(...) -> uint8_t (1 = continue, 0 = converged)
 *     HAADF/HAADF_ReconstructionEngine.h:120-129,
 *     HAADF/HAADF_ReconstructionEngine_UpdateVoxels.cpp:594-817
 *
 * Conventions
 *   - Every entry point returns int: 0 (or a documented non-negative value) on
 *     success, negative on error.  Codes mirror the reference: -1 bad input,
 *     -2 detector response failure, -999 cancelled
 *     (HAADF/HAADF_ReconstructionEngine.cpp:369-393,401,522-528).  Codes -100..-199
 *     are HIP runtime failures of this library.
 *   - The caller owns every host buffer.  The library owns device state from
 *     mbir_job_create until mbir_job_destroy.
 *   - A context may be driven from one host thread; mbir_job_cancel may be
 *     called from any thread.
 *   - Layouts are the reference's: volume [N_z][N_x][N_y] (y fastest),
 *     sinograms [N_theta][N_r][N_t] (t fastest)
 *     (GenericFilters/InitialReconstructionInitializer.cpp:158-160,
 *     Common/TomoArray.hpp:176-198).
 */
#ifndef MBIR_TEM_H_
#define MBIR_TEM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBIR_OK 0
#define MBIR_ERR_INPUT (-1)
#define MBIR_ERR_RESPONSE (-2)
#define MBIR_ERR_CANCELLED (-999)
#define MBIR_ERR_HIP (-100)
#define MBIR_ERR_NOMEM (-101)
#define MBIR_ERR_UNSUPPORTED (-3)
#define MBIR_ERR_BUDGET (-4)

typedef struct mbir_ctx mbir_ctx;
typedef struct mbir_job mbir_job;

/* Context options (no reference counterpart: device selection and the HBM
 * budget used for admission by the batch/serve paths). */
typedef struct mbir_opts {
  int32_t device;            /* HIP device ordinal */
  int32_t verbose;           /* 0 quiet, 1 progress lines on stderr */
  uint64_t hbm_budget_bytes; /* 0 = 90% of free device memory */
} mbir_opts;

/* Geometry produced by mbir_geometry_init; the fields of Sinogram and Geometry
 * (Reconstruction/ReconstructionStructures.h:68-97) the engine needs. */
typedef struct mbir_geometry {
  int32_t N_theta, N_r, N_t;
  int32_t N_x, N_y, N_z;
  double delta_r, delta_t, delta_xz, delta_xy;
  double R0, RMax, T0, TMax;
  double x0, y0, z0;
  double LengthX, LengthY, LengthZ;
} mbir_geometry;

/* Everything one reconstruction job needs: Sinogram + Geometry + TomoInputs +
 * AdvancedParameters + the HAADF_ForwardModel's initial nuisance arrays
 * (HAADF/HAADF_ReconstructionEngine.h:75-84, Reconstruction/ReconstructionStructures.h). */
typedef struct mbir_problem {
  mbir_geometry geom;
  const double* angles_deg; /* [N_theta] tilts in degrees (Sinogram::angles before execute) */
  const float* counts;      /* [N_theta][N_r][N_t] measured sinogram (Sinogram::counts) */
  const float* init_volume; /* [N_z][N_x][N_y] or NULL -> default_recon */
  double default_recon;     /* TomoInputs::defaultInitialRecon */
  const double* init_gain;     /* [N_theta] or NULL -> target_gain */
  const double* init_offset;   /* [N_theta] or NULL -> default_offset */
  const double* init_variance; /* [N_theta] or NULL -> default_variance */
  double target_gain;          /* TomoInputs::targetGain */
  double default_offset;       /* TomoInputs::defaultOffset */
  double default_variance;     /* TomoInputs::defaultVariance */
  double sigma_x;              /* TomoInputs::SigmaX */
  double mrf_q;                /* TomoInputs::p = diffuseness + 1 (MRF_Q); MRF_P = 2 */
  double mrf_c;                /* 1e-3 HAADF (HAADF/HAADF_QGGMRFPriorModel.cpp:15) */
  double stop_threshold;       /* TomoInputs::StopThreshold */
  int32_t num_outer;           /* TomoInputs::NumOuterIter */
  int32_t num_inner;           /* TomoInputs::NumIter */
  int32_t joint_estimation;    /* AdvancedParameters::JOINT_ESTIMATION */
  int32_t noise_model;         /* AdvancedParameters::NOISE_MODEL */
  int32_t zero_skipping;       /* AdvancedParameters::ZERO_SKIPPING */
  uint32_t seed;               /* update-order seed (the reference seeds from the clock) */
  /* ICD schedule: 0 = conflict-free groups (every step is equivalent to a
   * sequential ICD order); 1 = relaxed groups of rows_per_step spaced rows. */
  int32_t schedule;
  int32_t rows_per_step;
  /* sweep kernels: 0 = one launch per conflict-free step with the job's own kernel
   * (plain per-job GPU ICD); 1 = the batched step kernel (a batch of one job);
   * 2 = reserved for a whole sweep in one persistent launch (refused with -3: not
   * enabled in this build); 3 = the launches of mode 0 captured once in a HIP graph
   * and replayed every sweep (each step reads its lines from a fixed slot of a padded
   * schedule uploaded before the replay; same results as mode 0) */
  int32_t kernel_mode;
  /* Bright-field mode (BrightField/BFForwardModel.cpp, BFReconstructionEngine.cpp:575-872):
   * counts are transmitted counts; the measurement is -log(counts + bf_offset)
   * (processRawCounts :939-972), the weights (counts + bf_offset) / alpha (:328-385),
   * entries whose |e sqrt(w)| reaches the Bragg threshold T are rejected by the
   * selector; joint_estimation estimates the per-view offset (log dose) only and
   * noise_model is NOISE_ESTIMATION (one variance for all views); the first outer
   * iteration's inner sweeps run with the threshold 5e10 when num_inner > 1.  The
   * reference's BF prior uses mrf_c = 0.01 (BrightField/BF_QGGMRFPriorModel.cpp:54).
   * Needs schedule 0 and N_y <= 256; up to four detector rows per slice (the coarse levels
   * of the multi-resolution driver) in single-job sweeps; the batched sweep and the server
   * take bright-field jobs of one row per slice. */
  int32_t bf_mode;
  double bragg_T;     /* Bragg threshold T (BFReconstructionArgsParser.cpp:176, default 3) */
  double bragg_delta; /* delta of the Bragg rejector (:179, default 0.5) */
  double bf_offset;   /* added to the counts before the log (:182, default 0) */
  /* bright field: 1 = the reference's NHICD schedule (BrightField/BFConstants.h:51,
   * BFReconstructionEngine.cpp:641-884): each call updates either the next twentieth
   * of a seeded random list of all lines (homogeneous) or the lines whose 5 x 5
   * Hamming-filtered update magnitude exceeds its 95th percentile (non-homogeneous),
   * alternately, and the stop test compares the ROI update magnitude with the ROI
   * volume sum every 20 calls; 0 = full random-order sweeps */
  int32_t bf_nhicd;
  int32_t pad_;
  /* HAADF with a bright-field sinogram (the reference's BFTomoInputs / BFSinogram and
   * BF_Flag, HAADF/HAADF_ReconstructionEngine_Extra.cpp:89-175): optional raw
   * bright-field counts [N_theta][N_r][N_t]; every projection entry is scaled by
   * counts + BF_OFFSET (32768, HAADF/HAADFConstants.h:49) of its detector element, the
   * initial error is y - s Yest - mu (:718-727), and joint_estimation estimates the
   * offsets only: per view the weighted mean of e moves from e into mu, the gains stay
   * (:636-669).  Single-job paths only (mbir_job_run, mbir_job_sweep); not with
   * bf_mode.  NULL = none. */
  const float* bf_sinogram;
  /* AdvancedParameters::ESTIMATE_PRIOR (HAADF/HAADF_ReconstructionEngine.cpp:257, 807-813):
   * 1 = after every outer iteration that does not stop the run, re-estimate sigma_x from
   * the volume (mbir_job_estimate_prior).  HAADF single-job runs only (the server refuses
   * such jobs with -3). */
  int32_t estimate_prior;
  /* y-slabs of the conflict-free schedule (no reference counterpart; the reference splits
   * y into one slab per thread, HAADF/HAADF_ReconstructionEngine_UpdateVoxels.cpp:671-715):
   * 1 = each step updates whole lines; n >= 2 = the slices are cut into n slabs and a step
   * updates one conflict-free row class per slab, the classes of neighbouring slabs at
   * least two planes apart in z, so the units of a step still touch disjoint detector
   * rows and are never prior neighbours (every step equals a sequential ICD order);
   * 0 = chosen from the sizes.  Schedule 0 only; bright-field NHICD lists run whole lines. */
  int32_t y_slabs;
} mbir_problem;

/* Per-sweep statistics (updateVoxels, HAADF_ReconstructionEngine_UpdateVoxels.cpp:717-786) */
typedef struct mbir_sweep_stats {
  double sum_abs_update;   /* ROI sum |x_new - x_old| (AverageUpdate) */
  double sum_abs_value;    /* ROI sum |x_old| (AverageMagnitudeOfRecon) */
  int64_t zero_skipped;    /* voxels skipped by zero-skipping */
  int64_t updated;         /* voxels updated */
  int64_t nnz_updated;     /* sum over updated voxels of A-column entries x |VLR| */
  double kernel_ms;        /* device time from the sweep's first ICD launch to its last (timing on) */
  int32_t steps;           /* kernel steps in the sweep */
  int32_t status;          /* 1 = continue, 0 = converged (stop threshold met) */
} mbir_sweep_stats;

typedef struct mbir_run_stats {
  int32_t sweeps;
  int32_t outer_iterations;
  int32_t status; /* last sweep status */
  double seconds;
} mbir_run_stats;

/* ---- library / context ---- */
int32_t mbir_version(void);
int32_t mbir_create(const mbir_opts* opts, mbir_ctx** out);
int32_t mbir_destroy(mbir_ctx* ctx);

/* Geometry sizing of InitialReconstructionInitializer::execute
 * (GenericFilters/InitialReconstructionInitializer.cpp:98-167) and
 * MRCSinogramInitializer R0/T0 (GenericFilters/MRCSinogramInitializer.cpp:302-305).
 * delta_xz_in is TomoInputs::delta_xz as set by the multi-resolution driver
 * (HAADF/HAADF_MultiResolutionReconstruction.cpp:349-350), interpolate_factor is
 * TomoInputs::interpolateFactor (:262).  Host only. */
int32_t mbir_geometry_init(int32_t N_r, int32_t N_t, int32_t N_theta, double delta_r,
                           double delta_t, double thickness, double delta_xz_in,
                           double interpolate_factor, int32_t extend_object,
                           const double* angles_deg, mbir_geometry* out);

/* ---- one reconstruction job ----
 * mbir_job_create performs execute()'s setup (HAADF/HAADF_ReconstructionEngine.cpp:441-691):
 * voxel profiles, detector response, H_t and voxel-line responses, ROI mask, gain and
 * offset initialisation, the initial forward projection, error sinogram and weights. */
int32_t mbir_job_create(mbir_ctx* ctx, const mbir_problem* prob, mbir_job** out);
int32_t mbir_job_destroy(mbir_job* job);

/* One ICD sweep: updateVoxels(outer, inner, ...) (…_UpdateVoxels.cpp:594-817).
 * Returns 1 = continue, 0 = converged (stop test met), negative on error. */
int32_t mbir_job_sweep(mbir_job* job, int32_t outer, int32_t inner, mbir_sweep_stats* stats);

/* NHICD list helpers of the bright-field schedule (host only): the update-magnitude
 * map [N_z][N_x] filtered with the 5 x 5 Hamming window, zero outside
 * (BFReconstructionEngine::ComputeVSC, BrightField/BFReconstructionEngine_Extra.cpp:888-925),
 * and the threshold of the non-homogeneous list, the (n - n/20)-th smallest value
 * (SetNonHomThreshold, :934-987). */
int32_t mbir_nhicd_filter(const float* mag, int32_t N_z, int32_t N_x, double* filtered);
int32_t mbir_nhicd_threshold(const double* values, int32_t n, double* threshold);

/* Form of the batched line kernels (both forms compute identical results): per kernel
 * shape [N_y > 64][rows per slice > 1] and launch-size bucket (workgroups <= 256,
 * <= 512, <= 1024, <= 2048, more) 1 = the prefetching form (shorter line latency),
 * 0 = the form without prefetch (fewer registers).  No reference counterpart: the
 * reference has no batched kernels. */
typedef struct mbir_tuning {
  int32_t prefetch[2][2][5];
  double us_per_wg[2][2][5][2];  /* autotuning: mean launch time per workgroup [form], 0 = none */
  int32_t launches[2][2][5][2];  /* autotuning: launches timed [form] */
} mbir_tuning;

/* Autotuner (no reference counterpart): times `sweeps` batched sweeps of the jobs
 * with each form, launch by launch, and keeps for every shape and size bucket the
 * faster form in ctx (used by mbir_batch_sweep and servers created afterwards).  The
 * jobs advance by 2 x sweeps sweeps (results as with any form).  out may be NULL. */
int32_t mbir_autotune(mbir_ctx* ctx, mbir_job* const* jobs, int32_t n, const int32_t* outer,
                      const int32_t* inner, int32_t sweeps, mbir_tuning* out);
/* The context's current choice (default: prefetching up to 512 workgroups), and
 * setting a stored one. */
int32_t mbir_get_tuning(mbir_ctx* ctx, mbir_tuning* out);
int32_t mbir_set_tuning(mbir_ctx* ctx, const mbir_tuning* t);

/* Batched sweep: one ICD sweep of each of n jobs of ctx (no reference counterpart:
 * the reference runs one job per process).  Launch k runs conflict-free step k of
 * every job that has one, so the launches carry many jobs' lines at once; jobs stay
 * independent (a job's result is the same alone and in a batch).  Jobs need
 * units (y-slabs of a line) of at most 128 slices and at most 2 detector rows per slice.
 * outer/inner (may be NULL = 0)
 * are per-job iteration indices, stats (may be NULL) per-job results. */
int32_t mbir_batch_sweep(mbir_ctx* ctx, mbir_job* const* jobs, int32_t n, const int32_t* outer,
                         const int32_t* inner, mbir_sweep_stats* stats);

/* jointEstimation (HAADF/HAADF_ReconstructionEngine_Extra.cpp:419-687) when enabled,
 * then updateWeights (:795-881) when the noise model is on.  FP64 on the device. */
int32_t mbir_job_nuisance_update(mbir_job* job);

/* computeCost (HAADF/HAADF_ReconstructionEngine.cpp:1147-1364) in FP64 on the device.
 * terms3 (may be NULL) receives {data, prior, noise}. */
int32_t mbir_job_cost(mbir_job* job, double* total, double* terms3);

/* estimateSigmaX (HAADF/HAADF_ReconstructionEngine.cpp:1827-1945) then updatePriorModel
 * (HAADF/HAADF_QGGMRFPriorModel.cpp:24-31), FP64 on the device: estimate = the QGGMRF
 * value of the 13 forward neighbour pairs of every voxel in the planes
 * [(int) 0.25 N_z, (int) 0.75 N_z), differences scaled by the current sigma_x, divided by
 * N_x N_y (zEnd - zStart + 1); the new sigma_x = 5 (estimate q)^(1/q) replaces the prior's
 * constants for the following sweeps.  estimate and sigma_x (may be NULL) receive both.
 * -3 for bright-field jobs, -1 when the volume gives no scale (all zero). */
int32_t mbir_job_estimate_prior(mbir_job* job, double* estimate, double* sigma_x);
/* The prior scale the next sweep uses (TomoInputs::SigmaX). */
int32_t mbir_job_get_sigma_x(mbir_job* job, double* sigma_x);

/* The update order of the job's last sweep or NHICD call (no reference counterpart: the
 * reference draws its order inside UpdateYSlice, …_UpdateVoxels.cpp:213-231).  Returns the
 * number of conflict-free steps; step k holds units[step_ptr[k] .. step_ptr[k + 1]), each
 * unit three int32 (line = z N_x + x, first slice y0, slices ny), all of one step updated
 * at once; a step equals any sequential order of its units, and a unit is its even
 * slices, then its odd slices.  step_ptr needs cap_steps + 1 entries, units 3 cap_units;
 * either may be NULL. */
int32_t mbir_job_get_schedule(mbir_job* job, int32_t* step_ptr, int32_t cap_steps,
                              int32_t* units, int64_t cap_units);

/* Inspection of the bright-field NHICD schedule (no reference counterpart; the
 * reference prints its lists only in debug builds, BFReconstructionEngine.cpp:700-706):
 * with the log on, mbir_job_run records the lines of every NHICD call.
 * mbir_job_get_nhicd_log returns the number of calls of the last run and fills
 * call_ptr[0 .. calls] (call c's lines are lines[call_ptr[c] .. call_ptr[c + 1])) and
 * lines, as far as cap_calls and cap_lines allow; either may be NULL. */
int32_t mbir_job_set_nhicd_log(mbir_job* job, int32_t on);
int32_t mbir_job_get_nhicd_log(mbir_job* job, int64_t* call_ptr, int32_t cap_calls,
                               int32_t* lines, int64_t cap_lines);

/* execute()'s outer/inner loop (HAADF/HAADF_ReconstructionEngine.cpp:700-815).
 * costs (may be NULL) receives up to cap FP64 cost values, one after the setup,
 * one after every sweep and one after every nuisance step (COST_CALCULATE order). */
int32_t mbir_job_run(mbir_job* job, mbir_run_stats* stats, double* costs, int32_t cap,
                     int32_t* n_costs);

/* Ŷ = I0 * A x (HAADF/HAADF_ForwardProject.cpp:80-120) for volume (host,
 * [N_z][N_x][N_y]; NULL = the job's current volume), written as FP64 to out. */
int32_t mbir_job_forward_project(mbir_job* job, const float* volume, double* out);

/* mbir_job_forward_project without allocation.  No counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  The same kernels, with the FP64 projection accumulator and the uploaded
 * volume in a library-wide device array of 64 MiB per GPU (held until the process
 * ends): at most 4 Mi sinogram entries (N_theta N_r N_t) and, when volume is not NULL,
 * 8 Mi voxels; larger jobs get -1.  Calls no hipMalloc or hipFree.  Every call on one GPU
 * uses the same array, so calls that overlap on one GPU (from several host threads)
 * race; the accumulator is cleared by a kernel at each call. */
int32_t mbir_job_forward_project_noalloc(mbir_job* job, const float* volume, double* out);

/* A-matrix column (z, x) as computed on the device
 * (HAADF/HAADF_ReconstructionEngine.cpp:1369-1524).  Returns the entry count
 * (entries beyond cap are not written) or a negative error. */
int32_t mbir_job_get_column(mbir_job* job, int32_t z, int32_t x, uint32_t* index, float* values,
                            int32_t cap);

int32_t mbir_job_get_volume(mbir_job* job, float* out);
int32_t mbir_job_set_volume(mbir_job* job, const float* volume); /* also re-derives e */
/* The device-resident multi-resolution hand-over: fine's volume becomes coarse's
 * replicated 2 x 2 x 2 (coarse sizes = fine's N_z/2, N_x/2, N_y/2; the last plane of an
 * odd size and lines whose column is empty stay 0, as after a setup from the same
 * volume), then e is re-derived.  Both jobs on one device.  No counterpart in
 * the reference, which hands levels over through files: it replaces the level file
 * write (HAADF/HAADF_MultiResolutionReconstruction.cpp:211-229) and its interpolating
 * read (GenericFilters/InitialReconstructionBinReader.cpp:103-130). */
int32_t mbir_job_upsample_volume(mbir_job* fine, const mbir_job* coarse);
int32_t mbir_job_get_error_sinogram(mbir_job* job, float* out);
int32_t mbir_job_get_nuisance(mbir_job* job, double* I0, double* mu, double* alpha);
int32_t mbir_job_get_geometry(mbir_job* job, mbir_geometry* out);
/* detector response H[N_theta][64] (GenericFilters/DetectorResponse.cpp:66-127) */
int32_t mbir_job_get_detector_response(mbir_job* job, double* out);
/* Thread-safe cooperative cancel (AbstractFilter::setCancel); the next sweep
 * or run step returns MBIR_ERR_CANCELLED. */
int32_t mbir_job_cancel(mbir_job* job);
/* Bright field: the Bragg selector (1 kept, 0 rejected) and the weights w, both
 * [N_theta][N_r][N_t] (BFForwardModel m_Selector, m_Weight); -3 for other jobs. */
int32_t mbir_job_get_selector(mbir_job* job, uint8_t* out);
int32_t mbir_job_get_weights(mbir_job* job, float* out);
/* Bright field: the Bragg threshold the next sweep uses (the run loop sets it itself) */
int32_t mbir_job_set_bragg_threshold(mbir_job* job, double T);
/* device bytes held by the job (volume, sinograms, tables) */
int64_t mbir_job_device_bytes(const mbir_job* job);
/* 1 = time every ICD kernel step with HIP events on the job's stream */
int32_t mbir_job_set_timing(mbir_job* job, int32_t on);

/* ---- request server ----
 * No reference counterpart: the reference runs one reconstruction per process
 * (HAADF/HAADF_MultiResolutionReconstruction.cpp:410-436 drives one engine).  The
 * server takes jobs as they arrive, admits them in arrival order when their
 * device footprint fits the HBM budget and a slot is free, and runs every
 * admitted job's execute() loop (the one mbir_job_run runs) concurrently:
 * each launch carries the next conflict-free ICD step of every resident job, so
 * a job joins at the next launch after its setup and leaves when it converges.
 * A job's result is the same as mbir_job_run's for the same problem.  Jobs need
 * units of at most 128 slices and at most 2 detector rows per slice (else they fail
 * with -3). */
typedef struct mbir_server mbir_server;

typedef struct mbir_server_opts {
  int32_t max_jobs;          /* jobs resident on the device at once (0 = 64, at most 256) */
  int32_t setup_threads;     /* host threads that set admitted jobs up (0 = 2, at most 8) */
  uint64_t hbm_budget_bytes; /* device bytes for resident jobs (0 = the context's budget,
                                or 80% of the free device memory) */
  int32_t long_steps;        /* jobs with at least this many conflict-free steps per sweep
                                (N_z x spacing) run on a second launch stream, beside the
                                others (0 = 2048, < 0 = one stream for all) */
  int32_t long_jobs;         /* jobs on the second stream at once (0 = 4) */
} mbir_server_opts;

#define MBIR_JOB_QUEUED 0
#define MBIR_JOB_SETUP 1
#define MBIR_JOB_RUNNING 2
#define MBIR_JOB_DONE 3
#define MBIR_JOB_FAILED 4

typedef struct mbir_job_report {
  int32_t state;            /* MBIR_JOB_* */
  int32_t error;            /* 0, or the failure code (-999 cancelled, -4 over budget, ...) */
  int32_t sweeps;
  int32_t outer_iterations;
  int32_t status;           /* last sweep status (0 = stop threshold met) */
  int32_t pad;
  int64_t voxel_visits;     /* updated + zero-skipped voxels over all sweeps */
  int64_t updated;
  int64_t device_bytes;
  double t_submit, t_admit, t_start, t_done; /* seconds since the server was created */
} mbir_job_report;

typedef struct mbir_server_stats {
  int64_t ticks, launches, entries, workgroups;
  int32_t jobs_submitted, jobs_done, jobs_failed, peak_resident;
  uint64_t budget_bytes, reserved_bytes, peak_reserved_bytes;
  int32_t fatal_error;
  int32_t long_lane_jobs;   /* jobs that ran on the second launch stream */
} mbir_server_stats;

/* Upper bound of a job's device footprint, used for admission. */
int64_t mbir_problem_device_bytes(const mbir_problem* prob);
int32_t mbir_server_create(mbir_ctx* ctx, const mbir_server_opts* opts, mbir_server** out);
/* Copies the problem's host arrays; returns at once with a ticket. */
int32_t mbir_server_submit(mbir_server* srv, const mbir_problem* prob, uint64_t* ticket);
/* Waits for one job (timeout_s < 0: no limit).  0 = finished (rep->state DONE or
 * FAILED), 1 = timed out (rep holds the current state). */
int32_t mbir_server_wait(mbir_server* srv, uint64_t ticket, double timeout_s,
                         mbir_job_report* rep);
/* Waits for the next finished job not yet returned here: 0 and *ticket, or 1 on timeout. */
int32_t mbir_server_wait_any(mbir_server* srv, double timeout_s, uint64_t* ticket);
/* The finished job, for mbir_job_get_volume / _get_nuisance / ... (valid until release). */
int32_t mbir_server_job(mbir_server* srv, uint64_t ticket, mbir_job** job);
/* Frees a finished job's device state and returns its memory to the budget. */
int32_t mbir_server_release(mbir_server* srv, uint64_t ticket);
/* Thread-safe: a queued job fails at once, a running one at its next sweep end (-999). */
int32_t mbir_server_cancel(mbir_server* srv, uint64_t ticket);
int32_t mbir_server_get_stats(mbir_server* srv, mbir_server_stats* out);
/* Cancels what is still queued or running, then frees everything. */
int32_t mbir_server_destroy(mbir_server* srv);

#ifdef __cplusplus
}
#endif
#endif
