/*
 * mbir_oracle.h -- FP64 CPU restatement of OpenMBIR-TEM's HAADF reconstruction
 * hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path and the timed CPU baseline; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in openmbir-tem_amd/ links
 * or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's Code/MBIRLib/).  The reference computes in FP64 throughout
 * (Reconstruction/ReconstructionConstants.h:48, typedef double Real_t); so
 * does this oracle.  Pinning of each function (closed forms, invariants, brute
 * force) lives in tests/test_oracle_*.py.
 */
#ifndef MBIR_ORACLE_H_
#define MBIR_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference constants: HAADF_ReconstructionEngine.cpp:244-259 */
#define ORC_RESPONSE_BINS 64
#define ORC_PROFILE_RESOLUTION 1536
#define ORC_BEAM_RESOLUTION 512

/* Sinogram + Geometry + the TomoInputs fields the hot path reads
 * (Reconstruction/ReconstructionStructures.h:68-163). */
typedef struct orc_geom {
  int32_t N_theta, N_r, N_t;
  int32_t N_x, N_y, N_z;
  double delta_r, delta_t;
  double delta_xz, delta_xy;
  double R0, RMax, T0, TMax;
  double x0, y0, z0;
  double LengthX, LengthY, LengthZ;
} orc_geom;

/* QGGMRF::QGGMRF_Values (Reconstruction/QGGMRF_Functions.h:46-55) */
typedef struct orc_qggmrf {
  double MRF_P, SIGMA_X_P, MRF_Q, MRF_C, MRF_ALPHA, SIGMA_X_P_Q, SIGMA_X_Q;
} orc_qggmrf;

typedef struct orc_params {
  double sigma_x;        /* TomoInputs::SigmaX */
  double mrf_q;          /* TomoInputs::p = diffuseness + 1 */
  double mrf_c;          /* 0.001 HAADF (HAADF_QGGMRFPriorModel.cpp:15) */
  double target_gain;    /* TomoInputs::targetGain */
  double stop_threshold; /* TomoInputs::StopThreshold */
  int32_t num_outer;     /* TomoInputs::NumOuterIter */
  int32_t num_inner;     /* TomoInputs::NumIter */
  int32_t joint_estimation; /* AdvancedParameters::JOINT_ESTIMATION */
  int32_t noise_model;      /* AdvancedParameters::NOISE_MODEL */
  int32_t zero_skipping;    /* AdvancedParameters::ZERO_SKIPPING */
  int32_t compute_cost;     /* COST_CALCULATE */
  uint32_t seed;            /* order RNG seed (the reference seeds from the clock) */
  /* bright-field mode (BrightField/BFForwardModel.cpp, BFReconstructionEngine.cpp):
   * counts are raw transmitted counts; noise_model is NOISE_ESTIMATION there */
  int32_t bf_mode;
  double bragg_T;      /* Bragg threshold T (BFReconstructionArgsParser.cpp:176, default 3) */
  double bragg_delta;  /* delta of the Bragg rejector (:179, default 0.5) */
  double bf_offset;    /* added to the counts before the log (:182, default 0) */
  int32_t estimate_prior; /* AdvancedParameters::ESTIMATE_PRIOR (HAADF_ReconstructionEngine.cpp:807-813) */
  int32_t bf_nhicd;       /* bright field: the NHICD schedule (BrightField/BFConstants.h:51) */
} orc_params;

#define ORC_BF_MAX 1.0            /* BrightField/BFConstants.h */
#define ORC_DEF_BRAGG 5e10        /* DefBraggThreshold, BrightField/BFConstants.h */
#define ORC_BF_OFFSET 32768.0     /* BF_OFFSET, HAADF/HAADFConstants.h:49 */

/* ---- setup pieces (pure functions) ---- */
int orc_geometry_init(int32_t N_r, int32_t N_t, int32_t N_theta, double delta_r, double delta_t,
                      double thickness, double delta_xz_in, double interpolate_factor,
                      int32_t extend_object, const double* angles_deg, orc_geom* g);
void orc_voxel_profile(double* angles_inout, int32_t N_theta, double delta_xz, double* profile);
void orc_sincos(const double* angles_rad, int32_t n, double* cosv, double* sinv);
void orc_beam_profile(double delta_t, double* beam);
void orc_detector_response(const orc_geom* g, const double* profile, const double* beam,
                           double offsetR, double* H);
void orc_init_ht(const orc_geom* g, double offsetT, double* Ht);
int32_t orc_vlr(const orc_geom* g, const double* Ht, double offsetT, int32_t y,
                int32_t* t_index, double* values, int32_t cap);
int32_t orc_acol(const orc_geom* g, const double* cosv, const double* sinv, const double* H,
                 double offsetR, int32_t z, int32_t x, uint32_t* index, double* values, int32_t cap);
double orc_offset_r(const orc_geom* g);
double orc_offset_t(const orc_geom* g);

/* ---- QGGMRF (Reconstruction/QGGMRF_Functions.cpp) ---- */
void orc_qggmrf_init(double sigma_x, double mrf_q, double mrf_c, orc_qggmrf* v);
double orc_qggmrf_value(double delta, const orc_qggmrf* v);
double orc_qggmrf_derivative(double delta, const orc_qggmrf* v);
double orc_qggmrf_second_derivative_at_zero(const orc_qggmrf* v);
double orc_qggmrf_functional_substitution(double umin, double umax, double v,
                                          const uint8_t* bflag, const double* filter,
                                          const double* nbr, double theta1, double theta2,
                                          const orc_qggmrf* q);
void orc_filter(double* filter27);

/* ---- engine state: HAADF_ReconstructionEngine::execute() setup and loop ---- */
typedef struct orc_engine orc_engine;

orc_engine* orc_engine_create(const orc_geom* g, const double* angles_deg, const double* counts,
                              const double* init_gain, const double* init_offset,
                              const double* init_variance, const double* init_x,
                              double default_recon, const orc_params* p);
void orc_engine_destroy(orc_engine* e);
/* one ICD sweep over all voxels, serial; returns 1 = continue, 0 = converged.
 * stats: [0]=sum|dx| (ROI) [1]=sum|x_old| (ROI) [2]=zero-skipped visits [3]=updated visits */
int32_t orc_engine_sweep(orc_engine* e, int32_t outer, int32_t inner, double* stats);
/* the same sweep with the reference's y-slab thread decomposition */
int32_t orc_engine_sweep_threads(orc_engine* e, int32_t outer, int32_t inner, int32_t nthreads,
                                 double* stats);
/* the y-slab [y0, y1) of a sweep, one thread (a bounded part of a sweep, for timing) */
void orc_engine_sweep_slab(orc_engine* e, int32_t outer, int32_t inner, int32_t y0, int32_t y1,
                           double* stats);
/* voxel updates in a given order of units (line, y0, ny), even local slices before odd */
void orc_engine_update_units(orc_engine* e, const int32_t* units, int32_t n, int32_t outer,
                             int32_t inner, double* stats);
int32_t orc_engine_joint_estimation(orc_engine* e);
void orc_engine_update_weights(orc_engine* e);
/* estimateSigmaX then updatePriorModel; returns the estimate, the new SigmaX in *sigma_x */
double orc_engine_estimate_sigma_x(orc_engine* e, double* sigma_x);
double orc_engine_get_sigma_x(const orc_engine* e);
/* bright-field NHICD helpers: ComputeVSC's 5 x 5 Hamming filter and SetNonHomThreshold */
void orc_nhicd_filter(const double* mag, int32_t Nz, int32_t Nx, double* filt);
double orc_nhicd_threshold(const double* v, int32_t n);
/* the lines of every NHICD call of the last run (bf_nhicd): returns the call count */
int32_t orc_engine_nhicd_log(const orc_engine* e, int64_t* call_ptr, int32_t cap_calls,
                             int32_t* lines, int64_t cap_lines);
double orc_engine_cost(orc_engine* e);
double orc_engine_cost_terms(orc_engine* e, double* terms3);
/* full outer/inner loop of execute(); costs[] receives one value per cost evaluation */
int32_t orc_engine_run(orc_engine* e, double* costs, int32_t cap, int32_t* n_costs,
                       int32_t* n_sweeps);

/* getters (copies) */
void orc_engine_get_x(const orc_engine* e, double* x);
void orc_engine_set_x(orc_engine* e, const double* x);
void orc_engine_set_e(orc_engine* e, const double* err);
/* HAADF with a bright-field sinogram (BF_Flag): projection entries scaled by
 * bf_counts + ORC_BF_OFFSET per element, e re-derived, no gain/offset estimation */
void orc_engine_set_bf_sinogram(orc_engine* e, const double* bf_counts);
void orc_engine_get_e(const orc_engine* e, double* err);
void orc_engine_get_w(const orc_engine* e, double* w);
void orc_engine_get_yest(const orc_engine* e, double* yest);
void orc_engine_get_nuisance(const orc_engine* e, double* I0, double* mu, double* alpha);
void orc_engine_get_H(const orc_engine* e, double* H);
void orc_engine_get_profile(const orc_engine* e, double* profile);
int64_t orc_engine_nnz(const orc_engine* e);
int32_t orc_engine_get_col(const orc_engine* e, int32_t z, int32_t x, uint32_t* index,
                           double* values, int32_t cap);
/* A x for an arbitrary volume with the engine's columns and gains (I0 applied) */
void orc_engine_project(const orc_engine* e, const double* x, double* yest);
/* theta1/theta2 of voxel (z,x,y) against the current e, w (reference :343-369) */
void orc_engine_theta(const orc_engine* e, int32_t z, int32_t x, int32_t y, double* t1t2);

/* bright-field pieces (BrightField/BFForwardModel.cpp) */
void orc_engine_get_sel(const orc_engine* e, uint8_t* sel);
double orc_engine_get_bragg(const orc_engine* e);
void orc_engine_set_bragg(orc_engine* e, double T);
void orc_engine_update_selector(orc_engine* e);     /* updateSelector, :517-531 */
/* Bragg data term of one entry: e^2 w if |e sqrt(w)| < T, else
 * 2 T delta |e| sqrt(w) + T^2 (1 - 2 delta)  (forwardCost, :812-846, before the 1/2) */
double orc_bf_beta(double e, double w, double T, double delta);

/* MT19937 (Matsumoto & Nishimura 1998), as boost::mt19937 / std::mt19937 */
typedef struct orc_mt { uint32_t mt[624]; int32_t idx; } orc_mt;
void orc_mt_seed(orc_mt* m, uint32_t seed);
uint32_t orc_mt_next(orc_mt* m);

#ifdef __cplusplus
}
#endif
#endif
