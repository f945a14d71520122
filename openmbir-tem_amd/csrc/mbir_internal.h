// mbir_internal.h -- device-side job descriptor and kernel launchers shared by
// the HIP kernels (mbir_kernels.hip) and the host engine (mbir_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbir {

constexpr int kResponseBins = 64;     // DETECTOR_RESPONSE_BINS (HAADF_ReconstructionEngine.cpp:249)
constexpr int kProfileResolution = 1536;
constexpr int kBeamResolution = 512;
constexpr int kBlock = 256;           // threads per workgroup of the setup / FP64 kernels
constexpr int kWaves = kBlock / 64;
constexpr int kIcdBlock = 512;        // threads per ICD line workgroup (8 wave64)
constexpr int kIcdWaves = kIcdBlock / 64;
constexpr int kMaxNy = 256;           // voxels per line handled by one workgroup
constexpr int kMaxSlots = 2048;       // N_theta * NB per column
// zero entries behind the column table: a line's first batch reads rows past its own
// column (at most one column's slots plus a batch) without a bound per entry
constexpr int kColPad = kMaxSlots + 64;
constexpr int kMaxNB = 12;            // detector bins per view per column (voxels up to 5.5 delta_r)
constexpr int kVlrCap = 4;            // detector rows per voxel slice

// One scheduled line: its id and its column's place in the column table, so a
// line update starts without a dependent lookup.
struct LineRef {
  int line;
  int p0;  // first column-table entry
  int nc;  // column entries
  int pad;
};

// The units of one step carried in the kernel arguments (single-job launches of at
// most kInlineLines units): they arrive with the job descriptor in the launch's first
// memory round trip instead of one dependent load later.
// Each entry is 12 bytes, {line | nc << 20, p0, pad} (jobs with fewer than 2^20 lines,
// columns of fewer than 4096 entries); the launch takes the smaller of two capacities
// that holds the step (argument size changes a dependent launch's time by at most
// 0.3 us up to 3.5 KB on this device).
constexpr int kInlineSmall = 96;
constexpr int kInlineLines = 256;
template <int CAP>
struct StepLines {
  int3 l[CAP > 0 ? CAP : 1];
};

// A unit's share of its sweep's stop-test statistics, stored by the line update at the
// unit's schedule position (no atomics) and summed in schedule order once per sweep:
// sums of |x_new - x_old| and |x_old| over its ROI slices, and the counts zero-skipped
// + 2^10 updated + 2^20 detector rows of its updated slices.
struct UnitStat {
  float su, sv;
  int cnt;
  int pad;
};

struct SweepAccum {  // device accumulators of one sweep
  double sum_abs_update;
  double sum_abs_value;
  unsigned long long zero_skipped;
  unsigned long long updated;
  unsigned long long nnz_updated;
  unsigned long long pad[3];
};

// Everything a kernel needs about one job; passed by value.
struct DevJob {
  int N_theta, N_r, N_t, N_x, N_y, N_z;
  int NB;                // slots (candidate detector bins) per view
  int nslots;            // N_theta * NB
  int vlr_max;           // max detector rows of any slice (1 when delta_xy = delta_t)
  double x0, z0, delta_xz, delta_r, R0, RMax, offsetR;
  double inv_delta_r, inv_offsetR;
  const double* cosv;    // [N_theta]
  const double* sinv;    // [N_theta]
  const double* H;       // [N_theta][64] detector response
  const int* vlr_t0;     // [N_y] first detector row of slice y
  const int* vlr_n;      // [N_y] rows of slice y (<= kVlrCap)
  const float* vlr_val;  // [N_y][kVlrCap]
  const uint8_t* mask;   // [N_z][N_x] ROI
  const int* line_nnz;   // [N_z][N_x] A-column entries per line (0 = empty column)
  // compact A-matrix columns of all lines (built once on the device, FP32):
  // entries col_ptr[line] .. col_ptr[line + 1] - 1
  const int* col_ptr;    // [N_z * N_x + 1]
  const int* col_row;    // sinogram row base (th * N_r + r) * N_t
  const float* col_a;    // A value a
  const float* col_ka;   // I0[th] * a            (rebuilt when I0 changes)
  const float* col_kw;   // I0[th] * a * w[th]    (rebuilt when I0 or w changes)
  const double* line_kk; // [N_z * N_x] sum over the column of col_ka col_kw in FP64 (plain
                         // weights: the column factor of theta2; rebuilt with col_kw)
  float* x;              // [N_z][N_x][N_y]
  float* e;              // [N_theta][N_r][N_t] error sinogram
  const float* y;        // [N_theta][N_r][N_t] measurements
  float* wfull;          // [N_theta][N_r][N_t] per-entry weights (bright field), its sign the
                         // Bragg selector (< 0: rejected); nullptr = per-view weights
  const float* wview;    // [N_theta] per-view weight (identity noise model: 1/alpha)
  const double* I0;      // [N_theta]
  const double* mu;      // [N_theta]
  // QGGMRF (HAADF/HAADF_QGGMRFPriorModel.cpp:10-19), MRF_P = 2
  float mrf_q, mrf_c, sigma_p, sigma_pq, alpha, b0;  // b0 = P / (sigma^P c)
  float pq2, inv_sigma_p, inv_sigma_pq;            // 2 - q, 1 / sigma^P, 1 / sigma^(P-q)
  double mrf_qd, mrf_cd, sigma_pd, sigma_pqd;
  float bragg, bragg_dT;  // bright field: Bragg threshold T and delta T
  double bragg_d, delta_d;  // the same T and delta in FP64 (cost, nuisance updates)
  int e_atomic;          // 1: error-sinogram update by fp32 atomics (required when lines of
                         // a step may share detector bins: relaxed schedule), 0: plain
                         // read-modify-write (conflict-free steps only)
  int unit_max;          // slices of the job's longest schedule unit (N_y for whole lines):
                         // the line kernels' shape
  int col_slots;         // padded column entries of this job's longest line (a multiple of
                         // 4 P, P = 8 * 64 / TY streams): the line update's dynamic LDS
  uint32_t spare_b;      // M * 4: byte offset of the spare detector row (N_t zeros) that
                         // follows e and wfull, the target of padding column entries
  float* magmap;         // [N_z * N_x] per-line sum of |x_new - x_old| of the last update
                         // (NHICD), nullptr = not kept
  const float* bfs;      // HAADF with a bright-field sinogram: per-element scale (counts +
                         // BF_OFFSET), spare row 0; null without
  uint32_t nx_div;       // ceil(2^32 / N_x): line / N_x = umulhi(line, nx_div) for every
                         // line of the volume (checked on the host), 0 = divide
  const LineRef* sched;  // [N_z * N_x * y slabs] units of the current sweep, step by step
  UnitStat* ustat;       // per schedule position: the unit's statistics of the sweep
};


// One job's share of one batched launch: the lines of its current step.
struct BatchEntry {
  const LineRef* lines;  // lines of the step
  int nlines;
  int job;           // index into the launch's DevJob table
  int wg_off;        // first workgroup of this entry within the launch
  int flags;         // bit 0: first sweep, bit 1: zero skipping on
  int ypl, nv;       // kernel shape
  int fw;            // bright field (per-entry weights): kernels of their own
  int pos;           // schedule position of the entry's first unit in its job's sweep
};

// One job's share of one request-server launch: lines sched[line0 .. line0 + nlines)
// of job slot `job`, served by workgroups wg_off .. wg_off + nlines - 1.
struct TickEntry {
  int job;
  int line0;
  int nlines;
  int wg_off;
  int flags;  // bit 0: first sweep, bit 1: zero skipping on
};
constexpr int kTickMax = 48;  // entries per launch, passed by value as a kernel argument (1 kB:
                               // a larger argument block slows every launch)
struct TickArgs {
  int n;
  int slots;  // column entries (after padding) of the launch's longest line: dynamic LDS
  TickEntry e[kTickMax];
};

// Choice between the two forms of the batched line kernels (identical results):
// 1 = prefetching (shorter line latency), 0 = without prefetch (fewer registers), per
// kernel shape (ypl 1|2, nv 1|2) and launch size (workgroups) bucket.
constexpr int kTuneBuckets = 5;
inline int tune_bucket(int nwg)
{
  return nwg <= 256 ? 0 : nwg <= 512 ? 1 : nwg <= 1024 ? 2 : nwg <= 2048 ? 3 : 4;
}
struct LineTuning {
  int8_t pf[2][2][kTuneBuckets];
  LineTuning()
  {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int k = 0; k < kTuneBuckets; ++k) pf[a][b][k] = k <= 1 ? 1 : 0;  // <= 512: prefetch
  }
  int choose(int ypl, int nv, int nwg) const
  {
    return pf[ypl > 1 ? 1 : 0][nv > 1 ? 1 : 0][tune_bucket(nwg)];
  }
};

// ---- launchers (all asynchronous on `s`) ----
hipError_t launch_line_nnz(const DevJob& J, int* line_nnz, float* x, hipStream_t s);
hipError_t launch_forward_project(const DevJob& J, const float* vol, double* out, hipStream_t s);
hipError_t launch_init_error(const DevJob& J, const double* yest, hipStream_t s);
// fine volume (sizes N_z, N_x, N_y) from the coarse one of sizes N/2, replicated 2 x 2 x 2;
// lines with line_nnz 0 stay 0
hipError_t launch_upsample2(const float* coarse, float* fine, const int* line_nnz, int Nz, int Nx,
                            int Ny, hipStream_t s);
// host_lines: the same nlines entries in host memory; steps of at most kInlineLines
// lines then travel in the kernel arguments (not for graph capture, whose step slots
// are rewritten in device memory between replays).
// pos0: the schedule position of lines[0] (the units' statistics go to J.ustat[pos0 + i])
hipError_t launch_icd_step(const DevJob& J, const LineRef* lines, int nlines, int pos0, int outer,
                           int inner, int zero_skipping, hipStream_t s,
                           const LineRef* host_lines = nullptr);
// acc = the sums of the n units' statistics of sched / ustat (entries with line < 0 are
// padding), in schedule order; part holds 5 kReduceBlocks doubles
constexpr int kReduceBlocks = 128;
hipError_t launch_sweep_reduce(const LineRef* sched, const UnitStat* ustat, int n, SweepAccum* acc,
                               double* part, hipStream_t s);
void icd_shape(const DevJob& J, int max_vlr, int* ypl, int* nv);
// one step of several jobs of one kernel shape: entries [e0, e0 + ne) of ent,
// nwg workgroups in total
hipError_t launch_icd_batch_step(int ypl, int nv, const DevJob* jobs, const BatchEntry* ent,
                                 int e0, int ne, int nwg, int slots, int pf, hipStream_t s, int fw = 0);
hipError_t launch_build_columns(const DevJob& J, int* row, float* a, hipStream_t s);
hipError_t launch_column_weights(const DevJob& J, float* ka, float* kw, int nnz, hipStream_t s);
hipError_t launch_line_kk(const DevJob& J, double* kk, hipStream_t s);
hipError_t launch_export_column(const DevJob& J, int z, int x, float* vals, int* rows,
                                hipStream_t s);
// joint estimation / weights / cost (FP64)
hipError_t launch_joint_sums(const DevJob& J, double* sums6, hipStream_t s);
hipError_t launch_joint_solve(int N_theta, const double* sums6, double target_gain,
                              const double* I0_old, const double* mu_old, double* I0_new,
                              double* mu_new, hipStream_t s);
hipError_t launch_joint_apply(const DevJob& J, const double* I0_old, const double* mu_old,
                              const double* I0_new, const double* mu_new, hipStream_t s);
hipError_t launch_noise_update(const DevJob& J, double* alpha, float* wview, hipStream_t s);
// estimateSigmaX's prior sum (HAADF/HAADF_ReconstructionEngine.cpp:1827-1945) over the
// planes [z0, z1), differences scaled by sx, into *out; part holds 1024 doubles
hipError_t launch_sigma_sum(const DevJob& J, int z0, int z1, double sx, double* part, double* out,
                            hipStream_t s);
// HAADF given bright-field data: the offset-only branch of jointEstimation
// (HAADF/HAADF_ReconstructionEngine_Extra.cpp:636-669); nd holds 2 N_theta doubles
hipError_t launch_offset_only(const DevJob& J, double* nd, double* mu, hipStream_t s);
hipError_t launch_cost(const DevJob& J, const double* alpha, int noise_model, double* out3,
                       hipStream_t s);
// one launch of the request server: kernel shape (ypl, nv), nwg workgroups
hipError_t launch_icd_tick(int ypl, int nv, const DevJob* jobs, const TickArgs& args, int nwg,
                           int pf, hipStream_t s, int fw = 0);
hipError_t launch_fill_u32(unsigned int* p, unsigned int v, size_t n, hipStream_t s);
// out[0] += sum of the volume over the lines of the ROI mask (FP64); out zeroed by the caller
hipError_t launch_roi_sum(const DevJob& J, double* out, hipStream_t s);
// jobs[slot] = J, ordered on s (the descriptor travels as a kernel argument)
hipError_t launch_set_devjob(DevJob* jobs, int slot, const DevJob& J, hipStream_t s);
// bright field (BrightField/BFForwardModel.cpp): offset-only joint estimation
// (:390-441) with the selector refresh (:517-531), and the shared variance (:447-511)
hipError_t launch_bf_offset(const DevJob& J, double* nd, double* mu, hipStream_t s);
hipError_t launch_bf_noise(const DevJob& J, const double* alpha_old, double* alpha_new,
                           double* part, hipStream_t s);
hipError_t launch_fill_f32(float* p, float v, size_t n, hipStream_t s);
hipError_t launch_fill_f64(double* p, double v, size_t n, hipStream_t s);

}  // namespace mbir
