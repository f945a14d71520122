// mbir_host.h -- host-side state of the engine shared by the C-ABI (mbir_capi.cpp)
// and the request server (mbir_serve.cpp): the context, one job's device buffers
// and the sweep helpers both drive.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/mbir_tem.h"
#include "mbir_internal.h"

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      std::fprintf(stderr, "mbir: HIP error %s at %s:%d: %s\n", hipGetErrorName(_e),    \
                   __FILE__, __LINE__, #expr);                                           \
      return MBIR_ERR_HIP;                                                               \
    }                                                                                    \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count)
  {
    n = count;
    return hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};

using mbir::DevJob;
using mbir::SweepAccum;
using mbir::BatchEntry;
using mbir::LineRef;

struct mbir_ctx {
  int device = 0;
  int verbose = 0;
  uint64_t budget = 0;
  std::mutex mu;
  hipStream_t stream = nullptr;  // stream of batched launches
  DevBuf<DevJob> djobs;          // job table of a batched sweep
  DevBuf<BatchEntry> ents;       // entry table of a batched sweep
  mbir::LineTuning tune;         // form of the batched line kernels per shape and size
  // autotuning: when set, batched sweeps force this form (0 / 1) and time each launch
  struct Probe {
    int force = -1;
    std::vector<hipEvent_t> ev;       // two per launch
    std::vector<int> ypl, nv, nwg;    // per launch
  }* probe = nullptr;
  ~mbir_ctx()
  {
    djobs.release(); ents.release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct mbir_job {
  mbir_ctx* ctx = nullptr;
  mbir_problem prob{};
  mbir_geometry g{};
  hipStream_t stream = nullptr;
  DevJob J{};
  double offsetR = 0, offsetT = 0;
  int spacing = 1;  // conflict-free x spacing of lines in one step
  int ysplit = 1;   // y-slabs of the schedule: a step updates one unit (line, slab) per
                    // line of its row class in every slab (1 = whole lines)
  int zshift = 0;   // z offset between the row classes of neighbouring slabs in a step
  size_t sched_cap() const { return (size_t)g.N_z * g.N_x * ysplit; }
  std::vector<double> angles_rad, H;
  std::vector<double> I0h, muh, alphah;
  // device buffers
  DevBuf<float> x, e, y, wview, vlr_val, colv;
  DevBuf<double> cosv, sinv, Hd, I0, mu, alpha, I0b, mub, sums6, yest, cost, prj;
  DevBuf<int> vlr_t0, vlr_n, line_nnz, colr, col_ptr, col_row;
  DevBuf<LineRef> sched;
  DevBuf<float> col_a, col_ka, col_kw;
  DevBuf<double> line_kk;  // per line: the column factor of theta2 (plain weights)
  DevBuf<float> wfull;       // bright field: per-entry weights, sign = selector
  DevBuf<float> bfs;         // HAADF with a bright-field sinogram: per-element scale
  DevBuf<double> alpha_b;    // bright field: the next shared variance
  double bragg = 0;          // bright field: current Bragg threshold
  std::vector<int> nnz_h;  // per-line column entries (host copy)
  int64_t nnz_total = 0;
  DevBuf<uint8_t> mask;
  DevBuf<SweepAccum> acc;
  DevBuf<mbir::UnitStat> ustat;  // per schedule position: the unit's sweep statistics
  DevBuf<double> red;             // block sums of the statistics (5 kReduceBlocks)
  std::vector<LineRef> sched_h;
  std::vector<int> step_ptr;
  std::vector<int> cptr_h;  // column-table offsets of all lines (host copy)
  // bright-field NHICD: per-line update magnitudes, the seeded random list of all
  // lines, and the ROI mask on the host
  DevBuf<float> magmap;
  std::vector<float> magmap_h;
  std::vector<int> nh_all;
  std::vector<uint8_t> mask_h;
  // reach[d + N_z - 1]: the least x distance at which a line d rows away never shares a
  // detector bin with this one (every view) nor is a prior neighbour; subadditive in d
  std::vector<int> nh_reach;
  // the conflict-free steps of each homogeneous part (fixed lists: packed once)
  std::vector<std::vector<std::vector<int>>> nh_parts;
  int nh_log_on = 0;                 // record the lines of every NHICD call
  std::vector<int> nh_log;           // lines of all calls of the last run
  std::vector<int64_t> nh_log_ptr;   // call c: nh_log[nh_log_ptr[c] .. nh_log_ptr[c + 1])
  // kernel_mode 3: the sweep's step launches as a HIP graph over a padded schedule
  // (step k's lines at gsched[k * graph_lines], unused entries line = -1)
  DevBuf<LineRef> gsched;
  std::vector<LineRef> gsched_h;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  DevJob graph_J{};
  int graph_steps = 0, graph_lines = 0;
  std::atomic<int> cancel{0};
  int timing = 0;
  int sweep_counter = 0;
  double stop_threshold = 0;
  double sigma_x = 0;  // current prior scale (TomoInputs::SigmaX), re-estimated with estimate_prior
  int64_t bytes = 0;
  ~mbir_job()
  {
    x.release(); e.release(); y.release(); wview.release(); vlr_val.release(); colv.release();
    cosv.release(); sinv.release(); Hd.release(); I0.release(); mu.release(); alpha.release();
    I0b.release(); mub.release(); sums6.release(); yest.release(); cost.release(); prj.release();
    vlr_t0.release(); vlr_n.release(); line_nnz.release(); sched.release(); colr.release();
    mask.release(); acc.release(); ustat.release(); red.release(); col_ptr.release(); col_row.release();
    col_a.release(); col_ka.release(); col_kw.release(); wfull.release(); bfs.release(); alpha_b.release();
    gsched.release();
    magmap.release();
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace mbir {

// Sweep helpers (mbir_capi.cpp).  prepare_sweep builds the seeded step schedule
// of the job's next sweep, uploads it and clears the sweep accumulators on stream s;
// sweep_status applies the stop test to the accumulators (1 continue, 0 converged)
// and counts the sweep; nuisance_update_on enqueues joint estimation, the noise
// update and the column-weight refresh on s without waiting.
void build_schedule(mbir_job* j, bool drop_empty_steps);
int prepare_sweep_on(mbir_job* j, hipStream_t s, const LineRef* sched_src);
int sweep_status(mbir_job* j, const SweepAccum& a, double ms, mbir_sweep_stats* st);
int nuisance_update_on(mbir_job* j, hipStream_t s);
// Bright field: the Bragg threshold used by the next launches (DevJob fields).
void set_bragg(mbir_job* j, double T);
// QGGMRF constants of the prior scale sigma_x (DevJob fields).
void set_prior(mbir_job* j, double sigma_x);

}  // namespace mbir
