// mbir_capi.cpp -- host side of the MI355X MBIR engine: FP64 setup of the
// detector model, device state of one job, the ICD sweep schedule, and the
// extern "C" boundary declared in include/mbir_tem.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

#include "mbir_host.h"

using namespace mbir;

#define MBIR_VERSION 10000

namespace {

const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Host FP64 detector model (once per job).

// calculateVoxelProfile, HAADF/HAADF_ReconstructionEngine.cpp:957-1029; the angles are
// multiplied by pi/180 in place as in the reference (:978).
void voxel_profile(std::vector<double>& angles, double dxz, std::vector<double>& prof)
{
  const int n = (int)angles.size();
  prof.assign((size_t)n * kProfileResolution, 0.0);
  for (int i = 0; i < n; ++i) {
    angles[i] = angles[i] * (kPi / 180.0);
    double angle = angles[i];
    while (angle > kPi / 2) angle -= kPi / 2;
    while (angle < 0) angle += kPi / 2;
    const double maxv = angle <= kPi / 4 ? dxz / std::cos(angle) : dxz / std::cos(kPi / 2 - angle);
    const double c45 = std::cos(kPi / 4);
    const double d1 = c45 * std::cos(kPi / 4 - angle);
    const double d2 = c45 * std::fabs(std::cos(kPi / 4 + angle));
    const double lc = 1 - d1, ln = 1 - d2, rn = 1 + d2, rc = 1 + d1;
    for (int j = 0; j < kProfileResolution; ++j) {
      const double t = 2.0 * j / kProfileResolution;
      double v;
      if (t <= lc || t >= rc) v = 0;
      else if (t > rn) v = maxv * (rc - t) / (rc - rn);
      else if (t >= ln) v = maxv;
      else v = maxv * (t - lc) / (ln - lc);
      prof[(size_t)i * kProfileResolution + j] = v;
    }
  }
}

// DetectorParameters::initializeBeamProfile (GenericFilters/DetectorParameters.cpp:78-100)
// and DetectorResponse::execute (GenericFilters/DetectorResponse.cpp:66-127).
void detector_response(const mbir_geometry& g, const std::vector<double>& prof, double offsetR,
                       std::vector<double>& H)
{
  double beam[kBeamResolution], sum = 0;
  for (int i = 0; i < kBeamResolution; ++i) {
    beam[i] = 0.54 - 0.46 * std::cos((2.0 * kPi / kBeamResolution) * i);
    sum += beam[i];
  }
  for (int i = 0; i < kBeamResolution; ++i) beam[i] = beam[i] / sum / g.delta_t;
  const double r0 = -g.delta_r / 2, step = g.delta_r / kBeamResolution;
  const double tc = kProfileResolution / (2 * g.delta_xz);
  H.assign((size_t)g.N_theta * kResponseBins, 0.0);
  for (int k = 0; k < g.N_theta; ++k)
    for (int i = 0; i < kResponseBins; ++i) {
      const double rmin = i * offsetR - g.delta_xz;
      double s = 0;
      for (int p = 0; p < kBeamResolution; ++p) {
        const double r = r0 + p * step;
        if (r < rmin) continue;
        int pi = (int16_t)(int32_t)std::floor((r - rmin) * tc);
        if (pi < 0) pi = 0;
        if (pi >= kProfileResolution) pi = kProfileResolution - 1;
        s += prof[(size_t)k * kProfileResolution + pi] * beam[p];
      }
      H[(size_t)k * kResponseBins + i] = s;
    }
}

// H_t row of view 0 (HAADF/HAADF_ForwardModel.cpp:408-451) and the voxel-line
// responses (HAADF/HAADF_ReconstructionEngine_Extra.cpp:329-404).
int voxel_line_responses(const mbir_geometry& g, double offsetT, std::vector<int>& t0,
                         std::vector<int>& n, std::vector<float>& val)
{
  double Ht[kResponseBins];
  for (int i = 0; i < kResponseBins; ++i) {
    const double pct = i * offsetT;
    double v;
    if (g.delta_xy >= g.delta_t)
      v = pct <= (g.delta_xy / 2 - g.delta_t / 2) ? g.delta_t : -pct + g.delta_xy / 2 + g.delta_t / 2;
    else
      v = pct <= g.delta_t / 2 - g.delta_xy / 2 ? g.delta_xy : -pct + g.delta_xy / 2 + g.delta_t / 2;
    Ht[i] = v < 0 ? 0 : v;
  }
  t0.assign(g.N_y, 0);
  n.assign(g.N_y, 0);
  val.assign((size_t)g.N_y * kVlrCap, 0.f);
  for (int y = 0; y < g.N_y; ++y) {
    const double t = ((double)y + 0.5) * g.delta_xy + g.y0;
    const double tmin = (t - g.delta_xy / 2) > g.T0 ? t - g.delta_xy / 2 : g.T0;
    const double tmax = (t + g.delta_xy / 2) <= g.TMax ? t + g.delta_xy / 2 : g.TMax;
    int smin = (int16_t)(uint16_t)std::floor((tmin - g.T0) / g.delta_t);
    int smax = (int16_t)(uint16_t)std::floor((tmax - g.T0) / g.delta_t);
    if (smin < 0) smin = 0;
    if (smax >= g.N_t) smax = g.N_t - 1;
    int cnt = 0, first = -1, last = -1;
    for (int it = smin; it <= smax; ++it) {
      const double ct = ((double)it + 0.5) * g.delta_t + g.T0;
      const double dt = std::fabs(ct - t);
      const int idt = (int16_t)(uint16_t)std::floor(dt / offsetT);
      double pt = 0;
      if (idt < kResponseBins) {
        const double w3 = dt - idt * offsetT, w4 = (idt + 1) * offsetT - dt;
        const int nx = idt + 1 < kResponseBins ? idt + 1 : kResponseBins - 1;
        pt = (w4 / offsetT) * Ht[idt] + (w3 / offsetT) * Ht[nx];
      }
      if (pt != 0) {
        // the reference walks index[0] .. index[0]+count-1 (…_UpdateVoxels.cpp:349-350),
        // i.e. it treats the rows as contiguous
        if (cnt >= kVlrCap) return MBIR_ERR_UNSUPPORTED;
        if (first < 0) first = it;
        val[(size_t)y * kVlrCap + cnt] = (float)pt;
        last = it;
        cnt++;
      }
    }
    t0[y] = first < 0 ? 0 : first;
    n[y] = cnt;
    (void)last;
  }
  // slices must map to disjoint detector rows (holds whenever y0 = T0, the
  // reference's geometry, GenericFilters/InitialReconstructionInitializer.cpp:167)
  for (int y = 0; y + 1 < g.N_y; ++y)
    if (n[y] && n[y + 1] && t0[y] + n[y] > t0[y + 1]) return MBIR_ERR_UNSUPPORTED;
  return 0;
}


}  // namespace


// ---------------------------------------------------------------------------
extern "C" int32_t mbir_version(void) { return MBIR_VERSION; }

extern "C" int32_t mbir_create(const mbir_opts* opts, mbir_ctx** out)
{
  if (!out) return MBIR_ERR_INPUT;
  *out = nullptr;
  auto* c = new mbir_ctx();
  if (opts) {
    c->device = opts->device;
    c->verbose = opts->verbose;
    c->budget = opts->hbm_budget_bytes;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || c->device < 0 || c->device >= n) {
    delete c;
    return MBIR_ERR_HIP;
  }
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return MBIR_ERR_HIP;
  }
  *out = c;
  return MBIR_OK;
}

extern "C" int32_t mbir_destroy(mbir_ctx* ctx)
{
  delete ctx;
  return MBIR_OK;
}

extern "C" int32_t mbir_geometry_init(int32_t N_r, int32_t N_t, int32_t N_theta, double delta_r,
                                      double delta_t, double thickness, double delta_xz_in,
                                      double interp, int32_t extend_object,
                                      const double* angles_deg, mbir_geometry* g)
{
  if (!g || !angles_deg || N_r <= 0 || N_t <= 0 || N_theta <= 0 || delta_r <= 0 ||
      delta_t <= 0 || thickness <= 0 || delta_xz_in <= 0 || interp <= 0)
    return MBIR_ERR_INPUT;
  std::memset(g, 0, sizeof(*g));
  g->N_r = N_r;
  g->N_t = N_t;
  g->N_theta = N_theta;
  g->delta_r = delta_r;
  g->delta_t = delta_t;
  g->delta_xz = delta_r * delta_xz_in;  // :109
  g->delta_xy = g->delta_xz;
  double mx = 0;
  for (int i = 0; i < N_theta; ++i) mx = std::max(mx, std::fabs(angles_deg[i]));
  if (mx > 75.0) mx = 75.0;  // k_MaxAngleStretch
  mx = mx * kPi / 180.0;
  double LZ = thickness * 2.0;  // Z_STRETCH
  LZ /= (interp * delta_r);
  LZ = std::floor(LZ + 0.5) * interp * delta_r;
  double LX = N_r * delta_r;
  if (extend_object == 1) {
    LX = 0.6 * ((N_r * delta_r) / std::cos(mx)) + LZ * std::tan(mx);
    LX /= (interp * delta_r);
    LX = std::floor(LX + 0.5) * interp * delta_r;
  }
  const double LY = N_t * delta_t;
  g->N_x = (int32_t)(uint16_t)std::floor(LX / g->delta_xz);
  g->N_z = (int32_t)(uint16_t)std::floor(LZ / g->delta_xz);
  g->N_y = (int32_t)(uint16_t)std::floor(LY / g->delta_xy);
  g->LengthX = LX;
  g->LengthY = LY;
  g->LengthZ = LZ;
  g->x0 = -LX / 2;
  g->z0 = -LZ / 2;
  g->y0 = -LY / 2;
  g->R0 = -(N_r * delta_r) / 2;
  g->RMax = (N_r * delta_r) / 2;
  g->T0 = -(N_t * delta_t) / 2;
  g->TMax = (N_t * delta_t) / 2;
  return (g->N_x > 0 && g->N_y > 0 && g->N_z > 0) ? MBIR_OK : MBIR_ERR_INPUT;
}

namespace {

// Slabs of the automatic y split: whole lines up to 64 slices (a wave's lanes), else
// units of at most 64 slices, so that every job runs the one-lane-row form of the line
// update (its launch time grows with the slices per lane, while up to about 150
// workgroups of units cost no more than one of lines).
int auto_y_slabs(int Ny, int Nz)
{
  if (Nz < 4 || Ny <= 64) return 1;
  return (Ny + 63) / 64;
}

// The reference's random list (GenRandList, BrightField/BFReconstructionEngine_Extra.cpp:
// 1072-1105) of 0 .. n - 1 with a seeded MT19937 in place of its clock-seeded generator:
// entry t is Counter[Index], Index = next() % remaining, then Counter[Index] takes the last
// remaining counter.
std::vector<int> rand_list(int n, uint32_t seed)
{
  std::mt19937 rng(seed);
  std::vector<int> counter(n), out(n);
  for (int i = 0; i < n; ++i) counter[i] = i;
  int size = n;
  for (int t = 0; t < n; ++t) {
    const uint32_t idx = (uint32_t)rng() % (uint32_t)size;
    out[t] = counter[idx];
    counter[idx] = counter[size - 1];
    size--;
  }
  return out;
}

int setup_job(mbir_job* j)
{
  const mbir_problem& P = j->prob;
  const mbir_geometry& g = j->g;
  const int NT = g.N_theta;
  const size_t M = (size_t)NT * g.N_r * g.N_t;
  const size_t V = (size_t)g.N_z * g.N_x * g.N_y;
  const size_t L = (size_t)g.N_z * g.N_x;

  // ---- FP64 detector model on the host (execute(), :464-530) ----
  std::vector<double> prof;
  j->angles_rad.assign(P.angles_deg, P.angles_deg + NT);
  voxel_profile(j->angles_rad, g.delta_xz, prof);  // first call, :464
  j->offsetR = ((g.delta_xz / std::sqrt(3.0)) + g.delta_r / 2) / kResponseBins;  // :468
  j->offsetT = ((g.delta_xz / 2) + g.delta_t / 2) / kResponseBins;                // :469
  std::vector<double> cosv(NT), sinv(NT);
  for (int i = 0; i < NT; ++i) {  // DetectorParameters.cpp:60-73
    cosv[i] = std::cos(j->angles_rad[i]);
    sinv[i] = std::sin(j->angles_rad[i]);
  }
  {
    // second calculateVoxelProfile (:499) converts the radian angles once more;
    // that profile feeds the detector response
    std::vector<double> twice = j->angles_rad;
    voxel_profile(twice, g.delta_xz, prof);
  }
  detector_response(g, prof, j->offsetR, j->H);
  for (double h : j->H)
    if (!std::isfinite(h)) return MBIR_ERR_RESPONSE;
  std::vector<int> vt0, vn;
  std::vector<float> vval;
  int rc = voxel_line_responses(g, j->offsetT, vt0, vn, vval);
  if (rc < 0) return rc;

  // NB: detector bins a column can touch per view, floor((r+dxz-R0)/dr) - floor((r-dxz-R0)/dr) + 1
  const int NB = (int)std::ceil(2.0 * g.delta_xz / g.delta_r - 1e-12) + 1;
  if (NB > kMaxNB || NT * NB > kMaxSlots || g.N_y > kMaxNy || NT > kIcdBlock)
    return MBIR_ERR_UNSUPPORTED;
  // the line update addresses the error sinogram by 32-bit byte offsets
  if (((double)M + g.N_t) * 4.0 >= 4294967295.0) return MBIR_ERR_UNSUPPORTED;

  // conflict-free spacing: lines of one row dx apart never share a detector bin
  // if dx * |cos th| >= 2 * 64 * OffsetR for every view (H support radius).
  double cmin = 1.0;
  for (int i = 0; i < NT; ++i) cmin = std::min(cmin, std::fabs(cosv[i]));
  const double need = 2.0 * kResponseBins * j->offsetR / g.delta_xz;
  int S = cmin > 1e-6 ? (int)std::ceil(need / cmin - 1e-9) : g.N_x;
  S = std::max(2, std::min(S, g.N_x));
  j->spacing = S;
  if (P.bf_nhicd) {
    // lines d rows apart and dx >= 0 columns apart share no bin when
    // |dx cos th - d sin th| >= need for every view, i.e. dx lies above every view's
    // interval d tan th +- need / |cos th|; reach[d] is the least integer above them all
    // (at least 2: prior neighbours).  A maximum of terms d t_k + c_k with c_k >= 0 is
    // subadditive in d, so lines at increasing x, each at least reach[d] after the
    // previous one, are pairwise free of conflicts.
    const int Nz = g.N_z;
    j->nh_reach.assign((size_t)2 * Nz - 1, g.N_x);
    for (int d = -(Nz - 1); d <= Nz - 1; ++d) {
      double hi = 0.0;
      bool ok = true;
      for (int i = 0; i < NT && ok; ++i) {
        const double c = std::fabs(cosv[i]);
        if (c < 1e-6) ok = false;
        else hi = std::max(hi, d * sinv[i] / cosv[i] + need / c);
      }
      if (ok) j->nh_reach[(size_t)(d + Nz - 1)] =
          std::max(2, (int)std::min<double>(g.N_x, std::ceil(hi + 1e-9)));
    }
  }
  // y-slabs: units of a step lie in different slabs, the row classes of neighbouring slabs
  // zshift >= 2 planes apart (and N_z - zshift >= 2 across the wrap), so they are never
  // prior neighbours; slices of different slabs touch disjoint detector rows
  {
    int ns = P.y_slabs;
    if (ns == 0) ns = auto_y_slabs(g.N_y, g.N_z);
    ns = std::max(1, std::min(ns, g.N_y));
    if (P.schedule != 0 || P.bf_nhicd || g.N_z < 4) ns = 1;
    j->ysplit = ns;
    j->zshift = ns > 1 ? std::max(2, std::min(g.N_z / ns, g.N_z - 2)) : 0;
    j->J.unit_max = (g.N_y + ns - 1) / ns;
  }

  // ROI mask (…_Extra.cpp:227-248)
  std::vector<uint8_t> mask(L);
  for (int i = 0; i < g.N_z; ++i)
    for (int k = 0; k < g.N_x; ++k) {
      const double xx = g.x0 + ((double)k + 0.5) * g.delta_xz;
      const double zz = g.z0 + ((double)i + 0.5) * g.delta_xz;
      mask[(size_t)i * g.N_x + k] = xx >= -(g.N_r * g.delta_r) / 2 &&
                                    xx <= (g.N_r * g.delta_r) / 2 && zz >= -g.LengthZ / 2 &&
                                    zz <= g.LengthZ / 2;
    }

  // gainAndOffsetInitialization (…_Extra.cpp:280-307) and initial variances
  j->I0h.resize(NT);
  j->muh.resize(NT);
  j->alphah.resize(NT);
  double sum = 0;
  for (int k = 0; k < NT; ++k) {
    const double g0 = P.init_gain ? P.init_gain[k] : P.target_gain;
    j->I0h[k] = g0;
    j->muh[k] = P.init_offset ? P.init_offset[k] : P.default_offset;
    j->alphah[k] = P.init_variance ? P.init_variance[k] : P.default_variance;
    sum += g0;
  }
  sum /= NT;
  if (sum - P.target_gain > 1e-5)
    for (int k = 0; k < NT; ++k) j->I0h[k] *= P.target_gain / sum;
  for (int k = 0; k < NT; ++k)
    if (!(j->I0h[k] != 0.0) || (P.noise_model && !(j->alphah[k] > 0.0))) return MBIR_ERR_INPUT;

  // ---- device state ----
  hipStream_t s = j->stream;
  HIP_TRY(j->x.alloc(V));
  // a spare detector row behind e (and wfull): padding entries of the line update
  HIP_TRY(j->e.alloc(M + g.N_t));
  HIP_TRY(launch_fill_f32(j->e.p + M, 0.f, g.N_t, s));
  HIP_TRY(j->y.alloc(M));
  HIP_TRY(j->wview.alloc(NT));
  HIP_TRY(j->vlr_val.alloc((size_t)g.N_y * kVlrCap));
  HIP_TRY(j->vlr_t0.alloc(g.N_y));
  HIP_TRY(j->vlr_n.alloc(g.N_y));
  HIP_TRY(j->cosv.alloc(NT));
  HIP_TRY(j->sinv.alloc(NT));
  HIP_TRY(j->Hd.alloc((size_t)NT * kResponseBins));
  HIP_TRY(j->I0.alloc(NT));
  HIP_TRY(j->mu.alloc(NT));
  HIP_TRY(j->alpha.alloc(NT));
  HIP_TRY(j->I0b.alloc(NT));
  HIP_TRY(j->mub.alloc(NT));
  HIP_TRY(j->sums6.alloc((size_t)NT * 6));
  HIP_TRY(j->cost.alloc(4 + 2 * 1024));
  HIP_TRY(j->mask.alloc(L));
  HIP_TRY(j->line_nnz.alloc(L));
  HIP_TRY(j->sched.alloc(j->sched_cap()));
  HIP_TRY(j->ustat.alloc(j->sched_cap()));
  HIP_TRY(j->red.alloc(5 * kReduceBlocks));
  HIP_TRY(j->acc.alloc(1));
  HIP_TRY(j->colv.alloc(kMaxSlots));
  HIP_TRY(j->colr.alloc(kMaxSlots));
  j->bytes = (int64_t)(V * 4 + 2 * M * 4 + L * 9 + (size_t)NT * 64 * 8);

  // bright field: -log(counts + offset) and the count weights in FP64 on the host
  // (processRawCounts, BrightField/BFForwardModel.cpp:939-972; calculateMeasurementWeight
  // :328-385 with BF_MAX = 1)
  std::vector<float> bf_y, bf_w;
  if (P.bf_mode) {
    if (P.schedule != 0) return MBIR_ERR_UNSUPPORTED;
    bf_y.resize(M);
    bf_w.resize(M);
    const size_t per = (size_t)g.N_r * g.N_t;
    for (size_t i = 0; i < M; ++i) {
      const double c = (double)P.counts[i] + P.bf_offset;
      if (!(c > 0.0)) return MBIR_ERR_INPUT;  // negative counts (:956-960)
      const double yv = -std::log(c / 1.0);
      double w = 1.0 / std::exp(yv);
      if (P.noise_model) w /= j->alphah[i / per];
      bf_y[i] = (float)yv;
      bf_w[i] = (float)w;  // positive: every entry selected (:347)
    }
    HIP_TRY(j->wfull.alloc(M + g.N_t));
    HIP_TRY(launch_fill_f32(j->wfull.p + M, 0.f, g.N_t, s));
    HIP_TRY(j->alpha_b.alloc(NT));
    HIP_TRY(hipMemcpyAsync(j->wfull.p, bf_w.data(), M * sizeof(float), hipMemcpyHostToDevice, s));
    j->bytes += (int64_t)M * 4;
  }
  // HAADF with a bright-field sinogram (HAADF/HAADF_ReconstructionEngine_Extra.cpp:
  // 143-168): the per-element scale counts + BF_OFFSET, a zero spare row behind it
  std::vector<float> bfs_h;
  if (P.bf_sinogram) {
    if (P.bf_mode) return MBIR_ERR_INPUT;
    constexpr double kBfOffset = 32768.0;  // BF_OFFSET, HAADF/HAADFConstants.h:49
    bfs_h.resize(M);
    for (size_t i = 0; i < M; ++i) bfs_h[i] = (float)((double)P.bf_sinogram[i] + kBfOffset);
    HIP_TRY(j->bfs.alloc(M + g.N_t));
    HIP_TRY(launch_fill_f32(j->bfs.p + M, 0.f, g.N_t, s));
    HIP_TRY(hipMemcpyAsync(j->bfs.p, bfs_h.data(), M * sizeof(float), hipMemcpyHostToDevice, s));
    j->bytes += (int64_t)M * 4;
  }
  HIP_TRY(hipMemcpyAsync(j->y.p, P.bf_mode ? bf_y.data() : P.counts, M * sizeof(float),
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->cosv.p, cosv.data(), NT * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->sinv.p, sinv.data(), NT * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->Hd.p, j->H.data(), j->H.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->vlr_t0.p, vt0.data(), vt0.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->vlr_n.p, vn.data(), vn.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->vlr_val.p, vval.data(), vval.size() * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->mask.p, mask.data(), L, hipMemcpyHostToDevice, s));
  j->mask_h = mask;
  HIP_TRY(hipMemcpyAsync(j->I0.p, j->I0h.data(), NT * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->mu.p, j->muh.data(), NT * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(j->alpha.p, j->alphah.data(), NT * sizeof(double), hipMemcpyHostToDevice, s));
  // weights: identity noise model, w = 1 / alpha when the noise model is on (…_Extra.cpp:740-757)
  std::vector<float> wv(NT);
  for (int k = 0; k < NT; ++k) wv[k] = P.noise_model && !P.bf_mode ? (float)(1.0 / j->alphah[k]) : 1.0f;
  HIP_TRY(hipMemcpyAsync(j->wview.p, wv.data(), NT * sizeof(float), hipMemcpyHostToDevice, s));
  if (P.init_volume)
    HIP_TRY(hipMemcpyAsync(j->x.p, P.init_volume, V * sizeof(float), hipMemcpyHostToDevice, s));
  else
    HIP_TRY(launch_fill_f32(j->x.p, (float)P.default_recon, V, s));

  DevJob& J = j->J;
  J.N_theta = NT; J.N_r = g.N_r; J.N_t = g.N_t; J.N_x = g.N_x; J.N_y = g.N_y; J.N_z = g.N_z;
  {
    // line / N_x by a multiply-high: m = ceil(2^32 / N_x) = (2^32 + r) / N_x is exact
    // for every line l < N_z N_x when l r < 2^32 (the error l r / (N_x 2^32) then stays
    // below 1 / N_x)
    const uint64_t nx = (uint64_t)g.N_x, m = ((1ull << 32) + nx - 1) / nx;
    const uint64_t r = m * nx - (1ull << 32), lmax = (uint64_t)g.N_z * nx;
    J.nx_div = (m < (1ull << 32) && lmax * r < (1ull << 32)) ? (uint32_t)m : 0u;
  }
  J.NB = NB;
  J.nslots = NT * NB;
  J.vlr_max = 1;
  for (int c : vn) J.vlr_max = std::max(J.vlr_max, c);
  J.x0 = g.x0; J.z0 = g.z0; J.delta_xz = g.delta_xz; J.delta_r = g.delta_r;
  J.R0 = g.R0; J.RMax = g.RMax; J.offsetR = j->offsetR;
  J.inv_delta_r = 1.0 / g.delta_r; J.inv_offsetR = 1.0 / j->offsetR;
  J.cosv = j->cosv.p; J.sinv = j->sinv.p; J.H = j->Hd.p;
  J.vlr_t0 = j->vlr_t0.p; J.vlr_n = j->vlr_n.p; J.vlr_val = j->vlr_val.p;
  J.mask = j->mask.p; J.line_nnz = j->line_nnz.p;
  J.spare_b = (uint32_t)(M * 4);
  J.x = j->x.p; J.e = j->e.p; J.y = j->y.p; J.wfull = nullptr; J.wview = j->wview.p;
  J.I0 = j->I0.p; J.mu = j->mu.p;
  J.sched = j->sched.p; J.ustat = j->ustat.p;
  // the error-sinogram update where a stream's values are not kept in registers: a
  // plain read-modify-write on conflict-free steps (no two lines of a step share a
  // detector bin; 12 % faster per launch than fp32 atomics on a 141-view 512x64x256
  // job, twice as fast on steps of more units than compute units), fp32 atomics on the
  // relaxed schedule, whose lines may share bins
  J.e_atomic = P.schedule == 1 ? 1 : 0;
  J.magmap = nullptr;
  J.bfs = P.bf_sinogram ? j->bfs.p : nullptr;
  if (J.bfs && J.vlr_max > 2) return MBIR_ERR_UNSUPPORTED;  // scaled line update: <= 2 rows
  if (P.bf_mode) {
    J.wfull = j->wfull.p;
    J.e_atomic = 0;  // the selector refresh is a read-modify-write of its own
    set_bragg(j, P.bragg_T);
    if (P.bf_nhicd) {
      // NHICD: update magnitudes start at zero; one seeded random list of all lines
      // (BFReconstructionEngine.cpp:595-608)
      HIP_TRY(j->magmap.alloc(L));
      HIP_TRY(launch_fill_f32(j->magmap.p, 0.f, L, s));
      J.magmap = j->magmap.p;
      j->nh_all = rand_list((int)L, P.seed * 2654435761u + 17u);
      j->bytes += (int64_t)L * 4;
    }
  }
  set_prior(j, P.sigma_x);

  // columns: per-line nnz, zero lines with empty columns (:615-632), then the
  // compact column table of all lines (built once, FP32, with gains and weights folded in)
  HIP_TRY(launch_line_nnz(J, j->line_nnz.p, j->x.p, s));
  j->nnz_h.assign(L, 0);
  HIP_TRY(hipMemcpyAsync(j->nnz_h.data(), j->line_nnz.p, L * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::vector<int> cptr(L + 1, 0);
  int64_t tot = 0;
  for (size_t l = 0; l < L; ++l) {
    cptr[l] = (int)tot;
    tot += j->nnz_h[l];
  }
  if (tot >= 2147483647LL) return MBIR_ERR_UNSUPPORTED;
  cptr[L] = (int)tot;
  j->nnz_total = tot;
  {
    // LDS of the line update: the longest column padded to a multiple of 4 P entries,
    // P = 8 * 64 / TY streams for TY = min(64, the power of two >= the shortest unit)
    const int umin = g.N_y / j->ysplit;
    int ty = 1;
    while (ty < umin && ty < 64) ty <<= 1;
    const int q = 4 * kIcdWaves * (64 / ty);
    int nmax = 1;
    for (size_t l = 0; l < L; ++l) nmax = std::max(nmax, j->nnz_h[l]);
    J.col_slots = (nmax + q - 1) / q * q;
    if (J.col_slots > kMaxSlots) return MBIR_ERR_UNSUPPORTED;
  }
  j->cptr_h = cptr;
  HIP_TRY(j->col_ptr.alloc(L + 1));
  // the row table has kColPad zero entries behind it (the line kernels read a batch of
  // rows past a column's end without a bound per entry)
  HIP_TRY(j->col_row.alloc((size_t)tot + kColPad));
  HIP_TRY(launch_fill_u32(reinterpret_cast<unsigned int*>(j->col_row.p) + tot, 0u, kColPad, s));
  HIP_TRY(j->col_a.alloc((size_t)tot));
  HIP_TRY(j->col_ka.alloc((size_t)tot));
  HIP_TRY(j->col_kw.alloc((size_t)tot));
  j->bytes += tot * 16 + (int64_t)kColPad * 4 + (int64_t)(L + 1) * 4;
  HIP_TRY(hipMemcpyAsync(j->col_ptr.p, cptr.data(), (L + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  J.col_ptr = j->col_ptr.p;
  J.col_row = j->col_row.p;
  J.col_a = j->col_a.p;
  J.col_ka = j->col_ka.p;
  J.col_kw = j->col_kw.p;
  HIP_TRY(launch_build_columns(J, j->col_row.p, j->col_a.p, s));
  HIP_TRY(launch_column_weights(J, j->col_ka.p, j->col_kw.p, (int)tot, s));
  HIP_TRY(j->line_kk.alloc(L));
  j->bytes += (int64_t)L * 8;
  J.line_kk = j->line_kk.p;
  HIP_TRY(launch_line_kk(J, j->line_kk.p, s));
  // initial forward projection and error sinogram (:646-691)
  HIP_TRY(j->yest.alloc(M));
  HIP_TRY(launch_fill_f64(j->yest.p, 0.0, M, s));
  HIP_TRY(launch_forward_project(J, j->x.p, j->yest.p, s));
  HIP_TRY(launch_init_error(J, j->yest.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  j->yest.release();
  j->stop_threshold = P.stop_threshold;
  return MBIR_OK;
}

}  // namespace

namespace mbir {

// Step schedule of one sweep.  schedule 0: step (z, c) = lines {(z, c + S m)},
// S the conflict-free spacing, steps in a seeded random order.  schedule 1:
// rows_per_step rows spaced N_z / rows_per_step apart share a step.  Steps whose
// lines all have empty columns are dropped when drop_empty_steps is set (they
// update nothing).
void build_schedule(mbir_job* j, bool drop_empty_steps)
{
  const int Nz = j->g.N_z, Nx = j->g.N_x, S = j->spacing;
  const int R = j->prob.schedule == 1 ? std::max(1, std::min(j->prob.rows_per_step, Nz)) : 1;
  const int Zs = (Nz + R - 1) / R;
  std::vector<std::pair<int, int>> steps;
  for (int zb = 0; zb < Zs; ++zb)
    for (int c = 0; c < S; ++c) steps.push_back({zb, c});
  std::mt19937 rng(j->prob.seed + 7919u * (uint32_t)j->sweep_counter);
  std::shuffle(steps.begin(), steps.end(), rng);
  j->sched_h.clear();
  j->step_ptr.assign(1, 0);
  const int ns = j->ysplit, Ny = j->g.N_y;
  for (auto& st : steps) {
    for (int k = 0; k < R; ++k) {
      for (int sl = 0; sl < ns; ++sl) {
        // slab sl: slices [sl Ny / ns, (sl + 1) Ny / ns), the row class moved sl zshift
        // planes (a bijection of the classes, so each unit comes once per sweep)
        const int zc = st.first + k * Zs;
        if (zc >= Nz) continue;
        const int z = ns > 1 ? (zc + sl * j->zshift) % Nz : zc;
        const int ya = (int)((int64_t)sl * Ny / ns), yb = (int)((int64_t)(sl + 1) * Ny / ns);
        const int pad = ns > 1 ? (ya | ((yb - ya) << 16)) : 0;
        for (int x = st.second; x < Nx; x += S)
          if (j->nnz_h[(size_t)z * Nx + x] > 0) {  // (:245-249)
            const int l = z * Nx + x;
            j->sched_h.push_back(LineRef{l, j->cptr_h[l], j->nnz_h[l], pad});
          }
      }
    }
    if (drop_empty_steps && (int)j->sched_h.size() == j->step_ptr.back()) continue;
    j->step_ptr.push_back((int)j->sched_h.size());
  }
}

// Upload the sweep's schedule (from sched_src, default the host schedule) on stream s
// (the accumulators are formed from the units' statistics after the sweep's last step).
int prepare_sweep_on(mbir_job* j, hipStream_t s, const LineRef* sched_src)
{
  if (!sched_src) sched_src = j->sched_h.data();
  if (!j->sched_h.empty())
    HIP_TRY(hipMemcpyAsync(j->sched.p, sched_src, j->sched_h.size() * sizeof(LineRef),
                           hipMemcpyHostToDevice, s));
  return MBIR_OK;
}

// Stop test on a sweep's accumulators (…_UpdateVoxels.cpp:765-786); counts the sweep.
int sweep_status(mbir_job* j, const SweepAccum& a, double ms, mbir_sweep_stats* st)
{
  j->sweep_counter++;
  int status = 1;
  if (a.sum_abs_value > 0 && (a.sum_abs_update / a.sum_abs_value) < j->stop_threshold) status = 0;
  if (st) {
    st->sum_abs_update = a.sum_abs_update;
    st->sum_abs_value = a.sum_abs_value;
    st->zero_skipped = (int64_t)a.zero_skipped;
    st->updated = (int64_t)a.updated;
    st->nnz_updated = (int64_t)a.nnz_updated;
    st->kernel_ms = ms;
    st->steps = (int)j->step_ptr.size() - 1;
    st->status = status;
  }
  return status;
}

// QGGMRF constants of sigma_x (HAADF/HAADF_QGGMRFPriorModel.cpp:10-19 at setup,
// updatePriorModel :24-31 after a re-estimation), MRF_P = 2, MRF_ALPHA = 1.5.
void set_prior(mbir_job* j, double sigma_x)
{
  DevJob& J = j->J;
  const mbir_problem& P = j->prob;
  j->sigma_x = sigma_x;
  const double sp = std::pow(sigma_x, 2.0), spq = std::pow(sigma_x, 2.0 - P.mrf_q);
  J.mrf_q = (float)P.mrf_q; J.mrf_c = (float)P.mrf_c; J.sigma_p = (float)sp;
  J.sigma_pq = (float)spq; J.alpha = 1.5f; J.b0 = (float)(2.0 / (sp * P.mrf_c));
  J.pq2 = (float)(2.0 - P.mrf_q); J.inv_sigma_p = (float)(1.0 / sp); J.inv_sigma_pq = (float)(1.0 / spq);
  J.mrf_qd = P.mrf_q; J.mrf_cd = P.mrf_c; J.sigma_pd = sp; J.sigma_pqd = spq;
}

void set_bragg(mbir_job* j, double T)
{
  j->bragg = T;
  j->J.bragg = (float)T;
  j->J.bragg_dT = (float)(j->prob.bragg_delta * T);
  j->J.bragg_d = T;
  j->J.delta_d = j->prob.bragg_delta;
}

// jointEstimation then updateWeights (HAADF/HAADF_ReconstructionEngine.cpp:768-786)
// and the refresh of the gain- and weight-folded column values, enqueued on s.
int nuisance_update_on(mbir_job* j, hipStream_t s)
{
  const int NT = j->g.N_theta;
  if (j->J.wfull) {
    // bright field (BrightField/BFReconstructionEngine.cpp:841-870): the offset, then
    // the shared variance, each followed by the selector refresh
    if (j->prob.joint_estimation) HIP_TRY(launch_bf_offset(j->J, j->sums6.p, j->mu.p, s));
    if (j->prob.noise_model) {
      HIP_TRY(launch_bf_noise(j->J, j->alpha.p, j->alpha_b.p, j->cost.p + 4, s));
      std::swap(j->alpha.p, j->alpha_b.p);
    }
    return MBIR_OK;
  }
  // with a bright-field sinogram only the offsets are estimated: per view the weighted
  // mean of e moves from e into mu (HAADF/HAADF_ReconstructionEngine_Extra.cpp:636-669)
  if (j->prob.joint_estimation && j->J.bfs)
    HIP_TRY(launch_offset_only(j->J, j->sums6.p, j->mu.p, s));
  if (j->prob.joint_estimation && !j->J.bfs) {
    HIP_TRY(launch_joint_sums(j->J, j->sums6.p, s));
    HIP_TRY(launch_joint_solve(NT, j->sums6.p, j->prob.target_gain, j->I0.p, j->mu.p, j->I0b.p,
                               j->mub.p, s));
    HIP_TRY(launch_joint_apply(j->J, j->I0.p, j->mu.p, j->I0b.p, j->mub.p, s));
    std::swap(j->I0.p, j->I0b.p);
    std::swap(j->mu.p, j->mub.p);
    j->J.I0 = j->I0.p;
    j->J.mu = j->mu.p;
  }
  if (j->prob.noise_model) HIP_TRY(launch_noise_update(j->J, j->alpha.p, j->wview.p, s));
  HIP_TRY(launch_column_weights(j->J, j->col_ka.p, j->col_kw.p, (int)j->nnz_total, s));
  HIP_TRY(launch_line_kk(j->J, j->line_kk.p, s));
  return MBIR_OK;
}

}  // namespace mbir

namespace {

// Schedule of the job's next sweep, uploaded on the job stream.
int prepare_sweep(mbir_job* j, bool drop_empty_steps = false)
{
  build_schedule(j, drop_empty_steps);
  int rc = prepare_sweep_on(j, j->stream, nullptr);
  return rc < 0 ? rc : MBIR_OK;
}

// Read the accumulators and apply the stop test.
int finish_sweep(mbir_job* j, double ms, mbir_sweep_stats* st)
{
  SweepAccum a{};
  HIP_TRY(hipMemcpyAsync(&a, j->acc.p, sizeof(a), hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  return sweep_status(j, a, ms, st);
}

int run_sweep_steps(mbir_job* j, int outer, int inner, mbir_sweep_stats* st)
{
  hipStream_t s = j->stream;
  int rc = prepare_sweep(j);
  if (rc < 0) return rc;
  const int nsteps = (int)j->step_ptr.size() - 1;
  // timing: one event pair around the sweep's back-to-back launches on the job stream
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (j->timing) {
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
  }
  for (int k = 0; k < nsteps; ++k) {
    const int b = j->step_ptr[k], n = j->step_ptr[k + 1] - b;
    HIP_TRY(launch_icd_step(j->J, j->sched.p + b, n, b, outer, inner, j->prob.zero_skipping, s,
                            j->sched_h.data() + b));
  }
  if (j->timing) HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(launch_sweep_reduce(j->sched.p, j->ustat.p, (int)j->sched_h.size(), j->acc.p, j->red.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  if (j->timing) {
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return finish_sweep(j, ms, st);
}

// ---- bright-field NHICD (BrightField/BFReconstructionEngine.cpp:641-884,
// BFReconstructionEngine_Extra.cpp:476-1273) ----

// 5 x 5 Hamming window of the update-magnitude filter (BFReconstructionEngine.cpp:157-181)
constexpr double kNhHamming[5][5] = {{0.0013, 0.0086, 0.0159, 0.0086, 0.0013},
                                     {0.0086, 0.0581, 0.1076, 0.0581, 0.0086},
                                     {0.0159, 0.1076, 0.1993, 0.1076, 0.0159},
                                     {0.0086, 0.0581, 0.1076, 0.0581, 0.0086},
                                     {0.0013, 0.0086, 0.0159, 0.0086, 0.0013}};
constexpr int kNhParts = 20;  // k_NumHomogeniousIter = k_NumNonHomogeniousIter

// ComputeVSC (Extra.cpp:888-925): the magnitude map filtered with the window, zero
// outside the volume.  Cells with every tap inside take a loop without bound tests over
// the same 25 terms in the same order (bit for bit the bounded sum); large maps split
// their rows over a few host threads.
void nhicd_filter(const float* mag, int Nz, int Nx, double* out)
{
  auto cell = [&](int i, int j) {
    double f = 0;
    for (int p = -2; p <= 2; ++p)
      for (int q = -2; q <= 2; ++q)
        if (i + p >= 0 && i + p < Nz && j + q >= 0 && j + q < Nx)
          f += kNhHamming[p + 2][q + 2] * (double)mag[(size_t)(i + p) * Nx + (j + q)];
    return f;
  };
  auto rows = [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      double* o = out + (size_t)i * Nx;
      if (i < 2 || i + 2 >= Nz || Nx < 5) {
        for (int j = 0; j < Nx; ++j) o[j] = cell(i, j);
        continue;
      }
      for (int j = 0; j < 2; ++j) o[j] = cell(i, j);
      for (int j = Nx - 2; j < Nx; ++j) o[j] = cell(i, j);
      const float* r[5];
      for (int p = 0; p < 5; ++p) r[p] = mag + (size_t)(i - 2 + p) * Nx;
      for (int j = 2; j < Nx - 2; ++j) {
        double f = 0;
        for (int p = 0; p < 5; ++p)
          for (int q = 0; q < 5; ++q) f += kNhHamming[p][q] * (double)r[p][j - 2 + q];
        o[j] = f;
      }
    }
  };
  const int nt = (int64_t)Nz * Nx >= (1 << 16)
                     ? std::max(1, std::min(8, (int)std::thread::hardware_concurrency()))
                     : 1;
  if (nt == 1) {
    rows(0, Nz);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back(rows, (int)((int64_t)Nz * t / nt), (int)((int64_t)Nz * (t + 1) / nt));
  for (auto& x : th) x.join();
}

// SetNonHomThreshold (Extra.cpp:934-987): the (n - n/20)-th smallest filtered value.
// A histogram of the top 16 bits of an order-preserving key (the bit pattern, sign
// folded) finds the bucket holding that rank; nth_element over that bucket alone then
// returns the value nth_element over all values would.
double nhicd_threshold(std::vector<double> v)
{
  const size_t n = v.size();
  const size_t k = n - n / kNhParts;
  if (k >= n) return v.empty() ? 0.0 : *std::max_element(v.begin(), v.end());
  auto key = [](double d) {
    uint64_t u;
    std::memcpy(&u, &d, sizeof u);
    return (u >> 63) ? ~u : (u | (1ull << 63));
  };
  std::vector<uint32_t> hist(1u << 16, 0);
  for (double d : v) hist[key(d) >> 48]++;
  size_t below = 0;
  uint32_t b = 0;
  while (below + hist[b] <= k) below += hist[b++];
  std::vector<double> cand;
  cand.reserve(hist[b]);
  for (double d : v)
    if ((key(d) >> 48) == b) cand.push_back(d);
  std::nth_element(cand.begin(), cand.begin() + (long)(k - below), cand.end());
  return cand[k - below];
}

// Conflict-free steps of an NHICD list: chains in increasing x, each line at least
// reach[d] after the last line of its chain, d rows away (then at least as far from every
// earlier one), the first chain that takes it.  Lines of one row need at least S between
// them, as in a sweep's classes (z, x mod S), but lines of other rows far enough along x
// share the step too, so a call of a twentieth of the lines takes a few hundred steps
// where the classes would take most of a sweep's.  The lines are visited by x, then row,
// through a map of the list (each line at most once in a list); lines with empty columns
// are dropped.
std::vector<std::vector<int>> pack_chains(const mbir_job* j, const std::vector<int>& lines)
{
  const int Nx = j->g.N_x, Nz = j->g.N_z, off = Nz - 1;
  std::vector<uint8_t> in((size_t)Nx * Nz, 0);
  for (int l : lines)
    if (j->nnz_h[l] > 0) in[(size_t)l] = 1;
  std::vector<std::vector<int>> steps;
  std::vector<int> tx, tz;  // the last line of each chain
  for (int x = 0; x < Nx; ++x)
    for (int z = 0; z < Nz; ++z) {
      const int l = z * Nx + x;
      if (!in[(size_t)l]) continue;
      size_t k = 0;
      while (k < tx.size() && x - tx[k] < j->nh_reach[(size_t)(z - tz[k] + off)]) ++k;
      if (k == tx.size()) {
        steps.emplace_back();
        tx.push_back(x);
        tz.push_back(z);
      }
      steps[k].push_back(l);
      tx[k] = x;
      tz[k] = z;
    }
  return steps;
}

// The steps in a seeded random order as the sweep's schedule.
void schedule_steps(mbir_job* j, const std::vector<std::vector<int>>& steps)
{
  std::vector<int> order(steps.size());
  std::iota(order.begin(), order.end(), 0);
  std::mt19937 rng(j->prob.seed + 7919u * (uint32_t)j->sweep_counter);
  std::shuffle(order.begin(), order.end(), rng);
  j->sched_h.clear();
  j->step_ptr.assign(1, 0);
  for (int c : order) {
    for (int l : steps[c]) j->sched_h.push_back(LineRef{l, j->cptr_h[l], j->nnz_h[l], 0});
    j->step_ptr.push_back((int)j->sched_h.size());
  }
}

// One NHICD call: the given conflict-free steps in a seeded order, per-step launches.
// while_running (may be empty) is host work done after the launches are queued, while
// the device runs them; it must not touch the call's steps.
int run_list_steps(mbir_job* j, const std::vector<std::vector<int>>& steps, int outer, int inner,
                   const std::function<void()>& while_running)
{
  hipStream_t s = j->stream;
  schedule_steps(j, steps);
  int rc = prepare_sweep_on(j, s, nullptr);
  if (rc < 0) return rc;
  const int nsteps = (int)j->step_ptr.size() - 1;
  for (int k = 0; k < nsteps; ++k) {
    const int b = j->step_ptr[k], n = j->step_ptr[k + 1] - b;
    HIP_TRY(launch_icd_step(j->J, j->sched.p + b, n, b, outer, inner, j->prob.zero_skipping, s,
                            j->sched_h.data() + b));
  }
  HIP_TRY(launch_sweep_reduce(j->sched.p, j->ustat.p, (int)j->sched_h.size(), j->acc.p, j->red.p, s));
  if (while_running) while_running();
  return finish_sweep(j, 0.0, nullptr);
}

double roi_volume_sum(mbir_job* j)
{
  double* d = j->cost.p + 3;  // the cost buffer's spare fourth double
  if (launch_fill_f64(d, 0.0, 1, j->stream) != hipSuccess) return -1;
  if (launch_roi_sum(j->J, d, j->stream) != hipSuccess) return -1;
  double v = 0;
  if (hipMemcpyAsync(&v, d, sizeof(double), hipMemcpyDeviceToHost, j->stream) != hipSuccess ||
      hipStreamSynchronize(j->stream) != hipSuccess)
    return -1;
  return v;
}

int fetch_magmap(mbir_job* j)
{
  const size_t L = (size_t)j->g.N_z * j->g.N_x;
  j->magmap_h.resize(L);
  HIP_TRY(hipMemcpyAsync(j->magmap_h.data(), j->magmap.p, L * sizeof(float),
                         hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  return MBIR_OK;
}

// kernel_mode 3: the step launches of a sweep replayed from a HIP graph.  The graph
// is captured once per job (again when the job descriptor or the schedule's shape
// changes); a sweep uploads its permuted schedule into the padded slots, replays the
// graph and sums the units' statistics by padded slot.  The first sweep (zero skipping
// off) runs with direct launches.
int run_sweep_graph(mbir_job* j, int outer, int inner, mbir_sweep_stats* st)
{
  if (outer == 0 && inner == 0) return run_sweep_steps(j, outer, inner, st);
  hipStream_t s = j->stream;
  build_schedule(j, false);
  const int nsteps = (int)j->step_ptr.size() - 1;
  int width = 1;
  for (int k = 0; k < nsteps; ++k) width = std::max(width, j->step_ptr[k + 1] - j->step_ptr[k]);
  const bool stale0 = !j->graph_exec || nsteps != j->graph_steps || width != j->graph_lines;
  const size_t slots = (size_t)nsteps * width;
  if (j->gsched.n < slots) {
    j->gsched.release();
    HIP_TRY(j->gsched.alloc(slots));
  }
  if (j->ustat.n < slots) {  // statistics by padded slot (the descriptor changes: recapture)
    j->ustat.release();
    HIP_TRY(j->ustat.alloc(slots));
    j->J.ustat = j->ustat.p;
  }
  const bool stale = stale0 || std::memcmp(&j->graph_J, &j->J, sizeof(DevJob)) != 0;
  if (stale) {
    if (j->graph_exec) (void)hipGraphExecDestroy(j->graph_exec);
    if (j->graph) (void)hipGraphDestroy(j->graph);
    j->graph_exec = nullptr;
    j->graph = nullptr;
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int k = 0; k < nsteps && le == hipSuccess; ++k)
      le = launch_icd_step(j->J, j->gsched.p + (size_t)k * width, width, k * width, 1, 1,
                           j->prob.zero_skipping, s);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    HIP_TRY(le);
    HIP_TRY(ce);
    j->graph = g;
    HIP_TRY(hipGraphInstantiate(&j->graph_exec, g, nullptr, nullptr, 0));
    j->graph_J = j->J;
    j->graph_steps = nsteps;
    j->graph_lines = width;
  }
  j->gsched_h.assign(slots, LineRef{-1, 0, 0, 0});
  for (int k = 0; k < nsteps; ++k)
    for (int i = j->step_ptr[k]; i < j->step_ptr[k + 1]; ++i)
      j->gsched_h[(size_t)k * width + (i - j->step_ptr[k])] = j->sched_h[i];
  HIP_TRY(hipMemcpyAsync(j->gsched.p, j->gsched_h.data(), slots * sizeof(LineRef),
                         hipMemcpyHostToDevice, s));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (j->timing) {
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
  }
  HIP_TRY(hipGraphLaunch(j->graph_exec, s));
  if (j->timing) HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(launch_sweep_reduce(j->gsched.p, j->ustat.p, (int)slots, j->acc.p, j->red.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  if (j->timing) {
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return finish_sweep(j, ms, st);
}

// Batched sweep: launch k runs step k of every job that has one (each job's
// steps are its own conflict-free groups; jobs never share memory), so one
// launch carries the lines of many jobs and the per-launch cost is shared.
int run_batch(mbir_ctx* ctx, const std::vector<mbir_job*>& jobs, const int* outer,
              const int* inner, mbir_sweep_stats* stats)
{
  const int n = (int)jobs.size();
  if (n == 0) return MBIR_OK;
  std::vector<DevJob> dj(n);
  int max_steps = 0;
  for (int i = 0; i < n; ++i) {
    mbir_job* j = jobs[i];
    if (j->ctx != ctx) return MBIR_ERR_INPUT;
    if (j->cancel.load()) return MBIR_ERR_CANCELLED;
    // bright field: one detector row per slice, full sweeps (NHICD lists run alone);
    // HAADF with a bright-field sinogram: single-job paths only
    if (j->J.unit_max > 128 || j->J.vlr_max > 2 || j->J.bfs ||
        (j->J.wfull && (j->prob.bf_nhicd || j->J.vlr_max > 1)))
      return MBIR_ERR_UNSUPPORTED;
    int rc = prepare_sweep(j);
    if (rc < 0) return rc;
    dj[i] = j->J;
    dj[i].e_atomic = j->prob.schedule == 1 ? 1 : 0;
    max_steps = std::max(max_steps, (int)j->step_ptr.size() - 1);
  }
  // jobs ordered by kernel shape (bright-field jobs apart, one detector row per
  // slice) so that each step's entries form one run per shape
  auto shape_of = [&](const mbir_job* j, int* y, int* v, int* f) {
    icd_shape(j->J, j->J.vlr_max, y, v);
    *f = j->J.wfull ? 1 : 0;
    if (*f) *v = 1;
  };
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    int ya, na, fa, yb, nb, fb;
    shape_of(jobs[a], &ya, &na, &fa);
    shape_of(jobs[b], &yb, &nb, &fb);
    return fa != fb ? fa < fb : (ya != yb ? ya < yb : na < nb);
  });
  std::vector<BatchEntry> ent;
  std::vector<int> l_e0(max_steps), l_ne(max_steps), l_wg(max_steps);
  for (int k = 0; k < max_steps; ++k) {
    l_e0[k] = (int)ent.size();
    int wg = 0, cur_ypl = -1, cur_nv = -1, cur_fw = -1;
    for (int i : order) {
      mbir_job* j = jobs[i];
      if (k + 1 >= (int)j->step_ptr.size()) continue;
      const int b = j->step_ptr[k], cnt = j->step_ptr[k + 1] - b;
      if (cnt <= 0) continue;
      BatchEntry e{};
      e.lines = j->sched.p + b;
      e.nlines = cnt;
      e.job = i;
      e.wg_off = wg;
      const int o = outer ? outer[i] : 0, in = inner ? inner[i] : 0;
      e.flags = ((o == 0 && in == 0) ? 1 : 0) | (j->prob.zero_skipping ? 2 : 0);
      shape_of(j, &e.ypl, &e.nv, &e.fw);
      if (e.ypl != cur_ypl || e.nv != cur_nv || e.fw != cur_fw) {  // a new launch: offsets restart
        wg = 0;
        cur_ypl = e.ypl;
        cur_nv = e.nv;
        cur_fw = e.fw;
      }
      e.wg_off = wg;
      e.pos = b;
      ent.push_back(e);
      wg += cnt;
    }
    l_ne[k] = (int)ent.size() - l_e0[k];
    l_wg[k] = wg;
  }
  if (ctx->djobs.n < (size_t)n) {
    ctx->djobs.release();
    HIP_TRY(ctx->djobs.alloc(n));
  }
  if (ctx->ents.n < ent.size()) {
    ctx->ents.release();
    HIP_TRY(ctx->ents.alloc(ent.size()));
  }
  hipStream_t s = ctx->stream;
  for (mbir_job* j : jobs) HIP_TRY(hipStreamSynchronize(j->stream));  // schedules uploaded
  HIP_TRY(hipMemcpyAsync(ctx->djobs.p, dj.data(), n * sizeof(DevJob), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ctx->ents.p, ent.data(), ent.size() * sizeof(BatchEntry),
                         hipMemcpyHostToDevice, s));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timing = jobs[0]->timing != 0;
  if (timing) {
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
  }
  // each step: one launch per kernel shape present (entries of a step are grouped by shape)
  for (int k = 0; k < max_steps; ++k) {
    int e = l_e0[k];
    const int eend = l_e0[k] + l_ne[k];
    while (e < eend) {
      int f = e, wg = 0;
      int slots = 4;
      while (f < eend && ent[f].ypl == ent[e].ypl && ent[f].nv == ent[e].nv &&
             ent[f].fw == ent[e].fw) {
        slots = std::max(slots, jobs[ent[f].job]->J.col_slots);
        wg += ent[f++].nlines;
      }
      int pf = ctx->tune.choose(ent[e].ypl, ent[e].nv, wg);
      mbir_ctx::Probe* pr = ctx->probe;
      hipEvent_t a = nullptr, b = nullptr;
      if (pr) {
        pf = pr->force;
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, s));
      }
      HIP_TRY(launch_icd_batch_step(ent[e].ypl, ent[e].nv, ctx->djobs.p, ctx->ents.p, e, f - e,
                                    wg, slots, pf, s, ent[e].fw));
      if (pr) {
        HIP_TRY(hipEventRecord(b, s));
        pr->ev.push_back(a);
        pr->ev.push_back(b);
        pr->ypl.push_back(ent[e].ypl);
        pr->nv.push_back(ent[e].nv);
        pr->nwg.push_back(wg);
      }
      e = f;
    }
  }
  if (timing) HIP_TRY(hipEventRecord(e1, s));
  for (mbir_job* j : jobs)
    HIP_TRY(launch_sweep_reduce(j->sched.p, j->ustat.p, (int)j->sched_h.size(), j->acc.p, j->red.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  if (timing) {
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (ctx->verbose)
    std::fprintf(stderr, "mbir: batch of %d jobs, %d launches, %zu entries\n", n, max_steps,
                 ent.size());
  for (int i = 0; i < n; ++i) {
    int st = finish_sweep(jobs[i], ms, stats ? stats + i : nullptr);
    if (st < 0) return st;
    if (stats) stats[i].steps = max_steps;
  }
  return MBIR_OK;
}

int run_sweep(mbir_job* j, int outer, int inner, mbir_sweep_stats* st)
{
  if (j->cancel.load()) return MBIR_ERR_CANCELLED;
  // a whole sweep in one persistent launch is not offered: its first form did not
  // finish on the device (its step waits never saw the completion counts) and was
  // taken out; kernel_mode 2 is refused
  if (j->prob.kernel_mode == 2) return MBIR_ERR_UNSUPPORTED;
  if (j->prob.kernel_mode == 3) return run_sweep_graph(j, outer, inner, st);
  if (j->prob.kernel_mode == 1) {
    mbir_sweep_stats tmp{};
    std::vector<mbir_job*> one{j};
    int rc = run_batch(j->ctx, one, &outer, &inner, &tmp);
    if (rc < 0) return rc;
    if (st) *st = tmp;
    return tmp.status;
  }
  return run_sweep_steps(j, outer, inner, st);
}

int nuisance_update(mbir_job* j)
{
  int rc = nuisance_update_on(j, j->stream);
  if (rc < 0) return rc;
  HIP_TRY(hipStreamSynchronize(j->stream));
  return MBIR_OK;
}

int job_cost(mbir_job* j, double* total, double* terms3)
{
  hipStream_t s = j->stream;
  HIP_TRY(launch_cost(j->J, j->alpha.p, j->prob.noise_model, j->cost.p, s));
  double out[3];
  HIP_TRY(hipMemcpyAsync(out, j->cost.p, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (terms3) std::memcpy(terms3, out, sizeof(out));
  if (total) *total = out[0] + out[1] + (j->prob.noise_model ? out[2] : 0.0);
  return MBIR_OK;
}

// estimateSigmaX (HAADF/HAADF_ReconstructionEngine.cpp:1827-1945): the forward-pair prior
// sum of the planes [(int16) 0.25 N_z, (int16) 0.75 N_z), differences scaled by the current
// sigma_x, over N_x N_y (zEnd - zStart + 1) entries; then updatePriorModel
// (HAADF/HAADF_QGGMRFPriorModel.cpp:24-31): sigma_x = gamma (estimate q)^(1/q) with
// gamma = k_QGGMRF_Gamma = 5 (Reconstruction/ReconstructionConstants.h:78), and the
// prior constants of the new sigma_x.
int estimate_prior(mbir_job* j, double* estimate, double* sigma_x)
{
  hipStream_t s = j->stream;
  const int z0 = (int)(int16_t)(0.25 * j->g.N_z), z1 = (int)(int16_t)(0.75 * j->g.N_z);
  double sum = 0.0;
  if (z1 > z0) {
    HIP_TRY(launch_sigma_sum(j->J, z0, z1, j->sigma_x, j->cost.p + 4, j->cost.p + 3, s));
    HIP_TRY(hipMemcpyAsync(&sum, j->cost.p + 3, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  const double entries = (double)j->g.N_x * j->g.N_y * (double)(z1 - z0 + 1);
  const double est = sum / entries;
  const double q = j->prob.mrf_q;
  const double sx = 5.0 * std::pow(est * q, 1.0 / q);
  if (estimate) *estimate = est;
  if (!(sx > 0) || !std::isfinite(sx)) return MBIR_ERR_INPUT;  // a zero volume has no scale
  set_prior(j, sx);
  if (sigma_x) *sigma_x = sx;
  return MBIR_OK;
}

}  // namespace

extern "C" int32_t mbir_job_create(mbir_ctx* ctx, const mbir_problem* prob, mbir_job** out)
{
  if (!ctx || !prob || !out) return MBIR_ERR_INPUT;
  *out = nullptr;
  const mbir_geometry& g = prob->geom;
  if (g.N_theta <= 0 || g.N_r <= 0 || g.N_t <= 0 || g.N_x <= 0 || g.N_y <= 0 || g.N_z <= 0 ||
      !prob->angles_deg || !prob->counts || !(g.delta_r > 0) || !(g.delta_t > 0) ||
      !(g.delta_xz > 0) || !(g.delta_xy > 0) || !(prob->sigma_x > 0) || !(prob->mrf_c > 0) ||
      !(prob->mrf_q >= 1.0 && prob->mrf_q <= 2.0))
    return MBIR_ERR_INPUT;
  if (hipSetDevice(ctx->device) != hipSuccess) return MBIR_ERR_HIP;
  auto* j = new mbir_job();
  j->ctx = ctx;
  j->prob = *prob;
  j->g = g;
  j->prob.angles_deg = nullptr;  // host pointers are not kept
  j->prob.counts = nullptr;
  j->prob.init_volume = nullptr;
  j->prob.init_gain = j->prob.init_offset = j->prob.init_variance = nullptr;
  j->prob.bf_sinogram = nullptr;
  if (hipStreamCreate(&j->stream) != hipSuccess) {
    delete j;
    return MBIR_ERR_HIP;
  }
  mbir_problem tmp = *prob;
  std::swap(j->prob, tmp);
  int rc = setup_job(j);
  std::swap(j->prob, tmp);
  if (rc != MBIR_OK) {
    delete j;
    return rc;
  }
  *out = j;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_destroy(mbir_job* job)
{
  if (!job) return MBIR_ERR_INPUT;
  (void)hipStreamSynchronize(job->stream);
  delete job;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_sweep(mbir_job* job, int32_t outer, int32_t inner, mbir_sweep_stats* st)
{
  if (!job) return MBIR_ERR_INPUT;
  return run_sweep(job, outer, inner, st);
}

extern "C" int32_t mbir_batch_sweep(mbir_ctx* ctx, mbir_job* const* jobs, int32_t n,
                                    const int32_t* outer, const int32_t* inner,
                                    mbir_sweep_stats* stats)
{
  if (!ctx || !jobs || n < 0) return MBIR_ERR_INPUT;
  if (hipSetDevice(ctx->device) != hipSuccess) return MBIR_ERR_HIP;
  std::vector<mbir_job*> all(jobs, jobs + n);
  return run_batch(ctx, all, outer, inner, stats);
}

namespace {

void tuning_out(const LineTuning& t, mbir_tuning* out)
{
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < kTuneBuckets; ++k) out->prefetch[a][b][k] = t.pf[a][b][k];
}

}  // namespace

extern "C" int32_t mbir_nhicd_filter(const float* mag, int32_t N_z, int32_t N_x, double* out)
{
  if (!mag || !out || N_z <= 0 || N_x <= 0) return MBIR_ERR_INPUT;
  nhicd_filter(mag, N_z, N_x, out);
  return MBIR_OK;
}

extern "C" int32_t mbir_nhicd_threshold(const double* v, int32_t n, double* thr)
{
  if (!v || !thr || n <= 0) return MBIR_ERR_INPUT;
  *thr = nhicd_threshold(std::vector<double>(v, v + n));
  return MBIR_OK;
}

extern "C" int32_t mbir_get_tuning(mbir_ctx* ctx, mbir_tuning* out)
{
  if (!ctx || !out) return MBIR_ERR_INPUT;
  std::memset(out, 0, sizeof(*out));
  tuning_out(ctx->tune, out);
  return MBIR_OK;
}

extern "C" int32_t mbir_set_tuning(mbir_ctx* ctx, const mbir_tuning* t)
{
  if (!ctx || !t) return MBIR_ERR_INPUT;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < kTuneBuckets; ++k) ctx->tune.pf[a][b][k] = t->prefetch[a][b][k] ? 1 : 0;
  return MBIR_OK;
}

extern "C" int32_t mbir_autotune(mbir_ctx* ctx, mbir_job* const* jobs, int32_t n,
                                 const int32_t* outer, const int32_t* inner, int32_t sweeps,
                                 mbir_tuning* out)
{
  if (!ctx || !jobs || n <= 0 || sweeps <= 0) return MBIR_ERR_INPUT;
  if (hipSetDevice(ctx->device) != hipSuccess) return MBIR_ERR_HIP;
  std::vector<mbir_job*> all(jobs, jobs + n);
  double ms[2][2][kTuneBuckets][2] = {};
  double wgs[2][2][kTuneBuckets][2] = {};
  int cnt[2][2][kTuneBuckets][2] = {};
  int rc = MBIR_OK;
  for (int form = 1; form >= 0 && rc >= 0; --form) {
    mbir_ctx::Probe pr;
    pr.force = form;
    ctx->probe = &pr;
    for (int sw = 0; sw < sweeps && rc >= 0; ++sw) rc = run_batch(ctx, all, outer, inner, nullptr);
    ctx->probe = nullptr;
    for (size_t i = 0; i < pr.nwg.size(); ++i) {
      float t = 0;
      if (rc >= 0 && hipEventElapsedTime(&t, pr.ev[2 * i], pr.ev[2 * i + 1]) == hipSuccess) {
        const int a = pr.ypl[i] > 1 ? 1 : 0, b = pr.nv[i] > 1 ? 1 : 0, k = tune_bucket(pr.nwg[i]);
        ms[a][b][k][form] += t;
        wgs[a][b][k][form] += pr.nwg[i];
        cnt[a][b][k][form]++;
      }
    }
    for (hipEvent_t e : pr.ev) (void)hipEventDestroy(e);
  }
  if (rc < 0) return rc;
  LineTuning t = ctx->tune;
  if (out) std::memset(out, 0, sizeof(*out));
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < kTuneBuckets; ++k) {
        double us[2] = {0, 0};
        for (int f = 0; f < 2; ++f)
          if (cnt[a][b][k][f] > 0) us[f] = 1e3 * ms[a][b][k][f] / wgs[a][b][k][f];
        if (cnt[a][b][k][0] > 0 && cnt[a][b][k][1] > 0) t.pf[a][b][k] = us[1] <= us[0] ? 1 : 0;
        if (out)
          for (int f = 0; f < 2; ++f) {
            out->us_per_wg[a][b][k][f] = us[f];
            out->launches[a][b][k][f] = cnt[a][b][k][f];
          }
      }
  ctx->tune = t;
  if (out) tuning_out(t, out);
  return MBIR_OK;
}

extern "C" int32_t mbir_job_nuisance_update(mbir_job* job)
{
  if (!job) return MBIR_ERR_INPUT;
  if (job->cancel.load()) return MBIR_ERR_CANCELLED;
  return nuisance_update(job);
}

extern "C" int32_t mbir_job_cost(mbir_job* job, double* total, double* terms3)
{
  if (!job) return MBIR_ERR_INPUT;
  return job_cost(job, total, terms3);
}

extern "C" int32_t mbir_job_estimate_prior(mbir_job* job, double* estimate, double* sigma_x)
{
  if (!job) return MBIR_ERR_INPUT;
  if (job->J.wfull) return MBIR_ERR_UNSUPPORTED;  // the HAADF engine's step only
  if (job->cancel.load()) return MBIR_ERR_CANCELLED;
  return estimate_prior(job, estimate, sigma_x);
}

extern "C" int32_t mbir_job_get_schedule(mbir_job* job, int32_t* step_ptr, int32_t cap_steps,
                                         int32_t* units, int64_t cap_units)
{
  if (!job) return MBIR_ERR_INPUT;
  const int32_t steps = job->step_ptr.empty() ? 0 : (int32_t)job->step_ptr.size() - 1;
  if (step_ptr)
    for (int32_t k = 0; k <= steps && k <= cap_steps; ++k) step_ptr[k] = job->step_ptr[k];
  if (units) {
    const int64_t n = std::min<int64_t>((int64_t)job->sched_h.size(), cap_units);
    for (int64_t i = 0; i < n; ++i) {
      const LineRef& r = job->sched_h[i];
      const int ny = r.pad >> 16;
      units[3 * i] = r.line;
      units[3 * i + 1] = ny ? (r.pad & 0xffff) : 0;
      units[3 * i + 2] = ny ? ny : job->g.N_y;
    }
  }
  return steps;
}

extern "C" int32_t mbir_job_set_nhicd_log(mbir_job* job, int32_t on)
{
  if (!job) return MBIR_ERR_INPUT;
  job->nh_log_on = on != 0;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_get_nhicd_log(mbir_job* job, int64_t* call_ptr, int32_t cap_calls,
                                          int32_t* lines, int64_t cap_lines)
{
  if (!job) return MBIR_ERR_INPUT;
  const int32_t calls = job->nh_log_ptr.empty() ? 0 : (int32_t)job->nh_log_ptr.size() - 1;
  if (call_ptr)
    for (int32_t c = 0; c <= calls && c <= cap_calls; ++c) call_ptr[c] = job->nh_log_ptr[c];
  if (lines) {
    const int64_t n = std::min<int64_t>((int64_t)job->nh_log.size(), cap_lines);
    for (int64_t i = 0; i < n; ++i) lines[i] = job->nh_log[i];
  }
  return calls;
}

extern "C" int32_t mbir_job_get_sigma_x(mbir_job* job, double* sigma_x)
{
  if (!job || !sigma_x) return MBIR_ERR_INPUT;
  *sigma_x = job->sigma_x;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_run(mbir_job* j, mbir_run_stats* rs, double* costs, int32_t cap,
                                int32_t* n_costs)
{
  if (!j) return MBIR_ERR_INPUT;
  auto t0 = std::chrono::steady_clock::now();
  int nc = 0, ns = 0, status = 1, outer_done = 0;
  auto push_cost = [&]() -> int {
    if (!costs || nc >= cap) return 0;
    double c = 0;
    int rc = job_cost(j, &c, nullptr);
    if (rc < 0) return rc;
    costs[nc++] = c;
    return 0;
  };
  int rc = push_cost();
  if (rc < 0) return rc;
  int num_iter = j->prob.num_inner;
  if (j->J.wfull) {
    // bright field: BFReconstructionEngine::execute's loop (BrightField/
    // BFReconstructionEngine.cpp:575-872) with full random-order sweeps: the
    // threshold 5e10 for the first outer iteration's inner sweeps when NumIter > 1,
    // T from its last inner sweep on; the stop test ends the run before the nuisance
    // step, which runs only when NumOuterIter > 1
    const double kDefBragg = 5e10;  // DefBraggThreshold, BrightField/BFConstants.h
    const double des = j->prob.bragg_T;
    double tmp = num_iter > 1 ? kDefBragg : des;
    set_bragg(j, tmp);
    // NHICD state (BFReconstructionEngine.cpp:606-635): the call counter, the next
    // homogeneous part and the ROI volume sum of the last check
    const bool nhicd = j->prob.bf_nhicd != 0;
    j->nh_log.clear();
    j->nh_log_ptr.assign(1, 0);
    uint32_t eff = 0;
    int part = 0;
    double prev_mag = 0;
    for (int outer = 0; outer < j->prob.num_outer; ++outer) {
      if (outer > 0) {
        num_iter = 1;
        set_bragg(j, tmp);
      }
      for (int inner = 0; inner < num_iter; ++inner) {
        if (num_iter > 1 && inner == 0) set_bragg(j, kDefBragg);
        if (!nhicd) {
          status = run_sweep(j, outer, inner, nullptr);
        } else {
          if (j->cancel.load()) return MBIR_ERR_CANCELLED;
          std::vector<int> list;
          const std::vector<std::vector<int>>* steps = nullptr;
          std::vector<std::vector<int>> nh_steps;
          const size_t n = j->nh_all.size();
          if (eff % 2 == 0) {
            // homogeneous: the next twentieth of the random list, the last one longer
            // (:681-712); a part's lines never change, so its steps are packed once
            part %= kNhParts;
            size_t ne = n / kNhParts;
            if (part == kNhParts - 1) ne += n - ne * kNhParts;
            const size_t start = (n / kNhParts) * (size_t)part;
            list.assign(j->nh_all.begin() + (long)start, j->nh_all.begin() + (long)(start + ne));
            if (j->nh_parts.size() != (size_t)kNhParts) j->nh_parts.assign(kNhParts, {});
            if (j->nh_parts[part].empty()) j->nh_parts[part] = pack_chains(j, list);
            steps = &j->nh_parts[part];
            part++;
          } else {
            // non-homogeneous: lines whose filtered update magnitude exceeds the
            // threshold, in a seeded random order (Extra.cpp:571-591, 990-1025)
            if ((rc = fetch_magmap(j)) < 0) return rc;
            std::vector<double> filt(n);
            nhicd_filter(j->magmap_h.data(), j->g.N_z, j->g.N_x, filt.data());
            const double thr = nhicd_threshold(filt);
            for (size_t l = 0; l < n; ++l)
              if (filt[l] > thr) list.push_back((int)l);
            nh_steps = pack_chains(j, list);
            steps = &nh_steps;
          }
          if (j->nh_log_on) {
            j->nh_log.insert(j->nh_log.end(), list.begin(), list.end());
            j->nh_log_ptr.push_back((int64_t)j->nh_log.size());
          }
          // the next homogeneous part, when a later call may need it and it is not packed
          // yet, is packed while the device runs this call
          const bool later = inner + 1 < num_iter || outer + 1 < j->prob.num_outer;
          const int next = part % kNhParts;
          rc = run_list_steps(j, *steps, outer, inner, [&]() {
            if (!later || j->nh_parts.size() != (size_t)kNhParts || !j->nh_parts[next].empty())
              return;
            const size_t nn = n / kNhParts;
            const size_t s0 = nn * (size_t)next, ne = next == kNhParts - 1 ? n - s0 : nn;
            const std::vector<int> nl(j->nh_all.begin() + (long)s0,
                                      j->nh_all.begin() + (long)(s0 + ne));
            j->nh_parts[next] = pack_chains(j, nl);
          });
          if (rc < 0) return rc;
          // stop test every 20 calls: ROI update magnitude against the ROI volume sum
          // of 20 calls before (Extra.cpp:1217-1268; BFReconstructionEngine.cpp:739-750)
          status = 1;
          if (eff % kNhParts == 0 && eff > 0 && prev_mag > 0) {
            if ((rc = fetch_magmap(j)) < 0) return rc;
            double sum = 0;
            for (size_t l = 0; l < n; ++l)
              if (j->mask_h[l] == 1) sum += j->magmap_h[l];
            if (std::fabs(sum) / std::fabs(prev_mag) < j->stop_threshold) status = 0;
          }
          if (eff % kNhParts == 0) {
            prev_mag = roi_volume_sum(j);
            if (prev_mag < 0) return MBIR_ERR_HIP;  // the volume is non-negative
          }
          eff++;
        }
        if (status < 0) return status;
        ns++;
        if (status == 0) break;
        if ((rc = push_cost()) < 0) return rc;
        if (inner == num_iter - 1 && inner != 0) {
          tmp = des;
          set_bragg(j, tmp);
        }
      }
      if (j->cancel.load()) return MBIR_ERR_CANCELLED;
      outer_done = outer + 1;
      if (status == 0 && outer >= 1) break;
      if (j->prob.num_outer > 1) {
        const int je = j->prob.joint_estimation, nm = j->prob.noise_model;
        if (je) {
          j->prob.noise_model = 0;
          rc = nuisance_update(j);
          j->prob.noise_model = nm;
          if (rc < 0) return rc;
          if ((rc = push_cost()) < 0) return rc;
        }
        if (nm) {
          j->prob.joint_estimation = 0;
          rc = nuisance_update(j);
          j->prob.joint_estimation = je;
          if (rc < 0) return rc;
          if ((rc = push_cost()) < 0) return rc;
        }
      }
    }
    if (rs) {
      rs->sweeps = ns;
      rs->outer_iterations = outer_done;
      rs->status = status;
      rs->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (n_costs) *n_costs = nc;
    return MBIR_OK;
  }
  for (int outer = 0; outer < j->prob.num_outer; ++outer) {
    if (outer != 0) num_iter = 1;  // :712-715
    for (int inner = 0; inner < num_iter; ++inner) {
      status = run_sweep(j, outer, inner, nullptr);
      if (status < 0) return status;
      ns++;
      if ((rc = push_cost()) < 0) return rc;
      if (status == 0) break;
    }
    if (j->cancel.load()) return MBIR_ERR_CANCELLED;
    if (j->prob.joint_estimation || j->prob.noise_model) {
      if (j->prob.joint_estimation) {
        // joint estimation alone, then the noise update, so the cost after each is recorded
        const int nm = j->prob.noise_model;
        j->prob.noise_model = 0;
        rc = nuisance_update(j);
        j->prob.noise_model = nm;
        if (rc < 0) return rc;
        if ((rc = push_cost()) < 0) return rc;
      }
      if (j->prob.noise_model) {
        const int je = j->prob.joint_estimation;
        j->prob.joint_estimation = 0;
        rc = nuisance_update(j);
        j->prob.joint_estimation = je;
        if (rc < 0) return rc;
        if ((rc = push_cost()) < 0) return rc;
      }
    }
    outer_done = outer + 1;
    if (status == 0 && outer >= 1) break;  // :789-802
    if (j->prob.estimate_prior && (rc = estimate_prior(j, nullptr, nullptr)) < 0)
      return rc;  // ESTIMATE_PRIOR (:807-813)
  }
  if (rs) {
    rs->sweeps = ns;
    rs->outer_iterations = outer_done;
    rs->status = status;
    rs->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (n_costs) *n_costs = nc;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_forward_project(mbir_job* j, const float* volume, double* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  const size_t M = (size_t)j->g.N_theta * j->g.N_r * j->g.N_t;
  const size_t V = (size_t)j->g.N_z * j->g.N_x * j->g.N_y;
  hipStream_t s = j->stream;
  DevBuf<double> yest;
  DevBuf<float> vol;
  if (yest.alloc(M) != hipSuccess) return MBIR_ERR_NOMEM;
  const float* src = j->x.p;
  if (volume) {
    if (vol.alloc(V) != hipSuccess) {
      yest.release();
      return MBIR_ERR_NOMEM;
    }
    if (hipMemcpyAsync(vol.p, volume, V * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess) {
      yest.release();
      vol.release();
      return MBIR_ERR_HIP;
    }
    src = vol.p;
  }
  int rc = MBIR_OK;
  if (launch_fill_f64(yest.p, 0.0, M, s) != hipSuccess ||
      launch_forward_project(j->J, src, yest.p, s) != hipSuccess ||
      hipMemcpyAsync(out, yest.p, M * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    rc = MBIR_ERR_HIP;
  yest.release();
  vol.release();
  return rc;
}

extern "C" int32_t mbir_job_get_column(mbir_job* j, int32_t z, int32_t x, uint32_t* index,
                                       float* values, int32_t cap)
{
  if (!j || z < 0 || x < 0 || z >= j->g.N_z || x >= j->g.N_x) return MBIR_ERR_INPUT;
  // the entries the ICD kernels use: the device-built column table
  const size_t line = (size_t)z * j->g.N_x + x;
  int p[2];
  HIP_TRY(hipMemcpyAsync(p, j->col_ptr.p + line, 2 * sizeof(int), hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  const int n = p[1] - p[0];
  if (n > 0 && cap > 0) {
    const int m = std::min(n, cap);
    std::vector<int> r(m);
    std::vector<float> v(m);
    HIP_TRY(hipMemcpyAsync(r.data(), j->col_row.p + p[0], m * sizeof(int), hipMemcpyDeviceToHost, j->stream));
    HIP_TRY(hipMemcpyAsync(v.data(), j->col_a.p + p[0], m * sizeof(float), hipMemcpyDeviceToHost, j->stream));
    HIP_TRY(hipStreamSynchronize(j->stream));
    for (int k = 0; k < m; ++k) {
      if (index) index[k] = (uint32_t)(r[k] / j->g.N_t);
      if (values) values[k] = v[k];
    }
  }
  return n;
}

extern "C" int32_t mbir_job_get_volume(mbir_job* j, float* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  const size_t V = (size_t)j->g.N_z * j->g.N_x * j->g.N_y;
  HIP_TRY(hipMemcpyAsync(out, j->x.p, V * sizeof(float), hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  return MBIR_OK;
}

extern "C" int32_t mbir_job_set_volume(mbir_job* j, const float* vol)
{
  if (!j || !vol) return MBIR_ERR_INPUT;
  const size_t V = (size_t)j->g.N_z * j->g.N_x * j->g.N_y;
  const size_t M = (size_t)j->g.N_theta * j->g.N_r * j->g.N_t;
  hipStream_t s = j->stream;
  HIP_TRY(hipMemcpyAsync(j->x.p, vol, V * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(j->yest.alloc(M));
  HIP_TRY(launch_fill_f64(j->yest.p, 0.0, M, s));
  HIP_TRY(launch_forward_project(j->J, j->x.p, j->yest.p, s));
  HIP_TRY(launch_init_error(j->J, j->yest.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  j->yest.release();
  return MBIR_OK;
}

// The device-resident form of the multi-resolution hand-over: the coarse level's
// volume replicated 2 x 2 x 2 into this job's volume as the reference's reader does
// with the level file (HAADF/HAADF_MultiResolutionReconstruction.cpp:211-229,
// GenericFilters/InitialReconstructionBinReader.cpp:103-130), then e re-derived.
extern "C" int32_t mbir_job_upsample_volume(mbir_job* fine, const mbir_job* coarse)
{
  if (!fine || !coarse || fine == coarse) return MBIR_ERR_INPUT;
  if (fine->ctx->device != coarse->ctx->device) return MBIR_ERR_INPUT;
  const mbir_geometry& f = fine->g;
  const mbir_geometry& c = coarse->g;
  if (c.N_z != f.N_z / 2 || c.N_x != f.N_x / 2 || c.N_y != f.N_y / 2 || c.N_z < 1 || c.N_x < 1 ||
      c.N_y < 1)
    return MBIR_ERR_INPUT;
  const size_t M = (size_t)f.N_theta * f.N_r * f.N_t;
  hipStream_t s = fine->stream;
  HIP_TRY(hipStreamSynchronize(coarse->stream));  // the coarse volume is final
  HIP_TRY(launch_upsample2(coarse->x.p, fine->x.p, fine->line_nnz.p, f.N_z, f.N_x, f.N_y, s));
  HIP_TRY(fine->yest.alloc(M));
  HIP_TRY(launch_fill_f64(fine->yest.p, 0.0, M, s));
  HIP_TRY(launch_forward_project(fine->J, fine->x.p, fine->yest.p, s));
  HIP_TRY(launch_init_error(fine->J, fine->yest.p, s));
  HIP_TRY(hipStreamSynchronize(s));
  fine->yest.release();
  return MBIR_OK;
}

extern "C" int32_t mbir_job_get_error_sinogram(mbir_job* j, float* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  const size_t M = (size_t)j->g.N_theta * j->g.N_r * j->g.N_t;
  HIP_TRY(hipMemcpyAsync(out, j->e.p, M * sizeof(float), hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  return MBIR_OK;
}

extern "C" int32_t mbir_job_get_nuisance(mbir_job* j, double* I0, double* mu, double* alpha)
{
  if (!j) return MBIR_ERR_INPUT;
  const size_t n = (size_t)j->g.N_theta * sizeof(double);
  if (I0) HIP_TRY(hipMemcpyAsync(I0, j->I0.p, n, hipMemcpyDeviceToHost, j->stream));
  if (mu) HIP_TRY(hipMemcpyAsync(mu, j->mu.p, n, hipMemcpyDeviceToHost, j->stream));
  if (alpha) HIP_TRY(hipMemcpyAsync(alpha, j->alpha.p, n, hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  return MBIR_OK;
}

extern "C" int32_t mbir_job_get_geometry(mbir_job* j, mbir_geometry* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  *out = j->g;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_get_detector_response(mbir_job* j, double* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  std::memcpy(out, j->H.data(), j->H.size() * sizeof(double));
  return MBIR_OK;
}

extern "C" int32_t mbir_job_cancel(mbir_job* j)
{
  if (!j) return MBIR_ERR_INPUT;
  j->cancel.store(1);
  return MBIR_OK;
}

extern "C" int64_t mbir_job_device_bytes(const mbir_job* j) { return j ? j->bytes : -1; }

extern "C" int32_t mbir_job_set_timing(mbir_job* j, int32_t on)
{
  if (!j) return MBIR_ERR_INPUT;
  j->timing = on ? 1 : 0;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_get_selector(mbir_job* j, uint8_t* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  if (!j->J.wfull) return MBIR_ERR_UNSUPPORTED;
  const size_t M = (size_t)j->g.N_theta * j->g.N_r * j->g.N_t;
  std::vector<float> w(M);
  HIP_TRY(hipMemcpyAsync(w.data(), j->wfull.p, M * sizeof(float), hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  for (size_t i = 0; i < M; ++i) out[i] = w[i] >= 0.f ? 1 : 0;
  return MBIR_OK;
}

extern "C" int32_t mbir_job_get_weights(mbir_job* j, float* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  if (!j->J.wfull) return MBIR_ERR_UNSUPPORTED;
  const size_t M = (size_t)j->g.N_theta * j->g.N_r * j->g.N_t;
  HIP_TRY(hipMemcpyAsync(out, j->wfull.p, M * sizeof(float), hipMemcpyDeviceToHost, j->stream));
  HIP_TRY(hipStreamSynchronize(j->stream));
  for (size_t i = 0; i < M; ++i) out[i] = std::fabs(out[i]);
  return MBIR_OK;
}

extern "C" int32_t mbir_job_set_bragg_threshold(mbir_job* j, double T)
{
  if (!j || !(T > 0)) return MBIR_ERR_INPUT;
  if (!j->J.wfull) return MBIR_ERR_UNSUPPORTED;
  set_bragg(j, T);
  return MBIR_OK;
}
