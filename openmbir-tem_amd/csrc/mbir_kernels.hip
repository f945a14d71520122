// mbir_kernels.hip -- HIP kernels of the MBIR hot path for gfx950.
//
// Data layout in HBM (the reference's layouts, Common/TomoArray.hpp:176-198):
//   volume x [N_z][N_x][N_y] fp32, y fastest: one "line" (z, x) = N_y contiguous voxels
//   error sinogram e, measurements y [N_theta][N_r][N_t] fp32, t fastest
// Because a voxel at slice y projects only onto detector rows t in VLR(y)
// (HAADF/HAADF_ReconstructionEngine_Extra.cpp:329-404), the A-column of a line
// touches, per (theta, r) bin, a contiguous run of N_t floats.  Every kernel maps
// wavefront lanes onto that run, so each sinogram access is a coalesced 64-lane
// load/atomic.  The columns are computed once per job on the device, in FP64 from
// the detector response, and kept as a compact table (sinogram row, I0 a, I0 a w:
// 12 B per entry, about 2.1 N_theta entries per line) that each line update copies
// into LDS; the per-voxel VLR weights are applied from registers.
#include <hip/hip_runtime.h>

#include <cstring>
#include <algorithm>

#include "mbir_internal.h"

namespace mbir {

namespace {

// ---------------------------------------------------------------------------
// A-matrix column of line (z, x), view th.  Restates
// HAADF/HAADF_ReconstructionEngine.cpp:1402-1497 in FP64 (divisions by delta_r and
// OffsetR as multiplications by their reciprocals): r = x cos - z sin, candidate bins
// [floor((r - dxz - R0)/dr), floor((r + dxz - R0)/dr)] clipped to the detector,
// value = lerp of H[th] at |r - r_c| / OffsetR, kept if > 0.  av[k] is the value of
// bin jmin + k (0 = no entry); returns jmin.  All loads are issued unconditionally
// from clamped addresses so that they are in flight together.
__device__ inline int view_column(const DevJob& J, double xc, double zc, int th, double* av)
{
  const bool vth = th < J.N_theta;
  const int thc = vth ? th : 0;
  const double r = xc * J.cosv[thc] - zc * J.sinv[thc];
  const double rmin = r - J.delta_xz, rmax = r + J.delta_xz;
  int jmin = 0, jmax = -1;
  if (vth && !(rmax < J.R0 || rmin > J.RMax)) {
    jmin = (int)floor((rmin - J.R0) * J.inv_delta_r);
    jmax = (int)floor((rmax - J.R0) * J.inv_delta_r);
    if (jmax >= J.N_r) jmax = J.N_r - 1;
    if (jmin < 0) jmin = 0;
  }
  const double* Hth = J.H + (size_t)thc * kResponseBins;
  auto slot = [&](int k) {
    const int j = jmin + k;
    const double rc = J.R0 + ((double)j + 0.5) * J.delta_r;
    const double dr = fabs(r - rc);
    int idr = (int)floor(dr * J.inv_offsetR);
    const bool ok = j <= jmax && idr >= 0 && idr < kResponseBins;
    idr = ok ? idr : 0;
    const int iidx = idr + 1 < kResponseBins ? idr + 1 : kResponseBins - 1;
    const double h0 = Hth[idr], h1 = Hth[iidx];
    const double w1 = dr - idr * J.offsetR;
    const double w2 = (idr + 1) * J.offsetR - dr;
    const double f1 = (w2 * J.inv_offsetR) * h0 + (w1 * J.inv_offsetR) * h1;
    return (ok && f1 > 0) ? f1 : 0.0;
  };
  double a4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) a4[k] = slot(k);
#pragma unroll
  for (int k = 0; k < 4; ++k) av[k] = k < J.NB ? a4[k] : 0.0;
#pragma unroll
  for (int k = 4; k < kMaxNB; ++k) av[k] = 0.0;
  if (J.NB > 4) {  // coarse voxels (delta_xz > 1.5 delta_r): more bins per view
#pragma unroll
    for (int k = 4; k < kMaxNB; ++k) av[k] = k < J.NB ? slot(k) : 0.0;
  }
  return jmin;
}

// Column of line (z, x) slot by slot: store(slot, th, j, a) for slot = th * NB + k.
template <typename F>
__device__ inline void build_column(const DevJob& J, int z, int x, F&& store)
{
  const double xc = J.x0 + ((double)x + 0.5) * J.delta_xz;
  const double zc = J.z0 + ((double)z + 0.5) * J.delta_xz;
  for (int th = threadIdx.x; th < J.N_theta; th += blockDim.x) {
    double av[kMaxNB];
    const int jmin = view_column(J, xc, zc, th, av);
#pragma unroll
    for (int k = 0; k < kMaxNB; ++k)
      if (k < J.NB) store(th * J.NB + k, th, jmin + k, av[k]);
  }
}

__device__ inline double block_sum_d(double v, double* red)
{
  // deterministic tree reduction over kBlock threads
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// Kahan-compensated fp32 accumulation (Kahan 1965).
struct Kahan {
  float s = 0.f, c = 0.f;
  __device__ inline void add(float v)
  {
    const float yv = v - c;
    const float t = s + yv;
    c = (t - s) - yv;
    s = t;
  }
};

__device__ inline float filter_weight(int dy, int dz, int dx)
{
  // FILTER (HAADF/HAADF_ReconstructionEngine.cpp:266-294): faces 0.0523,
  // edges 0.0370, corners 0.0302
  const int zeros = (dy == 0) + (dz == 0) + (dx == 0);
  return zeros == 2 ? 0.0523f : (zeros == 1 ? 0.0370f : 0.0302f);
}

// b_j = rho'(d)/d of the QGGMRF potential with MRF_P = 2
// (Reconstruction/QGGMRF_Functions.cpp:90-112,118-147):
//   t1 = |d|^(P-Q) / sigma^(P-Q),  t3 = c + t1,
//   rho'(d)/d = (P - (P-Q) t1 / t3) / (t3 sigma^P);   P / (sigma^P c) at d = 0.
__device__ inline float qggmrf_b(float d, const DevJob& J)
{
  // |d|^(2-q) through the hardware log2 / exp2 (v_log_f32, v_exp_f32, ~1 ulp) and
  // reciprocals through v_rcp_f32 (1 ulp): the coefficient of a quadratic
  // surrogate, where 1e-6 relative error is far below the fp32 solve's own.
  // Evaluated for every d and selected (no branch: callers interleave several).
  const float a = fabsf(d) < 1e-30f ? 1e-30f : fabsf(d);
  const float t1 = __builtin_amdgcn_exp2f(J.pq2 * __builtin_amdgcn_logf(a)) * J.inv_sigma_pq;
  const float r3 = __builtin_amdgcn_rcpf(J.mrf_c + t1);
  const float b = (2.f - J.pq2 * t1 * r3) * r3 * J.inv_sigma_p;
  // b0 read first: with a member read in the conditional's arm the front end emits a
  // branch, the evaluation above sinks into it, and the callers' evaluations then run
  // one after another instead of interleaved
  const float b0 = J.b0;
  return d == 0.f ? b0 : b;
}

// Butterfly reduction inside aligned groups of G (1, 2, 4, 8 or 16) lanes with DPP
// lane moves (no LDS): quad_perm xor 1, quad_perm xor 2, row_half_mirror, row_mirror.
template <int CTRL>
__device__ inline float dpp_mov(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ inline void dpp_step(float& s1, float& s2, float& lo, float& hi)
{
  s1 += dpp_mov<CTRL>(s1);
  s2 += dpp_mov<CTRL>(s2);
  lo = fminf(lo, dpp_mov<CTRL>(lo));
  hi = fmaxf(hi, dpp_mov<CTRL>(hi));
}
__device__ inline void group_reduce(int G, float& s1, float& s2, float& lo, float& hi)
{
  if (G >= 2) dpp_step<0xB1>(s1, s2, lo, hi);   // quad_perm [1,0,3,2]
  if (G >= 4) dpp_step<0x4E>(s1, s2, lo, hi);   // quad_perm [2,3,0,1]
  if (G >= 8) dpp_step<0x141>(s1, s2, lo, hi);  // row_half_mirror
  if (G >= 16) dpp_step<0x140>(s1, s2, lo, hi); // row_mirror
}

// Sum over the 64 lanes of a wave in a fixed order by DPP moves (no LDS): prefix
// sums within rows of 16 (row_shr 1, 2, 4, 8), then row_bcast 15 and 31; the total
// ends in lane 63.  Every lane of the wave must take part.
template <int CTRL, int ROWS>
__device__ inline int dpp_upd(int v)
{
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}
__device__ inline float wave_sum(float v)
{
  v += __int_as_float(dpp_upd<0x111, 0xF>(__float_as_int(v)));
  v += __int_as_float(dpp_upd<0x112, 0xF>(__float_as_int(v)));
  v += __int_as_float(dpp_upd<0x114, 0xF>(__float_as_int(v)));
  v += __int_as_float(dpp_upd<0x118, 0xF>(__float_as_int(v)));
  v += __int_as_float(dpp_upd<0x142, 0xA>(__float_as_int(v)));
  v += __int_as_float(dpp_upd<0x143, 0xC>(__float_as_int(v)));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// Sum of lanes 0-7 (every other lane 0) by three DPP prefix steps, in lane 7.
__device__ inline float sum8(float v)
{
  v += __int_as_float(dpp_upd<0x111, 0xF>(__float_as_int(v)));
  v += __int_as_float(dpp_upd<0x112, 0xF>(__float_as_int(v)));
  v += __int_as_float(dpp_upd<0x114, 0xF>(__float_as_int(v)));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 7));
}
__device__ inline int sum8(int v)
{
  v += dpp_upd<0x111, 0xF>(v);
  v += dpp_upd<0x112, 0xF>(v);
  v += dpp_upd<0x114, 0xF>(v);
  return __builtin_amdgcn_readlane(v, 7);
}
template <int CTRL, int ROWS>
__device__ inline double dpp_upd(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = dpp_upd<CTRL, ROWS>((int)b), hi = dpp_upd<CTRL, ROWS>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double wave_sum(double v)
{
  v += dpp_upd<0x111, 0xF>(v);
  v += dpp_upd<0x112, 0xF>(v);
  v += dpp_upd<0x114, 0xF>(v);
  v += dpp_upd<0x118, 0xF>(v);
  v += dpp_upd<0x142, 0xA>(v);
  v += dpp_upd<0x143, 0xC>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline int wave_sum(int v)
{
  v += dpp_upd<0x111, 0xF>(v);
  v += dpp_upd<0x112, 0xF>(v);
  v += dpp_upd<0x114, 0xF>(v);
  v += dpp_upd<0x118, 0xF>(v);
  v += dpp_upd<0x142, 0xA>(v);
  v += dpp_upd<0x143, 0xC>(v);
  return __builtin_amdgcn_readlane(v, 63);
}

}  // namespace

// ---------------------------------------------------------------------------
// Per-line column entry counts; zeroes lines whose column is empty
// (HAADF/HAADF_ReconstructionEngine.cpp:619-629).
__global__ __launch_bounds__(kBlock) void k_line_nnz(DevJob J, int* line_nnz, float* x)
{
  __shared__ double red[kBlock];
  const int line = blockIdx.x;
  const int z = line / J.N_x, xi = line % J.N_x;
  int cnt = 0;
  build_column(J, z, xi, [&](int, int, int, double v) { cnt += (v > 0.0); });
  const double tot = block_sum_d((double)cnt, red);
  const int n = (int)tot;
  if (threadIdx.x == 0) line_nnz[line] = n;
  if (n == 0)
    for (int y = threadIdx.x; y < J.N_y; y += blockDim.x) x[(size_t)line * J.N_y + y] = 0.f;
}

// Ŷ[th][r][t] += I0[th] * a * vlr * x (HAADF/HAADF_ForwardProject.cpp:80-120),
// one workgroup per line, FP64 atomic accumulation.
__global__ __launch_bounds__(kBlock) void k_forward_project(DevJob J, const float* vol, double* out)
{
  __shared__ double s_ka[kMaxSlots];
  __shared__ int s_row[kMaxSlots];
  __shared__ float s_x[kMaxNy];
  const int line = blockIdx.x;
  if (J.line_nnz[line] == 0) return;
  const int z = line / J.N_x, xi = line % J.N_x;
  for (int y = threadIdx.x; y < J.N_y; y += blockDim.x) s_x[y] = vol[(size_t)line * J.N_y + y];
  build_column(J, z, xi, [&](int s, int th, int j, double v) {
    s_ka[s] = J.I0[th] * v;
    s_row[s] = v > 0.0 ? (th * J.N_r + j) * J.N_t : 0;
  });
  __syncthreads();
  // lanes over y, waves over slots
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int s = wave; s < J.nslots; s += kWaves) {
    const double ka = s_ka[s];
    if (ka == 0.0) continue;
    const int row = s_row[s];
    for (int y = lane; y < J.N_y; y += 64) {
      const double xv = s_x[y];
      if (xv == 0.f) continue;
      const int t0 = J.vlr_t0[y], n = J.vlr_n[y];
      for (int k = 0; k < n; ++k) {
        const double v = ka * (double)J.vlr_val[y * kVlrCap + k] * xv;
        atomicAdd(out + row + t0 + k, v);
      }
    }
  }
}

// e = y - Ŷ - mu  (HAADF/HAADF_ReconstructionEngine_Extra.cpp:718-721)
__global__ void k_init_error(DevJob J, const double* yest)
{
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
  const size_t per = (size_t)J.N_r * J.N_t;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int th = (int)(i / per);
    // with a bright-field sinogram: y - s Yest - mu (HAADF_ReconstructionEngine_Extra.cpp:724-727)
    const double ye = J.bfs ? (double)J.bfs[i] * yest[i] : yest[i];
    J.e[i] = (float)((double)J.y[i] - ye - J.mu[th]);
  }
}

// Export column (z, x): per slot value and sinogram row base.
__global__ __launch_bounds__(kBlock) void k_export_column(DevJob J, int z, int x, float* vals,
                                                          int* rows)
{
  build_column(J, z, x, [&](int s, int th, int j, double v) {
    vals[s] = (float)v;
    rows[s] = th * J.N_r + j;
  });
}

// Compact columns of all lines into the table (col_ptr from the per-line counts):
// one workgroup per line, one thread per view, order fixed by a wave scan.
__global__ __launch_bounds__(kIcdBlock) void k_build_columns(DevJob J, int* row, float* a)
{
  __shared__ int wsum[kIcdWaves];
  const int line = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int z = line / J.N_x, xi = line % J.N_x;
  const double xc = J.x0 + ((double)xi + 0.5) * J.delta_xz;
  const double zc = J.z0 + ((double)z + 0.5) * J.delta_xz;
  double av[kMaxNB];
  const int jmin = view_column(J, xc, zc, tid, av);
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < kMaxNB; ++k) cnt += av[k] > 0.0 ? 1 : 0;
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int pos = J.col_ptr[line] + incl - cnt;
  for (int w = 0; w < wave; ++w) pos += wsum[w];
  if (tid < J.N_theta && cnt > 0) {
    const int rb = tid * J.N_r + jmin;
#pragma unroll
    for (int k = 0; k < kMaxNB; ++k)
      if (av[k] > 0.0) {
        row[pos] = (rb + k) * J.N_t;
        a[pos] = (float)av[k];
        ++pos;
      }
  }
}

// Fold the gains and the per-view weights into the column table:
// ka = I0[th] a (kConst0 of …_UpdateVoxels.cpp:347), kw = ka w[th].
__global__ void k_column_weights(DevJob J, float* ka, float* kw, int nnz)
{
  const int per = J.N_r * J.N_t;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
    const int th = J.col_row[i] / per;
    const double k0 = J.I0[th] * (double)J.col_a[i];
    ka[i] = (float)k0;
    kw[i] = (float)(k0 * (double)J.wview[th]);
  }
}

// The column factor of theta2 per line with plain weights: sum over the line's column
// of (I0 a)(I0 a w), exact products summed in FP64 (lanes stride the entries, then a
// fixed-order DPP sum); one wave per line.
__global__ __launch_bounds__(kBlock) void k_line_kk(DevJob J, const int* col_ptr, double* kk, int L)
{
  const int lane = threadIdx.x & 63;
  const int l = (int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
  if (l >= L) return;  // (uniform per wave)
  double a = 0.0;
  for (int i = col_ptr[l] + lane; i < col_ptr[l + 1]; i += 64)
    a = __fma_rn((double)J.col_ka[i], (double)J.col_kw[i], a);
  a = wave_sum(a);
  if (lane == 63) kk[l] = a;
}

// ---------------------------------------------------------------------------
// ICD update of one line (z, x): all N_y voxels of the line.  Restates
// UpdateYSlice::execute for one line (HAADF/HAADF_ReconstructionEngine_UpdateVoxels.cpp:260-455):
// neighbourhood gather (:281-305), zero skipping (:311-338), theta1/theta2
// (:343-369), bounds (:47-63,370), QGGMRF functional substitution
// (Reconstruction/QGGMRF_Functions.cpp:153-206), positivity (:394-400), ROI
// statistics (:412-421) and the error-sinogram update (:423-449).
//
// Slices y are independent in the data term (disjoint VLR rows); the prior
// couples y and y +- 1, so even slices are solved first and odd slices second,
// which makes one line update equal to a sequential order of its voxels.
//   1./2./3. one memory round trip brings the 3x3 block of lines around (z, x)
//      (one wave per line, lanes along y) and the line's column (from the table
//      built at setup) into LDS, and the sinogram rows of each lane's first theta
//      batch, whose error-sinogram loads are then in flight while LDS is written;
//   4. theta1/theta2: the column is cut into contiguous streams of entries, one
//      or more per wave; lanes run along the contiguous detector row t, U entries
//      in flight per lane, addresses as 32-bit byte offsets from a scalar base,
//      four entries' offsets and gains per 16-byte LDS read; Kahan-compensated
//      fp32 partial sums combined in FP64 (lane shuffles within a wave, LDS
//      across waves);
//   5. the voxel solve of each parity runs on lane groups: the 26 neighbour
//      terms of one voxel are spread over G lanes and combined by DPP moves;
//   6. sweep statistics: per-wave shares formed in the solve, summed by one thread
//      in a fixed order, one set of global atomics per line;
//   7. e -= I0 a vlr dx: from the registers of step 2 when a stream fits one batch,
//      else a plain read-modify-write (no two lines of a conflict-free step share a
//      detector bin), or fp32 atomics on the relaxed schedule; all round alike;

// Byte-offset access through a uniform base pointer: base in scalar registers plus
// a 32-bit offset in one vector register (global load/store with a scalar base).
// Accesses through pointers typed as global memory (address space 1) compile to
// global_* instructions.  The job descriptors of the batched kernels are read from
// device memory, and through their plain (generic) pointers the compiler emits flat
// accesses, which count against the LDS wait counter too: every LDS wait of the line
// update would then also wait for all of its loads in flight.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ inline gptr<T> gp(T* p)
{
  return (gptr<T>)p;
}
template <class T>
__device__ inline gptr<const T> gp(const T* p)
{
  return (gptr<const T>)p;
}

__device__ inline float ldo(const float* base, uint32_t boff)
{
  return *(gptr<const float>)((gptr<const char>)base + boff);
}
__device__ inline gptr<float> pto(float* base, uint32_t boff)
{
  return (gptr<float>)((gptr<char>)base + boff);
}
// A line reference read through a global pointer (field by field: a struct lvalue in
// another address space does not bind to the copy constructor)
__device__ inline LineRef load_ref(const LineRef* p)
{
  const gptr<const LineRef> q = gp(p);
  return LineRef{q->line, q->p0, q->nc, q->pad};
}
template <class T>
__device__ inline void atomic_add_g(T* p, T v)
{
  (void)__hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-line LDS of the line update for lines of at most 64 YPL voxels.  The column
// (row offsets and gains) lives in dynamic LDS sized per launch (IcdCol).
template <int YPL>
struct IcdSmem {
  static constexpr int NY = 64 * YPL;
  // lines (z+dz, x+dx) over the unit's slices y = -1 .. ny (0 outside the volume);
  // nbl[4] is the unit's own line, where the even phase writes its new values
  float nbl[9][NY + 2];
  float dx[NY];
  // per slice: the prior terms of the 24 neighbours in the other eight lines (not
  // written by the launch, so formed before the solve): sum b x_j, sum b, min, max
  float ps1[NY], ps2[NY], plo[NY], phi[NY];
  // theta partial sums per wave and slice (FP64)
  // (16-byte aligned: the solve reads the partials of two adjacent slices at once)
  alignas(16) double pd1[kIcdWaves][NY];
  alignas(16) double pd2[kIcdWaves][NY];
  // per wave: ROI sums of |x_new - x_old| and |x_old| (fp32, at most 256 non-negative
  // terms per line), and the counts zero-skipped + 2^10 updated + 2^20 detector rows
  // of the updated slices
  float wsu[kIcdWaves], wsv[kIcdWaves], wmg[kIcdWaves];  // wmg: |x_new - x_old|, all slices
  int wcnt[kIcdWaves];
  unsigned char vst[NY];      // per slice: 1 updated, 2 zero-skipped
};

// The line's column in dynamic LDS, entries padded to a multiple of four per stream:
// slots entries each of rb (byte offset of the entry's detector row, (th N_r + r) N_t 4),
// ka = I0[th] a and kw = I0[th] a w[th].  Every line of a launch has at most `slots`
// entries after padding (the host sizes the launch by its jobs' col_slots).
struct IcdCol {
  uint32_t* rb;
  float* ka;
  float* kw;
};
extern __shared__ __align__(16) uint32_t icd_dyn[];
__device__ inline IcdCol icd_col(int slots)
{
  return IcdCol{icd_dyn, reinterpret_cast<float*>(icd_dyn + slots),
                reinterpret_cast<float*>(icd_dyn + 2 * slots)};
}
// dynamic LDS bytes of a launch whose lines need at most `slots` padded entries
inline size_t icd_col_bytes(int slots) { return (size_t)slots * 12; }

// Waves per SIMD the batched line kernels are compiled for: the form that fills the
// device at least 6 (80 VGPRs), the prefetching form 5, wide shapes 4, and no more
// than the LDS of a CU (160 KiB) allows with 8-wave workgroups.
template <int YPL, int NV>
constexpr int icd_waves_per_eu(bool pf)
{
  // the static part and a column of 1024 entries (273 views at 3 bins per view)
  const int by_lds = 2 * (160 * 1024 / ((int)sizeof(IcdSmem<YPL>) + 1024 * 12));
  const int want = YPL * NV >= 4 ? 4 : (pf ? 5 : 6);
  return by_lds < want ? by_lds : want;
}

// PF: issue the first theta batch's loads before the LDS staging (shorter line
// latency, more registers: the form for launches that do not fill the device).
// SB: HAADF with a bright-field sinogram, every entry's element scaled by J.bfs.
// KR (with PF, YPL NV <= 4): keep the first batch's values in registers through the
// solve and write the new error sinogram from them when a stream fits one batch.
template <int YPL, int NV, bool FW, bool PF, bool SB = false, bool KR = false>
__device__ void icd_line(const DevJob& J, const LineRef& ref, int first, int zs_on, int pos,
                         IcdSmem<YPL>& sm, const IcdCol col)
{
  // entries per lane and batch, a multiple of four (about 24 loads in flight per lane)
  constexpr int UR = 24 / (YPL * NV);
  constexpr int U = UR < 4 ? 4 : (UR + 3) / 4 * 4;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int Ny = J.N_y;
  const int line = __builtin_amdgcn_readfirstlane(ref.line);
  const int z = J.nx_div ? (int)__umulhi((unsigned)line, J.nx_div) : line / J.N_x;
  const int xi = line - z * J.N_x;
  const bool roi = gp(J.mask)[line] != 0;  // used by the solve
  // plain weights: theta2 of slice y is (sum_k vlr_yk^2) times the line's column factor
  // sum (I0 a)(I0 a w), formed when the gains or weights change
  const double K = (!FW && !SB) ? gp(J.line_kk)[line] : 0.0;
  // the unit: slices y0 .. y0 + ny - 1 of the line (ref.pad = y0 | ny << 16; 0 = the
  // whole line).  Units of one launch lie on disjoint detector rows and are never prior
  // neighbours, so the slices just outside a unit are read, not written, by this launch.
  const int upad = __builtin_amdgcn_readfirstlane(ref.pad);
  const int ny = (upad >> 16) ? (upad >> 16) : Ny;
  const int y0 = (upad >> 16) ? (upad & 0xffff) : 0;
  const size_t lbase = (size_t)line * Ny + y0;
  const bool zs_active = zs_on && !first;
  const int p0 = __builtin_amdgcn_readfirstlane(ref.p0);
  const int nc = __builtin_amdgcn_readfirstlane(ref.nc);

  // ---- 1. lanes run along the TY-wide (power of two >= N_y, at most 64) slice index;
  // a wave holds 64 / TY streams of column entries, stream p owns the contiguous
  // entries [p C, p C + cnt) of the column; the entries [nc, P C) are padding that
  // point at the spare detector row behind the error sinogram with zero gains, so
  // whole groups of four entries are loaded and updated without per-entry tests ----
  // lg = log2 of the power of two >= ny, at most 6
  const int lg = ny <= 1 ? 0 : min(6, 32 - __clz(ny - 1));
  const int TY = 1 << lg, RPW = 64 >> lg, sub = lane >> lg, yl = lane & (TY - 1);
  const int P = kIcdWaves * RPW, p = wave * RPW + sub;
  const int C = (((nc + P - 1) >> (3 + 6 - lg)) + 3) & ~3;  // entries per stream, 4 | C
  const int ntot = P * C;                                     // <= kMaxSlots
  const int s0 = p * C;
  const int cnt = min(max(nc - s0, 0), C);
  // the lane's voxel-line response, loaded here with everything else and unpacked
  // (vtb, vv, vok) once the column is staged
  int vt0_[YPL], vn_[YPL];
  float vval_[YPL][NV];
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy) {
    const int y = yl + TY * jy;
    const int yc = y < ny ? y0 + y : 0;
    vt0_[jy] = gp(J.vlr_t0)[yc];
    vn_[jy] = gp(J.vlr_n)[yc];
#pragma unroll
    for (int k = 0; k < NV; ++k) vval_[jy][k] = gp(J.vlr_val)[yc * kVlrCap + k];
  }

  // ---- 2. one round trip: the 3x3 neighbour lines (wave w: line w, wave 0 also
  // line 8, lanes along y), the column table (for LDS) and, PF, the detector rows of
  // this lane's first batch.  Every load here is unconditional (addresses clamped
  // into the arrays, values selected afterwards) and nothing computes on a loaded
  // value before the last load is issued: a load under a per-lane branch, or one
  // whose value is used at once, is waited for before the next load goes out ----
  float nbv[2][YPL], nbh[2];
  bool nbin[2][YPL], nbhin[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h == 1 && wave != 0) break;  // (uniform) the ninth line is wave 0's
    const int l = wave + 8 * h;
    const int zz = z + l / 3 - 1, xx = xi + l % 3 - 1;
    const bool in = l < 9 && zz >= 0 && zz < J.N_z && xx >= 0 && xx < J.N_x;
    const gptr<const float> src =
        gp(J.x) + ((size_t)(in ? zz : 0) * J.N_x + (in ? xx : 0)) * Ny + (in ? y0 : 0);
#pragma unroll
    for (int jy = 0; jy < YPL; ++jy) {
      const int y = lane + 64 * jy;
      nbv[h][jy] = src[y < ny ? y : ny - 1];
      nbin[h][jy] = in && y < ny;
    }
    // the slice below (even lanes) and above (odd lanes) the unit, 0 outside the volume
    const int hy = (lane & 1) ? ny : -1;
    nbhin[h] = in && y0 + hy >= 0 && y0 + hy < Ny;
    nbh[h] = src[nbhin[h] ? hy : 0];
  }
  constexpr int kColIt = kMaxSlots / kIcdBlock;
  uint32_t rw_[kColIt];
  float ka_[kColIt], kw_[kColIt];
#pragma unroll
  for (int it = 0; it < kColIt; ++it) {
    if (it == 0 || it * kIcdBlock < nc) {  // uniform; a column has at least one entry
      const int i = tid + it * kIcdBlock;
      const int ic = p0 + (i < nc ? i : nc - 1);
      rw_[it] = (uint32_t)gp(J.col_row)[ic];
      ka_[it] = gp(J.col_ka)[ic];
      kw_[it] = gp(J.col_kw)[ic];
    }
  }
  // (PF) the rows of the stream's first batch, read straight from the table without a
  // per-entry bound: entries past the stream's count belong to the next stream, the
  // next line or the zero pad behind the table (p0 + s0 + u < p0 + P C + U), so every
  // read is in bounds; they are used only as addresses of loads whose values meet
  // zero gains, never as store targets
  uint32_t r0[PF ? U : 1];
  if constexpr (PF) {
#pragma unroll
    for (int u = 0; u < U; ++u) r0[u] = (uint32_t)gp(J.col_row)[p0 + s0 + u];
  }
  // the solving lane's slices 2 tid (even phase) and 2 tid + 1 (odd phase): detector
  // rows and voxel-line response values (theta2 = column factor * sum_k vlr^2)
  int vn_s[2];
  float vl_s[2][NV];
#pragma unroll
  for (int ph = 0; ph < 2; ++ph) {
    const int ys = 2 * tid + ph;
    const int yc = ys < ny ? y0 + ys : 0;
    vn_s[ph] = gp(J.vlr_n)[yc];
#pragma unroll
    for (int k = 0; k < NV; ++k) vl_s[ph][k] = gp(J.vlr_val)[yc * kVlrCap + k];
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int l = wave + 8 * h;
    if (l < 9) {
#pragma unroll
      for (int jy = 0; jy < YPL; ++jy) {
        const int y = lane + 64 * jy;
        if (y < ny) sm.nbl[l][y + 1] = nbin[h][jy] ? nbv[h][jy] : 0.f;
      }
      if (lane == 0) sm.nbl[l][0] = nbhin[h] ? nbh[h] : 0.f;
      if (lane == 1) sm.nbl[l][ny + 1] = nbhin[h] ? nbh[h] : 0.f;
    }
  }
#pragma unroll
  for (int it = 0; it < kColIt; ++it) {
    const int i = tid + it * kIcdBlock;
    if (it * kIcdBlock < ntot && i < ntot) {
      const bool in = i < nc;
      col.rb[i] = in ? rw_[it] * 4u : J.spare_b;
      col.ka[i] = in ? ka_[it] : 0.f;
      col.kw[i] = in ? kw_[it] : 0.f;
    }
  }
  uint32_t vtb[YPL][NV];  // byte offset of the lane's detector row position t
  float vv[YPL][NV];      // 0 where the lane has no entry: it then adds nothing
  bool vok[YPL][NV];
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy) {
    const int n = yl + TY * jy < ny ? vn_[jy] : 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      vok[jy][k] = k < n;
      vv[jy][k] = k < n ? vval_[jy][k] : 0.f;
      vtb[jy][k] = k < n ? (uint32_t)(vt0_[jy] + k) * 4u : 0u;
    }
  }
  // ---- 3. (PF) the first batch's error-sinogram loads, in flight across the barrier ----
  // four entries of this lane's stream: row offsets and gains by 16-byte LDS reads
  auto rows4 = [&](int i) { return *reinterpret_cast<const uint4*>(col.rb + s0 + i); };
  auto ka4 = [&](int i) { return *reinterpret_cast<const float4*>(col.ka + s0 + i); };
  auto kw4 = [&](int i) { return *reinterpret_cast<const float4*>(col.kw + s0 + i); };
  auto load_entry = [&](uint32_t rb, float (&ev)[YPL][NV], float (&wf)[YPL][NV]) {
#pragma unroll
    for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        ev[jy][k] = ldo(J.e, rb + vtb[jy][k]);
        if constexpr (FW) wf[jy][k] = ldo(J.wfull, rb + vtb[jy][k]);
        if constexpr (SB) wf[jy][k] = ldo(J.bfs, rb + vtb[jy][k]);  // the element's scale
      }
  };
  // the batch [base, base + U) of this lane's stream, rows from LDS
  auto load_batch = [&](int base, float (&ev)[U][YPL][NV], float (&wf)[U][YPL][NV]) {
#pragma unroll
    for (int q = 0; q < U / 4; ++q) {
      if (base + 4 * q < cnt) {
        const uint4 r = rows4(base + 4 * q);
        load_entry(r.x, ev[4 * q], wf[4 * q]);
        load_entry(r.y, ev[4 * q + 1], wf[4 * q + 1]);
        load_entry(r.z, ev[4 * q + 2], wf[4 * q + 2]);
        load_entry(r.w, ev[4 * q + 3], wf[4 * q + 3]);
      }
    }
  };
  // entries u >= cnt read some detector row of the error sinogram (in bounds); their
  // gains in the LDS column are zero, so they add nothing (and are never stored)
  float ev0[PF ? U : 1][YPL][NV], wf0[PF ? U : 1][YPL][NV];
  if constexpr (PF) {
#pragma unroll
    for (int u = 0; u < U; ++u) load_entry(r0[u] * 4u, ev0[u], wf0[u]);
  }
  __syncthreads();  // A: neighbour lines and column visible
  float* const smx = &sm.nbl[4][1];  // this unit's slices, smx[-1] and smx[ny] outside it

  // ---- 3b. the prior terms of each slice's 24 neighbours in the other eight lines
  // (QGGMRF_Functions.cpp:90-147 with the FILTER weights): those lines are not
  // updated by this launch, and a slice's own value changes only in its own phase,
  // so these terms are formed for every slice here, while the error-sinogram loads
  // are in flight.  GP = 8 / YPL lanes per slice, each with 8 / GP of the lines and
  // all three dy, summed by DPP in the group; the two neighbours in the line itself are
  // added in the solve ----
  {
    constexpr int GP = 8 / YPL, NL = 8 / GP;  // lanes per slice, lines per lane
    const int vy = tid / GP, gq = tid % GP;
    float s1 = 0.f, s2 = 0.f, lo = 0.f, hi = 0.f;  // lo/hi include the zeroed centre
    if (vy < ny) {
      const float v = smx[vy];
#pragma unroll
      for (int m = 0; m < NL; ++m) {
        const int l8 = gq + m * GP;  // one of the eight other lines, all three dy
        const int l = l8 < 4 ? l8 : l8 + 1;
        const int dz = l / 3 - 1, dxl = l % 3 - 1;
        const bool lin = z + dz >= 0 && z + dz < J.N_z && xi + dxl >= 0 && xi + dxl < J.N_x;
        const float* src = &sm.nbl[l][vy + 1];
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
          const bool ex = lin && y0 + vy + dy >= 0 && y0 + vy + dy < Ny;
          const float t = src[dy];
          const float nb = ex ? t : 0.f;
          lo = fminf(lo, nb);
          hi = fmaxf(hi, nb);
          const float b = filter_weight(dy, dz, dxl) * qggmrf_b(v - nb, J);
          s1 = ex ? s1 + b * nb : s1;
          s2 = ex ? s2 + b : s2;
        }
      }
    }
    group_reduce(GP, s1, s2, lo, hi);
    if (vy < ny && gq == 0) {
      sm.ps1[vy] = s1;
      sm.ps2[vy] = s2;
      sm.plo[vy] = lo;
      sm.phi[vy] = hi;
    }
  }

  // ---- 4. theta1 / theta2 (:343-369): every slice of the line (a zero-skipped
  // slice's sums are simply unused).  Plain weights: theta1 of slice y is
  // sum_k vlr_yk sum_entries (I0 a w) e, the inner sums per lane as exact products
  // accumulated in FP64 (no compensation needed); theta2 comes from the column
  // factor above.  Bright field: both per entry, Kahan-compensated fp32. ----
  Kahan t1[YPL], t2[YPL];
  double a1[YPL][NV], a2[YPL][NV];
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
    for (int k = 0; k < NV; ++k) a1[jy][k] = a2[jy][k] = 0.0;
  auto accumulate = [&](int base, const auto& ev, const auto& wf) {
    if constexpr (SB) {
      // scaled entries (HAADF_ReconstructionEngine_UpdateVoxels.cpp:354-362): theta1
      // sums (e s) kw, theta2 sums s^2 (ka kw), per lane in FP64 (wf holds s)
#pragma unroll
      for (int q = 0; q < U / 4; ++q) {
        if (base + 4 * q < cnt) {
          const float4 a4 = ka4(base + 4 * q);
          const float4 w4 = kw4(base + 4 * q);
          const float kav[4] = {a4.x, a4.y, a4.z, a4.w};
          const float kwv[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const double kd = (double)kwv[t], kk2 = (double)kav[t] * kd;
#pragma unroll
            for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
              for (int k = 0; k < NV; ++k) {
                const double sd = (double)wf[4 * q + t][jy][k];
                a1[jy][k] = __fma_rn((double)ev[4 * q + t][jy][k] * sd, kd, a1[jy][k]);
                a2[jy][k] = __fma_rn(sd * sd, kk2, a2[jy][k]);
              }
          }
        }
      }
    } else if constexpr (!FW) {
#pragma unroll
      for (int q = 0; q < U / 4; ++q) {
        if (base + 4 * q < cnt) {
          const float4 w4 = kw4(base + 4 * q);
          const float kwv[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const double kd = (double)kwv[t];
#pragma unroll
            for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
              for (int k = 0; k < NV; ++k)
                a1[jy][k] = __fma_rn((double)ev[4 * q + t][jy][k], kd, a1[jy][k]);
          }
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < U / 4; ++q) {
        if (base + 4 * q < cnt) {
          const float4 a4 = ka4(base + 4 * q);
          const float kav[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int u = 4 * q + t;
#pragma unroll
            for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
              for (int k = 0; k < NV; ++k) {
                // w per entry, its sign the Bragg selector: a rejected entry (w < 0)
                // weighs delta T / (|e| sqrt(|w|)) more
                // (BrightField/BFForwardModel.cpp:851-892); padding and lanes without
                // an entry have pe = 0 and add nothing
                const float pe = kav[t] * vv[jy][k];
                const float w = wf[u][jy][k];
                const float aw = fabsf(w);
                const float pw = (w >= 0.f || pe == 0.f)
                                     ? pe * aw
                                     : pe * aw * (J.bragg_dT / (fabsf(ev[u][jy][k]) * sqrtf(aw)));
                t2[jy].add(pe * pw);
                t1[jy].add(ev[u][jy][k] * pw);
              }
          }
        }
      }
    }
  };
  int base0 = 0;
  if constexpr (PF) {
    accumulate(0, ev0, wf0);
    base0 = U;
    // the first batch's loads have all been consumed (or were padding); saying so
    // here, where no store is pending yet, keeps the compiler from waiting on them
    // again before each later store (stores share the load counter on this
    // architecture, so such a wait would stall on the acknowledgement of earlier
    // stores): s_waitcnt vmcnt(0)
    __builtin_amdgcn_s_waitcnt(0x0F70);
  }
  for (int base = base0; base < cnt; base += U) {
    float ev[U][YPL][NV], wf[U][YPL][NV];
    load_batch(base, ev, wf);
    accumulate(base, ev, wf);
  }
  // FP64 partials, summed first over the streams of a wave that hold the same slice
  // (lanes yl, yl + TY, ...), then over the waves
  double d1[YPL], d2[YPL];
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy) {
    if constexpr (FW) {
      d1[jy] = (double)t1[jy].s - (double)t1[jy].c;
      d2[jy] = (double)t2[jy].s - (double)t2[jy].c;
    } else {
      d1[jy] = 0.0;
      d2[jy] = 0.0;
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        d1[jy] = __fma_rn((double)vv[jy][k], a1[jy][k], d1[jy]);
        if constexpr (SB) d2[jy] = __fma_rn((double)vv[jy][k] * vv[jy][k], a2[jy][k], d2[jy]);
      }
    }
  }
  for (int m = TY; m < 64; m <<= 1)
#pragma unroll
    for (int jy = 0; jy < YPL; ++jy) {
      d1[jy] += __shfl_xor(d1[jy], m);
      if constexpr (FW || SB) d2[jy] += __shfl_xor(d2[jy], m);
    }
  if (sub == 0)
#pragma unroll
    for (int jy = 0; jy < YPL; ++jy) {
      sm.pd1[wave][yl + TY * jy] = d1[jy];
      if constexpr (FW || SB) sm.pd2[wave][yl + TY * jy] = d2[jy];
    }
  __syncthreads();  // D

  // ---- 5. voxel solve (:311-421), even slices then odd slices, one lane per slice
  // (lane tid: slices 2 tid and 2 tid + 1): theta1 / theta2 from the per-wave
  // partials, the prior terms from 3b plus the two neighbours in the line itself, the
  // only ones the even phase changes ----
  const int nsw = __builtin_amdgcn_readfirstlane((((ny + 1) >> 1) + 63) >> 6);  // solving waves
  // this lane's statistics over both phases: |dx| and |x| in the ROI, |dx|
  // everywhere, and the counts zero-skipped + 2^10 updated + 2^20 detector rows of
  // the updated slices
  // both phases' theta1 / theta2 first (slices 2 tid and 2 tid + 1, one 16-byte LDS
  // read per wave partial): independent of the solve order, so the odd slice's sums
  // are off its critical path
  double T1s[2] = {0.0, 0.0}, T2s[2] = {0.0, 0.0};
  if (2 * tid < ny) {
#pragma unroll
    for (int w = 0; w < kIcdWaves; ++w) {
      const double2 p1 = *reinterpret_cast<const double2*>(&sm.pd1[w][2 * tid]);
      T1s[0] += p1.x;
      T1s[1] += p1.y;
    }
    if constexpr (FW || SB) {
#pragma unroll
      for (int w = 0; w < kIcdWaves; ++w) {
        const double2 p2 = *reinterpret_cast<const double2*>(&sm.pd2[w][2 * tid]);
        T2s[0] += p2.x;
        T2s[1] += p2.y;
      }
    } else {
#pragma unroll
      for (int ph = 0; ph < 2; ++ph) {
        double vsq = 0.0;  // sum_k vlr^2 of the slice
#pragma unroll
        for (int k = 0; k < NV; ++k)
          if (k < vn_s[ph]) vsq = __fma_rn((double)vl_s[ph][k], (double)vl_s[ph][k], vsq);
        T2s[ph] = K * vsq;
      }
    }
    T1s[0] = -T1s[0];  // THETA1 *= -1 (:369)
    T1s[1] = -T1s[1];
  }
  float lsu = 0.f, lsv = 0.f, lmg = 0.f;
  int lcnt = 0;
  // the update of slice y (parity ph) from its in-line neighbours xl (y - 1) and xr
  // (y + 1): writes x, the slice's state and dx; returns the new value (the old one
  // when zero-skipped)
  auto solve = [&](int y, int ph, float xl, float xr) -> float {
    const double T1 = T1s[ph], T2 = T2s[ph];
    const float v = smx[y];
    float s1 = sm.ps1[y], s2 = sm.ps2[y], lo = sm.plo[y], hi = sm.phi[y];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int dy = 2 * h - 1;
      const bool ex = y0 + y + dy >= 0 && y0 + y + dy < Ny;
      const float nb = ex ? (h ? xr : xl) : 0.f;
      lo = fminf(lo, nb);
      hi = fmaxf(hi, nb);
      const float b = filter_weight(dy, 0, 0) * qggmrf_b(v - nb, J);
      s1 = ex ? s1 + b * nb : s1;
      s2 = ex ? s2 + b : s2;
    }
    float xnew = v;
    int vst = 2;
    if (zs_active && v == 0.f && !(hi > 0.f)) {
      // zero-skipped (:311-338)
    } else {
      // the surrogate's minimiser and the clip bounds: two independent divisions
      const double c = (double)v - T1 / (T2 != 0.0 ? T2 : 1.0);
      const double dlo = T2 != 0.0 ? fmin((double)lo, c) : (double)lo;
      const double dhi = T2 != 0.0 ? fmax((double)hi, c) : (double)hi;
      const double u = ((double)s1 + T2 * (double)v - T1) / ((double)s2 + T2);
      double nv = (double)v + (double)J.alpha * (u - (double)v);
      nv = nv < dlo ? dlo : (nv > dhi ? dhi : nv);
      if (nv < 0.0) nv = 0.0;  // POSITIVITY_CONSTRAINT (:394-400)
      xnew = (float)nv;
      gp(J.x)[lbase + y] = xnew;
      vst = 1;
      lmg += fabsf(xnew - v);
      if (roi) {
        lsu += fabsf(xnew - v);
        lsv += fabsf(v);
      }
    }
    lcnt += vst == 2 ? 1 : (1 << 10) + ((NV == 1 ? 1 : vn_s[ph]) << 20);
    sm.vst[y] = (unsigned char)vst;
    sm.dx[y] = xnew - v;
    return xnew;
  };
  // this unit's statistics (:412-421): |dx| and |x| in the ROI, the counts, |dx|
  float su_u = 0.f, sv_u = 0.f, mg_u = 0.f;
  int cn_u = 0;
  if constexpr (YPL <= 2) {  // one wave holds every pair of slices (ny <= 128)
    if (wave == 0) {
      // even slice 2 lane, then odd slice 2 lane + 1, whose upper neighbour is the next
      // lane's new even value (or the slice above the unit): no barrier between
      const int ye = 2 * lane, yo = ye + 1;
      float xe = 0.f;
      if (ye < ny) xe = solve(ye, 0, smx[ye - 1], smx[ye + 1]);
      const float xe_next = __shfl_down(xe, 1);
      if (yo < ny) solve(yo, 1, xe, yo + 1 < ny ? xe_next : smx[yo + 1]);
      // sums in a fixed order
      su_u = wave_sum(lsu);
      sv_u = wave_sum(lsv);
      mg_u = J.magmap ? wave_sum(lmg) : 0.f;
      cn_u = wave_sum(lcnt);
    }
    __syncthreads();  // G: x, dx and the slice states visible
  } else {
    for (int phase = 0; phase < 2; ++phase) {
      const int y = 2 * tid + phase;  // within the unit
      if (y < ny) smx[y] = solve(y, phase, smx[y - 1], smx[y + 1]);
      if (phase == 1 && wave < nsw) {
        // this wave's share of the sweep statistics, summed in a fixed order
        const float wsu = wave_sum(lsu), wsv = wave_sum(lsv);
        const float wmg = J.magmap ? wave_sum(lmg) : 0.f;
        const int wcnt = wave_sum(lcnt);
        if (lane == 0) {
          sm.wsu[wave] = wsu;
          sm.wsv[wave] = wsv;
          sm.wmg[wave] = wmg;
          sm.wcnt[wave] = wcnt;
        }
      }
      __syncthreads();  // F, G
    }
    if (wave == 0) {
      su_u = sum8(lane < nsw ? sm.wsu[lane] : 0.f);
      sv_u = sum8(lane < nsw ? sm.wsv[lane] : 0.f);
      mg_u = J.magmap ? sum8(lane < nsw ? sm.wmg[lane] : 0.f) : 0.f;
      cn_u = sum8(lane < nsw ? sm.wcnt[lane] : 0);
    }
  }

  // ---- 6. sweep statistics, stored at the unit's schedule position (one 16-byte
  // store; the sweep's sums are formed in schedule order after its last step, so no two
  // units meet at one address); issued before the error-sinogram update where the
  // line's latency counts (PF), after it where the device is full ----
  auto line_stats = [&]() {
    if (wave == 0 && lane == 0) {
      typedef float f32x4 __attribute__((ext_vector_type(4)));
      *(gptr<f32x4>)(J.ustat + pos) = f32x4{su_u, sv_u, __int_as_float(cn_u), 0.f};
      // the line's update magnitude (BFUpdateYSlice.cpp:171, 303-304); NHICD lists run
      // whole lines, so the unit is the line
      if (J.magmap) gp(J.magmap)[line] = mg_u;
    }
  };
  if constexpr (PF) line_stats();

  // ---- 7. e -= I0 a (vlr dx) (:423-449): the lane's vlr dx once, then per entry
  // one product and one add (no fma), in every form ----
  float d[YPL], vd[YPL][NV];
  bool upd[YPL];
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy) {
    const int y = yl + TY * jy;
    d[jy] = y < ny ? sm.dx[y] : 0.f;
    upd[jy] = y < ny && sm.vst[y] == 1;
#pragma unroll
    for (int k = 0; k < NV; ++k) vd[jy][k] = __fmul_rn(vv[jy][k], d[jy]);
  }
  auto enew = [&](float e, float ka, int jy, int k) {
    return __fadd_rn(e, -__fmul_rn(ka, vd[jy][k]));
  };
  // scaled entries (..._UpdateVoxels.cpp:438-446): e -= s (I0 a vlr dx)
  auto enew_s = [&](float e, float ka, int jy, int k, float sc) {
    return __fadd_rn(e, -__fmul_rn(__fmul_rn(ka, vd[jy][k]), sc));
  };
  // PF with one batch per stream on a conflict-free step: the stream's error-sinogram
  // values (and the bright-field weights or the scales) are still in registers from
  // the theta pass, and no other line of the launch touches these bins, so the new
  // values go out as plain stores with no second round trip.  Not on the relaxed
  // schedule (e_atomic), whose lines may share bins: a plain store there would drop
  // another line's update.
  bool from_regs = false;
  if constexpr (PF && KR && YPL * NV <= (FW || SB ? 2 : 4)) {  // (wider shapes would spill)
    if (C <= U && !J.e_atomic) {  // uniform
      from_regs = true;
      // per lane and row: the slice changed (updated, bright field) and has that row
      bool pr[YPL][NV];
#pragma unroll
      for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
        for (int k = 0; k < NV; ++k) pr[jy][k] = (FW ? upd[jy] : d[jy] != 0.f) && vok[jy][k];
      // the stores of entries [4 q, 4 q + 4) that lie below cnt (every one in a full group)
      auto store4 = [&](int q, bool full, int jy, int k) {
        const float4 a4 = ka4(4 * q);
        const float aa[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int u = 4 * q + t;
          if (!full && u >= cnt) continue;
          const uint32_t rb = r0[u] * 4u + vtb[jy][k];
          if constexpr (FW) {
            const float en = enew(ev0[u][jy][k], aa[t], jy, k);
            const float aw = fabsf(wf0[u][jy][k]);
            *pto(J.e, rb) = en;
            *pto(J.wfull, rb) = fabsf(en * sqrtf(aw)) < J.bragg ? aw : -aw;
          } else if constexpr (SB) {
            *pto(J.e, rb) = enew_s(ev0[u][jy][k], aa[t], jy, k, wf0[u][jy][k]);
          } else {
            *pto(J.e, rb) = enew(ev0[u][jy][k], aa[t], jy, k);
          }
        }
      };
#pragma unroll
      for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          if (pr[jy][k]) {  // one mask for the whole stream
#pragma unroll
            for (int q = 0; q < U / 4; ++q) {
              if (4 * q + 4 <= cnt)
                store4(q, true, jy, k);
              else if (4 * q < cnt)
                store4(q, false, jy, k);
            }
          }
        }
    }
  }
  // padding entries (zero gains, spare row) are updated like the others: they leave
  // the spare row at zero
  for (int base = 0; !from_regs && base < cnt; base += U) {
    if constexpr (FW) {
      // bright field: every updated slice (even with dx = 0) refreshes the selector
      // of the entries it touches, |e sqrt(w)| < T
      // (BFForwardModel::updateErrorSinogram, :898-932); plain read-modify-write
      float ev[U][YPL][NV], wv[U][YPL][NV];
      load_batch(base, ev, wv);
#pragma unroll
      for (int q = 0; q < U / 4; ++q) {
        if (base + 4 * q < cnt) {
          const uint4 r4 = rows4(base + 4 * q);
          const float4 a4 = ka4(base + 4 * q);
          const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
          const float aa[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
              for (int k = 0; k < NV; ++k)
                if (upd[jy] && vok[jy][k]) {
                  const int u = 4 * q + t;
                  const float en = enew(ev[u][jy][k], aa[t], jy, k);
                  const float aw = fabsf(wv[u][jy][k]);
                  *pto(J.e, rr[t] + vtb[jy][k]) = en;
                  *pto(J.wfull, rr[t] + vtb[jy][k]) = fabsf(en * sqrtf(aw)) < J.bragg ? aw : -aw;
                }
        }
      }
    } else if (J.e_atomic) {
      // relaxed schedule (lines of a step may share detector bins): fp32 atomics
#pragma unroll
      for (int q = 0; q < U / 4; ++q) {
        if (base + 4 * q < cnt) {
          const uint4 r4 = rows4(base + 4 * q);
          const float4 a4 = ka4(base + 4 * q);
          const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
          const float aa[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
              for (int k = 0; k < NV; ++k)
                if (d[jy] != 0.f && vok[jy][k]) {
                  float p = __fmul_rn(aa[t], vd[jy][k]);
                  if constexpr (SB) p = __fmul_rn(p, ldo(J.bfs, rr[t] + vtb[jy][k]));
                  (void)__hip_atomic_fetch_add(pto(J.e, rr[t] + vtb[jy][k]), -p, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                }
        }
      }
    } else {
      // full batches: read-modify-write, every load of a batch in flight before its
      // stores (lanes whose slices did not change load and store nothing)
      float ev[U][YPL][NV], sv[SB ? U : 1][YPL][NV];
      if constexpr (YPL * NV == 1) {
        // one detector row per lane: one lane mask for the whole batch
        if (d[0] != 0.f && vok[0][0]) {
#pragma unroll
          for (int q = 0; q < U / 4; ++q) {
            if (base + 4 * q < cnt) {
              const uint4 r4 = rows4(base + 4 * q);
              const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                ev[4 * q + t][0][0] = ldo(J.e, rr[t] + vtb[0][0]);
                if constexpr (SB) sv[4 * q + t][0][0] = ldo(J.bfs, rr[t] + vtb[0][0]);
              }
            }
          }
#pragma unroll
          for (int q = 0; q < U / 4; ++q) {
            if (base + 4 * q < cnt) {
              const uint4 r4 = rows4(base + 4 * q);
              const float4 a4 = ka4(base + 4 * q);
              const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
              const float aa[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                if constexpr (SB)
                  *pto(J.e, rr[t] + vtb[0][0]) =
                      enew_s(ev[4 * q + t][0][0], aa[t], 0, 0, sv[4 * q + t][0][0]);
                else
                  *pto(J.e, rr[t] + vtb[0][0]) = enew(ev[4 * q + t][0][0], aa[t], 0, 0);
              }
            }
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < U / 4; ++q) {
          if (base + 4 * q < cnt) {
            const uint4 r4 = rows4(base + 4 * q);
            const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
              for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
                for (int k = 0; k < NV; ++k)
                  if (d[jy] != 0.f && vok[jy][k]) {
                    ev[4 * q + t][jy][k] = ldo(J.e, rr[t] + vtb[jy][k]);
                    if constexpr (SB) sv[4 * q + t][jy][k] = ldo(J.bfs, rr[t] + vtb[jy][k]);
                  }
          }
        }
#pragma unroll
        for (int q = 0; q < U / 4; ++q) {
          if (base + 4 * q < cnt) {
            const uint4 r4 = rows4(base + 4 * q);
            const float4 a4 = ka4(base + 4 * q);
            const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
            const float aa[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
              for (int jy = 0; jy < YPL; ++jy)
#pragma unroll
                for (int k = 0; k < NV; ++k)
                  if (d[jy] != 0.f && vok[jy][k]) {
                    if constexpr (SB)
                      *pto(J.e, rr[t] + vtb[jy][k]) =
                          enew_s(ev[4 * q + t][jy][k], aa[t], jy, k, sv[4 * q + t][jy][k]);
                    else
                      *pto(J.e, rr[t] + vtb[jy][k]) = enew(ev[4 * q + t][jy][k], aa[t], jy, k);
                  }
          }
        }
      }
    }
  }

  if constexpr (!PF) line_stats();
}

// INL: the step's lines come in the kernel arguments (`inl`, exactly one per
// workgroup, no padding), so no branch precedes the argument loads and the line's
// entry arrives with the job descriptor.
template <int YPL, int NV, bool FW = false, bool SB = false, int CAP = 0>
__global__ __launch_bounds__(kIcdBlock) void k_icd_step(DevJob J, const LineRef* lines, int nlines,
                                                        int pos0, int flags, StepLines<CAP> inl)
{
  __shared__ IcdSmem<YPL> sm;
  LineRef ref;
  if constexpr (CAP > 0) {
    const int3 c = inl.l[blockIdx.x];
    ref = LineRef{c.x & 0xfffff, c.y, (int)((unsigned)c.x >> 20), c.z};
    // the arguments the line's first loads depend on, fetched in the same round trip
    // as its entry (an empty asm that takes them as operands: no instruction)
    asm volatile(";" ::"s"(ref.line), "s"(ref.p0), "s"(ref.nc), "s"(J.N_x), "s"(J.N_y), "s"(J.NB),
                 "s"(J.N_z), "s"(J.mask), "s"(J.vlr_t0), "s"(J.vlr_n), "s"(J.col_row),
                 "s"(J.col_ka), "s"(J.col_kw), "s"(J.x), "s"(J.e), "s"(J.col_slots),
                 "s"(J.spare_b), "s"(ref.pad), "s"(J.line_kk), "s"(J.nx_div));
  } else {
    if ((int)blockIdx.x >= nlines) return;
    ref = lines[blockIdx.x];
    if (__builtin_amdgcn_readfirstlane(ref.line) < 0) return;  // padding of a graph's step slot
  }
  icd_line<YPL, NV, FW, true, SB, true>(J, ref, flags & 1, (flags >> 1) & 1, pos0 + (int)blockIdx.x,
                                        sm, icd_col(J.col_slots));
}

// A step with more units than the device has compute units: the prefetching form
// without the register-held batch, compiled for two resident workgroups per compute
// unit (at most 128 registers), so the step's workgroups run in one round.
template <int YPL, int NV, bool FW = false, bool SB = false>
__global__ __launch_bounds__(kIcdBlock) __attribute__((amdgpu_waves_per_eu(4)))
void k_icd_step_wide(DevJob J, const LineRef* lines, int nlines, int pos0, int flags)
{
  __shared__ IcdSmem<YPL> sm;
  if ((int)blockIdx.x >= nlines) return;
  const LineRef ref = lines[blockIdx.x];
  if (__builtin_amdgcn_readfirstlane(ref.line) < 0) return;  // padding of a graph's step slot
  icd_line<YPL, NV, FW, true, SB, false>(J, ref, flags & 1, (flags >> 1) & 1,
                                         pos0 + (int)blockIdx.x, sm, icd_col(J.col_slots));
}

// One conflict-free step of every job of a batch that has kernel shape
// (YPL, NV): entries [e0, e0 + ne) of the sweep's entry table; workgroup g serves
// entry e with wg_off[e] <= g < wg_off[e] + nlines[e].  Jobs never share memory,
// so one launch needs no ordering between them.
template <int YPL, int NV, bool PF, bool FW = false>
__global__ __launch_bounds__(kIcdBlock) __attribute__((amdgpu_waves_per_eu(icd_waves_per_eu<YPL, NV>(PF || FW))))
void k_icd_batch_step(const DevJob* __restrict__ jobs,
                      const BatchEntry* __restrict__ ent, int e0, int ne, int slots)
{
  __shared__ IcdSmem<YPL> sm;
  const int g = blockIdx.x;
  int lo = e0, hi = e0 + ne - 1;
  while (lo < hi) {  // last entry with wg_off <= g
    const int mid = (lo + hi + 1) >> 1;
    if (ent[mid].wg_off <= g) lo = mid;
    else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const BatchEntry& E = ent[lo];
  const int i = g - E.wg_off;
  if (i < 0 || i >= E.nlines) return;
  const DevJob J = jobs[__builtin_amdgcn_readfirstlane(E.job)];
  const LineRef ref = load_ref(E.lines + i);
  icd_line<YPL, NV, FW, PF>(J, ref, E.flags & 1, (E.flags >> 1) & 1,
                            __builtin_amdgcn_readfirstlane(E.pos) + i, sm, icd_col(slots));
}

// One launch of the request server: entries travel by value in the kernel
// arguments (no host-to-device copy per launch), each naming a job slot and a
// contiguous run of its current sweep's schedule.  FW: bright-field jobs (per-entry
// weights and selector, one detector row per slice).
template <int YPL, int NV, bool PF, bool FW = false>
__global__ __launch_bounds__(kIcdBlock) __attribute__((amdgpu_waves_per_eu(icd_waves_per_eu<YPL, NV>(PF || FW))))
void k_icd_tick(const DevJob* __restrict__ jobs, const TickArgs args)
{
  __shared__ IcdSmem<YPL> sm;
  const int g = blockIdx.x;
  int lo = 0, hi = args.n - 1;
  while (lo < hi) {  // last entry with wg_off <= g
    const int mid = (lo + hi + 1) >> 1;
    if (args.e[mid].wg_off <= g) lo = mid;
    else hi = mid - 1;
  }
  // the entry is uniform over the workgroup: keep its fields and the job
  // descriptor's address in scalar registers
  lo = __builtin_amdgcn_readfirstlane(lo);
  const int job = __builtin_amdgcn_readfirstlane(args.e[lo].job);
  const int i = g - __builtin_amdgcn_readfirstlane(args.e[lo].wg_off);
  const int nl = __builtin_amdgcn_readfirstlane(args.e[lo].nlines);
  const int l0 = __builtin_amdgcn_readfirstlane(args.e[lo].line0);
  const int flags = __builtin_amdgcn_readfirstlane(args.e[lo].flags);
  if (i < 0 || i >= nl) return;
  const DevJob J = jobs[job];
  const LineRef ref = load_ref(J.sched + l0 + i);
  icd_line<YPL, NV, FW, PF>(J, ref, flags & 1, (flags >> 1) & 1, l0 + i, sm, icd_col(args.slots));
}

// jobs[slot] = J: a job descriptor written in stream order.
__global__ void k_set_devjob(DevJob* jobs, int slot, const DevJob J)
{
  if (threadIdx.x == 0) jobs[slot] = J;
}

// ---------------------------------------------------------------------------
// Joint gain/offset estimation, FP64 per-view sums
// (HAADF/HAADF_ReconstructionEngine_Extra.cpp:443-498): with Ŷ = (y - e - mu)/I0,
// sums6[th] = {Σ y w, Σ w, Σ Ŷ w, Σ Ŷ w y, Σ y² w, Σ Ŷ² w}.
__global__ __launch_bounds__(kBlock) void k_joint_sums(DevJob J, double* sums6)
{
  __shared__ double red[kBlock];
  const int th = blockIdx.x;
  const size_t per = (size_t)J.N_r * J.N_t, base = (size_t)th * per;
  const double I0 = J.I0[th], mu = J.mu[th];
  const double wv = (double)J.wview[th];
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0;
  for (size_t k = threadIdx.x; k < per; k += blockDim.x) {
    const double yv = J.y[base + k];
    const double ye = (yv - (double)J.e[base + k] - mu) / I0;
    const double w = J.wfull ? (double)J.wfull[base + k] : wv;
    a0 += yv * w;
    a1 += w;
    a2 += ye * w;
    a3 += ye * w * yv;
    a4 += yv * yv * w;
    a5 += ye * ye * w;
  }
  double r[6] = {a0, a1, a2, a3, a4, a5};
  for (int i = 0; i < 6; ++i) {
    const double t = block_sum_d(r[i], red);
    if (threadIdx.x == 0) sums6[th * 6 + i] = t;
  }
}

// Lagrange-constrained 2x2 solves for all views (…_Extra.cpp:550-582).
__global__ __launch_bounds__(kBlock) void k_joint_solve(int N, const double* sums6, double target,
                                                        double* I0_new, double* mu_new)
{
  __shared__ double red[kBlock];
  double s1 = 0, s2 = 0;
  for (int th = threadIdx.x; th < N; th += blockDim.x) {
    const double* q = sums6 + th * 6;
    const double num = q[0], den = q[1], a = q[2], b = q[3], d = q[5];
    const double d1 = num / den, d2 = a / den;
    const double D = d - a * d2;
    s1 += 1.0 / D;
    s2 += (b - a * d1) / D;
  }
  const double sum1 = block_sum_d(s1, red);
  const double sum2 = block_sum_d(s2, red);
  const double lambda = (-N * target + sum2) / sum1;
  for (int th = threadIdx.x; th < N; th += blockDim.x) {
    const double* q = sums6 + th * 6;
    const double num = q[0], den = q[1], a = q[2], b = q[3], d = q[5];
    const double d1 = num / den, d2 = a / den;
    const double I = (-1 * lambda - a * d1 + b) / (d - a * d2);
    I0_new[th] = I;
    mu_new[th] = d1 - d2 * I;
  }
}

// e = y - mu_new - I0_new * Ŷ with Ŷ = (y - e - mu_old) / I0_old (…_Extra.cpp:601-614).
__global__ void k_joint_apply(DevJob J, const double* I0o, const double* muo, const double* I0n,
                              const double* mun)
{
  const size_t per = (size_t)J.N_r * J.N_t, M = per * J.N_theta;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int th = (int)(i / per);
    const double yv = J.y[i];
    const double ye = (yv - (double)J.e[i] - muo[th]) / I0o[th];
    J.e[i] = (float)(yv - mun[th] - I0n[th] * ye);
  }
}

// Offset-only estimation of the HAADF engine given bright-field data
// (HAADF/HAADF_ReconstructionEngine_Extra.cpp:636-669): per view the weighted mean
// a = Σ e w / Σ w of the error sinogram, FP64, one block per view.  Under the
// identity noise model the weight of every entry of a view is its w[th].
__global__ __launch_bounds__(kBlock) void k_offset_sums(DevJob J, double* nd)
{
  __shared__ double red[kBlock];
  const int th = blockIdx.x;
  const size_t per = (size_t)J.N_r * J.N_t, base = (size_t)th * per;
  const double w = (double)J.wview[th];
  double num = 0, den = 0;
  for (size_t k = threadIdx.x; k < per; k += blockDim.x) {
    num += (double)J.e[base + k] * w;
    den += w;
  }
  num = block_sum_d(num, red);
  den = block_sum_d(den, red);
  if (threadIdx.x == 0) {
    nd[2 * th] = num;
    nd[2 * th + 1] = den;
  }
}

// e -= a_th, mu_th += a_th (:656-665)
__global__ void k_offset_apply(DevJob J, const double* nd, double* mu)
{
  const size_t per = (size_t)J.N_r * J.N_t, M = per * J.N_theta;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int th = (int)(i / per);
    const double a = nd[2 * th] / nd[2 * th + 1];
    J.e[i] = (float)((double)J.e[i] - a);
    if (i == (size_t)th * per) mu[th] += a;
  }
}

// updateWeights, identity noise model (…_Extra.cpp:795-866): alpha = mean e², w = 1/alpha.
__global__ __launch_bounds__(kBlock) void k_noise_update(DevJob J, double* alpha, float* wview)
{
  __shared__ double red[kBlock];
  const int th = blockIdx.x;
  const size_t per = (size_t)J.N_r * J.N_t, base = (size_t)th * per;
  double s = 0;
  for (size_t k = threadIdx.x; k < per; k += blockDim.x) {
    const double ev = J.e[base + k];
    s += ev * ev;
  }
  s = block_sum_d(s, red);
  if (threadIdx.x == 0) {
    const double a = s / (double)(J.N_r * J.N_t);
    alpha[th] = a;
    wview[th] = (float)(1.0 / a);
  }
}

// computeCost (HAADF/HAADF_ReconstructionEngine.cpp:1147-1364), FP64: per-block
// partials of the data term and the 13-forward-pair QGGMRF prior.
__device__ inline double qggmrf_value_d(double d, const DevJob& J)
{
  const double a = fabs(d);
  return ((a * a) / J.sigma_pd) / (J.mrf_cd + pow(a, 2.0 - J.mrf_qd) / J.sigma_pqd);
}

// The QGGMRF terms of voxel v's 13 forward neighbour pairs, each difference scaled by
// s: the prior sum of computeCost (HAADF/HAADF_ReconstructionEngine.cpp:1147-1364,
// s = 1) and of estimateSigmaX (:1827-1945, s = sigma_x), with FILTER weights 0.0523
// (faces), 0.0370 (edges), 0.0302 (corners).
__device__ inline double forward_pairs(const DevJob& J, size_t v, double s)
{
  const int Ny = J.N_y, Nx = J.N_x, Nz = J.N_z;
  const int k = (int)(v % Ny);
  const int j = (int)((v / Ny) % Nx);
  const int i = (int)(v / ((size_t)Ny * Nx));
  const double xv = J.x[v];
  double prior = 0.0;
#define XV(a, b, c) (double)J.x[((size_t)(a) * Nx + (b)) * Ny + (c)]
#define ACC(a, b, c, f) prior += (f) * qggmrf_value_d((xv - XV(a, b, c)) * s, J)
  const double fF = 0.0523, fE = 0.0370, fC = 0.0302;
  if (k + 1 < Ny) ACC(i, j, k + 1, fF);
  if (j + 1 < Nx) {
    if (k - 1 >= 0) ACC(i, j + 1, k - 1, fE);
    ACC(i, j + 1, k, fF);
    if (k + 1 < Ny) ACC(i, j + 1, k + 1, fE);
  }
  if (i + 1 < Nz) {
    if (j - 1 >= 0) ACC(i + 1, j - 1, k, fE);
    ACC(i + 1, j, k, fF);
    if (j + 1 < Nx) ACC(i + 1, j + 1, k, fE);
    if (j - 1 >= 0) {
      if (k - 1 >= 0) ACC(i + 1, j - 1, k - 1, fC);
      if (k + 1 < Ny) ACC(i + 1, j - 1, k + 1, fC);
    }
    if (k - 1 >= 0) ACC(i + 1, j, k - 1, fE);
    if (j + 1 < Nx) {
      if (k - 1 >= 0) ACC(i + 1, j + 1, k - 1, fC);
      if (k + 1 < Ny) ACC(i + 1, j + 1, k + 1, fC);
    }
    if (k + 1 < Ny) ACC(i + 1, j, k + 1, fE);
  }
#undef ACC
#undef XV
  return prior;
}

__global__ __launch_bounds__(kBlock) void k_cost_partials(DevJob J, double* part)
{
  __shared__ double red[kBlock];
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t, per = (size_t)J.N_r * J.N_t;
  const size_t V = (size_t)J.N_z * J.N_x * J.N_y;
  const size_t g0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x, gs = (size_t)gridDim.x * blockDim.x;
  double data = 0, prior = 0;
  for (size_t i = g0; i < M; i += gs) {
    const double ev = J.e[i];
    if (J.wfull) {
      // BFForwardModel::forwardCost (:812-846): Bragg rejector beyond T
      const double w = (double)J.wfull[i];
      if (w >= 0.0) {
        data += ev * ev * w;
      } else {
        const double T = J.bragg_d, dl = J.delta_d;
        data += 2.0 * T * dl * fabs(ev) * sqrt(-w) + T * T * (1.0 - 2.0 * dl);
      }
    } else {
      data += ev * ev * (double)J.wview[i / per];
    }
  }
  for (size_t v = g0; v < V; v += gs) prior += forward_pairs(J, v, 1.0);
  data = block_sum_d(data, red);
  prior = block_sum_d(prior, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2 + 0] = data;
    part[blockIdx.x * 2 + 1] = prior;
  }
}

__global__ __launch_bounds__(kBlock) void k_cost_final(DevJob J, const double* part, int nparts,
                                                       const double* alpha, int noise_model,
                                                       double* out3)
{
  __shared__ double red[kBlock];
  double d = 0, p = 0, n = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    d += part[b * 2 + 0];
    p += part[b * 2 + 1];
  }
  const double per = (double)J.N_r * (double)J.N_t;
  if (noise_model && !J.wfull)
    for (int th = threadIdx.x; th < J.N_theta; th += blockDim.x) {
      const double w = (double)J.wview[th];
      if (w != 0.0) n += per * log(2.0 * 3.14159265358979323846 * (1.0 / w));
    }
  d = block_sum_d(d, red);
  p = block_sum_d(p, red);
  n = block_sum_d(n, red);
  if (threadIdx.x == 0) {
    out3[0] = d / 2;
    out3[1] = p;
    out3[2] = n / 2;
    if (noise_model && J.wfull) {
      // bright field: (N_theta N_r N_t / 2) log(alpha[0]), integer halving (:837-841)
      const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
      out3[2] = (double)(M / 2) * log(alpha[0]);
    }
  }
}

// estimateSigmaX (HAADF/HAADF_ReconstructionEngine.cpp:1827-1945): the forward-pair
// prior sum, differences scaled by sigma_x, over the planes z0 <= z < z1; FP64 block
// partials, then one block adds them in a fixed order.
__global__ __launch_bounds__(kBlock) void k_sigma_partials(DevJob J, int z0, int z1, double sx,
                                                           double* part)
{
  __shared__ double red[kBlock];
  const size_t plane = (size_t)J.N_x * J.N_y;
  const size_t v0 = (size_t)z0 * plane, v1 = (size_t)z1 * plane;
  double acc = 0.0;
  for (size_t v = v0 + blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < v1;
       v += (size_t)gridDim.x * blockDim.x)
    acc += forward_pairs(J, v, sx);
  acc = block_sum_d(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kBlock) void k_sum_final(const double* part, int nparts, double* out)
{
  __shared__ double red[kBlock];
  double a = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) a += part[b];
  a = block_sum_d(a, red);
  if (threadIdx.x == 0) *out = a;
}

// Bright field, offset (log dose) per view: the minimiser of the selector's
// half-quadratic surrogate, rejected entries weighing delta T / (|e| sqrt(w))
// (BFForwardModel::jointEstimation, :396-418).  FP64 sums, one block per view.
__global__ __launch_bounds__(kBlock) void k_bf_offset_sums(DevJob J, double* nd)
{
  __shared__ double red[kBlock];
  const int th = blockIdx.x;
  const size_t per = (size_t)J.N_r * J.N_t, base = (size_t)th * per;
  const double Td = J.bragg_d * J.delta_d;
  double num = 0, den = 0;
  for (size_t k = threadIdx.x; k < per; k += blockDim.x) {
    const double ev = J.e[base + k], w = J.wfull[base + k];
    if (w >= 0.0) {
      num += ev * w;
      den += w;
    } else {
      num += (Td * ev / fabs(ev)) * sqrt(-w);
      den += (Td * sqrt(-w)) / fabs(ev);
    }
  }
  num = block_sum_d(num, red);
  den = block_sum_d(den, red);
  if (threadIdx.x == 0) {
    nd[2 * th] = num;
    nd[2 * th + 1] = den;
  }
}

// e -= a_th, mu_th += a_th (:420-431), then updateSelector: e^2 w < T^2 (:517-531).
__global__ void k_bf_offset_apply(DevJob J, const double* nd, double* mu)
{
  const size_t per = (size_t)J.N_r * J.N_t, M = per * J.N_theta;
  const double T2 = J.bragg_d * J.bragg_d;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int th = (int)(i / per);
    const double a = nd[2 * th] / nd[2 * th + 1];
    const float en = (float)((double)J.e[i] - a);
    const float aw = fabsf(J.wfull[i]);
    J.e[i] = en;
    J.wfull[i] = (double)en * (double)en * (double)aw < T2 ? aw : -aw;
    if (i == (size_t)th * per) mu[th] += a;
  }
}

// Shared variance (:447-511): with the unscaled weights w_u = w alpha_th,
// alpha = (sum_sel e^2 w_u + delta T sum_rej |e| sqrt(alpha_th w_u)) / M.
__global__ __launch_bounds__(kBlock) void k_bf_noise_partials(DevJob J, const double* alpha,
                                                              double* part)
{
  __shared__ double red[kBlock];
  const size_t per = (size_t)J.N_r * J.N_t, M = per * J.N_theta;
  double s1 = 0, s2 = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const double at = alpha[i / per];
    const double ev = J.e[i], w = J.wfull[i];
    const double wu = fabs(w) * at;
    if (w >= 0.0) s1 += ev * ev * wu;
    else s2 += fabs(ev) * sqrt(at * wu);
  }
  s1 = block_sum_d(s1, red);
  s2 = block_sum_d(s2, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

__global__ __launch_bounds__(kBlock) void k_bf_noise_final(DevJob J, const double* part,
                                                           int nparts, double* alpha_new)
{
  __shared__ double red[kBlock];
  double s1 = 0, s2 = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    s1 += part[2 * b];
    s2 += part[2 * b + 1];
  }
  s1 = block_sum_d(s1, red);
  s2 = block_sum_d(s2, red);
  const double M = (double)J.N_theta * (double)J.N_r * (double)J.N_t;
  const double upd = (s1 + (J.delta_d * J.bragg_d) * s2) / M;
  for (int th = threadIdx.x; th < J.N_theta; th += blockDim.x) alpha_new[th] = upd;
}

// w = w_u / alpha_new, then updateSelector (:503-510, :517-531).
__global__ void k_bf_noise_apply(DevJob J, const double* alpha_old, const double* alpha_new)
{
  const size_t per = (size_t)J.N_r * J.N_t, M = per * J.N_theta;
  const double T2 = J.bragg_d * J.bragg_d;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int th = (int)(i / per);
    const double wn = (fabs((double)J.wfull[i]) * alpha_old[th]) / alpha_new[th];
    const double ev = J.e[i];
    const float aw = (float)wn;
    J.wfull[i] = ev * ev * wn < T2 ? aw : -aw;
  }
}

__global__ void k_fill_f32(float* p, float v, size_t n)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
// Sum of the volume over the ROI lines (BFReconstructionEngine_Extra.cpp:1216-1232).
__global__ __launch_bounds__(kBlock) void k_roi_sum(DevJob J, double* out)
{
  __shared__ double red[kBlock];
  double s = 0.0;
  const int L = J.N_z * J.N_x;
  for (int l = blockIdx.x; l < L; l += gridDim.x) {
    if (!J.mask[l]) continue;  // uniform over the block
    for (int y = threadIdx.x; y < J.N_y; y += blockDim.x) s += J.x[(size_t)l * J.N_y + y];
  }
  s = block_sum_d(s, red);
  if (threadIdx.x == 0 && s != 0.0) atomicAdd(out, s);
}

__global__ void k_fill_u32(unsigned int* p, unsigned int v, size_t n)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, size_t n)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
static int grid_for(size_t n)
{
  size_t g = (n + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_line_nnz(const DevJob& J, int* line_nnz, float* x, hipStream_t s)
{
  hipLaunchKernelGGL(k_line_nnz, dim3(J.N_z * J.N_x), dim3(kBlock), 0, s, J, line_nnz, x);
  return hipGetLastError();
}

hipError_t launch_forward_project(const DevJob& J, const float* vol, double* out, hipStream_t s)
{
  hipLaunchKernelGGL(k_forward_project, dim3(J.N_z * J.N_x), dim3(kBlock), 0, s, J, vol, out);
  return hipGetLastError();
}

// fine[z][x][y] = coarse[z/2][x/2][y/2] for z < 2 (N_z/2), x < 2 (N_x/2), y < 2 (N_y/2),
// 0 elsewhere (the last plane of an odd size); sizes are the fine volume's
// (GenericFilters/InitialReconstructionBinReader.cpp:103-130).  Lines whose column is
// empty stay 0, as the setup leaves them (k_line_nnz).
__global__ __launch_bounds__(kBlock) void k_upsample2(const float* __restrict__ coarse,
                                                      float* __restrict__ fine,
                                                      const int* __restrict__ line_nnz, int Nz,
                                                      int Nx, int Ny)
{
  const int ze = Nz / 2, xe = Nx / 2, ye = Ny / 2;
  const size_t V = (size_t)Nz * Nx * Ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < V;
       i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i % Ny), x = (int)((i / Ny) % Nx), z = (int)(i / ((size_t)Ny * Nx));
    float v = 0.f;
    if (z < 2 * ze && x < 2 * xe && y < 2 * ye && line_nnz[(size_t)z * Nx + x] > 0)
      v = coarse[((size_t)(z >> 1) * xe + (x >> 1)) * ye + (y >> 1)];
    fine[i] = v;
  }
}

hipError_t launch_upsample2(const float* coarse, float* fine, const int* line_nnz, int Nz, int Nx,
                            int Ny, hipStream_t s)
{
  const size_t V = (size_t)Nz * Nx * Ny;
  hipLaunchKernelGGL(k_upsample2, dim3(grid_for(V)), dim3(kBlock), 0, s, coarse, fine, line_nnz,
                     Nz, Nx, Ny);
  return hipGetLastError();
}

hipError_t launch_init_error(const DevJob& J, const double* yest, hipStream_t s)
{
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
  hipLaunchKernelGGL(k_init_error, dim3(grid_for(M)), dim3(kBlock), 0, s, J, yest);
  return hipGetLastError();
}

hipError_t launch_build_columns(const DevJob& J, int* row, float* a, hipStream_t s)
{
  hipLaunchKernelGGL(k_build_columns, dim3(J.N_z * J.N_x), dim3(kIcdBlock), 0, s, J, row, a);
  return hipGetLastError();
}

hipError_t launch_column_weights(const DevJob& J, float* ka, float* kw, int nnz, hipStream_t s)
{
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_column_weights, dim3(grid_for((size_t)nnz)), dim3(kBlock), 0, s, J, ka,
                     kw, nnz);
  return hipGetLastError();
}

hipError_t launch_line_kk(const DevJob& J, double* kk, hipStream_t s)
{
  const int L = J.N_z * J.N_x, per = kBlock / 64;
  hipLaunchKernelGGL(k_line_kk, dim3((L + per - 1) / per), dim3(kBlock), 0, s, J, J.col_ptr, kk, L);
  return hipGetLastError();
}

hipError_t launch_export_column(const DevJob& J, int z, int x, float* vals, int* rows,
                                hipStream_t s)
{
  hipLaunchKernelGGL(k_export_column, dim3(1), dim3(kBlock), 0, s, J, z, x, vals, rows);
  return hipGetLastError();
}

void icd_shape(const DevJob& J, int max_vlr, int* ypl, int* nv)
{
  const int ny = J.unit_max > 0 ? J.unit_max : J.N_y;
  *ypl = ny <= 64 ? 1 : (ny <= 128 ? 2 : 4);
  *nv = max_vlr <= 1 ? 1 : (max_vlr == 2 ? 2 : 4);
}

template <int YPL, int CAP>
static void launch_step_nv(const DevJob& J, const LineRef* lines, int nlines, int pos0, int flags,
                           hipStream_t s, const StepLines<CAP>& inl)
{
  const size_t dyn = icd_col_bytes(J.col_slots);
  const dim3 grid(nlines), block(kIcdBlock);
  if (J.wfull && J.vlr_max <= 1)  // bright field: per-entry weights and selector
    hipLaunchKernelGGL((k_icd_step<YPL, 1, true, false, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
  else if (J.wfull && J.vlr_max <= 2)  // bright field, two detector rows per slice
    hipLaunchKernelGGL((k_icd_step<YPL, 2, true, false, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
  else if (J.wfull)  // bright field, three or four rows per slice (the coarsest levels)
    hipLaunchKernelGGL((k_icd_step<YPL, 4, true, false, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
  else if (J.bfs && J.vlr_max <= 1)  // HAADF with a bright-field sinogram
    hipLaunchKernelGGL((k_icd_step<YPL, 1, false, true, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
  else if (J.bfs)
    hipLaunchKernelGGL((k_icd_step<YPL, 2, false, true, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
  else if (J.vlr_max <= 1)
    hipLaunchKernelGGL((k_icd_step<YPL, 1, false, false, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
  else if (J.vlr_max == 2)
    hipLaunchKernelGGL((k_icd_step<YPL, 2, false, false, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
  else
    hipLaunchKernelGGL((k_icd_step<YPL, 4, false, false, CAP>), grid, block, dyn, s, J, lines,
                       nlines, pos0, flags, inl);
}

template <int CAP>
static void launch_step_ypl(const DevJob& J, const LineRef* lines, int nlines, int pos0,
                            int flags, hipStream_t s, const StepLines<CAP>& inl)
{
  int ypl, nv;
  icd_shape(J, J.vlr_max, &ypl, &nv);
  if (ypl == 1)
    launch_step_nv<1, CAP>(J, lines, nlines, pos0, flags, s, inl);
  else if (ypl == 2)
    launch_step_nv<2, CAP>(J, lines, nlines, pos0, flags, s, inl);
  else
    launch_step_nv<4, CAP>(J, lines, nlines, pos0, flags, s, inl);
}

template <int YPL>
static void launch_step_wide(const DevJob& J, const LineRef* lines, int nlines, int pos0,
                             int flags, hipStream_t s)
{
  const size_t dyn = icd_col_bytes(J.col_slots);
  const dim3 grid(nlines), block(kIcdBlock);
#define MBIR_WIDE(NV, FW, SB)                                                                 \
  hipLaunchKernelGGL((k_icd_step_wide<YPL, NV, FW, SB>), grid, block, dyn, s, J, lines, nlines, \
                     pos0, flags)
  if (J.wfull && J.vlr_max <= 1)
    MBIR_WIDE(1, true, false);
  else if (J.wfull && J.vlr_max <= 2)
    MBIR_WIDE(2, true, false);
  else if (J.wfull)
    MBIR_WIDE(4, true, false);
  else if (J.bfs && J.vlr_max <= 1)
    MBIR_WIDE(1, false, true);
  else if (J.bfs)
    MBIR_WIDE(2, false, true);
  else if (J.vlr_max <= 1)
    MBIR_WIDE(1, false, false);
  else if (J.vlr_max == 2)
    MBIR_WIDE(2, false, false);
  else
    MBIR_WIDE(4, false, false);
#undef MBIR_WIDE
}

// compute units of the current device (all devices of a process alike), once
static int device_cus()
{
  // read once (a function-local static: initialised once even when the server's setup
  // and tick threads call this together)
  static const int n = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      return v;
    return 256;
  }();
  return n;
}

// the wide form within 128 registers (lane rows times detector rows at most 4)
static bool wide_ok(const DevJob& J)
{
  int ypl, nv;
  icd_shape(J, J.vlr_max, &ypl, &nv);
  return ypl * nv <= 4;
}

// kernel arguments of the step kernel: the job descriptor, the entries and a few scalars
// must stay within the 4 KB argument limit
static_assert(sizeof(DevJob) + sizeof(StepLines<kInlineLines>) + 64 <= 4096,
              "step kernel arguments exceed 4 KB");

template <int CAP>
static void launch_inline(const DevJob& J, const LineRef* lines, const LineRef* host_lines,
                          int nlines, int pos0, int flags, hipStream_t s)
{
  StepLines<CAP> inl;  // entries past nlines are never read
  for (int i = 0; i < nlines; ++i)
    inl.l[i] = make_int3(host_lines[i].line | (host_lines[i].nc << 20), host_lines[i].p0,
                         host_lines[i].pad);
  launch_step_ypl<CAP>(J, lines, nlines, pos0, flags, s, inl);
}

hipError_t launch_icd_step(const DevJob& J, const LineRef* lines, int nlines, int pos0, int outer,
                           int inner, int zero_skipping, hipStream_t s, const LineRef* host_lines)
{
  if (nlines <= 0) return hipSuccess;
  const int flags = ((outer == 0 && inner == 0) ? 1 : 0) | (zero_skipping ? 2 : 0);
  // the units travel in the arguments for jobs of fewer than 2^20 lines and columns of
  // fewer than 4096 entries (12 bytes each)
  bool inl = host_lines && nlines <= kInlineLines && (int64_t)J.N_z * J.N_x < (1 << 20);
  for (int i = 0; inl && i < nlines; ++i) inl = host_lines[i].nc < 4096;
  if (inl && nlines <= kInlineSmall)
    launch_inline<kInlineSmall>(J, lines, host_lines, nlines, pos0, flags, s);
  else if (inl && nlines <= device_cus())
    launch_inline<kInlineLines>(J, lines, host_lines, nlines, pos0, flags, s);
  else if (nlines > device_cus() && wide_ok(J)) {
    int ypl, nv;
    icd_shape(J, J.vlr_max, &ypl, &nv);
    if (ypl == 1)
      launch_step_wide<1>(J, lines, nlines, pos0, flags, s);
    else if (ypl == 2)
      launch_step_wide<2>(J, lines, nlines, pos0, flags, s);
    else
      launch_step_wide<4>(J, lines, nlines, pos0, flags, s);
  } else {
    static const StepLines<0> none{};
    launch_step_ypl<0>(J, lines, nlines, pos0, flags, s, none);
  }
  return hipGetLastError();
}

// The sweep's stop-test sums (…_UpdateVoxels.cpp:417-421, 765-770) from its units' shares
// in schedule order, FP64: block b sums positions [b chunk, (b + 1) chunk), then one
// block adds the blocks' sums in order (the same sums whichever unit finished first).
__global__ __launch_bounds__(kBlock) void k_sweep_partials(const LineRef* sched, const UnitStat* us,
                                                           int n, int chunk, double* part)
{
  __shared__ double red[kBlock];
  double su = 0, sv = 0, nz = 0, nu = 0, nn = 0;
  const int b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const LineRef r = sched[i];
    if (r.line < 0) continue;  // a graph step's padding
    const UnitStat u = us[i];
    su += (double)u.su;
    sv += (double)u.sv;
    nz += (double)(u.cnt & 1023);
    nu += (double)((u.cnt >> 10) & 1023);
    nn += (double)r.nc * (double)(u.cnt >> 20);
  }
  const double v[5] = {su, sv, nz, nu, nn};
  for (int q = 0; q < 5; ++q) {
    const double t = block_sum_d(v[q], red);
    if (threadIdx.x == 0) part[blockIdx.x * 5 + q] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_sweep_final(const double* part, int nb, SweepAccum* acc)
{
  if (threadIdx.x != 0) return;
  double t[5] = {0, 0, 0, 0, 0};
  for (int b = 0; b < nb; ++b)
    for (int q = 0; q < 5; ++q) t[q] += part[b * 5 + q];
  SweepAccum a{};
  a.sum_abs_update = t[0];
  a.sum_abs_value = t[1];
  a.zero_skipped = (unsigned long long)t[2];  // counts: exact in FP64 below 2^53
  a.updated = (unsigned long long)t[3];
  a.nnz_updated = (unsigned long long)t[4];
  *acc = a;
}

hipError_t launch_sweep_reduce(const LineRef* sched, const UnitStat* ustat, int n, SweepAccum* acc,
                               double* part, hipStream_t s)
{
  const int chunk = std::max(2048, (n + kReduceBlocks - 1) / kReduceBlocks);
  const int nb = std::max(1, (n + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_sweep_partials, dim3(nb), dim3(kBlock), 0, s, sched, ustat, n, chunk, part);
  hipLaunchKernelGGL(k_sweep_final, dim3(1), dim3(64), 0, s, (const double*)part, nb, acc);
  return hipGetLastError();
}

// The caller picks the form of the line update (pf: prefetching; both compute the same
// values) from its tuning table.
hipError_t launch_icd_batch_step(int ypl, int nv, const DevJob* jobs, const BatchEntry* ent,
                                 int e0, int ne, int nwg, int slots, int pf, hipStream_t s, int fw)
{
  if (ne <= 0 || nwg <= 0) return hipSuccess;
  const size_t dyn = icd_col_bytes(slots);
  if (fw) {  // bright field: one detector row per slice
    if (nv != 1 || ypl > 2) return hipErrorInvalidValue;
#define MBIR_BATCHF(Y)                                                                        \
  if (pf)                                                                                     \
    hipLaunchKernelGGL((k_icd_batch_step<Y, 1, true, true>), dim3(nwg), dim3(kIcdBlock), dyn, \
                       s, jobs, ent, e0, ne, slots);                                          \
  else                                                                                        \
    hipLaunchKernelGGL((k_icd_batch_step<Y, 1, false, true>), dim3(nwg), dim3(kIcdBlock), dyn,\
                       s, jobs, ent, e0, ne, slots);
    if (ypl == 1) {
      MBIR_BATCHF(1)
    } else {
      MBIR_BATCHF(2)
    }
#undef MBIR_BATCHF
    return hipGetLastError();
  }
  // the batched forms cover units of at most 128 slices and two detector rows per slice
  if (ypl > 2 || nv > 2) return hipErrorInvalidValue;
#define MBIR_BATCH(Y, N)                                                                      \
  if (pf)                                                                                     \
    hipLaunchKernelGGL((k_icd_batch_step<Y, N, true>), dim3(nwg), dim3(kIcdBlock), dyn, s,    \
                       jobs, ent, e0, ne, slots);                                             \
  else                                                                                        \
    hipLaunchKernelGGL((k_icd_batch_step<Y, N, false>), dim3(nwg), dim3(kIcdBlock), dyn, s,   \
                       jobs, ent, e0, ne, slots);
  if (ypl == 1 && nv == 1) {
    MBIR_BATCH(1, 1)
  } else if (ypl == 1) {
    MBIR_BATCH(1, 2)
  } else if (nv == 1) {
    MBIR_BATCH(2, 1)
  } else {
    MBIR_BATCH(2, 2)
  }
#undef MBIR_BATCH
  return hipGetLastError();
}

hipError_t launch_icd_tick(int ypl, int nv, const DevJob* jobs, const TickArgs& args, int nwg,
                           int pf, hipStream_t s, int fw)
{
  if (args.n <= 0 || args.n > kTickMax || nwg <= 0) return hipSuccess;
  const size_t dyn = icd_col_bytes(args.slots);
  if (fw) {  // bright field: one detector row per slice
    if (nv != 1 || ypl > 2) return hipErrorInvalidValue;
#define MBIR_TICKF(Y)                                                                         \
  if (pf)                                                                                     \
    hipLaunchKernelGGL((k_icd_tick<Y, 1, true, true>), dim3(nwg), dim3(kIcdBlock), dyn, s,    \
                       jobs, args);                                                           \
  else                                                                                        \
    hipLaunchKernelGGL((k_icd_tick<Y, 1, false, true>), dim3(nwg), dim3(kIcdBlock), dyn, s,   \
                       jobs, args);
    if (ypl == 1) {
      MBIR_TICKF(1)
    } else {
      MBIR_TICKF(2)
    }
#undef MBIR_TICKF
    return hipGetLastError();
  }
  // the server's forms cover units of at most 128 slices and two detector rows per slice
  if (ypl > 2 || nv > 2) return hipErrorInvalidValue;
#define MBIR_TICK(Y, N)                                                                        \
  if (pf)                                                                                      \
    hipLaunchKernelGGL((k_icd_tick<Y, N, true>), dim3(nwg), dim3(kIcdBlock), dyn, s, jobs,      \
                       args);                                                                  \
  else                                                                                         \
    hipLaunchKernelGGL((k_icd_tick<Y, N, false>), dim3(nwg), dim3(kIcdBlock), dyn, s, jobs,     \
                       args);
  if (ypl == 1 && nv == 1) {
    MBIR_TICK(1, 1)
  } else if (ypl == 1) {
    MBIR_TICK(1, 2)
  } else if (nv == 1) {
    MBIR_TICK(2, 1)
  } else {
    MBIR_TICK(2, 2)
  }
#undef MBIR_TICK
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int* p, unsigned int v, size_t n, hipStream_t s);

hipError_t launch_set_devjob(DevJob* jobs, int slot, const DevJob& J, hipStream_t s)
{
  hipLaunchKernelGGL(k_set_devjob, dim3(1), dim3(64), 0, s, jobs, slot, J);
  return hipGetLastError();
}

hipError_t launch_roi_sum(const DevJob& J, double* out, hipStream_t s)
{
  const int L = J.N_z * J.N_x;
  hipLaunchKernelGGL(k_roi_sum, dim3(std::max(1, std::min(L, 1024))), dim3(kBlock), 0, s, J, out);
  return hipGetLastError();
}

hipError_t launch_joint_sums(const DevJob& J, double* sums6, hipStream_t s)
{
  hipLaunchKernelGGL(k_joint_sums, dim3(J.N_theta), dim3(kBlock), 0, s, J, sums6);
  return hipGetLastError();
}

hipError_t launch_joint_solve(int N_theta, const double* sums6, double target_gain,
                              const double*, const double*, double* I0_new, double* mu_new,
                              hipStream_t s)
{
  hipLaunchKernelGGL(k_joint_solve, dim3(1), dim3(kBlock), 0, s, N_theta, sums6, target_gain,
                     I0_new, mu_new);
  return hipGetLastError();
}

hipError_t launch_joint_apply(const DevJob& J, const double* I0_old, const double* mu_old,
                              const double* I0_new, const double* mu_new, hipStream_t s)
{
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
  hipLaunchKernelGGL(k_joint_apply, dim3(grid_for(M)), dim3(kBlock), 0, s, J, I0_old, mu_old,
                     I0_new, mu_new);
  return hipGetLastError();
}

hipError_t launch_noise_update(const DevJob& J, double* alpha, float* wview, hipStream_t s)
{
  hipLaunchKernelGGL(k_noise_update, dim3(J.N_theta), dim3(kBlock), 0, s, J, alpha, wview);
  return hipGetLastError();
}

hipError_t launch_cost(const DevJob& J, const double* alpha, int noise_model, double* out3,
                       hipStream_t s)
{
  // out3 must have room for 3 + 2 * 1024 doubles (partials follow the result)
  const int nparts = 1024;
  double* part = out3 + 4;
  hipLaunchKernelGGL(k_cost_partials, dim3(nparts), dim3(kBlock), 0, s, J, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cost_final, dim3(1), dim3(kBlock), 0, s, J, part, nparts, alpha,
                     noise_model, out3);
  return hipGetLastError();
}

hipError_t launch_sigma_sum(const DevJob& J, int z0, int z1, double sx, double* part, double* out,
                            hipStream_t s)
{
  const int nparts = 1024;
  hipLaunchKernelGGL(k_sigma_partials, dim3(nparts), dim3(kBlock), 0, s, J, z0, z1, sx, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, s, (const double*)part, nparts, out);
  return hipGetLastError();
}

hipError_t launch_offset_only(const DevJob& J, double* nd, double* mu, hipStream_t s)
{
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
  hipLaunchKernelGGL(k_offset_sums, dim3(J.N_theta), dim3(kBlock), 0, s, J, nd);
  hipLaunchKernelGGL(k_offset_apply, dim3(grid_for(M)), dim3(kBlock), 0, s, J,
                     (const double*)nd, mu);
  return hipGetLastError();
}

hipError_t launch_bf_offset(const DevJob& J, double* nd, double* mu, hipStream_t s)
{
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
  hipLaunchKernelGGL(k_bf_offset_sums, dim3(J.N_theta), dim3(kBlock), 0, s, J, nd);
  hipLaunchKernelGGL(k_bf_offset_apply, dim3(grid_for(M)), dim3(kBlock), 0, s, J,
                     (const double*)nd, mu);
  return hipGetLastError();
}

hipError_t launch_bf_noise(const DevJob& J, const double* alpha_old, double* alpha_new,
                           double* part, hipStream_t s)
{
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
  const int nb = std::min(grid_for(M), 1024);
  hipLaunchKernelGGL(k_bf_noise_partials, dim3(nb), dim3(kBlock), 0, s, J, alpha_old, part);
  hipLaunchKernelGGL(k_bf_noise_final, dim3(1), dim3(kBlock), 0, s, J, (const double*)part, nb,
                     alpha_new);
  hipLaunchKernelGGL(k_bf_noise_apply, dim3(grid_for(M)), dim3(kBlock), 0, s, J, alpha_old,
                     (const double*)alpha_new);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int* p, unsigned int v, size_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, size_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_f64(double* p, double v, size_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, p, v, n);
  return hipGetLastError();
}

}  // namespace mbir
