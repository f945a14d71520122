// mbir_kernels.hip -- HIP kernels of the MBIR hot path for gfx950.
//
// Data layout in HBM (the reference's layouts, Common/TomoArray.hpp:176-198):
//   volume x [N_z][N_x][N_y] fp32, y fastest: one "line" (z, x) = N_y contiguous voxels
//   error sinogram e, measurements y [N_theta][N_r][N_t] fp32, t fastest
// Because a voxel at slice y projects only onto detector rows t in VLR(y)
// (HAADF/HAADF_ReconstructionEngine_Extra.cpp:329-404), the A-column of a line
// touches, per (theta, r) bin, a contiguous run of N_t floats.  Every kernel maps
// wavefront lanes onto that run, so each sinogram access is a coalesced 64-lane
// load/atomic, and the column itself is computed on the fly in LDS from the
// detector response (never stored in HBM).
#include <hip/hip_runtime.h>

#include "mbir_internal.h"

namespace mbir {

namespace {

// ---------------------------------------------------------------------------
// A-matrix column of line (z, x) into LDS slots.  Slot (th, k) is detector bin
// jmin(th) + k.  Restates HAADF/HAADF_ReconstructionEngine.cpp:1402-1497 in FP64:
// r = x cos - z sin, bins [floor((r - dxz - R0)/dr), floor((r + dxz - R0)/dr)]
// clipped to the detector, value = lerp of H[th] at |r - r_c| / OffsetR, kept
// if > 0.  Empty slots get value 0.
template <typename F>
__device__ inline void build_column(const DevJob& J, int z, int x, F&& store)
{
  const double xc = J.x0 + ((double)x + 0.5) * J.delta_xz;
  const double zc = J.z0 + ((double)z + 0.5) * J.delta_xz;
  for (int th = threadIdx.x; th < J.N_theta; th += blockDim.x) {
    const double r = xc * J.cosv[th] - zc * J.sinv[th];
    const double rmin = r - J.delta_xz;
    const double rmax = r + J.delta_xz;
    int jmin = 0, jmax = -1;
    if (!(rmax < J.R0 || rmin > J.RMax)) {
      jmin = (int)floor((rmin - J.R0) / J.delta_r);
      jmax = (int)floor((rmax - J.R0) / J.delta_r);
      if (jmax >= J.N_r) jmax = J.N_r - 1;
      if (jmin < 0) jmin = 0;
    }
    const double* Hth = J.H + (size_t)th * kResponseBins;
    for (int k = 0; k < J.NB; ++k) {
      const int j = jmin + k;
      double val = 0.0;
      if (j <= jmax) {
        const double rc = J.R0 + ((double)j + 0.5) * J.delta_r;
        const double dr = fabs(r - rc);
        const int idr = (int)floor(dr / J.offsetR);
        if (idr >= 0 && idr < kResponseBins) {
          const double w1 = dr - idr * J.offsetR;
          const double w2 = (idr + 1) * J.offsetR - dr;
          const int iidx = idr + 1 < kResponseBins ? idr + 1 : kResponseBins - 1;
          const double f1 = (w2 / J.offsetR) * Hth[idr] + (w1 / J.offsetR) * Hth[iidx];
          if (f1 > 0) val = f1;
        }
      }
      store(th * J.NB + k, th, j, val);
    }
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
  if (d == 0.f) return J.b0;
  const float a = fabsf(d);
  const float t1 = exp2f((2.f - J.mrf_q) * log2f(a)) / J.sigma_pq;
  const float t3 = J.mrf_c + t1;
  return (2.f - (2.f - J.mrf_q) * t1 / t3) / (t3 * J.sigma_p);
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
    J.e[i] = (float)((double)J.y[i] - yest[i] - J.mu[th]);
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

// ---------------------------------------------------------------------------
// One ICD step: every workgroup updates all N_y voxels of one line (z, x) of the
// step's line list.  Restates UpdateYSlice::execute for one line
// (HAADF/HAADF_ReconstructionEngine_UpdateVoxels.cpp:260-455):
//   neighbourhood gather (:281-305), zero skipping (:311-338), theta1/theta2
//   (:343-369), bounds (:47-63,370), QGGMRF functional substitution
//   (Reconstruction/QGGMRF_Functions.cpp:153-206), positivity (:394-400),
//   ROI statistics (:412-421), error-sinogram update (:423-449).
// Slices y are independent in the data term (disjoint VLR rows); the prior
// couples y and y +- 1, so even slices are solved first and odd slices second,
// which makes the step equal to a sequential order.  Lines within one step are
// chosen by the host schedule; the e update uses fp32 atomics so concurrent lines
// that share a detector bin never lose an update.
template <int YPL>
__global__ __launch_bounds__(kBlock) void k_icd_step(DevJob J, const int* lines, int nlines,
                                                     int first_sweep, int zs_on, int TY,
                                                     SweepAccum* acc)
{
  __shared__ float s_ka[kMaxSlots];
  __shared__ int s_row[kMaxSlots];
  __shared__ float s_x[kMaxNy];
  __shared__ float s_dx[kMaxNy];
  __shared__ unsigned char s_cand[kMaxNy];
  __shared__ unsigned char s_need[kMaxNy];
  __shared__ float s_p1h[kBlock * YPL], s_p1l[kBlock * YPL], s_p2h[kBlock * YPL], s_p2l[kBlock * YPL];
  __shared__ double red[kBlock];

  if ((int)blockIdx.x >= nlines) return;
  const int line = lines[blockIdx.x];
  const int nnz_line = J.line_nnz[line];
  if (nnz_line == 0) return;  // the reference skips empty columns (:245-249)
  const int z = line / J.N_x, xi = line % J.N_x;
  const int tid = threadIdx.x;
  const int Ny = J.N_y;
  const size_t lbase = (size_t)line * Ny;

  // ---- 1. current values of this line and the 24 out-of-line neighbours ----
  float v = 0.f;
  float nb[27];
  unsigned int inb = 0;  // bit i set: neighbour i exists (BOUNDARYFLAG)
  if (tid < Ny) {
    v = J.x[lbase + tid];
    s_x[tid] = v;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int i = (dy + 1) * 9 + (dz + 1) * 3 + (dx + 1);
          nb[i] = 0.f;
          const int yy = tid + dy, zz = z + dz, xx = xi + dx;
          if (yy >= 0 && yy < Ny && zz >= 0 && zz < J.N_z && xx >= 0 && xx < J.N_x) {
            inb |= 1u << i;
            if (dz != 0 || dx != 0) nb[i] = J.x[((size_t)zz * J.N_x + xx) * Ny + yy];
          }
        }
  }
  // ---- 2. column in LDS ----
  build_column(J, z, xi, [&](int s, int th, int j, double a) {
    s_ka[s] = (float)(J.I0[th] * a);
    s_row[s] = a > 0.0 ? (th * J.N_r + j) * J.N_t : 0;
  });
  __syncthreads();
  // zero-skip candidates with the pre-step in-line neighbours
  bool cand = false;
  if (tid < Ny) {
    if (zs_on && !first_sweep && v == 0.f) {
      cand = true;
#pragma unroll
      for (int i = 0; i < 27; ++i)
        if (i != 13 && i != 4 && i != 22 && nb[i] > 0.f) cand = false;
      for (int dy = -1; dy <= 1; dy += 2) {
        const int yy = tid + dy;
        if (yy >= 0 && yy < Ny && s_x[yy] > 0.f) cand = false;
      }
    }
    s_cand[tid] = cand;
  }
  __syncthreads();
  if (tid < Ny) {
    bool need = !s_cand[tid];
    if ((tid & 1) && !need) {  // odd slices see the even slices' new values
      if (tid - 1 >= 0 && !s_cand[tid - 1]) need = true;
      if (tid + 1 < Ny && !s_cand[tid + 1]) need = true;
    }
    s_need[tid] = need;
  }
  __syncthreads();

  // ---- 3. theta1 / theta2: lanes over y, (wave, sub-row) over slots ----
  const int lane = tid & 63, wave = tid >> 6;
  const int RPW = 64 / TY;
  const int sub = lane / TY, yl = lane % TY;
  const int P = kWaves * RPW;
  const int p = wave * RPW + sub;
  Kahan t1[YPL], t2[YPL];
  int vt0[YPL], vn[YPL];
  float vv[YPL][kVlrCap];
  bool act[YPL];
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy) {
    const int y = yl + TY * jy;
    act[jy] = y < Ny && s_need[y];
    vt0[jy] = 0;
    vn[jy] = 0;
    if (y < Ny) {
      vt0[jy] = J.vlr_t0[y];
      vn[jy] = J.vlr_n[y];
#pragma unroll
      for (int k = 0; k < kVlrCap; ++k) vv[jy][k] = J.vlr_val[y * kVlrCap + k];
    }
  }
  for (int s = p; s < J.nslots; s += P) {
    const float ka = s_ka[s];
    if (ka == 0.f) continue;
    const int row = s_row[s];
    const float wv = J.wview[s / J.NB];
#pragma unroll
    for (int jy = 0; jy < YPL; ++jy) {
      if (!act[jy]) continue;
      for (int k = 0; k < vn[jy]; ++k) {
        const int idx = row + vt0[jy] + k;
        const float pe = ka * vv[jy][k];
        const float w = J.wfull ? J.wfull[idx] : wv;
        const float ev = J.e[idx];
        t2[jy].add(pe * pe * w);
        t1[jy].add(ev * pe * w);
      }
    }
  }
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy) {
    const int y = yl + TY * jy;
    if (y < TY * YPL) {
      const int o = p * (TY * YPL) + y;
      s_p1h[o] = t1[jy].s;
      s_p1l[o] = -t1[jy].c;
      s_p2h[o] = t2[jy].s;
      s_p2l[o] = -t2[jy].c;
    }
  }
  __syncthreads();

  // ---- 4. voxel solve, even slices then odd slices ----
  double T1 = 0.0, T2 = 0.0;
  if (tid < Ny) {
    for (int q = 0; q < P; ++q) {
      const int o = q * (TY * YPL) + tid;
      T1 += (double)s_p1h[o] + (double)s_p1l[o];
      T2 += (double)s_p2h[o] + (double)s_p2l[o];
    }
    T1 = -T1;  // THETA1 *= -1 (:369)
  }
  float xnew = v;
  bool skipped = false, updated = false;
  for (int phase = 0; phase < 2; ++phase) {
    if (tid < Ny && (tid & 1) == phase) {
      // in-line neighbours y +- 1 from LDS (odd slices see the even slices' new values)
      nb[4] = 0.f;   // (dy=-1, dz=0, dx=0)
      nb[22] = 0.f;  // (dy=+1, dz=0, dx=0)
      if (tid - 1 >= 0) nb[4] = s_x[tid - 1];
      if (tid + 1 < Ny) nb[22] = s_x[tid + 1];
      nb[13] = 0.f;  // NEIGHBORHOOD[1][1][1] = 0 (:305)
      bool zs = false;
      if (zs_on && !first_sweep && v == 0.f) {
        zs = true;
#pragma unroll
        for (int i = 0; i < 27; ++i)
          if (nb[i] > 0.f) zs = false;
      }
      if (zs) {
        skipped = true;
      } else {
        // bounds: min/max over the 27 entries, widened by v - theta1/theta2 (:47-63)
        double lo = nb[0], hi = nb[0];
#pragma unroll
        for (int i = 1; i < 27; ++i) {
          lo = fmin(lo, (double)nb[i]);
          hi = fmax(hi, (double)nb[i]);
        }
        if (T2 != 0.0) {
          const double c = (double)v - T1 / T2;
          lo = fmin(lo, c);
          hi = fmax(hi, c);
        }
        float s1 = 0.f, s2 = 0.f;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
              const int i = (dy + 1) * 9 + (dz + 1) * 3 + (dx + 1);
              if (i == 13 || !(inb & (1u << i))) continue;
              const float b = filter_weight(dy, dz, dx) * qggmrf_b(v - nb[i], J);
              s1 += b * nb[i];
              s2 += b;
            }
        const double u = ((double)s1 + T2 * (double)v - T1) / ((double)s2 + T2);
        double nv = (double)v + (double)J.alpha * (u - (double)v);
        nv = nv < lo ? lo : (nv > hi ? hi : nv);
        if (nv < 0.0) nv = 0.0;  // POSITIVITY_CONSTRAINT (:394-400)
        xnew = (float)nv;
        updated = true;
      }
      s_x[tid] = xnew;
    }
    __syncthreads();
  }
  if (tid < Ny) {
    s_dx[tid] = xnew - v;
    if (updated) J.x[lbase + tid] = xnew;
  }
  __syncthreads();

  // ---- 5. e -= I0 a vlr dx (:423-449), fp32 atomics ----
#pragma unroll
  for (int jy = 0; jy < YPL; ++jy) {
    const int y = yl + TY * jy;
    act[jy] = y < Ny && s_dx[y] != 0.f;
  }
  for (int s = p; s < J.nslots; s += P) {
    const float ka = s_ka[s];
    if (ka == 0.f) continue;
    const int row = s_row[s];
#pragma unroll
    for (int jy = 0; jy < YPL; ++jy) {
      if (!act[jy]) continue;
      const float d = s_dx[yl + TY * jy];
      for (int k = 0; k < vn[jy]; ++k)
        unsafeAtomicAdd(J.e + row + vt0[jy] + k, -(ka * vv[jy][k] * d));
    }
  }

  // ---- 6. sweep statistics ----
  const bool roi = J.mask[line] != 0;
  double su = 0.0, sv = 0.0, nz = 0.0, nu = 0.0, nn = 0.0;
  if (tid < Ny) {
    if (updated) {
      nu = 1.0;
      nn = (double)nnz_line * (double)J.vlr_n[tid];
      if (roi) {
        su = fabs((double)xnew - (double)v);
        sv = fabs((double)v);
      }
    }
    if (skipped) nz = 1.0;
  }
  su = block_sum_d(su, red);
  sv = block_sum_d(sv, red);
  nz = block_sum_d(nz, red);
  nu = block_sum_d(nu, red);
  nn = block_sum_d(nn, red);
  if (tid == 0) {
    atomicAdd(&acc->sum_abs_update, su);
    atomicAdd(&acc->sum_abs_value, sv);
    atomicAdd(&acc->zero_skipped, (unsigned long long)nz);
    atomicAdd(&acc->updated, (unsigned long long)nu);
    atomicAdd(&acc->nnz_updated, (unsigned long long)nn);
  }
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

__global__ __launch_bounds__(kBlock) void k_cost_partials(DevJob J, double* part)
{
  __shared__ double red[kBlock];
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t, per = (size_t)J.N_r * J.N_t;
  const size_t V = (size_t)J.N_z * J.N_x * J.N_y;
  const size_t g0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x, gs = (size_t)gridDim.x * blockDim.x;
  double data = 0, prior = 0;
  for (size_t i = g0; i < M; i += gs) {
    const double ev = J.e[i];
    const double w = J.wfull ? (double)J.wfull[i] : (double)J.wview[i / per];
    data += ev * ev * w;
  }
  const int Ny = J.N_y, Nx = J.N_x, Nz = J.N_z;
  for (size_t v = g0; v < V; v += gs) {
    const int k = (int)(v % Ny);
    const int j = (int)((v / Ny) % Nx);
    const int i = (int)(v / ((size_t)Ny * Nx));
    const double xv = J.x[v];
#define XV(a, b, c) (double)J.x[((size_t)(a) * Nx + (b)) * Ny + (c)]
#define ACC(a, b, c, f) prior += (f) * qggmrf_value_d(xv - XV(a, b, c), J)
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
  }
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
  }
}

__global__ void k_fill_f32(float* p, float v, size_t n)
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

hipError_t launch_init_error(const DevJob& J, const double* yest, hipStream_t s)
{
  const size_t M = (size_t)J.N_theta * J.N_r * J.N_t;
  hipLaunchKernelGGL(k_init_error, dim3(grid_for(M)), dim3(kBlock), 0, s, J, yest);
  return hipGetLastError();
}

hipError_t launch_export_column(const DevJob& J, int z, int x, float* vals, int* rows,
                                hipStream_t s)
{
  hipLaunchKernelGGL(k_export_column, dim3(1), dim3(kBlock), 0, s, J, z, x, vals, rows);
  return hipGetLastError();
}

hipError_t launch_icd_step(const DevJob& J, const int* lines, int nlines, int outer, int inner,
                           int zero_skipping, SweepAccum* acc, hipStream_t s)
{
  if (nlines <= 0) return hipSuccess;
  const int first = (outer == 0 && inner == 0) ? 1 : 0;
  int TY = 64, YPL = 1;
  if (J.N_y <= 64) {
    TY = 1;
    while (TY < J.N_y) TY <<= 1;
  } else {
    YPL = (J.N_y + 63) / 64;
  }
  if (YPL == 1)
    hipLaunchKernelGGL(k_icd_step<1>, dim3(nlines), dim3(kBlock), 0, s, J, lines, nlines, first,
                       zero_skipping, TY, acc);
  else if (YPL == 2)
    hipLaunchKernelGGL(k_icd_step<2>, dim3(nlines), dim3(kBlock), 0, s, J, lines, nlines, first,
                       zero_skipping, TY, acc);
  else
    hipLaunchKernelGGL(k_icd_step<4>, dim3(nlines), dim3(kBlock), 0, s, J, lines, nlines, first,
                       zero_skipping, TY, acc);
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
