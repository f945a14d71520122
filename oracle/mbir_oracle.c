/*
 * mbir_oracle.c -- FP64 CPU restatement of OpenMBIR-TEM's HAADF MBIR hot path.
 *
 * TEST INFRASTRUCTURE: the checker for the HIP product path and the timed CPU
 * baseline ("port").  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  Citations are relative to the reference's
 * Code/MBIRLib/ directory.  Pinning of every function is in tests/test_oracle_*.py.
 */
#define _GNU_SOURCE
#include "mbir_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_PI_2
#define M_PI_2 1.57079632679489661923
#endif
#ifndef M_PI_4
#define M_PI_4 0.78539816339744830962
#endif

#define IDX3(i, j, k) ((9 * (i)) + (3 * (j)) + (k)) /* ReconstructionConstants.h:44-45 */

/* ------------------------------------------------------------------------- */
/* MT19937: the published generator of Matsumoto & Nishimura (1998).  The
 * reference draws its update order from boost::mt19937 through
 * uniform_int<uint32_t>(0, 2^32-1), i.e. the raw 32-bit output
 * (HAADF/HAADF_ReconstructionEngine_UpdateVoxels.cpp:152-162,220). */
void orc_mt_seed(orc_mt* m, uint32_t seed)
{
  m->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    m->mt[i] = 1812433253u * (m->mt[i - 1] ^ (m->mt[i - 1] >> 30)) + (uint32_t)i;
  m->idx = 624;
}

uint32_t orc_mt_next(orc_mt* m)
{
  if (m->idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (m->mt[i] & 0x80000000u) | (m->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = m->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      m->mt[i] = v;
    }
    m->idx = 0;
  }
  uint32_t y = m->mt[m->idx++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

/* ------------------------------------------------------------------------- */
/* a1: geometry sizing.
 * GenericFilters/InitialReconstructionInitializer.cpp:98-167 (delta_xz = delta_r *
 * delta_xz at :109, LengthZ rounding :121-124, extend_object :129-135, N_x/N_y/N_z
 * :144-146, origins :164-167) and GenericFilters/MRCSinogramInitializer.cpp:302-305
 * (R0, RMax, T0, TMax).  delta_xz_in is the driver's 2^(L-1-i)*finalRes
 * (HAADF/HAADF_MultiResolutionReconstruction.cpp:349-350); interpolate_factor is
 * 2^(L-1)*finalRes (:262). */
int orc_geometry_init(int32_t N_r, int32_t N_t, int32_t N_theta, double delta_r, double delta_t,
                      double thickness, double delta_xz_in, double interpolate_factor,
                      int32_t extend_object, const double* angles_deg, orc_geom* g)
{
  memset(g, 0, sizeof(*g));
  if (N_r <= 0 || N_t <= 0 || N_theta <= 0 || delta_r <= 0 || delta_t <= 0) return -1;
  g->N_r = N_r;
  g->N_t = N_t;
  g->N_theta = N_theta;
  g->delta_r = delta_r;
  g->delta_t = delta_t;
  g->delta_xz = delta_r * delta_xz_in;
  g->delta_xy = g->delta_xz;
  double mx = fabs(angles_deg[0]);
  for (int i = 1; i < N_theta; ++i)
    if (fabs(angles_deg[i]) > mx) mx = fabs(angles_deg[i]);
  if (mx > 75.0) mx = 75.0; /* MBIR::Constants::k_MaxAngleStretch */
  mx = mx * M_PI / 180.0;
  double LengthZ = thickness * 2.0; /* Z_STRETCH = 2 */
  LengthZ /= (interpolate_factor * delta_r);
  LengthZ = floor(LengthZ + 0.5) * interpolate_factor * delta_r;
  double LengthX = N_r * delta_r;
  if (extend_object == 1) {
    LengthX = 0.6 * ((N_r * delta_r) / cos(mx)) + LengthZ * tan(mx); /* X_SHRINK_FACTOR */
    LengthX /= (interpolate_factor * delta_r);
    LengthX = floor(LengthX + 0.5) * interpolate_factor * delta_r;
  }
  double LengthY = N_t * delta_t; /* (yEnd - yStart + 1) * delta_t */
  g->N_x = (int32_t)(uint16_t)floor(LengthX / g->delta_xz);
  g->N_z = (int32_t)(uint16_t)floor(LengthZ / g->delta_xz);
  g->N_y = (int32_t)(uint16_t)floor(LengthY / g->delta_xy);
  g->LengthX = LengthX;
  g->LengthY = LengthY;
  g->LengthZ = LengthZ;
  g->x0 = -LengthX / 2;
  g->z0 = -LengthZ / 2;
  g->y0 = -LengthY / 2;
  g->R0 = -(N_r * delta_r) / 2;
  g->RMax = (N_r * delta_r) / 2;
  g->T0 = -(N_t * delta_t) / 2;
  g->TMax = (N_t * delta_t) / 2;
  return (g->N_x > 0 && g->N_y > 0 && g->N_z > 0) ? 0 : -1;
}

/* OffsetR / OffsetT: HAADF/HAADF_ReconstructionEngine.cpp:468-469 */
double orc_offset_r(const orc_geom* g)
{
  return ((g->delta_xz / sqrt(3.0)) + g->delta_r / 2) / ORC_RESPONSE_BINS;
}
double orc_offset_t(const orc_geom* g)
{
  return ((g->delta_xz / 2) + g->delta_t / 2) / ORC_RESPONSE_BINS;
}

/* a2: trapezoidal voxel profile.  HAADF/HAADF_ReconstructionEngine.cpp:957-1029.
 * angles_inout is multiplied by pi/180 in place, exactly as the reference does to
 * m_Sinogram->angles (:978); the engine calls this twice (:464, :499). */
void orc_voxel_profile(double* angles_inout, int32_t N_theta, double delta_xz, double* profile)
{
  for (int i = 0; i < N_theta; ++i) {
    angles_inout[i] = angles_inout[i] * (M_PI / 180.0);
    double angle = angles_inout[i];
    while (angle > M_PI_2) angle -= M_PI_2;
    while (angle < 0) angle += M_PI_2;
    double MaxValLineIntegral;
    if (angle <= M_PI_4)
      MaxValLineIntegral = delta_xz / cos(angle);
    else
      MaxValLineIntegral = delta_xz / cos(M_PI_2 - angle);
    double temp = cos(M_PI_4);
    double dist1 = temp * cos((M_PI_4 - angle));
    double dist2 = temp * fabs((cos((M_PI_4 + angle))));
    double LeftCorner = 1 - dist1, LeftNear = 1 - dist2;
    double RightNear = 1 + dist2, RightCorner = 1 + dist1;
    for (int j = 0; j < ORC_PROFILE_RESOLUTION; ++j) {
      double t = 2.0 * j / ORC_PROFILE_RESOLUTION;
      double v;
      if (t <= LeftCorner || t >= RightCorner)
        v = 0;
      else if (t > RightNear)
        v = MaxValLineIntegral * (RightCorner - t) / (RightCorner - RightNear);
      else if (t >= LeftNear)
        v = MaxValLineIntegral;
      else
        v = MaxValLineIntegral * (t - LeftCorner) / (LeftNear - LeftCorner);
      profile[(size_t)i * ORC_PROFILE_RESOLUTION + j] = v;
    }
  }
}

/* a3: GenericFilters/DetectorParameters.cpp:60-73 */
void orc_sincos(const double* angles_rad, int32_t n, double* cosv, double* sinv)
{
  for (int i = 0; i < n; ++i) {
    cosv[i] = cos(angles_rad[i]);
    sinv[i] = sin(angles_rad[i]);
  }
}

/* a3: Hamming beam, GenericFilters/DetectorParameters.cpp:78-100 */
void orc_beam_profile(double delta_t, double* beam)
{
  double sum = 0;
  for (int i = 0; i < ORC_BEAM_RESOLUTION; ++i) {
    beam[i] = 0.54 - 0.46 * cos((2.0 * M_PI / ORC_BEAM_RESOLUTION) * i);
    sum = sum + beam[i];
  }
  for (int i = 0; i < ORC_BEAM_RESOLUTION; ++i) {
    beam[i] /= sum;
    beam[i] /= delta_t;
  }
}

/* a4: detector response H[theta][k], GenericFilters/DetectorResponse.cpp:66-127
 * (only the j = 0 t-displacement is computed by the reference). */
void orc_detector_response(const orc_geom* g, const double* profile, const double* beam,
                           double offsetR, double* H)
{
  double beamWidth = g->delta_r; /* setBeamWidth(delta_r), HAADF_ReconstructionEngine.cpp:470 */
  double r0 = -(beamWidth) / 2;
  double StepSize = beamWidth / ORC_BEAM_RESOLUTION;
  double TempConst = (ORC_PROFILE_RESOLUTION) / (2 * g->delta_xz);
  for (int k = 0; k < g->N_theta; ++k) {
    for (int i = 0; i < ORC_RESPONSE_BINS; ++i) {
      double ProfileCenterR = i * offsetR;
      double rmin = ProfileCenterR - g->delta_xz;
      double sum = 0;
      for (int p = 0; p < ORC_BEAM_RESOLUTION; ++p) {
        double r = r0 + p * StepSize;
        if (r < rmin) continue;
        int32_t ProfileIndex = (int16_t)(int32_t)floor((r - rmin) * TempConst);
        if (ProfileIndex < 0) ProfileIndex = 0;
        if (ProfileIndex >= ORC_PROFILE_RESOLUTION) ProfileIndex = ORC_PROFILE_RESOLUTION - 1;
        sum += profile[(size_t)k * ORC_PROFILE_RESOLUTION + ProfileIndex] * beam[p];
      }
      H[(size_t)k * ORC_RESPONSE_BINS + i] = sum;
    }
  }
}

/* a5: H_t, HAADF/HAADF_ForwardModel.cpp:408-451 */
void orc_init_ht(const orc_geom* g, double OffsetT, double* Ht)
{
  for (int k = 0; k < g->N_theta; ++k) {
    for (int i = 0; i < ORC_RESPONSE_BINS; ++i) {
      double PCT = i * OffsetT;
      double v;
      if (g->delta_xy >= g->delta_t) {
        if (PCT <= ((g->delta_xy / 2) - (g->delta_t / 2)))
          v = g->delta_t;
        else
          v = -1 * PCT + (g->delta_xy / 2) + g->delta_t / 2;
      } else {
        if (PCT <= g->delta_t / 2 - g->delta_xy / 2)
          v = g->delta_xy;
        else
          v = -PCT + (g->delta_xy / 2) + g->delta_t / 2;
      }
      if (v < 0) v = 0;
      Ht[(size_t)k * ORC_RESPONSE_BINS + i] = v;
    }
  }
}

/* a5: voxel-line response of slice y, HAADF/HAADF_ReconstructionEngine_Extra.cpp:329-404.
 * Uses view 0 of H_t (:380).  Returns the number of entries written. */
int32_t orc_vlr(const orc_geom* g, const double* Ht, double OffsetT, int32_t y,
                int32_t* t_index, double* values, int32_t cap)
{
  double t = ((double)y + 0.5) * g->delta_xy + g->y0;
  double tmin = (t - g->delta_xy / 2) > g->T0 ? t - g->delta_xy / 2 : g->T0;
  double tmax = (t + g->delta_xy / 2) <= g->TMax ? t + g->delta_xy / 2 : g->TMax;
  int32_t smin = (int16_t)(uint16_t)floor((tmin - g->T0) / g->delta_t);
  int32_t smax = (int16_t)(uint16_t)floor((tmax - g->T0) / g->delta_t);
  if (smin < 0) smin = 0;
  if (smax >= g->N_t) smax = g->N_t - 1;
  int32_t count = 0;
  for (int it = smin; it <= smax; ++it) {
    double center_t = ((double)it + 0.5) * g->delta_t + g->T0;
    double delta_t = fabs(center_t - t);
    int32_t idt = (int16_t)(uint16_t)floor(delta_t / OffsetT);
    double PT;
    if (idt < ORC_RESPONSE_BINS) {
      double w3 = delta_t - (double)(idt)*OffsetT;
      double w4 = ((double)idt + 1) * OffsetT - delta_t;
      int32_t ttmp = idt + 1 < ORC_RESPONSE_BINS ? idt + 1 : ORC_RESPONSE_BINS - 1;
      PT = (w4 / OffsetT) * Ht[idt] + (w3 / OffsetT) * Ht[ttmp];
    } else {
      PT = 0;
    }
    if (PT != 0) {
      if (count < cap) {
        t_index[count] = it;
        values[count] = PT;
      }
      count++;
    }
  }
  return count;
}

/* a6: one A-matrix column, HAADF/HAADF_ReconstructionEngine.cpp:1369-1524
 * (identical logic in Common/AMatrixCol.cpp:82-239).  Entries with value > 0
 * are kept, index = theta * N_r + r. */
int32_t orc_acol(const orc_geom* g, const double* cosv, const double* sinv, const double* H,
                 double OffsetR, int32_t row, int32_t col, uint32_t* index, double* values,
                 int32_t cap)
{
  double x = g->x0 + ((double)col + 0.5) * g->delta_xz;
  double z = g->z0 + ((double)row + 0.5) * g->delta_xz;
  int32_t count = 0;
  for (int i = 0; i < g->N_theta; ++i) {
    double r = x * cosv[i] - z * sinv[i];
    double rmin = r - g->delta_xz;
    double rmax = r + g->delta_xz;
    if (rmax < g->R0 || rmin > g->RMax) continue;
    int32_t index_min = (int32_t)floor(((rmin - g->R0) / g->delta_r));
    int32_t index_max = (int32_t)floor((rmax - g->R0) / g->delta_r);
    if (index_max >= g->N_r) index_max = g->N_r - 1;
    if (index_min < 0) index_min = 0;
    int32_t BaseIndex = i * g->N_r;
    for (int j = index_min; j <= index_max; ++j) {
      double R_Center = (g->R0 + (((double)j) + 0.5) * (g->delta_r));
      double delta_r = fabs(r - R_Center);
      int32_t idr = (int32_t)floor((delta_r / OffsetR));
      if (idr >= 0 && idr < ORC_RESPONSE_BINS) {
        double w1 = delta_r - idr * OffsetR;
        double w2 = (idr + 1) * OffsetR - delta_r;
        int32_t iidx = idr + 1 < ORC_RESPONSE_BINS ? idr + 1 : ORC_RESPONSE_BINS - 1;
        double f1 = (w2 / OffsetR) * H[(size_t)i * ORC_RESPONSE_BINS + idr] +
                    (w1 / OffsetR) * H[(size_t)i * ORC_RESPONSE_BINS + iidx];
        if (f1 > 0) {
          if (count < cap) {
            values[count] = f1;
            index[count] = (uint32_t)(BaseIndex + j);
          }
          count++;
        }
      }
    }
  }
  return count;
}

/* ------------------------------------------------------------------------- */
/* QGGMRF.  Reconstruction/QGGMRF_Functions.cpp:82-206; constants
 * HAADF/HAADF_QGGMRFPriorModel.cpp:10-19. */
void orc_qggmrf_init(double sigma_x, double mrf_q, double mrf_c, orc_qggmrf* v)
{
  v->MRF_P = 2;
  v->MRF_Q = mrf_q;
  v->MRF_C = mrf_c;
  v->MRF_ALPHA = 1.5;
  v->SIGMA_X_P = pow(sigma_x, v->MRF_P);
  v->SIGMA_X_P_Q = pow(sigma_x, (v->MRF_P - v->MRF_Q));
  v->SIGMA_X_Q = pow(sigma_x, v->MRF_Q);
}

/* QGGMRF_Functions.cpp:82-85 */
double orc_qggmrf_value(double delta, const orc_qggmrf* q)
{
  return ((pow(fabs(delta), q->MRF_P) / q->SIGMA_X_P) /
          (q->MRF_C + pow(fabs(delta), q->MRF_P - q->MRF_Q) / q->SIGMA_X_P_Q));
}

/* QGGMRF_Functions.cpp:90-104 */
double orc_qggmrf_derivative(double delta, const orc_qggmrf* q)
{
  double temp1 = pow(fabs(delta), q->MRF_P - q->MRF_Q) / (q->SIGMA_X_P_Q);
  double temp2 = pow(fabs(delta), q->MRF_P - 1);
  double temp3 = q->MRF_C + temp1;
  double mag = (temp2 / (temp3 * q->SIGMA_X_P)) * (q->MRF_P - ((q->MRF_P - q->MRF_Q) * temp1) / (temp3));
  return delta < 0 ? -mag : mag;
}

/* QGGMRF_Functions.cpp:109-112 */
double orc_qggmrf_second_derivative_at_zero(const orc_qggmrf* q)
{
  return q->MRF_P / (q->SIGMA_X_P * q->MRF_C);
}

static double clip(double x, double a, double b) /* HAADF_ReconstructionEngine.cpp:102-105 */
{
  return (x < a) ? a : ((x > b) ? b : x);
}

/* QGGMRF_Functions.cpp:118-147 (ComputeParameters) and :153-206
 * (FunctionalSubstitution, QGGMRF_ITER = 1, QGGMRF_Functions.h:46).  The
 * POSITIVITY_CONSTRAINT clip of umin at :165-168 is not compiled in that
 * translation unit (the macro is defined only in HAADF/HAADFConstants.h). */
double orc_qggmrf_functional_substitution(double umin, double umax, double v,
                                          const uint8_t* bflag, const double* filter,
                                          const double* nbr, double theta1, double theta2,
                                          const orc_qggmrf* q)
{
  double refValue = v, temp1 = 0, temp2 = 0;
  double params[26];
  int count = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        if ((i != 1 || j != 1 || k != 1) && bflag[IDX3(i, j, k)] == 1) {
          double Delta0 = refValue - nbr[IDX3(i, j, k)];
          if (Delta0 != 0)
            params[count] = orc_qggmrf_derivative(Delta0, q) / (Delta0);
          else
            params[count] = orc_qggmrf_second_derivative_at_zero(q);
          count++;
        }
  count = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        if ((i != 1 || j != 1 || k != 1) && bflag[IDX3(i, j, k)] == 1) {
          double temp_const = filter[IDX3(i, j, k)] * params[count];
          temp1 += temp_const * nbr[IDX3(i, j, k)];
          temp2 += temp_const;
          count++;
        }
  double u = (temp1 + (theta2 * v) - theta1) / (temp2 + theta2);
  refValue = clip(refValue + q->MRF_ALPHA * (u - refValue), umin, umax);
  return refValue;
}

/* FILTER, HAADF/HAADF_ReconstructionEngine.cpp:266-294 */
void orc_filter(double* F)
{
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        int ones = (i == 1) + (j == 1) + (k == 1);
        F[IDX3(i, j, k)] = ones == 3 ? 0.0 : ones == 2 ? 0.0523 : ones == 1 ? 0.0370 : 0.0302;
      }
}

/* ------------------------------------------------------------------------- */
struct orc_engine {
  orc_geom g;
  orc_params p;
  orc_qggmrf q;
  double filter[27];
  double offsetR, offsetT;
  double* angles_rad; /* after the first calculateVoxelProfile() */
  double* cosv;
  double* sinv;
  double* profile; /* second profile (F3 quirk) -> fed to DetectorResponse */
  double* beam;
  double* H;
  double* Ht;
  /* VLR per y */
  int32_t* vlr_count;
  int32_t* vlr_t;  /* [N_y * 8] */
  double* vlr_val; /* [N_y * 8] */
  /* columns, CSR over lines z*N_x + x */
  int64_t* col_ptr;
  uint32_t* col_idx;
  double* col_val;
  uint8_t* mask; /* ROI [N_z][N_x] */
  double* x;     /* [N_z][N_x][N_y] */
  double* counts;
  double* e;
  double* w;
  double* yest;
  double* I0;
  double* mu;
  double* alpha;
  double* init_variance;
  /* bright-field sinogram of the HAADF engine (BF_Flag): bright-field counts +
   * BF_OFFSET per detector element, scaling every projection entry; NULL without */
  double* bfs;
  double stop_threshold;
  int32_t sweep_counter;
  uint8_t* sel;    /* bright field: Bragg selector [N_theta][N_r][N_t] */
  double bragg;    /* bright field: current Bragg threshold */
  /* bright-field NHICD: the lines of every call (call c: nh_log[nh_ptr[c] .. nh_ptr[c+1])) */
  int32_t* nh_log;
  int64_t* nh_ptr;
  int32_t nh_calls, nh_cap_calls;
  int64_t nh_cap_lines;
};

#define VLR_CAP 8

static size_t sino_size(const orc_geom* g) { return (size_t)g->N_theta * g->N_r * g->N_t; }
static size_t vol_size(const orc_geom* g) { return (size_t)g->N_z * g->N_x * g->N_y; }

void orc_engine_destroy(orc_engine* e)
{
  if (!e) return;
  free(e->angles_rad); free(e->cosv); free(e->sinv); free(e->profile); free(e->beam);
  free(e->H); free(e->Ht); free(e->vlr_count); free(e->vlr_t); free(e->vlr_val);
  free(e->col_ptr); free(e->col_idx); free(e->col_val); free(e->mask); free(e->x);
  free(e->counts); free(e->e); free(e->w); free(e->yest); free(e->I0); free(e->mu);
  free(e->alpha); free(e->init_variance); free(e->sel); free(e->bfs);
  free(e->nh_log); free(e->nh_ptr);
  free(e);
}

/* Forward projection: HAADF/HAADF_ForwardProject.cpp:80-120, one z at a time in
 * order (the serial build of HAADF_ReconstructionEngine.cpp:671-680). */
static void project_into(const orc_engine* e, const double* x, double* yest)
{
  const orc_geom* g = &e->g;
  memset(yest, 0, sino_size(g) * sizeof(double));
  for (int z = 0; z < g->N_z; ++z)
    for (int xx = 0; xx < g->N_x; ++xx) {
      int64_t line = (int64_t)z * g->N_x + xx;
      int64_t b = e->col_ptr[line], en = e->col_ptr[line + 1];
      if (en - b <= 0) continue;
      for (int y = 0; y < g->N_y; ++y) {
        double xv = x[(size_t)line * g->N_y + y];
        for (int64_t q = b; q < en; ++q) {
          int32_t it = (int32_t)(e->col_idx[q] / (uint32_t)g->N_r);
          int32_t ir = (int32_t)(e->col_idx[q] % (uint32_t)g->N_r);
          for (int k = 0; k < e->vlr_count[y]; ++k) {
            int32_t t = e->vlr_t[y * VLR_CAP + k];
            double ttmp = e->I0[it] * (e->col_val[q] * e->vlr_val[y * VLR_CAP + k] * xv);
            yest[((size_t)it * g->N_r + ir) * g->N_t + t] += ttmp;
          }
        }
      }
    }
}

void orc_engine_project(const orc_engine* e, const double* x, double* yest) { project_into(e, x, yest); }

orc_engine* orc_engine_create(const orc_geom* g, const double* angles_deg, const double* counts,
                              const double* init_gain, const double* init_offset,
                              const double* init_variance, const double* init_x,
                              double default_recon, const orc_params* p)
{
  orc_engine* e = (orc_engine*)calloc(1, sizeof(orc_engine));
  e->g = *g;
  e->p = *p;
  int NT = g->N_theta;
  size_t M = sino_size(g), V = vol_size(g);
  /* execute() order, HAADF/HAADF_ReconstructionEngine.cpp:441-691 */
  e->angles_rad = (double*)malloc(sizeof(double) * NT);
  memcpy(e->angles_rad, angles_deg, sizeof(double) * NT);
  e->profile = (double*)malloc(sizeof(double) * NT * ORC_PROFILE_RESOLUTION);
  orc_voxel_profile(e->angles_rad, NT, g->delta_xz, e->profile); /* :464, degrees -> radians */
  e->offsetR = orc_offset_r(g);
  e->offsetT = orc_offset_t(g);
  e->cosv = (double*)malloc(sizeof(double) * NT);
  e->sinv = (double*)malloc(sizeof(double) * NT);
  orc_sincos(e->angles_rad, NT, e->cosv, e->sinv); /* :471 */
  e->beam = (double*)malloc(sizeof(double) * ORC_BEAM_RESOLUTION);
  orc_beam_profile(g->delta_t, e->beam); /* :473 */
  orc_qggmrf_init(p->sigma_x, p->mrf_q, p->mrf_c, &e->q); /* :477 */
  orc_filter(e->filter);
  {
    /* second calculateVoxelProfile() (:499): converts the already-radian angles
     * again (quirk F3); its profile feeds the detector response. */
    double* tmp = (double*)malloc(sizeof(double) * NT);
    memcpy(tmp, e->angles_rad, sizeof(double) * NT);
    orc_voxel_profile(tmp, NT, g->delta_xz, e->profile);
    free(tmp);
  }
  e->H = (double*)malloc(sizeof(double) * NT * ORC_RESPONSE_BINS);
  orc_detector_response(g, e->profile, e->beam, e->offsetR, e->H); /* :512-530 */
  /* ROI mask, HAADF/HAADF_ReconstructionEngine_Extra.cpp:227-248 */
  e->mask = (uint8_t*)malloc((size_t)g->N_z * g->N_x);
  for (int i = 0; i < g->N_z; ++i)
    for (int j = 0; j < g->N_x; ++j) {
      double xx = g->x0 + ((double)j + 0.5) * g->delta_xz;
      double zz = g->z0 + ((double)i + 0.5) * g->delta_xz;
      int in = xx >= -(g->N_r * g->delta_r) / 2 && xx <= (g->N_r * g->delta_r) / 2 &&
               zz >= -g->LengthZ / 2 && zz <= g->LengthZ / 2;
      e->mask[(size_t)i * g->N_x + j] = (uint8_t)in;
    }
  /* gainAndOffsetInitialization, ..._Extra.cpp:280-307 */
  e->I0 = (double*)malloc(sizeof(double) * NT);
  e->mu = (double*)malloc(sizeof(double) * NT);
  e->alpha = (double*)malloc(sizeof(double) * NT);
  e->init_variance = (double*)malloc(sizeof(double) * NT);
  double sum = 0;
  for (int k = 0; k < NT; ++k) {
    e->I0[k] = init_gain[k];
    e->mu[k] = init_offset[k];
    sum += e->I0[k];
    e->init_variance[k] = init_variance[k];
  }
  sum /= NT;
  if (sum - p->target_gain > 1e-5) {
    double temp = p->target_gain / sum;
    for (int k = 0; k < NT; ++k) e->I0[k] = init_gain[k] * temp;
  }
  /* H_t and VLR: :582-586, :596-604, :634 */
  e->Ht = (double*)malloc(sizeof(double) * NT * ORC_RESPONSE_BINS);
  orc_init_ht(g, e->offsetT, e->Ht);
  e->vlr_count = (int32_t*)calloc(g->N_y, sizeof(int32_t));
  e->vlr_t = (int32_t*)calloc((size_t)g->N_y * VLR_CAP, sizeof(int32_t));
  e->vlr_val = (double*)calloc((size_t)g->N_y * VLR_CAP, sizeof(double));
  for (int y = 0; y < g->N_y; ++y) {
    int32_t c = orc_vlr(g, e->Ht, e->offsetT, y, e->vlr_t + y * VLR_CAP, e->vlr_val + y * VLR_CAP, VLR_CAP);
    e->vlr_count[y] = c < VLR_CAP ? c : VLR_CAP;
  }
  /* object init (InitialReconstructionInitializer.cpp:73-90) */
  e->x = (double*)malloc(sizeof(double) * V);
  for (size_t i = 0; i < V; ++i) e->x[i] = init_x ? init_x[i] : default_recon;
  /* A-columns (:615-632) */
  size_t L = (size_t)g->N_z * g->N_x;
  e->col_ptr = (int64_t*)malloc(sizeof(int64_t) * (L + 1));
  int32_t cap = 9 * NT + 16;
  uint32_t* tidx = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  double* tval = (double*)malloc(sizeof(double) * cap);
  size_t nnz_cap = L * (size_t)(3 * NT) + 16, nnz = 0;
  e->col_idx = (uint32_t*)malloc(sizeof(uint32_t) * nnz_cap);
  e->col_val = (double*)malloc(sizeof(double) * nnz_cap);
  for (int z = 0; z < g->N_z; ++z)
    for (int xx = 0; xx < g->N_x; ++xx) {
      size_t line = (size_t)z * g->N_x + xx;
      int32_t c = orc_acol(g, e->cosv, e->sinv, e->H, e->offsetR, z, xx, tidx, tval, cap);
      if (c > cap) c = cap;
      if (nnz + c > nnz_cap) {
        nnz_cap = (nnz + c) * 2;
        e->col_idx = (uint32_t*)realloc(e->col_idx, sizeof(uint32_t) * nnz_cap);
        e->col_val = (double*)realloc(e->col_val, sizeof(double) * nnz_cap);
      }
      e->col_ptr[line] = (int64_t)nnz;
      memcpy(e->col_idx + nnz, tidx, sizeof(uint32_t) * c);
      memcpy(e->col_val + nnz, tval, sizeof(double) * c);
      nnz += c;
      if (c == 0)
        for (int y = 0; y < g->N_y; ++y) e->x[line * g->N_y + y] = 0.0;
    }
  e->col_ptr[L] = (int64_t)nnz;
  free(tidx);
  free(tval);
  /* forward projection (:646-684) and calculateMeasurementWeight (..._Extra.cpp:693-771) */
  e->counts = (double*)malloc(sizeof(double) * M);
  memcpy(e->counts, counts, sizeof(double) * M);
  if (p->bf_mode) {
    /* processRawCounts, BrightField/BFForwardModel.cpp:939-972 (run before the
     * setup above in the reference; nothing above reads the counts) */
    for (size_t i = 0; i < M; ++i) e->counts[i] = -log((e->counts[i] + p->bf_offset) / ORC_BF_MAX);
    e->sel = (uint8_t*)malloc(M);
    e->bragg = p->bragg_T;
  }
  e->yest = (double*)malloc(sizeof(double) * M);
  e->e = (double*)malloc(sizeof(double) * M);
  e->w = (double*)malloc(sizeof(double) * M);
  project_into(e, e->x, e->yest);
  for (int it = 0; it < NT; ++it) {
    if (p->noise_model) e->alpha[it] = e->init_variance[it];
    for (int ir = 0; ir < g->N_r; ++ir)
      for (int t = 0; t < g->N_t; ++t) {
        size_t idx = ((size_t)it * g->N_r + ir) * g->N_t + t;
        e->e[idx] = e->counts[idx] - e->yest[idx] - e->mu[it];
        double wv = 1.0; /* IDENTITY_NOISE_MODEL, HAADF/HAADFConstants.h */
        if (p->bf_mode) {
          /* BF calculateMeasurementWeight, BFForwardModel.cpp:328-385 */
          wv = ORC_BF_MAX / exp(e->counts[idx]);
          e->sel[idx] = 1;
        }
        if (p->noise_model) wv /= e->alpha[it];
        e->w[idx] = wv;
      }
  }
  e->stop_threshold = p->stop_threshold;
  return e;
}

/* One voxel of UpdateYSlice::execute, HAADF/HAADF_ReconstructionEngine_UpdateVoxels.cpp:260-455.
 * Returns 1 if updated, 0 if zero-skipped; accumulates ROI sums. */
static int update_voxel(orc_engine* e, int32_t j_new, int32_t k_new, int32_t i, int32_t outer,
                        int32_t inner, double* avg_update, double* avg_mag)
{
  const orc_geom* g = &e->g;
  double NB[27];
  uint8_t BF[27];
  memset(NB, 0, sizeof(NB));
  memset(BF, 0, sizeof(BF));
  for (int p = -1; p <= 1; ++p)
    for (int q = -1; q <= 1; ++q)
      for (int r = -1; r <= 1; ++r)
        if (i + p >= 0 && i + p < g->N_y)
          if (j_new + q >= 0 && j_new + q < g->N_z) {
            if (k_new + r >= 0 && k_new + r < g->N_x) {
              NB[IDX3(p + 1, q + 1, r + 1)] =
                e->x[((size_t)(q + j_new) * g->N_x + (r + k_new)) * g->N_y + (p + i)];
              BF[IDX3(p + 1, q + 1, r + 1)] = 1;
            } else {
              BF[IDX3(p + 1, q + 1, r + 1)] = 0;
            }
          }
  NB[IDX3(1, 1, 1)] = 0.0;
  size_t line = (size_t)j_new * g->N_x + k_new;
  double v = e->x[line * g->N_y + i];
  double THETA1 = 0.0, THETA2 = 0.0;
  int ZSFlag = 1;
  if (e->p.zero_skipping == 1) {
    ZSFlag = 1;
    if (v == 0.0 && (inner > 0 || outer > 0)) {
      for (int a = 0; a < 27; ++a)
        if (NB[a] > 0.0) { ZSFlag = 0; break; }
    } else {
      ZSFlag = 0;
    }
  } else {
    ZSFlag = 0;
  }
  if (ZSFlag) return 0;
  int64_t b = e->col_ptr[line], en = e->col_ptr[line + 1];
  for (int64_t q = b; q < en; ++q) {
    int32_t it = (int32_t)(e->col_idx[q] / (uint32_t)g->N_r);
    int32_t ir = (int32_t)(e->col_idx[q] % (uint32_t)g->N_r);
    double kConst0 = e->I0[it] * (e->col_val[q]);
    for (int k = 0; k < e->vlr_count[i]; ++k) {
      size_t eidx = ((size_t)it * g->N_r + ir) * g->N_t + e->vlr_t[i * VLR_CAP + k];
      double PE = kConst0 * e->vlr_val[i * VLR_CAP + k];
      if (e->bfs) PE *= e->bfs[eidx]; /* BF_Flag, ..._UpdateVoxels.cpp:354-362 */
      if (e->p.bf_mode && e->sel[eidx] == 0) {
        /* rejected entry: BFForwardModel::computeTheta, :874-882 */
        double QC = (e->p.bragg_delta * e->bragg) / (fabs(e->e[eidx]) * sqrt(e->w[eidx]));
        THETA2 += QC * (PE * PE * e->w[eidx]);
        THETA1 += QC * (e->e[eidx] * PE * e->w[eidx]);
      } else {
        THETA2 += (PE * PE * e->w[eidx]);
        THETA1 += (e->e[eidx] * PE * e->w[eidx]);
      }
    }
  }
  THETA1 *= -1;
  /* find_min_max, :47-63 */
  double low = NB[0], high = NB[0];
  for (int a = 0; a < 27; ++a) {
    if (NB[a] < low) low = NB[a];
    if (NB[a] > high) high = NB[a];
  }
  if (THETA2 != 0) {
    double c = v - (THETA1 / THETA2);
    low = (low > c ? c : low);
    high = (high < c ? c : high);
  }
  double U = orc_qggmrf_functional_substitution(low, high, v, BF, e->filter, NB, THETA1, THETA2, &e->q);
  if (U < 0.0) U = 0.0; /* POSITIVITY_CONSTRAINT, :394-400 */
  e->x[line * g->N_y + i] = U;
  if (e->mask[line] == 1) {
    *avg_update += fabs(U - v);
    *avg_mag += fabs(v);
  }
  double dx = U - v;
  for (int64_t q = b; q < en; ++q) {
    int32_t it = (int32_t)(e->col_idx[q] / (uint32_t)g->N_r);
    int32_t ir = (int32_t)(e->col_idx[q] % (uint32_t)g->N_r);
    for (int k = 0; k < e->vlr_count[i]; ++k) {
      size_t eidx = ((size_t)it * g->N_r + ir) * g->N_t + e->vlr_t[i * VLR_CAP + k];
      double kConst2 = (e->I0[it] * (e->col_val[q] * e->vlr_val[i * VLR_CAP + k] * dx));
      if (e->bfs) kConst2 *= e->bfs[eidx]; /* BF_Flag, ..._UpdateVoxels.cpp:438-446 */
      e->e[eidx] -= kConst2;
      if (e->p.bf_mode) /* selector refresh, BFForwardModel::updateErrorSinogram :924-929 */
        e->sel[eidx] = fabs(e->e[eidx] * sqrt(e->w[eidx])) < e->bragg ? 1 : 0;
    }
  }
  return 1;
}

typedef struct slab_args {
  orc_engine* e;
  int32_t y0, y1, outer, inner;
  uint32_t seed;
  double avg_update, avg_mag, zeros, visits;
} slab_args;

/* UpdateYSlice::execute for one y-slab, random line order, :149-231 */
static void* run_slab(void* vp)
{
  slab_args* a = (slab_args*)vp;
  orc_engine* e = a->e;
  const orc_geom* g = &e->g;
  orc_mt rng;
  orc_mt_seed(&rng, a->seed);
  int32_t ArraySize = g->N_x * g->N_z;
  int32_t* Counter = (int32_t*)malloc(sizeof(int32_t) * ArraySize);
  for (int32_t j = 0; j < ArraySize; ++j) Counter[j] = j;
  for (int32_t j = 0; j < g->N_z; ++j)
    for (int32_t k = 0; k < g->N_x; ++k) {
      uint32_t Index = orc_mt_next(&rng) % (uint32_t)ArraySize;
      int32_t k_new = Counter[Index] % g->N_x;
      int32_t j_new = Counter[Index] / g->N_x;
      Counter[Index] = Counter[ArraySize - 1];
      ArraySize--;
      size_t line = (size_t)j_new * g->N_x + k_new;
      if (e->col_ptr[line + 1] - e->col_ptr[line] <= 0) continue;
      for (int32_t i = a->y0; i < a->y1; ++i) {
        if (update_voxel(e, j_new, k_new, i, a->outer, a->inner, &a->avg_update, &a->avg_mag))
          a->visits += 1;
        else
          a->zeros += 1;
      }
    }
  free(Counter);
  return NULL;
}

/* updateVoxels stop test, :765-786 */
static int32_t finish_sweep(orc_engine* e, double au, double am, double* stats, double zeros,
                            double visits)
{
  if (stats) {
    stats[0] = au;
    stats[1] = am;
    stats[2] = zeros;
    stats[3] = visits;
  }
  e->sweep_counter++;
  if (am > 0 && (au / am) < e->stop_threshold) return 0; /* THRESHOLD_REDUCTION_FACTOR = 1 */
  return 1;
}

int32_t orc_engine_sweep(orc_engine* e, int32_t outer, int32_t inner, double* stats)
{
  slab_args a;
  memset(&a, 0, sizeof(a));
  a.e = e;
  a.y0 = 0;
  a.y1 = e->g.N_y;
  a.outer = outer;
  a.inner = inner;
  a.seed = e->p.seed + 7919u * (uint32_t)e->sweep_counter;
  run_slab(&a);
  return finish_sweep(e, a.avg_update, a.avg_mag, stats, a.zeros, a.visits);
}

/* A given update order: units u = (line, y0, ny) in sequence, each the slices y0 .. y0 + ny - 1
 * of line (z, x) = (line / N_x, line % N_x), even local slices first, then odd ones (the
 * prior couples slices y and y +- 1 only, so this is one sequential ICD order of
 * UpdateYSlice's voxel updates, ..._UpdateVoxels.cpp:260-455).  Units whose line has an
 * empty column are skipped (:245-249).  stats as orc_engine_sweep's (no stop test). */
void orc_engine_update_units(orc_engine* e, const int32_t* units, int32_t n, int32_t outer,
                             int32_t inner, double* stats)
{
  const orc_geom* g = &e->g;
  double au = 0, am = 0, zeros = 0, visits = 0;
  for (int32_t u = 0; u < n; ++u) {
    int32_t line = units[3 * u], y0 = units[3 * u + 1], ny = units[3 * u + 2];
    if (e->col_ptr[line + 1] - e->col_ptr[line] <= 0) continue;
    for (int phase = 0; phase < 2; ++phase)
      for (int32_t yl = phase; yl < ny; yl += 2) {
        if (update_voxel(e, line / g->N_x, line % g->N_x, y0 + yl, outer, inner, &au, &am))
          visits += 1;
        else
          zeros += 1;
      }
  }
  if (stats) {
    stats[0] = au;
    stats[1] = am;
    stats[2] = zeros;
    stats[3] = visits;
  }
}

/* One thread's share of a sweep: UpdateYSlice::execute over the y-slab [y0, y1) only
 * (:149-231), the line order drawn as orc_engine_sweep's.  Does not count as a sweep (no
 * stop test); for timing a bounded part of a sweep. */
void orc_engine_sweep_slab(orc_engine* e, int32_t outer, int32_t inner, int32_t y0, int32_t y1,
                           double* stats)
{
  slab_args a;
  memset(&a, 0, sizeof(a));
  a.e = e;
  a.y0 = y0 < 0 ? 0 : y0;
  a.y1 = y1 > e->g.N_y ? e->g.N_y : y1;
  a.outer = outer;
  a.inner = inner;
  a.seed = e->p.seed + 7919u * (uint32_t)e->sweep_counter;
  run_slab(&a);
  if (stats) {
    stats[0] = a.avg_update;
    stats[1] = a.avg_mag;
    stats[2] = a.zeros;
    stats[3] = a.visits;
  }
}

/* y split round-robin into per-thread counts, one slab per thread, :670-723 */
int32_t orc_engine_sweep_threads(orc_engine* e, int32_t outer, int32_t inner, int32_t nthreads,
                                 double* stats)
{
  if (nthreads < 1) nthreads = 1;
  int32_t* ycount = (int32_t*)calloc(nthreads, sizeof(int32_t));
  int t = 0;
  for (int y = 0; y < e->g.N_y; ++y) {
    ycount[t]++;
    if (++t == nthreads) t = 0;
  }
  slab_args* args = (slab_args*)calloc(nthreads, sizeof(slab_args));
  pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
  int32_t ys = 0;
  for (int i = 0; i < nthreads; ++i) {
    args[i].e = e;
    args[i].y0 = ys;
    args[i].y1 = ys + ycount[i];
    ys += ycount[i];
    args[i].outer = outer;
    args[i].inner = inner;
    args[i].seed = e->p.seed + 7919u * (uint32_t)e->sweep_counter + 104729u * (uint32_t)i;
    if (args[i].y1 > args[i].y0) pthread_create(&th[i], NULL, run_slab, &args[i]);
  }
  double au = 0, am = 0, zz = 0, vv = 0;
  for (int i = 0; i < nthreads; ++i) {
    if (args[i].y1 > args[i].y0) pthread_join(th[i], NULL);
    au += args[i].avg_update;
    am += args[i].avg_mag;
    zz += args[i].zeros;
    vv += args[i].visits;
  }
  free(ycount);
  free(args);
  free(th);
  return finish_sweep(e, au, am, stats, zz, vv);
}

/* BF jointEstimation: per-view offset (log dose) only, BrightField/BFForwardModel.cpp:390-441 */
static void bf_joint_estimation(orc_engine* e)
{
  const orc_geom* g = &e->g;
  size_t per = (size_t)g->N_r * g->N_t;
  const double Td = e->bragg * e->p.bragg_delta;
  for (int it = 0; it < g->N_theta; ++it) {
    double num_sum = 0, den_sum = 0;
    for (size_t k = 0; k < per; ++k) {
      size_t idx = it * per + k;
      double ev = e->e[idx], w = e->w[idx];
      if (e->sel[idx] == 1) {
        num_sum += ev * w;
        den_sum += w;
      } else {
        num_sum += (Td * ev / fabs(ev)) * sqrt(w);
        den_sum += (Td * sqrt(w)) / fabs(ev);
      }
    }
    double alpha = num_sum / den_sum;
    for (size_t k = 0; k < per; ++k) e->e[it * per + k] -= alpha;
    e->mu[it] += alpha;
  }
}

/* updateSelector, BrightField/BFForwardModel.cpp:517-531 */
void orc_engine_update_selector(orc_engine* e)
{
  size_t M = sino_size(&e->g);
  for (size_t i = 0; i < M; ++i)
    e->sel[i] = e->e[i] * e->e[i] * e->w[i] < e->bragg * e->bragg ? 1 : 0;
}

/* BF updateWeights: one variance for all views, BrightField/BFForwardModel.cpp:447-511 */
static void bf_update_weights(orc_engine* e)
{
  const orc_geom* g = &e->g;
  size_t per = (size_t)g->N_r * g->N_t;
  for (int it = 0; it < g->N_theta; ++it)
    for (size_t k = 0; k < per; ++k) e->w[it * per + k] *= e->alpha[it];
  double sum1 = 0, sum2 = 0;
  for (int it = 0; it < g->N_theta; ++it)
    for (size_t k = 0; k < per; ++k) {
      size_t idx = it * per + k;
      if (e->sel[idx] == 1) sum1 += e->e[idx] * e->e[idx] * e->w[idx];
      else sum2 += fabs(e->e[idx]) * sqrt(e->alpha[it] * e->w[idx]);
    }
  double update = (sum1 + (e->p.bragg_delta * e->bragg) * sum2) / ((double)g->N_theta * g->N_r * g->N_t);
  for (int it = 0; it < g->N_theta; ++it) e->alpha[it] = update;
  for (int it = 0; it < g->N_theta; ++it)
    for (size_t k = 0; k < per; ++k) e->w[it * per + k] /= e->alpha[it];
}

double orc_bf_beta(double ev, double w, double T, double delta)
{
  if (ev * ev * w < T * T) return ev * ev * w;  /* the selector's test */
  return 2 * T * delta * fabs(ev) * sqrt(w) + T * T * (1 - 2 * delta);
}

/* jointEstimation given bright-field data: offsets only,
 * HAADF/HAADF_ReconstructionEngine_Extra.cpp:636-669.  Per view alpha = sum e w / sum w
 * (:641-652), then e -= alpha over the view (:654-660) and mu += alpha (:662). */
static void offset_only_estimation(orc_engine* e)
{
  const orc_geom* g = &e->g;
  size_t per = (size_t)g->N_r * g->N_t;
  for (int it = 0; it < g->N_theta; ++it) {
    double num_sum = 0, den_sum = 0;
    for (size_t k = 0; k < per; ++k) {
      num_sum += (e->e[it * per + k] * e->w[it * per + k]);
      den_sum += e->w[it * per + k];
    }
    double alpha = num_sum / den_sum;
    for (size_t k = 0; k < per; ++k) e->e[it * per + k] -= alpha;
    e->mu[it] += alpha;
  }
}

/* jointEstimation, HAADF/HAADF_ReconstructionEngine_Extra.cpp:419-687: the gain + offset
 * branch (:429-635) without bright-field data, the offset-only branch (:636-669) with it */
int32_t orc_engine_joint_estimation(orc_engine* e)
{
  if (e->p.bf_mode) {
    bf_joint_estimation(e);
    return 0;
  }
  if (e->bfs) {
    offset_only_estimation(e);
    return 0;
  }
  const orc_geom* g = &e->g;
  int NT = g->N_theta;
  size_t per = (size_t)g->N_r * g->N_t;
  for (int it = 0; it < NT; ++it)
    for (size_t k = 0; k < per; ++k) {
      size_t idx = it * per + k;
      double ttmp = e->counts[idx] - e->e[idx] - e->mu[it];
      e->yest[idx] = ttmp / e->I0[it];
    }
  double* Q0 = (double*)malloc(sizeof(double) * NT);
  double* Q1 = (double*)malloc(sizeof(double) * NT);
  double* b0 = (double*)malloc(sizeof(double) * NT);
  double* d1 = (double*)malloc(sizeof(double) * NT);
  double* d2 = (double*)malloc(sizeof(double) * NT);
  for (int it = 0; it < NT; ++it) {
    double a = 0, b = 0, c = 0, d = 0, num = 0, den = 0;
    for (size_t k = 0; k < per; ++k) {
      size_t idx = it * per + k;
      double y = e->counts[idx], w = e->w[idx], ye = e->yest[idx];
      num += (y * w);
      den += (w);
      a += (ye * w);
      b += (ye * w * y);
      c += (y * y * w);
      d += (ye * ye * w);
    }
    (void)c;
    b0[it] = b;
    Q0[it] = d;
    Q1[it] = a;
    d1[it] = num / den;
    d2[it] = a / den;
  }
  double sum1 = 0, sum2 = 0;
  for (int it = 0; it < NT; ++it) {
    sum1 += (1.0 / (Q0[it] - Q1[it] * d2[it]));
    sum2 += ((b0[it] - Q1[it] * d1[it]) / (Q0[it] - Q1[it] * d2[it]));
  }
  double lambda = (-NT * e->p.target_gain + sum2) / sum1;
  for (int it = 0; it < NT; ++it) {
    double NewI = (-1 * lambda - Q1[it] * d1[it] + b0[it]) / (Q0[it] - Q1[it] * d2[it]);
    e->I0[it] = NewI;
    e->mu[it] = d1[it] - d2[it] * e->I0[it];
  }
  for (int it = 0; it < NT; ++it)
    for (size_t k = 0; k < per; ++k) {
      size_t idx = it * per + k;
      e->e[idx] = e->counts[idx] - e->mu[it] - (e->I0[it] * e->yest[idx]);
    }
  free(Q0); free(Q1); free(b0); free(d1); free(d2);
  return 0;
}

/* updateWeights, IDENTITY_NOISE_MODEL branch, ..._Extra.cpp:795-881 */
void orc_engine_update_weights(orc_engine* e)
{
  if (e->p.bf_mode) {
    bf_update_weights(e);
    return;
  }
  const orc_geom* g = &e->g;
  size_t per = (size_t)g->N_r * g->N_t;
  for (int it = 0; it < g->N_theta; ++it) {
    double sum = 0;
    for (size_t k = 0; k < per; ++k) e->w[it * per + k] = 1.0;
    for (size_t k = 0; k < per; ++k) {
      double ev = e->e[it * per + k];
      sum += (ev * ev * e->w[it * per + k]);
    }
    sum /= (g->N_r * g->N_t);
    e->alpha[it] = sum;
    for (size_t k = 0; k < per; ++k) e->w[it * per + k] = 1.0 / e->alpha[it];
  }
}

/* estimateSigmaX, HAADF/HAADF_ReconstructionEngine.cpp:1827-1945: the 13 forward
 * neighbour pairs of every voxel in the planes zStart = (int16) 0.25 N_z <= i < zEnd =
 * (int16) 0.75 N_z, each difference multiplied by SigmaX before QGGMRF::Value, weighted by
 * FILTER, summed and divided by N_x N_y (zEnd - zStart + 1) (:1938-1939).  Then
 * updatePriorModel, HAADF/HAADF_QGGMRFPriorModel.cpp:24-31: SigmaX = gamma (est q)^(1/q),
 * gamma = k_QGGMRF_Gamma = 5 (Reconstruction/ReconstructionConstants.h:78), and the
 * SIGMA_X_* constants of the new SigmaX. */
double orc_engine_estimate_sigma_x(orc_engine* e, double* sigma_x)
{
  const orc_geom* g = &e->g;
  const double fF = 0.0523, fE = 0.0370, fC = 0.0302;
  const double sx = e->p.sigma_x;
  int16_t zStart = (int16_t)(0.25 * g->N_z);
  int16_t zEnd = (int16_t)(0.75 * g->N_z);
  double temp = 0, delta;
#define XV(a, b, c) e->x[((size_t)(a) * g->N_x + (b)) * g->N_y + (c)]
#define ACC(a, b, c, f)                                  \
  do {                                                   \
    delta = XV(i, j, k) - XV(a, b, c);                   \
    delta *= sx;                                         \
    temp += (f) * orc_qggmrf_value(delta, &e->q);        \
  } while (0)
  for (int i = zStart; i < zEnd; ++i)
    for (int j = 0; j < g->N_x; ++j)
      for (int k = 0; k < g->N_y; ++k) {
        if (k + 1 < g->N_y) ACC(i, j, k + 1, fF);
        if (j + 1 < g->N_x) {
          if (k - 1 >= 0) ACC(i, j + 1, k - 1, fE);
          ACC(i, j + 1, k, fF);
          if (k + 1 < g->N_y) ACC(i, j + 1, k + 1, fE);
        }
        if (i + 1 < g->N_z) {
          if (j - 1 >= 0) ACC(i + 1, j - 1, k, fE);
          ACC(i + 1, j, k, fF);
          if (j + 1 < g->N_x) ACC(i + 1, j + 1, k, fE);
          if (j - 1 >= 0) {
            if (k - 1 >= 0) ACC(i + 1, j - 1, k - 1, fC);
            if (k + 1 < g->N_y) ACC(i + 1, j - 1, k + 1, fC);
          }
          if (k - 1 >= 0) ACC(i + 1, j, k - 1, fE);
          if (j + 1 < g->N_x) {
            if (k - 1 >= 0) ACC(i + 1, j + 1, k - 1, fC);
            if (k + 1 < g->N_y) ACC(i + 1, j + 1, k + 1, fC);
          }
          if (k + 1 < g->N_y) ACC(i + 1, j, k + 1, fE);
        }
      }
#undef ACC
#undef XV
  double NumEntries = (double)g->N_x * g->N_y * (zEnd - zStart + 1);
  double est = temp / NumEntries;
  double SigmaX = 5.0 * pow(est * e->q.MRF_Q, 1.0 / e->q.MRF_Q);
  e->q.SIGMA_X_P = pow(SigmaX, e->q.MRF_P);
  e->q.SIGMA_X_P_Q = pow(SigmaX, (e->q.MRF_P - e->q.MRF_Q));
  e->q.SIGMA_X_Q = pow(SigmaX, e->q.MRF_Q);
  e->p.sigma_x = SigmaX;
  if (sigma_x) *sigma_x = SigmaX;
  return est;
}
double orc_engine_get_sigma_x(const orc_engine* e) { return e->p.sigma_x; }

/* computeCost, HAADF/HAADF_ReconstructionEngine.cpp:1147-1364 (QGGMRF branch). */
double orc_engine_cost_terms(orc_engine* e, double* terms3)
{
  const orc_geom* g = &e->g;
  size_t M = sino_size(g);
  double cost = 0;
  if (e->p.bf_mode) {
    /* BFForwardModel::forwardCost, :812-846 */
    const double T = e->bragg, d = e->p.bragg_delta;
    for (size_t i = 0; i < M; ++i) {
      double ev = e->e[i];
      if (e->sel[i] == 1) cost += ev * ev * e->w[i];
      else cost += 2 * T * d * fabs(ev) * sqrt(e->w[i]) + T * T * (1 - 2 * d);
    }
  } else {
    for (size_t i = 0; i < M; ++i) cost += e->e[i] * e->e[i] * e->w[i];
  }
  cost /= 2;
  double data = cost;
  double temp = 0;
  const double fF = 0.0523, fE = 0.0370, fC = 0.0302;
#define XV(a, b, c) e->x[((size_t)(a) * g->N_x + (b)) * g->N_y + (c)]
#define ACC(a, b, c, f) temp += (f) * orc_qggmrf_value(XV(i, j, k) - XV(a, b, c), &e->q)
  for (int i = 0; i < g->N_z; ++i)
    for (int j = 0; j < g->N_x; ++j)
      for (int k = 0; k < g->N_y; ++k) {
        if (k + 1 < g->N_y) ACC(i, j, k + 1, fF);
        if (j + 1 < g->N_x) {
          if (k - 1 >= 0) ACC(i, j + 1, k - 1, fE);
          ACC(i, j + 1, k, fF);
          if (k + 1 < g->N_y) ACC(i, j + 1, k + 1, fE);
        }
        if (i + 1 < g->N_z) {
          if (j - 1 >= 0) ACC(i + 1, j - 1, k, fE);
          ACC(i + 1, j, k, fF);
          if (j + 1 < g->N_x) ACC(i + 1, j + 1, k, fE);
          if (j - 1 >= 0) {
            if (k - 1 >= 0) ACC(i + 1, j - 1, k - 1, fC);
            if (k + 1 < g->N_y) ACC(i + 1, j - 1, k + 1, fC);
          }
          if (k - 1 >= 0) ACC(i + 1, j, k - 1, fE);
          if (j + 1 < g->N_x) {
            if (k - 1 >= 0) ACC(i + 1, j + 1, k - 1, fC);
            if (k + 1 < g->N_y) ACC(i + 1, j + 1, k + 1, fC);
          }
          if (k + 1 < g->N_y) ACC(i + 1, j, k + 1, fE);
        }
      }
#undef ACC
#undef XV
  cost += temp;
  double noise = 0;
  if (e->p.bf_mode && e->p.noise_model) {
    /* (N_theta N_r N_t / 2) log(alpha[0]), integer halving as in :839 */
    noise = (double)(M / 2) * log(e->alpha[0]);
    cost += noise;
  } else if (e->p.noise_model) {
    for (size_t i = 0; i < M; ++i)
      if (e->w[i] != 0) noise += log(2 * M_PI * (1.0 / e->w[i]));
    noise /= 2;
    cost += noise;
  }
  if (terms3) {
    terms3[0] = data;
    terms3[1] = temp;
    terms3[2] = noise;
  }
  return cost;
}

double orc_engine_cost(orc_engine* e) { return orc_engine_cost_terms(e, NULL); }

/* ---- bright-field NHICD (non-homogeneous ICD) ---- */

/* GenRandList, BrightField/BFReconstructionEngine_Extra.cpp:1072-1105: output t is the
 * input entry Counter[Index], Index = next() % remaining, then Counter[Index] takes the
 * last remaining counter (the same draw as UpdateYSlice's voxel order). */
static void nh_rand_list(const int32_t* in, int32_t n, orc_mt* rng, int32_t* out)
{
  int32_t* counter = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  for (int32_t j = 0; j < n; ++j) counter[j] = j;
  int32_t size = n;
  for (int32_t t = 0; t < n; ++t) {
    uint32_t idx = orc_mt_next(rng) % (uint32_t)size;
    out[t] = in[counter[idx]];
    counter[idx] = counter[size - 1];
    size--;
  }
  free(counter);
}

/* m_HammingWindow, BrightField/BFReconstructionEngine.cpp:157-181 */
static const double kNhWindow[5][5] = {{0.0013, 0.0086, 0.0159, 0.0086, 0.0013},
                                       {0.0086, 0.0581, 0.1076, 0.0581, 0.0086},
                                       {0.0159, 0.1076, 0.1993, 0.1076, 0.0159},
                                       {0.0086, 0.0581, 0.1076, 0.0581, 0.0086},
                                       {0.0013, 0.0086, 0.0159, 0.0086, 0.0013}};
#define ORC_NH_PARTS 20 /* k_NumHomogeniousIter = k_NumNonHomogeniousIter */

/* ComputeVSC, BrightField/BFReconstructionEngine_Extra.cpp:888-925 */
void orc_nhicd_filter(const double* mag, int32_t Nz, int32_t Nx, double* filt)
{
  for (int i = 0; i < Nz; ++i)
    for (int j = 0; j < Nx; ++j) {
      double f = 0;
      for (int p = -2; p <= 2; ++p)
        for (int q = -2; q <= 2; ++q)
          if (i + p >= 0 && i + p < Nz && j + q >= 0 && j + q < Nx)
            f += kNhWindow[p + 2][q + 2] * mag[(size_t)(i + p) * Nx + (j + q)];
      filt[(size_t)i * Nx + j] = f;
    }
}

static int cmp_double(const void* a, const void* b)
{
  double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* SetNonHomThreshold, :934-1010: RandomizedSelect(A, 0, n - 1, n - n/20) is the order
 * statistic at 0-based position n - n/20 (:1124-1147) */
double orc_nhicd_threshold(const double* v, int32_t n)
{
  double* t = (double*)malloc(sizeof(double) * n);
  memcpy(t, v, sizeof(double) * n);
  qsort(t, n, sizeof(double), cmp_double);
  int32_t k = n - n / ORC_NH_PARTS;
  double r = t[k < n ? k : n - 1];
  free(t);
  return r;
}

static void nh_log_call(orc_engine* e, const int32_t* lines, int32_t n)
{
  if (e->nh_calls + 2 > e->nh_cap_calls) {
    e->nh_cap_calls = 2 * e->nh_cap_calls + 64;
    e->nh_ptr = (int64_t*)realloc(e->nh_ptr, sizeof(int64_t) * e->nh_cap_calls);
    if (e->nh_calls == 0) e->nh_ptr[0] = 0;
  }
  int64_t b = e->nh_ptr[e->nh_calls];
  if (b + n > e->nh_cap_lines) {
    e->nh_cap_lines = 2 * (b + n) + 1024;
    e->nh_log = (int32_t*)realloc(e->nh_log, sizeof(int32_t) * e->nh_cap_lines);
  }
  memcpy(e->nh_log + b, lines, sizeof(int32_t) * n);
  e->nh_ptr[++e->nh_calls] = b + n;
}

/* The lines of every NHICD call of the last run: returns the number of calls; call_ptr
 * (cap_calls + 1 entries) and lines (cap_lines) receive what fits. */
int32_t orc_engine_nhicd_log(const orc_engine* e, int64_t* call_ptr, int32_t cap_calls,
                             int32_t* lines, int64_t cap_lines)
{
  for (int32_t c = 0; c <= e->nh_calls && c <= cap_calls; ++c)
    call_ptr[c] = e->nh_calls ? e->nh_ptr[c] : 0;
  int64_t n = e->nh_calls ? e->nh_ptr[e->nh_calls] : 0;
  if (lines && n > 0) memcpy(lines, e->nh_log, sizeof(int32_t) * (n < cap_lines ? n : cap_lines));
  return e->nh_calls;
}

/* One NHICD call, BFReconstructionEngine_Extra.cpp:476-880 with BFUpdateYSlice.cpp:147-335
 * on one y-slab (the serial path): the list's magnitudes are reset (:640-646), the list is
 * put in a random order (GenRandList, :677), and every slice of each line with a non-empty
 * column is updated in that order, the line's magnitude accumulating |x_new - x_old|
 * (BFUpdateYSlice.cpp:303-304).  The stop test (stopCriteria, :1236-1270) on calls that
 * are multiples of 20: ROI magnitude sum / the ROI volume sum of 20 calls before. */
static int32_t nh_call(orc_engine* e, const int32_t* list, int32_t n, int32_t outer,
                       int32_t inner, uint32_t eff, double prev_mag, double* mag)
{
  const orc_geom* g = &e->g;
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  orc_mt rng;
  orc_mt_seed(&rng, e->p.seed + 7919u * eff + 1u);
  nh_rand_list(list, n, &rng, order);
  for (int32_t t = 0; t < n; ++t) mag[list[t]] = 0.0;
  for (int32_t t = 0; t < n; ++t) {
    int32_t line = order[t];
    int32_t j_new = line / g->N_x, k_new = line % g->N_x;
    if (e->col_ptr[line + 1] - e->col_ptr[line] <= 0) continue;
    for (int32_t i = 0; i < g->N_y; ++i) {
      double au = 0, am = 0;
      double old = e->x[(size_t)line * g->N_y + i];
      update_voxel(e, j_new, k_new, i, outer, inner, &au, &am);
      mag[line] += fabs(e->x[(size_t)line * g->N_y + i] - old);
    }
  }
  free(order);
  int32_t status = 1;
  if (eff % ORC_NH_PARTS == 0 && eff > 0 && prev_mag > 0) {
    double sum = 0;
    for (size_t l = 0; l < (size_t)g->N_z * g->N_x; ++l)
      if (e->mask[l] == 1) sum += mag[l];
    if (fabs(sum) / fabs(prev_mag) < e->stop_threshold) status = 0;
  }
  return status;
}

/* roiVolumeSum, BFReconstructionEngine_Extra.cpp:1216-1232 */
static double roi_volume_sum(const orc_engine* e)
{
  const orc_geom* g = &e->g;
  double sum = 0;
  for (size_t l = 0; l < (size_t)g->N_z * g->N_x; ++l)
    if (e->mask[l] == 1)
      for (int i = 0; i < g->N_y; ++i) sum += e->x[l * g->N_y + i];
  return sum;
}

/* BFReconstructionEngine::execute() loop, BrightField/BFReconstructionEngine.cpp:575-872,
 * with regular (full, random-order) sweeps in place of the NHICD voxel lists:
 * the Bragg threshold is DefBraggThreshold for the first outer iteration's inner
 * sweeps when NumIter > 1, and the desired T after its last inner sweep; the stop
 * test ends the run before the nuisance step; the nuisance step runs only when
 * NumOuterIter > 1 (offset, then the shared variance, each followed by the selector). */
static int32_t bf_run(orc_engine* e, double* costs, int32_t cap, int32_t* n_costs, int32_t* n_sweeps)
{
  int32_t nc = 0, ns = 0;
#define PUSH_COST()                                       \
  do {                                                    \
    if (e->p.compute_cost) {                              \
      double cv = orc_engine_cost(e);                     \
      if (costs && nc < cap) costs[nc] = cv;              \
      nc++;                                               \
    }                                                     \
  } while (0)
  PUSH_COST();
  int32_t num_iter = e->p.num_inner;
  const double des = e->p.bragg_T;
  double tmp = num_iter > 1 ? ORC_DEF_BRAGG : des;
  e->bragg = tmp;
  /* NHICD state (BFReconstructionEngine.cpp:595-635): the random list of all lines, the
   * update-magnitude map, the next homogeneous part, the call counter and the ROI volume
   * sum of the last check */
  const int32_t L = e->g.N_z * e->g.N_x;
  int32_t* all = NULL;
  int32_t* list = NULL;
  double* mag = NULL;
  double* filt = NULL;
  uint32_t eff = 0, part = 0;
  double prev_mag = 0;
  e->nh_calls = 0;
  if (e->p.bf_nhicd) {
    int32_t* reg = (int32_t*)malloc(sizeof(int32_t) * L);
    for (int32_t l = 0; l < L; ++l) reg[l] = l; /* GenRegularList: z-major (x, z) pairs */
    all = (int32_t*)malloc(sizeof(int32_t) * L);
    orc_mt rng;
    orc_mt_seed(&rng, e->p.seed * 2654435761u + 17u);
    nh_rand_list(reg, L, &rng, all);
    free(reg);
    list = (int32_t*)malloc(sizeof(int32_t) * L);
    mag = (double*)calloc(L, sizeof(double));
    filt = (double*)malloc(sizeof(double) * L);
  }
  for (int outer = 0; outer < e->p.num_outer; ++outer) {
    if (outer > 0) {
      num_iter = 1;
      e->bragg = tmp;
    }
    int32_t status = 1;
    for (int inner = 0; inner < num_iter; ++inner) {
      if (num_iter > 1 && inner == 0) e->bragg = ORC_DEF_BRAGG;
      if (e->p.bf_nhicd) {
        int32_t n = 0;
        if (eff % 2 == 0) {
          /* homogeneous: the next twentieth of the list, the last one longer (:681-712) */
          part %= ORC_NH_PARTS;
          int32_t ne = L / ORC_NH_PARTS;
          if (part == ORC_NH_PARTS - 1) ne += L - ne * ORC_NH_PARTS;
          int32_t start = (L / ORC_NH_PARTS) * (int32_t)part;
          for (int32_t t = 0; t < ne; ++t) list[n++] = all[start + t];
          part++;
        } else {
          /* non-homogeneous: the lines whose filtered magnitude exceeds the threshold,
           * in raster order (GenNonHomList, Extra.cpp:1013-1033) */
          orc_nhicd_filter(mag, e->g.N_z, e->g.N_x, filt);
          double thr = orc_nhicd_threshold(filt, L);
          for (int32_t l = 0; l < L; ++l)
            if (filt[l] > thr) list[n++] = l;
        }
        nh_log_call(e, list, n);
        status = nh_call(e, list, n, outer, inner, eff, prev_mag, mag);
        if (eff % ORC_NH_PARTS == 0) prev_mag = roi_volume_sum(e); /* :730-735 */
        eff++;
      } else {
        double st[4];
        status = orc_engine_sweep(e, outer, inner, st);
      }
      ns++;
      if (status == 0) break;
      PUSH_COST();
      if (inner == num_iter - 1 && inner != 0) {
        tmp = des;
        e->bragg = tmp;
      }
    }
    if (status == 0 && outer >= 1) break;
    if (e->p.num_outer > 1) {
      if (e->p.joint_estimation) {
        bf_joint_estimation(e);
        orc_engine_update_selector(e);
        PUSH_COST();
      }
      if (e->p.noise_model) {
        bf_update_weights(e);
        orc_engine_update_selector(e);
        PUSH_COST();
      }
    }
  }
  free(all);
  free(list);
  free(mag);
  free(filt);
#undef PUSH_COST
  if (n_costs) *n_costs = nc;
  if (n_sweeps) *n_sweeps = ns;
  return 0;
}

/* execute() outer/inner loop, HAADF/HAADF_ReconstructionEngine.cpp:700-815, with
 * COST_CALCULATE evaluations (:700-702, UpdateVoxels.cpp:750-763, :780-787). */
int32_t orc_engine_run(orc_engine* e, double* costs, int32_t cap, int32_t* n_costs,
                       int32_t* n_sweeps)
{
  if (e->p.bf_mode) return bf_run(e, costs, cap, n_costs, n_sweeps);
  int32_t nc = 0, ns = 0;
#define PUSH_COST()                                       \
  do {                                                    \
    if (e->p.compute_cost) {                              \
      double cv = orc_engine_cost(e);                     \
      if (costs && nc < cap) costs[nc] = cv;              \
      nc++;                                               \
    }                                                     \
  } while (0)
  PUSH_COST();
  int32_t num_iter = e->p.num_inner;
  for (int outer = 0; outer < e->p.num_outer; ++outer) {
    if (outer != 0) num_iter = 1;
    int32_t status = 1;
    for (int inner = 0; inner < num_iter; ++inner) {
      double st[4];
      status = orc_engine_sweep(e, outer, inner, st);
      ns++;
      PUSH_COST();
      if (status == 0) break;
    }
    if (e->p.joint_estimation) {
      orc_engine_joint_estimation(e);
      PUSH_COST(); /* calculateCost inside jointEstimation, ..._Extra.cpp:616-623 */
    }
    if (e->p.noise_model) {
      orc_engine_update_weights(e);
      PUSH_COST();
    }
    if (status == 0 && outer >= 1) break;
    if (e->p.estimate_prior) orc_engine_estimate_sigma_x(e, NULL); /* :807-813 */
  }
#undef PUSH_COST
  if (n_costs) *n_costs = nc;
  if (n_sweeps) *n_sweeps = ns;
  return 0;
}

void orc_engine_get_x(const orc_engine* e, double* x) { memcpy(x, e->x, sizeof(double) * vol_size(&e->g)); }
void orc_engine_set_x(orc_engine* e, const double* x) { memcpy(e->x, x, sizeof(double) * vol_size(&e->g)); }
/* The error sinogram as given (a state to continue from, e.g. another engine's);
 * the volume and e are no longer tied by e = y - I0 A x - mu afterwards */
void orc_engine_set_e(orc_engine* e, const double* err) { memcpy(e->e, err, sizeof(double) * sino_size(&e->g)); }

/* The HAADF engine's bright-field sinogram (HAADF/HAADF_ReconstructionEngine_Extra.cpp:
 * 89-175): s = counts + BF_OFFSET (HAADF/HAADFConstants.h:49) per element, then the
 * error sinogram e = y - s Yest - mu (:718-727) */
void orc_engine_set_bf_sinogram(orc_engine* e, const double* bf_counts)
{
  const orc_geom* g = &e->g;
  size_t M = sino_size(g);
  if (!e->bfs) e->bfs = (double*)malloc(sizeof(double) * M);
  for (size_t i = 0; i < M; ++i) e->bfs[i] = bf_counts[i] + ORC_BF_OFFSET;
  size_t per = (size_t)g->N_r * g->N_t;
  for (size_t i = 0; i < M; ++i)
    e->e[i] = e->counts[i] - e->bfs[i] * e->yest[i] - e->mu[i / per];
}
void orc_engine_get_e(const orc_engine* e, double* v) { memcpy(v, e->e, sizeof(double) * sino_size(&e->g)); }
void orc_engine_get_w(const orc_engine* e, double* v) { memcpy(v, e->w, sizeof(double) * sino_size(&e->g)); }
void orc_engine_get_yest(const orc_engine* e, double* v) { memcpy(v, e->yest, sizeof(double) * sino_size(&e->g)); }
void orc_engine_get_nuisance(const orc_engine* e, double* I0, double* mu, double* alpha)
{
  size_t n = sizeof(double) * e->g.N_theta;
  if (I0) memcpy(I0, e->I0, n);
  if (mu) memcpy(mu, e->mu, n);
  if (alpha) memcpy(alpha, e->alpha, n);
}
void orc_engine_get_H(const orc_engine* e, double* H)
{
  memcpy(H, e->H, sizeof(double) * e->g.N_theta * ORC_RESPONSE_BINS);
}
void orc_engine_get_profile(const orc_engine* e, double* p)
{
  memcpy(p, e->profile, sizeof(double) * e->g.N_theta * ORC_PROFILE_RESOLUTION);
}
int64_t orc_engine_nnz(const orc_engine* e) { return e->col_ptr[(size_t)e->g.N_z * e->g.N_x]; }
int32_t orc_engine_get_col(const orc_engine* e, int32_t z, int32_t x, uint32_t* index,
                           double* values, int32_t cap)
{
  size_t line = (size_t)z * e->g.N_x + x;
  int64_t b = e->col_ptr[line], en = e->col_ptr[line + 1];
  int32_t n = (int32_t)(en - b);
  for (int32_t i = 0; i < n && i < cap; ++i) {
    index[i] = e->col_idx[b + i];
    values[i] = e->col_val[b + i];
  }
  return n;
}
void orc_engine_theta(const orc_engine* e, int32_t z, int32_t x, int32_t y, double* t1t2)
{
  const orc_geom* g = &e->g;
  size_t line = (size_t)z * g->N_x + x;
  double T1 = 0, T2 = 0;
  for (int64_t q = e->col_ptr[line]; q < e->col_ptr[line + 1]; ++q) {
    int32_t it = (int32_t)(e->col_idx[q] / (uint32_t)g->N_r);
    int32_t ir = (int32_t)(e->col_idx[q] % (uint32_t)g->N_r);
    double k0 = e->I0[it] * e->col_val[q];
    for (int k = 0; k < e->vlr_count[y]; ++k) {
      size_t eidx = ((size_t)it * g->N_r + ir) * g->N_t + e->vlr_t[y * VLR_CAP + k];
      double PE = k0 * e->vlr_val[y * VLR_CAP + k];
      double QC = 1.0;
      if (e->p.bf_mode && e->sel[eidx] == 0)
        QC = (e->p.bragg_delta * e->bragg) / (fabs(e->e[eidx]) * sqrt(e->w[eidx]));
      T2 += QC * (PE * PE * e->w[eidx]);
      T1 += QC * (e->e[eidx] * PE * e->w[eidx]);
    }
  }
  t1t2[0] = -T1;
  t1t2[1] = T2;
}

void orc_engine_get_sel(const orc_engine* e, uint8_t* sel)
{
  if (e->sel) memcpy(sel, e->sel, sino_size(&e->g));
}
double orc_engine_get_bragg(const orc_engine* e) { return e->bragg; }
void orc_engine_set_bragg(orc_engine* e, double T) { e->bragg = T; }
