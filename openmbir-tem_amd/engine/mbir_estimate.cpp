// mbir_estimate.cpp -- the reference's initial-parameter estimators as host-side
// per-view reductions (include/mbir_estimate.h).
#include "mbir_estimate.h"

#include <cmath>
#include <vector>

extern "C" int32_t mbir_initial_offsets(const double* counts, int32_t N_theta, int32_t N_r,
                                        int32_t N_t, const double* angles_deg, double* offsets,
                                        double* fit)
{
  if (!counts || !angles_deg || !offsets || N_theta <= 0 || N_r <= 0 || N_t <= 0) return -1;
  const size_t per = (size_t)N_r * N_t;
  // normal equations of  mean[th] ~ a g[th] + b,  g = 1 / cos(theta)
  double sgg = 0, sg = 0, s1 = 0, sgy = 0, sy = 0;
  for (int th = 0; th < N_theta; ++th) {
    double sum = 0;
    const double* c = counts + (size_t)th * per;
    for (size_t i = 0; i < per; ++i) sum += c[i];
    const double mean = sum / (double)per;
    const double gk = 1.0 / std::cos(angles_deg[th] * M_PI / 180);
    sgg += gk * gk;
    sg += gk;
    s1 += 1.0;
    sgy += gk * mean;
    sy += mean;
  }
  const double det = sgg * s1 - sg * sg;
  if (!(std::fabs(det) > 1e-12 * sgg * s1)) return -2;
  const double a = (s1 * sgy - sg * sy) / det;
  const double b = (sgg * sy - sg * sgy) / det;
  for (int th = 0; th < N_theta; ++th) offsets[th] = b;
  if (fit) {
    fit[0] = a;
    fit[1] = b;
  }
  return 0;
}

extern "C" int32_t mbir_sigma_x_estimate(const double* counts, int32_t N_theta, int32_t N_r,
                                         int32_t N_t, const double* angles_deg, int32_t x0,
                                         int32_t x1, int32_t y0, int32_t y1, double target_gain,
                                         double thickness, int32_t use_bf_offset, double bf_offset,
                                         double* sigma_x)
{
  if (!counts || !angles_deg || !sigma_x || N_theta <= 0 || !(thickness > 0) ||
      !(target_gain != 0))
    return -1;
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (x1 < 0 || x1 > N_r) x1 = N_r;
  if (y1 < 0 || y1 > N_t) y1 = N_t;
  if (x0 >= x1 || y0 >= y1) return -1;
  const double n = (double)(x1 - x0) * (y1 - y0);
  double sum1 = 0;
  for (int th = 0; th < N_theta; ++th) {
    const double* c = counts + (size_t)th * N_r * N_t;
    auto val = [&](int x, int y, bool* ok) {
      const double v = c[(size_t)x * N_t + y];
      if (!use_bf_offset) {
        *ok = true;
        return v;
      }
      const double t = v + bf_offset;
      *ok = t > 0;
      return *ok ? std::log(t) : 0.0;
    };
    // mean absolute deviation of the view (:61-100)
    double mean = 0;
    bool ok;
    for (int x = x0; x < x1; ++x)
      for (int y = y0; y < y1; ++y) mean += val(x, y, &ok);
    mean /= n;
    double dev = 0;
    for (int x = x0; x < x1; ++x)
      for (int y = y0; y < y1; ++y) {
        const double v = val(x, y, &ok);
        if (ok) dev += std::fabs(v - mean);
      }
    dev /= n;
    // (:219-227)
    sum1 += (dev / target_gain) * std::cos(angles_deg[th] * M_PI / 180) / thickness;
  }
  *sigma_x = sum1 / N_theta;
  return 0;
}
