/* mbir_estimate.h -- initial parameter estimators of the reference (host-side per-view
 * reductions over the measured data, no device needed).  Reference paths are relative
 * to the reference's Code/MBIRLib/.  Every entry point returns >= 0 on success and a
 * negative code on failure. */
#ifndef MBIR_ESTIMATE_H
#define MBIR_ESTIMATE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ComputeInitialOffsets::execute (GenericFilters/ComputeInitialOffsets.cpp:58-186):
 * the per-view mean of counts [N_theta][N_r][N_t] fitted by least squares as
 * a / cos(theta) + b; every view's initial offset is b.  The fitted (a, b) go to
 * fit[2] when fit is not NULL.  -1 bad input, -2 singular fit (fewer than two
 * distinct 1/cos(theta)). */
int32_t mbir_initial_offsets(const double* counts, int32_t N_theta, int32_t N_r, int32_t N_t,
                             const double* angles_deg, double* offsets, double* fit);

/* SigmaXEstimation::execute (GenericFilters/SigmaXEstimation.cpp:137-234): for every
 * view the mean absolute deviation of the image columns [x0, x1) and rows [y0, y1)
 * (of log(counts + bf_offset) over the positive entries when use_bf_offset), divided
 * by target_gain, times cos(theta) / thickness, averaged over the views.  counts is a
 * tilt series [N_theta][N_r][N_t] (r = image column, t = image row); x1 or y1 < 0
 * means the full width.  -1 bad input. */
int32_t mbir_sigma_x_estimate(const double* counts, int32_t N_theta, int32_t N_r, int32_t N_t,
                              const double* angles_deg, int32_t x0, int32_t x1, int32_t y0,
                              int32_t y1, double target_gain, double thickness,
                              int32_t use_bf_offset, double bf_offset, double* sigma_x);

#ifdef __cplusplus
}
#endif
#endif
