// mbir_io_capi.cpp -- extern "C" entry points of include/mbir_io.h over MBIRIO.
#include "mbir_io.h"

#include "MBIRIO.h"

using namespace mbirtem;

extern "C" int32_t mbir_mrc_read_info(const char* path, mbir_mrc_info* info, float* tilts,
                                      float* pixel_size)
{
  if (!path || !info) return -1;
  io::MRCInfo h;
  const int rc = io::readMRCHeader(path, &h);
  if (rc < 0) return rc;
  info->nx = h.nx; info->ny = h.ny; info->nz = h.nz;
  info->mode = h.mode; info->next = h.next; info->nlabels = h.nlabels;
  info->fei = h.fei ? 1 : 0;
  info->amin = h.amin; info->amax = h.amax; info->amean = h.amean;
  for (int z = 0; z < h.nz && h.fei; ++z) {
    if (tilts) tilts[z] = h.tilt[z];
    if (pixel_size) pixel_size[z] = h.pixelsize[z];
  }
  return 0;
}

extern "C" int32_t mbir_mrc_read_tilt_series(const char* path, const mbir_tilt_select* sel,
                                             mbir_tilt_dims* dims, double* counts, double* angles,
                                             int32_t* good_views)
{
  if (!path || !sel || !dims) return -1;
  io::TiltSeriesSelect s;
  s.useSubvolume = sel->use_subvolume != 0;
  s.xStart = sel->x_start; s.xEnd = sel->x_end;
  s.yStart = sel->y_start; s.yEnd = sel->y_end;
  s.zStart = sel->z_start; s.zEnd = sel->z_end;
  s.interpolateFactor = sel->interpolate_factor;
  for (int i = 0; i < sel->n_excluded && sel->excluded_views; ++i) s.excludedViews.push_back(sel->excluded_views[i]);
  for (int i = 0; i < sel->n_tilts && sel->tilts; ++i) s.tilts.push_back(sel->tilts[i]);
  io::TiltSeries t;
  const int rc = io::readMRCTiltSeries(path, s, &t);
  if (rc < 0) return rc;
  dims->N_theta = t.N_theta; dims->N_r = t.N_r; dims->N_t = t.N_t;
  dims->x_start = t.xStart; dims->x_end = t.xEnd;
  dims->y_start = t.yStart; dims->y_end = t.yEnd;
  dims->delta_r = t.delta_r; dims->delta_t = t.delta_t;
  if (counts)
    for (size_t i = 0; i < t.counts.size(); ++i) counts[i] = t.counts[i];
  if (angles)
    for (int i = 0; i < t.N_theta; ++i) angles[i] = t.angles[i];
  if (good_views)
    for (int i = 0; i < t.N_theta; ++i) good_views[i] = t.goodViews[i];
  return 0;
}

extern "C" int32_t mbir_mrc_write_volume(const char* path, const float* vol, int32_t N_z,
                                         int32_t N_x, int32_t N_y, int32_t x0, int32_t x1,
                                         double length_x)
{
  if (!path) return -1;
  const int rc = io::writeMRCVolume(path, vol, N_z, N_x, N_y, x0, x1, length_x);
  return rc < 0 ? rc : 0;
}

extern "C" int32_t mbir_original_x_dims(int32_t N_x, double x0, double delta_xz, int32_t N_r,
                                        double delta_r, int32_t* crop_start, int32_t* crop_end)
{
  if (!crop_start || !crop_end) return -1;
  int s = 0, e = 0;
  io::originalXDims(N_x, x0, delta_xz, N_r, delta_r, &s, &e);
  *crop_start = s;
  *crop_end = e;
  return 0;
}

extern "C" int32_t mbir_write_reconstructed_sinogram(const char* path, int32_t N_theta, int32_t N_r,
                                                     int32_t N_t, const double* counts,
                                                     const float* e, const double* I0,
                                                     const double* mu)
{
  if (!path || !counts || !e || !I0 || !mu) return -1;
  const int rc = io::writeReconstructedSinogram(path, N_theta, N_r, N_t, counts, e, I0, mu);
  return rc < 0 ? rc : 0;
}

extern "C" int32_t mbir_write_raw_volume(const char* path, const float* vol, int32_t N_z,
                                         int32_t N_x, int32_t N_y)
{
  if (!path || !vol) return -1;
  const int rc = io::writeRawVolume(path, vol, N_z, N_x, N_y);
  return rc < 0 ? rc : 0;
}

extern "C" int32_t mbir_read_raw_volume(const char* path, double* vol, int32_t N_z, int32_t N_x,
                                        int32_t N_y, int32_t interp)
{
  if (!path || !vol) return -1;
  const int rc = io::readRawVolume(path, vol, N_z, N_x, N_y, interp);
  return rc < 0 ? rc : 0;
}
