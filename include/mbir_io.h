/* mbir_io.h -- C-ABI of the engine's files: the MRC tilt series in, the MRC volume and
 * the binary side outputs out (no device needed).  Every entry point returns >= 0 on
 * success and a negative code on failure.  Reference paths are relative to the
 * reference's Code/MBIRLib/. */
#ifndef MBIR_IO_H
#define MBIR_IO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mbir_mrc_info {
  int32_t nx, ny, nz;   /* columns, rows, sections */
  int32_t mode;         /* 0 u8, 1 i16, 2 f32, 6 u16 */
  int32_t next;         /* extended header bytes */
  int32_t nlabels;
  int32_t fei;          /* 1: FEI-style extended header (tilt, pixel size per section) */
  float amin, amax, amean;
} mbir_mrc_info;

/* MRCReader::readHeader (IOFilters/MRCReader.cpp:99-173): the header, and when the
 * extended header is FEI-style the per-section tilt angle and pixel size (arrays of
 * nz floats, may be NULL).  -1 cannot open, -2 short header, -3 short extended header. */
int32_t mbir_mrc_read_info(const char* path, mbir_mrc_info* info, float* tilts,
                           float* pixel_size);

typedef struct mbir_tilt_select {
  int32_t use_subvolume;
  int32_t x_start, x_end, y_start, y_end, z_start, z_end;
  double interpolate_factor;
  const int32_t* excluded_views; /* file sections left out */
  int32_t n_excluded;
  const float* tilts;            /* degrees per used view; NULL: the FEI header's */
  int32_t n_tilts;
} mbir_tilt_select;

typedef struct mbir_tilt_dims {
  int32_t N_theta, N_r, N_t;
  int32_t x_start, x_end, y_start, y_end; /* after the sub-volume adjustments */
  double delta_r, delta_t;                /* FEI pixel size in nm, else 1 */
} mbir_tilt_dims;

/* MRCSinogramInitializer::execute (GenericFilters/MRCSinogramInitializer.cpp:80-305):
 * the used part of a tilt series as counts [N_theta][N_r][N_t] (r = file column,
 * t = file row), the tilt angles (degrees) and the file sections used.  With
 * counts == NULL only dims is filled.  -5 unsupported mode, -6 no tilt angles,
 * -7 empty or out-of-range selection. */
int32_t mbir_mrc_read_tilt_series(const char* path, const mbir_tilt_select* sel,
                                  mbir_tilt_dims* dims, double* counts, double* angles,
                                  int32_t* good_views);

/* MRCWriter::write (IOFilters/MRCWriter.cpp:172-290): volume [N_z][N_x][N_y] as an
 * MRC float volume over columns [x0, x1), sections from N_z - 1 down to 0, FEI
 * extended header with pixel size length_x, amin/amax/amean at byte 76. */
int32_t mbir_mrc_write_volume(const char* path, const float* vol, int32_t N_z, int32_t N_x,
                              int32_t N_y, int32_t x0, int32_t x1, double length_x);

/* HAADF_ReconstructionEngine::computeOriginalXDims
 * (HAADF/HAADF_ReconstructionEngine_Extra.cpp:252-275). */
int32_t mbir_original_x_dims(int32_t N_x, double x0, double delta_xz, int32_t N_r,
                             double delta_r, int32_t* crop_start, int32_t* crop_end);

/* SinogramBinWriter::execute (IOFilters/SinogramBinWriter.cpp:66-130):
 * ((counts - e) * I0[th] + mu[th]) as doubles in th, r, t order. */
int32_t mbir_write_reconstructed_sinogram(const char* path, int32_t N_theta, int32_t N_r,
                                          int32_t N_t, const double* counts, const float* e,
                                          const double* I0, const double* mu);

/* RawGeometryWriter::execute (IOFilters/RawGeometryWriter.cpp:55-91): volume
 * [N_z][N_x][N_y] as doubles in y, x, z order (z fastest). */
int32_t mbir_write_raw_volume(const char* path, const float* vol, int32_t N_z, int32_t N_x,
                              int32_t N_y);

/* InitialReconstructionBinReader::execute
 * (GenericFilters/InitialReconstructionBinReader.cpp:95-150): the same layout into
 * [N_z][N_x][N_y]; interp = 1 reads a half-size volume and replicates each voxel
 * 2 x 2 x 2.  -2 short file. */
int32_t mbir_read_raw_volume(const char* path, double* vol, int32_t N_z, int32_t N_x,
                             int32_t N_y, int32_t interp);

#ifdef __cplusplus
}
#endif
#endif
