// MBIRIO.h -- the engine's files: the MRC tilt series in, the MRC volume and the
// binary side outputs out, in the formats the reference engine reads and writes.
//
// Reference files (Code/MBIRLib/):
//   MRC header layout (IMOD 2.6.20+, 1024 bytes) and the FEI extended header
//       IOFilters/MRCHeader.h:40-216
//   MRC header / data reading              IOFilters/MRCReader.cpp:99-173, 205-368
//   tilt series -> sinogram (sub-volume, excluded views, x/y ranges)
//                                          GenericFilters/MRCSinogramInitializer.cpp:63-305
//   MRC volume writing (z reversed, amin/amax/amean patched at byte 76)
//                                          IOFilters/MRCWriter.cpp:90-290
//   per-view gains / offsets / variances   IOFilters/NuisanceParamWriter.cpp:78-147,
//                                          IOFilters/GainsOffsetsReader.cpp:64-230
//   ReconstructedSinogram.bin              IOFilters/SinogramBinWriter.cpp:66-130
//   ReconstructedObject.bin / initial reconstruction (y, x, z order, z fastest)
//                                          IOFilters/RawGeometryWriter.cpp:55-91,
//                                          GenericFilters/InitialReconstructionBinReader.cpp:95-150
//   x crop of the written volume           HAADF/HAADF_ReconstructionEngine_Extra.cpp:252-275
// Every function returns >= 0 on success and a negative code on failure.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mbirtem {
namespace io {

// What the engine uses of an MRC file's header.
struct MRCInfo {
  int32_t nx = 0, ny = 0, nz = 0;  // columns, rows, sections
  int32_t mode = 0;                // 0 u8, 1 i16, 2 f32, 6 u16
  int32_t next = 0;                // extended header bytes
  float amin = 0, amax = 0, amean = 0;
  int32_t nlabels = 0;
  std::vector<std::string> labels;  // 10 labels of 80 characters
  bool fei = false;                 // FEI-style extended header present
  std::vector<float> tilt;          // per section: a_tilt (FEI)
  std::vector<float> pixelsize;     // per section: pixel size (FEI, metres)
};

// -1 cannot open, -2 short header, -3 short extended header
int readMRCHeader(const std::string& path, MRCInfo* info);

// The part of a tilt series the reconstruction uses (the reference's TomoInputs
// fields of the same names): sections (views) zStart..zEnd minus excludedViews,
// detector columns xStart..xEnd (r) and rows yStart..yEnd (t).
struct TiltSeriesSelect {
  bool useSubvolume = false;
  int xStart = 0, xEnd = 0, yStart = 0, yEnd = 0, zStart = 0, zEnd = 0;
  double interpolateFactor = 1;
  std::vector<int> excludedViews;
  std::vector<float> tilts;  // per used view; empty: the FEI header's a_tilt
};

struct TiltSeries {
  int N_theta = 0, N_r = 0, N_t = 0;
  double delta_r = 1, delta_t = 1;   // from the FEI pixel size (nm) when present
  std::vector<double> counts;        // [N_theta][N_r][N_t]
  std::vector<double> angles;        // degrees
  std::vector<int> goodViews;        // file sections used, in order
  int xStart = 0, xEnd = 0, yStart = 0, yEnd = 0;  // after the reference's adjustments
  int fileX = 0, fileY = 0, fileZ = 0;
};

// -1..-3 header, -4 short data, -5 unsupported mode, -6 no tilt angles, -7 bad range
int readMRCTiltSeries(const std::string& path, const TiltSeriesSelect& sel, TiltSeries* out);

// Volume [N_z][N_x][N_y] -> MRC (mode 2), columns x in [x0, x1), rows y, sections z
// from N_z - 1 down to 0; the FEI extended header carries pixel size = length_x.
int writeMRCVolume(const std::string& path, const float* vol, int N_z, int N_x, int N_y, int x0,
                   int x1, double length_x);

// First and last x index of the volume whose centre lies over the detector.
void originalXDims(int N_x, double x0, double delta_xz, int N_r, double delta_r, int* crop_start,
                   int* crop_end);

// n doubles, raw.
int writeDoubles(const std::string& path, const double* v, size_t n);
// n doubles; -1 cannot open, -2 fewer than n values
int readDoubles(const std::string& path, double* v, size_t n);

// ReconstructedSinogram.bin: ((counts - e) * I0[th] + mu[th]) as doubles, th, r, t order.
int writeReconstructedSinogram(const std::string& path, int N_theta, int N_r, int N_t,
                               const double* counts, const float* e, const double* I0,
                               const double* mu);

// ReconstructedObject.bin: volume [N_z][N_x][N_y] as doubles in y, x, z order (z fastest).
int writeRawVolume(const std::string& path, const float* vol, int N_z, int N_x, int N_y);
// The same file back into [N_z][N_x][N_y]; interp = 1 reads a volume of half the size
// in each direction and replicates every voxel 2 x 2 x 2 (multi-resolution hand-over).
int readRawVolume(const std::string& path, double* vol, int N_z, int N_x, int N_y, int interp);

}  // namespace io
}  // namespace mbirtem
