// MBIRIO.cpp -- MRC and binary files of the engine (formats: see MBIRIO.h).
// Multi-byte fields are little-endian, at the byte offsets of the IMOD MRC layout.
#include "MBIRIO.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

namespace mbirtem {
namespace io {

namespace {

constexpr int kHeaderBytes = 1024;
constexpr int kFeiBytes = 128;  // 14 floats + 72 unused bytes per section
constexpr int kFeiPixelSize = 44;

struct File {
  FILE* f = nullptr;
  File(const std::string& p, const char* mode) : f(std::fopen(p.c_str(), mode)) {}
  ~File()
  {
    if (f) std::fclose(f);
  }
  explicit operator bool() const { return f != nullptr; }
};

template <typename T>
T get(const unsigned char* b, int off)
{
  T v;
  std::memcpy(&v, b + off, sizeof(T));
  return v;
}
template <typename T>
void put(unsigned char* b, int off, T v)
{
  std::memcpy(b + off, &v, sizeof(T));
}

// Labels that mark an FEI-style extended header (IOFilters/MRCReader.cpp:119-160).
bool fei_label(const std::string& l)
{
  return l.find("Fei Company") != std::string::npos || l.find("EIC project") != std::string::npos ||
         l.find("MCAP project, MDG, Copyright 2013") != std::string::npos;
}

int parse_header(FILE* f, MRCInfo* h)
{
  unsigned char b[kHeaderBytes];
  if (std::fread(b, 1, kHeaderBytes, f) != (size_t)kHeaderBytes) return -2;
  h->nx = get<int32_t>(b, 0);
  h->ny = get<int32_t>(b, 4);
  h->nz = get<int32_t>(b, 8);
  h->mode = get<int32_t>(b, 12);
  h->amin = get<float>(b, 76);
  h->amax = get<float>(b, 80);
  h->amean = get<float>(b, 84);
  h->next = get<int32_t>(b, 92);
  h->nlabels = get<int32_t>(b, 220);
  h->labels.assign(10, std::string());
  for (int i = 0; i < 10; ++i) {
    const char* s = reinterpret_cast<const char*>(b + 224 + 80 * i);
    h->labels[i] = std::string(s, strnlen(s, 80));
  }
  if (h->nx <= 0 || h->ny <= 0 || h->nz <= 0 || h->next < 0) return -2;
  std::vector<unsigned char> ext((size_t)h->next);
  if (h->next > 0 && std::fread(ext.data(), 1, ext.size(), f) != ext.size()) return -3;
  h->fei = fei_label(h->labels[0]) && (size_t)h->next >= (size_t)kFeiBytes * h->nz;
  h->tilt.clear();
  h->pixelsize.clear();
  if (h->fei) {
    for (int z = 0; z < h->nz; ++z) {
      h->tilt.push_back(get<float>(ext.data(), kFeiBytes * z));
      h->pixelsize.push_back(get<float>(ext.data(), kFeiBytes * z + kFeiPixelSize));
    }
  }
  return 1;
}

}  // namespace

int readMRCHeader(const std::string& path, MRCInfo* info)
{
  File f(path, "rb");
  if (!f) return -1;
  return parse_header(f.f, info);
}

int readMRCTiltSeries(const std::string& path, const TiltSeriesSelect& sel, TiltSeries* out)
{
  File f(path, "rb");
  if (!f) return -1;
  MRCInfo h;
  int rc = parse_header(f.f, &h);
  if (rc < 0) return rc;
  const long data0 = kHeaderBytes + (long)h.next;
  size_t tsize = 0;
  switch (h.mode) {
    case 1: case 6: tsize = 2; break;
    case 2: tsize = 4; break;
    default: return -5;  // MRCSinogramInitializer.cpp:270-279
  }
  TiltSeries& T = *out;
  T.fileX = h.nx;
  T.fileY = h.ny;
  T.fileZ = h.nz;
  int vmin[3] = {0, 0, 0}, vmax[3] = {h.nx - 1, h.ny - 1, h.nz - 1};
  int xs = 0, xe = h.nx - 1, ys = 0, ye = h.ny - 1;
  if (sel.useSubvolume) {
    // sub-volume symmetric about the centre column, rows a multiple of 3 x the
    // interpolation factor (MRCSinogramInitializer.cpp:105-198)
    xs = sel.xStart; xe = sel.xEnd; ys = sel.yStart; ye = sel.yEnd;
    vmin[0] = xs; vmin[1] = ys; vmin[2] = sel.zStart;
    vmax[0] = xe; vmax[1] = ye; vmax[2] = sel.zEnd;
    const double centre = (double)(h.nx / 2);
    double left = centre - vmin[0], right = vmax[0] - centre + 1;
    if (left < 0) {
      vmin[0] = (int)(centre + left);
      xs = vmin[0];
      left = -left;
    }
    if (right < 0) {
      vmax[0] = (int)(centre - right);
      xe = vmax[0];
      right = -right;
    }
    if (left != right) {
      if (left < right) {
        double tx = centre + left - 1;
        if (tx >= h.nx) tx = h.nx - 1;
        vmax[0] = (int)tx;
        xe = vmax[0];
      } else {
        double tx = centre - right;
        if (tx < 0) tx = 0;
        vmin[0] = (int)tx;
        xs = vmin[0];
      }
    }
    const int16_t disty = (int16_t)(ye - ys + 1);
    const int16_t period = (int16_t)(sel.interpolateFactor * 3);
    const int16_t rem = period != 0 ? (int16_t)(disty % period) : 0;
    if (rem != 0) {
      const int16_t add = (int16_t)(sel.interpolateFactor * 3 - rem);
      if (ye + add < h.ny) ye += add;
      else ye = h.ny - 1;
    }
    vmax[1] = ye;
  }
  for (int k = 0; k < 3; ++k) {
    const int n = k == 0 ? h.nx : (k == 1 ? h.ny : h.nz);
    if (vmin[k] < 0 || vmax[k] >= n || vmin[k] > vmax[k]) return -7;
  }
  T.xStart = xs; T.xEnd = xe; T.yStart = ys; T.yEnd = ye;
  T.N_r = vmax[0] - vmin[0] + 1;
  T.N_t = vmax[1] - vmin[1] + 1;
  if (h.fei && !h.pixelsize.empty()) {  // :220-224
    T.delta_r = h.pixelsize[0] * 1.0e9;
    T.delta_t = h.pixelsize[0] * 1.0e9;
  }
  T.goodViews.clear();
  for (int i = vmin[2]; i <= vmax[2]; ++i) {  // :227-244
    bool good = true;
    for (int v : sel.excludedViews) good = good && v != i;
    if (good) T.goodViews.push_back(i);
  }
  T.N_theta = (int)T.goodViews.size();
  if (T.N_theta == 0) return -7;
  T.angles.assign(T.N_theta, 0.0);
  for (int z = 0; z < T.N_theta; ++z) {
    if (!sel.tilts.empty()) {
      if ((int)sel.tilts.size() < T.N_theta) return -6;
      T.angles[z] = sel.tilts[z];
    } else if (h.fei) {
      T.angles[z] = h.tilt[T.goodViews[z]];
    } else {
      return -6;
    }
  }
  // counts[th][r = x][t = y] from sections [z][y][x] (:63-74)
  T.counts.assign((size_t)T.N_theta * T.N_r * T.N_t, 0.0);
  std::vector<unsigned char> row((size_t)T.N_r * tsize);
  for (int z = 0; z < T.N_theta; ++z) {
    const long sec = T.goodViews[z];
    for (int y = 0; y < T.N_t; ++y) {
      const long off = data0 + (long)tsize * (((long)sec * h.ny + (vmin[1] + y)) * h.nx + vmin[0]);
      if (std::fseek(f.f, off, SEEK_SET) != 0) return -4;
      if (std::fread(row.data(), 1, row.size(), f.f) != row.size()) return -4;
      for (int x = 0; x < T.N_r; ++x) {
        double v = 0;
        if (h.mode == 1) v = get<int16_t>(row.data(), 2 * x);
        else if (h.mode == 6) v = get<uint16_t>(row.data(), 2 * x);
        else v = get<float>(row.data(), 4 * x);
        T.counts[((size_t)z * T.N_r + x) * T.N_t + y] = v;
      }
    }
  }
  return 1;
}

void originalXDims(int N_x, double x0, double delta_xz, int N_r, double delta_r, int* crop_start,
                   int* crop_end)
{
  // HAADF/HAADF_ReconstructionEngine_Extra.cpp:252-275
  const double half = (N_r * delta_r) / 2;
  int s = 0, e = N_x;
  for (int j = 0; j < N_x; ++j) {
    const double x = x0 + ((double)j + 0.5) * delta_xz;
    if (x >= -half && x <= half) {
      s = j;
      break;
    }
  }
  for (int j = N_x - 1; j >= 0; --j) {
    const double x = x0 + ((double)j + 0.5) * delta_xz;
    if (x >= -half && x <= half) {
      e = j;
      break;
    }
  }
  *crop_start = s;
  *crop_end = e;
}

int writeMRCVolume(const std::string& path, const float* vol, int N_z, int N_x, int N_y, int x0,
                   int x1, double length_x)
{
  if (!vol || N_z <= 0 || N_x <= 0 || N_y <= 0 || x0 < 0 || x1 > N_x || x0 >= x1) return -2;
  File f(path, "wb");
  if (!f) return -1;
  const int nx = x1 - x0, ny = N_y, nz = N_z;
  unsigned char b[kHeaderBytes];
  std::memset(b, 0, sizeof(b));
  // the header of IOFilters/MRCWriter.cpp:90-168 with the written dimensions (:199-212)
  put<int32_t>(b, 0, nx);
  put<int32_t>(b, 4, ny);
  put<int32_t>(b, 8, nz);
  put<int32_t>(b, 12, 2);  // float
  put<int32_t>(b, 28, nx);
  put<int32_t>(b, 32, ny);
  put<int32_t>(b, 36, nz);
  put<float>(b, 40, (float)nx);
  put<float>(b, 44, (float)ny);
  put<float>(b, 48, (float)nz);
  put<float>(b, 52, 90.f);
  put<float>(b, 56, 90.f);
  put<float>(b, 60, 90.f);
  put<int32_t>(b, 64, 1);
  put<int32_t>(b, 68, 2);
  put<int32_t>(b, 72, 3);
  put<int32_t>(b, 92, kFeiBytes * nz);
  std::memcpy(b + 208, "MAP ", 4);
  b[212] = 68;  // little-endian stamp
  b[213] = 65;
  // label 0 marks the FEI-style extended header for readers that key on it
  const char* labels[2] = {"Fei Company (C) Copyright 2003",
                           "Reconstruction by OpenMBIR-TEM (MI355X engine)"};
  put<int32_t>(b, 220, 2);
  for (int i = 0; i < 2; ++i) std::strncpy(reinterpret_cast<char*>(b + 224 + 80 * i), labels[i], 79);
  if (std::fwrite(b, 1, kHeaderBytes, f.f) != (size_t)kHeaderBytes) return -3;
  unsigned char fei[kFeiBytes];
  std::memset(fei, 0, sizeof(fei));
  put<float>(fei, kFeiPixelSize, (float)length_x);
  for (int z = 0; z < nz; ++z)
    if (std::fwrite(fei, 1, kFeiBytes, f.f) != (size_t)kFeiBytes) return -3;
  // sections z = N_z - 1 .. 0, rows y, columns x; the statistics as the reference
  // forms them: a float running sum for the mean, the maximum starting from 0
  // (the smallest positive double rounded to float) and the minimum from +inf
  std::vector<float> slice((size_t)nx * ny);
  float mean = 0.f;
  float dmax = (float)std::numeric_limits<double>::min();
  float dmin = (float)std::numeric_limits<double>::max();
  size_t count = 0;
  for (int z = N_z - 1; z >= 0; --z) {
    size_t i = 0;
    for (int y = 0; y < N_y; ++y)
      for (int x = x0; x < x1; ++x) {
        const double d = vol[((size_t)z * N_x + x) * N_y + y];
        slice[i] = (float)d;
        mean += slice[i];
        ++count;
        if (d > dmax) dmax = (float)d;
        if (d < dmin) dmin = (float)d;
        ++i;
      }
    if (std::fwrite(slice.data(), sizeof(float), slice.size(), f.f) != slice.size()) return -3;
  }
  mean = mean / (float)count;
  if (std::fseek(f.f, 76, SEEK_SET) != 0) return -3;
  const float st[3] = {dmin, dmax, mean};
  if (std::fwrite(st, sizeof(float), 3, f.f) != 3) return -3;
  return 1;
}

int writeDoubles(const std::string& path, const double* v, size_t n)
{
  File f(path, "wb");
  if (!f) return -1;
  return std::fwrite(v, sizeof(double), n, f.f) == n ? 1 : -3;
}

int readDoubles(const std::string& path, double* v, size_t n)
{
  File f(path, "rb");
  if (!f) return -1;
  return std::fread(v, sizeof(double), n, f.f) == n ? 1 : -2;
}

int writeReconstructedSinogram(const std::string& path, int N_theta, int N_r, int N_t,
                               const double* counts, const float* e, const double* I0,
                               const double* mu)
{
  File f(path, "wb");
  if (!f) return -1;
  std::vector<double> row((size_t)N_t);
  for (int th = 0; th < N_theta; ++th)
    for (int r = 0; r < N_r; ++r) {
      const size_t b = ((size_t)th * N_r + r) * N_t;
      for (int t = 0; t < N_t; ++t) {
        double v = counts[b + t] - (double)e[b + t];  // HAADF_ReconstructionEngine.cpp:858-867
        v *= I0[th];
        v += mu[th];
        row[t] = v;
      }
      if (std::fwrite(row.data(), sizeof(double), row.size(), f.f) != row.size()) return -3;
    }
  return 1;
}

int writeRawVolume(const std::string& path, const float* vol, int N_z, int N_x, int N_y)
{
  File f(path, "wb");
  if (!f) return -1;
  std::vector<double> col((size_t)N_z);
  for (int y = 0; y < N_y; ++y)
    for (int x = 0; x < N_x; ++x) {
      for (int z = 0; z < N_z; ++z) col[z] = vol[((size_t)z * N_x + x) * N_y + y];
      if (std::fwrite(col.data(), sizeof(double), col.size(), f.f) != col.size()) return -3;
    }
  return 1;
}

int readRawVolume(const std::string& path, double* vol, int N_z, int N_x, int N_y, int interp)
{
  File f(path, "rb");
  if (!f) return -1;
  if (interp == 1) {
    // half-size volume, every voxel replicated 2 x 2 x 2
    // (GenericFilters/InitialReconstructionBinReader.cpp:103-130)
    const int ye = N_y / 2, xe = N_x / 2, ze = N_z / 2;
    std::fill(vol, vol + (size_t)N_z * N_x * N_y, 0.0);
    for (int y = 0; y < ye; ++y)
      for (int x = 0; x < xe; ++x)
        for (int z = 0; z < ze; ++z) {
          double v;
          if (std::fread(&v, sizeof(double), 1, f.f) != 1) return -2;
          for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q)
              for (int r = 0; r < 2; ++r)
                vol[((size_t)(2 * z + p) * N_x + (2 * x + q)) * N_y + (2 * y + r)] = v;
        }
    return 1;
  }
  for (int y = 0; y < N_y; ++y)
    for (int x = 0; x < N_x; ++x)
      for (int z = 0; z < N_z; ++z) {
        double v;
        if (std::fread(&v, sizeof(double), 1, f.f) != 1) return -2;
        vol[((size_t)z * N_x + x) * N_y + y] = v;
      }
  return 1;
}

}  // namespace io
}  // namespace mbirtem
