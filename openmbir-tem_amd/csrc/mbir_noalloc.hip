// mbir_noalloc.hip -- the forward projection for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops): the same kernels as
// mbir_job_forward_project, with the FP64 projection accumulator and the uploaded
// volume in a library-wide device array instead of per-call allocations.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mbir_host.h"

namespace {

// 64 MiB per GPU, held from the library's load until the process ends: the first
// kNoallocEntries doubles accumulate the projection, the rest holds the volume
// (kNoallocVoxels floats).
constexpr size_t kNoallocBytes = 64ull << 20;
constexpr size_t kNoallocEntries = 4ull << 20;  // N_theta N_r N_t at most (32 MiB of FP64)
constexpr size_t kNoallocVoxels = 8ull << 20;   // N_z N_x N_y at most (32 MiB of float)
static_assert(kNoallocEntries * sizeof(double) + kNoallocVoxels * sizeof(float) <= kNoallocBytes,
              "scratch layout");

__device__ __align__(256) double scratch[kNoallocBytes / sizeof(double)];

// The array's device address, looked up once per device with that device current.
// In host code `scratch` names a host-side stand-in, never the device memory.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

int scratch_on(int device, double** out)
{
  if (device < 0 || device >= kMaxDevices) return MBIR_ERR_INPUT;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    HIP_TRY(e);
    g_ptr[device] = static_cast<double*>(p);
  }
  *out = g_ptr[device];
  return MBIR_OK;
}

}  // namespace

extern "C" int32_t mbir_job_forward_project_noalloc(mbir_job* j, const float* volume, double* out)
{
  if (!j || !out) return MBIR_ERR_INPUT;
  const size_t M = (size_t)j->g.N_theta * j->g.N_r * j->g.N_t;
  const size_t V = (size_t)j->g.N_z * j->g.N_x * j->g.N_y;
  if (M > kNoallocEntries || (volume && V > kNoallocVoxels)) {
    std::fprintf(stderr,
                 "mbir: mbir_job_forward_project_noalloc: %zu sinogram entries / %zu voxels exceed "
                 "the scratch (%zu / %zu)\n",
                 M, V, kNoallocEntries, kNoallocVoxels);
    return MBIR_ERR_INPUT;
  }
  double* base = nullptr;
  int rc = scratch_on(j->ctx->device, &base);
  if (rc < 0) return rc;
  double* yest = base;
  float* vol = reinterpret_cast<float*>(base + kNoallocEntries);
  hipStream_t s = j->stream;
  const float* src = j->x.p;
  if (volume) {
    HIP_TRY(hipMemcpyAsync(vol, volume, V * sizeof(float), hipMemcpyHostToDevice, s));
    src = vol;
  }
  // the accumulator keeps the last call's sums: cleared by a kernel (no memset node
  // for callers that capture the stream)
  HIP_TRY(mbir::launch_fill_f64(yest, 0.0, M, s));
  HIP_TRY(mbir::launch_forward_project(j->J, src, yest, s));
  HIP_TRY(hipMemcpyAsync(out, yest, M * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return MBIR_OK;
}
