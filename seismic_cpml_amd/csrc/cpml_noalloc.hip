// Allocation-free stability norm: cpml_velocity_norm_noalloc.
//
// The same quantity as cpml_gpu_velocity_norm (max over the grid of
// sqrt(vx^2 + vy^2 [+ vz^2]), the stability check of 2D2:716-726 and
// 3D:1183-1195), for callers that cannot allocate device memory at call time
// (code under HIP graph capture, tight serving loops): the caller passes the
// velocity arrays and a device double for the result, and the per-block
// partial maxima live in a file-scope device array instead of a buffer the
// library allocates.  Nothing here calls hipMalloc or hipFree.
//
// Every GPU holds its own copy of `scratch` for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress returns on
// that device (looked up once per device on the first call made with that
// device current).  All calls on one GPU share the array: calls that overlap
// on one GPU (several host threads, or streams that run concurrently) race
// on it and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "cpml_internal.h"

namespace {

// per-block partial maxima: one launch uses at most kNoAllocBlocks blocks
constexpr int kNoAllocBlocks = 1024;
constexpr int kNoAllocThreads = 256;
constexpr int kMaxDevices = 64;
// largest accepted grid (points); the kernels index with 64-bit offsets
constexpr double kNoAllocMaxPoints = 68719476736.0;  // 2^36

static __device__ __align__(256) double scratch[kNoAllocBlocks];

double *g_scratch[kMaxDevices] = {};
std::once_flag g_once[kMaxDevices];
hipError_t g_lookup[kMaxDevices] = {};

// max of the squared norm over this block's share of the points, written to
// part[blockIdx.x] (NaN wins so that a blow-up is reported)
template <typename T>
__global__ __launch_bounds__(kNoAllocThreads) void k_vnorm_partial(const T *vx, const T *vy,
                                                                   const T *vz, long n,
                                                                   double *part) {
  double m = 0.0;
  for (long q = blockIdx.x * (long)kNoAllocThreads + threadIdx.x; q < n;
       q += (long)gridDim.x * kNoAllocThreads) {
    T s = vx[q] * vx[q] + vy[q] * vy[q];
    if (vz) s = s + vz[q] * vz[q];
    const double d = (double)s;
    m = (d > m || d != d) ? d : m;
  }
  __shared__ double sm[kNoAllocThreads];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = kNoAllocThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = sm[threadIdx.x + s];
      if (o > sm[threadIdx.x] || o != o) sm[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(kNoAllocThreads) void k_vnorm_final(const double *part, int nblocks,
                                                                 double *out) {
  double m = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kNoAllocThreads) {
    const double o = part[b];
    m = (o > m || o != o) ? o : m;
  }
  __shared__ double sm[kNoAllocThreads];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = kNoAllocThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = sm[threadIdx.x + s];
      if (o > sm[threadIdx.x] || o != o) sm[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sqrt(sm[0]);
}

// device address of this GPU's scratch array (looked up once per device)
hipError_t scratch_ptr(int dev, double **p) {
  std::call_once(g_once[dev], [dev] {
    void *a = nullptr;
    g_lookup[dev] = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    g_scratch[dev] = static_cast<double *>(a);
  });
  *p = g_scratch[dev];
  return g_lookup[dev];
}

int fail_na(int code, const char *msg) {
  cpml::set_thread_error(msg);  // cpml_gpu_last_error(NULL, ...) returns it
  return code;
}

}  // namespace

extern "C" int cpml_velocity_norm_noalloc(int prec, int ndim, int NX, int NY, int NZ,
                                          const void *d_vx, const void *d_vy, const void *d_vz,
                                          double *d_vmax, void *stream) {
  if (prec != 4 && prec != 8) return fail_na(CPML_ERR_ARG, "prec must be 4 or 8");
  if (ndim != 2 && ndim != 3) return fail_na(CPML_ERR_ARG, "ndim must be 2 or 3");
  if (NX < 1 || NY < 1 || (ndim == 3 && NZ < 1))
    return fail_na(CPML_ERR_ARG, "NX, NY (and NZ in 3D) must be >= 1");
  if (!d_vx || !d_vy || (ndim == 3 && !d_vz) || !d_vmax)
    return fail_na(CPML_ERR_ARG, "velocity arrays and the result pointer must be given");
  const double points = (double)NX * NY * (ndim == 3 ? NZ : 1);
  if (points > kNoAllocMaxPoints)
    return fail_na(CPML_ERR_ARG, "grid larger than 2^36 points, the limit of this entry point");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail_na(CPML_ERR_NODEVICE, "no current HIP device");
  if (dev < 0 || dev >= kMaxDevices) return fail_na(CPML_ERR_ARG, "device ordinal above 63");
  double *part = nullptr;
  if (scratch_ptr(dev, &part) != hipSuccess || !part)
    return fail_na(CPML_ERR_HIP, "hipGetSymbolAddress of the scratch array failed");
  const long n = (long)points;
  long nb = (n + kNoAllocThreads - 1) / kNoAllocThreads;
  if (nb > kNoAllocBlocks) nb = kNoAllocBlocks;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (prec == 4)
    hipLaunchKernelGGL(k_vnorm_partial<float>, dim3((unsigned)nb), dim3(kNoAllocThreads), 0, s,
                       static_cast<const float *>(d_vx), static_cast<const float *>(d_vy),
                       ndim == 3 ? static_cast<const float *>(d_vz) : nullptr, n, part);
  else
    hipLaunchKernelGGL(k_vnorm_partial<double>, dim3((unsigned)nb), dim3(kNoAllocThreads), 0, s,
                       static_cast<const double *>(d_vx), static_cast<const double *>(d_vy),
                       ndim == 3 ? static_cast<const double *>(d_vz) : nullptr, n, part);
  hipLaunchKernelGGL(k_vnorm_final, dim3(1), dim3(kNoAllocThreads), 0, s, part, (int)nb, d_vmax);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_na(CPML_ERR_HIP, hipGetErrorString(e));
  return CPML_OK;
}
