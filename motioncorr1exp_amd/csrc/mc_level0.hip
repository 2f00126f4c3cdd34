// mc_level0.hip -- kernels behind the cufunc.h-compatible entry points
// (padded row-major spectrum layout), small helpers, and the FindShift entry
// point that allocates nothing (mc_find_shift_noalloc).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/mc_abi.h"
#include "mc_device.h"
#include "mc_kernels.h"

// Records the text returned by mc_last_error() (defined in mc_align.cpp).
int mc_internal_set_error(int code, const char* msg);

namespace mc {

// [kx][ky] (H rows of n)  ->  [ky][kx] (n rows of H), 32 x 32 LDS tiles.
__global__ __launch_bounds__(256) void k_transpose(const float2* __restrict__ src, int rows,
                                                   int cols, float2* __restrict__ dst) {
  __shared__ float2 t[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    if (r < rows && c < cols) t[j][tx] = src[(size_t)r * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + tx;  // dst[c][r]
    if (r < rows && c < cols) dst[(size_t)c * rows + r] = t[tx][j];
  }
}

static hipError_t transpose(const float2* src, int rows, int cols, float2* dst, hipStream_t st) {
  dim3 g((cols + 31) / 32, (rows + 31) / 32);
  hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, st, src, rows, cols, dst);
  return hipGetLastError();
}

hipError_t launch_spec_to_padded(const float2* spec, int n, float2* padded, hipStream_t st) {
  return transpose(spec, n / 2 + 1, n, padded, st);
}
hipError_t launch_padded_to_spec(const float2* padded, int n, float2* spec, hipStream_t st) {
  return transpose(padded, n, n / 2 + 1, spec, st);
}

// cuShiftCC (cufunc.cu:343-361): FP32 phase, exactly as the reference.
__global__ void k_mask_shift_cc(float2* dfft, const float2* dsum, const MaskElem* mask,
                                float shx, float shy, int size) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= size) return;
  const MaskElem pos = mask[id];
  const float sh = shx * pos.x + shy * pos.y;
  const float2 ph = make_float2(cosf(sh), sinf(sh));
  float2 v = cmulc(dsum[id], cmul(dfft[id], ph));
  v.x /= size;
  v.x *= pos.z;
  v.y /= size;
  v.y *= pos.z;
  dfft[id] = v;
}

// cuShift (cufunc.cu:372-385)
__global__ void k_mask_shift(float2* dfft, const MaskElem* mask, float shx, float shy,
                             int size) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= size) return;
  const MaskElem pos = mask[id];
  const float sh = shx * pos.x + shy * pos.y;
  dfft[id] = cmul(dfft[id], make_float2(cosf(sh), sinf(sh)));
}

// cuFFTLogModulus (cufunc.cu:178-183)
__global__ void k_log_modulus(float* dmod, const float2* dfft, float scale, int size) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= size) return;
  const float2 v = dfft[id];
  dmod[id] = logf(1.f + sqrtf(v.x * v.x + v.y * v.y) * scale);
}

hipError_t launch_mask_shift_cc(float2* dfft, const float2* dsum, const MaskElem* mask,
                                float shx, float shy, int n, hipStream_t st) {
  const int size = (n / 2 + 1) * n;
  hipLaunchKernelGGL(k_mask_shift_cc, dim3((size + 255) / 256), dim3(256), 0, st, dfft, dsum,
                     mask, shx, shy, size);
  return hipGetLastError();
}
hipError_t launch_mask_shift(float2* dfft, const MaskElem* mask, float shx, float shy, int n,
                             hipStream_t st) {
  const int size = (n / 2 + 1) * n;
  hipLaunchKernelGGL(k_mask_shift, dim3((size + 255) / 256), dim3(256), 0, st, dfft, mask, shx,
                     shy, size);
  return hipGetLastError();
}
hipError_t launch_log_modulus(float* dmod, const float2* dfft, int n, float scale,
                              hipStream_t st) {
  const int size = (n / 2 + 1) * n;
  hipLaunchKernelGGL(k_log_modulus, dim3((size + 255) / 256), dim3(256), 0, st, dmod, dfft,
                     scale, size);
  return hipGetLastError();
}

__global__ void k_dc_gather(const float2* spectra, size_t stride, int nframes, double scale,
                            double* out) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f < nframes) out[f] = (double)spectra[(size_t)f * stride].x * scale;
}
hipError_t launch_dc_gather(const float2* spectra, size_t stride, int nframes, double scale,
                            double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_dc_gather, dim3((nframes + 255) / 256), dim3(256), 0, st, spectra, stride,
                     nframes, scale, out);
  return hipGetLastError();
}

}  // namespace mc

// ------------------------------------------------ FindShift without allocation
// Device memory of mc_find_shift_noalloc: the box copied out of the caller's
// map (MC_NOALLOC_MAX_BOX^2 floats), then the peak record of the search (the
// result triple and the per-slice maxima, find_peaks_out_elems(1) doubles).
// One copy per GPU, held until the process ends.  Every element a call reads
// is written earlier in the same call, so nothing needs clearing between
// calls; calls that overlap on one GPU race on it.
namespace {
constexpr size_t kNoallocBoxBytes = sizeof(float) * MC_NOALLOC_MAX_BOX * MC_NOALLOC_MAX_BOX;
constexpr size_t kNoallocPeakDoubles = 32;
constexpr size_t kNoallocBytes = kNoallocBoxBytes + sizeof(double) * kNoallocPeakDoubles;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

// Device address of `scratch` on the current device, looked up once per
// device with that device current (the host-side name is not a device
// address).
static int noalloc_scratch(unsigned char** out) {
  static std::mutex mu;
  static unsigned char* ptrs[64] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return mc_internal_set_error(MC_ERR_DEVICE, hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return mc_internal_set_error(MC_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptrs[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      char buf[256];
      snprintf(buf, sizeof buf, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return mc_internal_set_error(MC_ERR_DEVICE, buf);
    }
    ptrs[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = ptrs[dev];
  return MC_OK;
}

// mc_find_shift (FindShift, cufunc.cu:396-497) with its box copy and peak
// record in `scratch`: the same copy, kernels and read-backs, so the same
// results, and no hipMalloc or hipFree.
extern "C" int mc_find_shift_noalloc(const float* dsrc, int n, float* hboxmap, int box,
                                     float* sx, float* sy, float* cc, int wnoise) {
  if (!dsrc || box < 16 || box > n)
    return mc_internal_set_error(MC_ERR_ARG, "bad box");
  if (box > MC_NOALLOC_MAX_BOX) {
    char buf[160];
    snprintf(buf, sizeof buf, "box %d exceeds MC_NOALLOC_MAX_BOX (%d) of mc_find_shift_noalloc",
             box, MC_NOALLOC_MAX_BOX);
    return mc_internal_set_error(MC_ERR_ARG, buf);
  }
  if (mc::find_peaks_out_elems(1) > kNoallocPeakDoubles)
    return mc_internal_set_error(MC_ERR_ARG, "peak record exceeds the fixed scratch");
  unsigned char* base = nullptr;
  const int rc = noalloc_scratch(&base);
  if (rc != MC_OK) return rc;
  float* dbox = reinterpret_cast<float*>(base);
  double* dres = reinterpret_cast<double*>(base + kNoallocBoxBytes);
  const int ori = (n - box) / 2;
  hipError_t e = hipMemcpy2D(dbox, sizeof(float) * box, dsrc + (size_t)ori * (n + 2) + ori,
                             sizeof(float) * (n + 2), sizeof(float) * box, box,
                             hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = mc::launch_find_peaks(dbox, 1, box, n, wnoise, dres, nullptr);
  double r[3] = {0, 0, 0};
  if (e == hipSuccess) e = hipMemcpy(r, dres, sizeof r, hipMemcpyDeviceToHost);
  if (e == hipSuccess && hboxmap)
    e = hipMemcpy(hboxmap, dbox, sizeof(float) * (size_t)box * box, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof buf, "FindShift (no allocation): %s", hipGetErrorString(e));
    return mc_internal_set_error(MC_ERR_DEVICE, buf);
  }
  if (sx) *sx = (float)r[0];
  if (sy) *sy = (float)r[1];
  if (cc) *cc = (float)r[2];
  return MC_OK;
}
