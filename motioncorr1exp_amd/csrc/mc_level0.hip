// mc_level0.hip -- kernels behind the cufunc.h-compatible entry points
// (padded row-major spectrum layout) and small helpers.
#include <hip/hip_runtime.h>

#include "mc_device.h"
#include "mc_kernels.h"

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

