// mc_sum.hip -- Fourier phase-shift-and-sum of a stack of spectra.
//
// Reference: cuShift (cufunc.cu:421-443) multiplies a frame spectrum by
// e^{i (shx x + shy y')}, shx = sx 2 pi/n, and cuAdd (cufunc.cu:292-301)
// accumulates it into the sum (iteration sums DFAlign.cu:974-989, final sum
// DFAlign.cu:1143-1267).  One pass here reads every frame once and writes
// the sum once: the F shifted spectra are never materialised.  The phase
// (sx kx + sy ky')/n is reduced to [0, 1) in FP64 before the FP32 sincospi,
// so its error does not grow with the shift or the frequency.
#include <hip/hip_runtime.h>

#include "mc_device.h"
#include "mc_kernels.h"

namespace mc {

__global__ __launch_bounds__(256) void k_shift_sum(const float2* __restrict__ spectra,
                                                   size_t spec_stride, int nframes, int n,
                                                   const double2* __restrict__ shifts,
                                                   float2* __restrict__ sum,
                                                   float2* __restrict__ sum2, int f2_last) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  const int kx = (int)(e / n), ky = (int)(e % n);
  const int yy = ky < n / 2 ? ky : ky - n;
  const double inv_n = 1.0 / n;
  float2 acc = make_float2(0.f, 0.f), acc2 = make_float2(0.f, 0.f);
  for (int f = 0; f < nframes; f++) {
    const double2 s = shifts[f];
    if (s.x != s.x) continue;  // NaN: frame not summed
    const float2 v = spectra[(size_t)f * spec_stride + e];
    double u = (s.x * kx + s.y * yy) * inv_n;
    u -= floor(u);
    float sn, cs;
    sincospif(2.0f * (float)u, &sn, &cs);
    const float2 t = cmul(v, make_float2(cs, sn));
    acc = cadd(acc, t);
    if (sum2 && f <= f2_last) acc2 = cadd(acc2, t);
  }
  sum[e] = acc;
  if (sum2) sum2[e] = acc2;
}

hipError_t launch_shift_sum(const float2* spectra, size_t spec_stride, int nframes, int n,
                            const double2* shifts, float2* sum, float2* sum2, int f2_last,
                            hipStream_t st) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  hipLaunchKernelGGL(k_shift_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     spectra, spec_stride, nframes, n, shifts, sum, sum2, f2_last);
  return hipGetLastError();
}

// ------------------------------------------- elementwise (cufunc.cu:292-312)
__global__ void k_add(float* __restrict__ dst, const float* __restrict__ src, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] += src[i];
}
__global__ void k_scale(float* __restrict__ dst, float s, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] *= s;
}

hipError_t launch_add(float* dst, const float* src, size_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}
hipError_t launch_scale(float* dst, float s, size_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, s, n);
  return hipGetLastError();
}

}  // namespace mc
