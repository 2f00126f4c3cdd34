// synth_gpu.hip -- on-device synthetic movies for the batched benchmark and
// tests (libmc_synth_gpu.so; not part of libmc).
//
// Phantom: g(x) = a * sum_j cos(2 pi (k_j . x + p_j)) over nw plane waves
// whose frequencies the caller draws from P(k) = 1/(1+(k/0.05)^2) (a
// spectral approximation of a Gaussian random field, a = sqrt(2/nw) so that
// var g = 1); intensity I = dose * exp(0.5 g - 0.125).  Frame f holds
// Poisson(I(x - t_f)) with t_f the caller's drift.  The phantom is a
// continuous function, so the shift of every frame is exact.  Poisson draws
// use a counter-based hash of (seed, frame, pixel) and a bounded inverse-CDF
// search (k <= 255, so uint8 output never wraps).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

constexpr int MAXW = 1024;

__global__ __launch_bounds__(256) void k_synth(void* out, int dtype, int nx, int ny,
                                               const double2* __restrict__ drift,
                                               const double* __restrict__ waves, int nw, float amp,
                                               float dose, uint64_t seed) {
  __shared__ double w[3 * MAXW];
  for (int i = threadIdx.x; i < 3 * nw; i += 256) w[i] = waves[i];
  __syncthreads();
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y;
  const int f = blockIdx.z;
  if (x >= nx) return;
  const double2 t = drift[f];
  const double X = x - t.x, Y = y - t.y;
  float g = 0.f;
  for (int j = 0; j < nw; j++) {
    double u = w[3 * j] * X + w[3 * j + 1] * Y + w[3 * j + 2];
    u -= floor(u);
    g += cospif(2.0f * (float)u);
  }
  const float lam = dose * expf(0.5f * amp * g - 0.125f);
  const size_t idx = ((size_t)f * ny + y) * nx + x;
  const uint64_t z = mix64(seed ^ mix64(idx));
  const double uu = ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double p = exp(-(double)lam), c = p;
  int k = 0;
  while (uu > c && k < 255) {
    k++;
    p *= (double)lam / k;
    c += p;
  }
  if (dtype == 0)
    static_cast<uint8_t*>(out)[idx] = (uint8_t)k;
  else
    static_cast<float*>(out)[idx] = (float)k;
}

}  // namespace

// Writes nz frames of nx x ny (dtype 0: uint8, 2: float32) to dev_out.
// drift: 2*nz doubles (t_x, t_y per frame); waves: 3*nw doubles (k_x, k_y in
// cycles/px, phase in cycles), nw <= 1024.  Returns 0 or a HIP error code.
extern "C" int mcs_gpu_movie(void* dev_out, int dtype, int nx, int ny, int nz,
                             const double* drift, const double* waves, int nw, float dose,
                             unsigned long long seed) {
  if (!dev_out || nx <= 0 || ny <= 0 || nz <= 0 || ny > 65535 || nz > 65535 || nw <= 0 ||
      nw > MAXW || (dtype != 0 && dtype != 2))
    return (int)hipErrorInvalidValue;
  double* dd = nullptr;
  double* dw = nullptr;
  hipError_t e = hipMalloc(&dd, sizeof(double) * 2 * nz);
  if (e == hipSuccess) e = hipMalloc(&dw, sizeof(double) * 3 * nw);
  if (e == hipSuccess) e = hipMemcpy(dd, drift, sizeof(double) * 2 * nz, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dw, waves, sizeof(double) * 3 * nw, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const float amp = sqrtf(2.0f / nw);
    hipLaunchKernelGGL(k_synth, dim3((nx + 255) / 256, ny, nz), dim3(256), 0, 0, dev_out, dtype,
                       nx, ny, reinterpret_cast<const double2*>(dd), dw, nw, amp, dose,
                       (uint64_t)seed);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (dd) (void)hipFree(dd);
  if (dw) (void)hipFree(dw);
  return (int)e;
}
