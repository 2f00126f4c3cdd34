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

// Frames are consumed in groups of U whose loads are issued together (the
// loop over frames is otherwise a chain of dependent load-use latencies);
// the accumulation order is the frame order either way.
template <int U>
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
  for (int f0 = 0; f0 < nframes; f0 += U) {
    float2 v[U];
    double2 sh[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      sh[u] = f0 + u < nframes ? shifts[f0 + u] : make_double2(NAN, NAN);
      v[u] = sh[u].x == sh[u].x ? spectra[(size_t)(f0 + u) * spec_stride + e] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int f = f0 + u;
      const double2 s = sh[u];
      if (s.x != s.x) continue;  // NaN: frame not summed (or past the end)
      double w = (s.x * kx + s.y * yy) * inv_n;
      w -= floor(w);
      float sn, cs;
      sincospif(2.0f * (float)w, &sn, &cs);
      const float2 t = cmul(v[u], make_float2(cs, sn));
      acc = cadd(acc, t);
      if (sum2 && f <= f2_last) acc2 = cadd(acc2, t);
    }
  }
  sum[e] = acc;
  if (sum2) sum2[e] = acc2;
}

hipError_t launch_shift_sum(const float2* spectra, size_t spec_stride, int nframes, int n,
                            const double2* shifts, float2* sum, float2* sum2, int f2_last,
                            hipStream_t st) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  hipLaunchKernelGGL(k_shift_sum<8>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     spectra, spec_stride, nframes, n, shifts, sum, sum2, f2_last);
  return hipGetLastError();
}

// Several sums in one launch (one per blockIdx.y): job j = (first slot,
// frame count, first shift, destination slot) sums its frames into its
// destination slot of the same spectrum array.  Destination slots are
// disjoint from every job's frame slots.
__global__ __launch_bounds__(256) void k_shift_sum_multi(float2* spectra, size_t spec_stride,
                                                         int n, const int4* __restrict__ jobs,
                                                         const double2* __restrict__ shifts) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  const int4 jb = jobs[blockIdx.y];
  const int kx = (int)(e / n), ky = (int)(e % n);
  const int yy = ky < n / 2 ? ky : ky - n;
  const double inv_n = 1.0 / n;
  float2 acc = make_float2(0.f, 0.f);
  constexpr int U = 8;
  for (int f0 = 0; f0 < jb.y; f0 += U) {
    float2 v[U];
    double2 sh[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      sh[u] = f0 + u < jb.y ? shifts[jb.z + f0 + u] : make_double2(NAN, NAN);
      v[u] = sh[u].x == sh[u].x ? spectra[(size_t)(jb.x + f0 + u) * spec_stride + e]
                                : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const double2 s = sh[u];
      if (s.x != s.x) continue;
      double w = (s.x * kx + s.y * yy) * inv_n;
      w -= floor(w);
      float sn, cs;
      sincospif(2.0f * (float)w, &sn, &cs);
      acc = cadd(acc, cmul(v[u], make_float2(cs, sn)));
    }
  }
  spectra[(size_t)jb.w * spec_stride + e] = acc;
}

hipError_t launch_shift_sum_multi(float2* spectra, size_t spec_stride, int n, const int4* jobs,
                                  int njobs, const double2* shifts, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (njobs > 65535) return hipErrorInvalidValue;
  const size_t total = (size_t)(n / 2 + 1) * n;
  hipLaunchKernelGGL(k_shift_sum_multi, dim3((unsigned)((total + 255) / 256), (unsigned)njobs),
                     dim3(256), 0, st, spectra, spec_stride, n, jobs, shifts);
  return hipGetLastError();
}

// ------------------------------------------------ sharpness and its adjoint
// J(s) = sum_k w(k) |S(k)|^2 with S = sum_f F_f e^{i phi_f}, phi_f(k) =
// 2 pi (s_fx kx + s_fy ky')/n, w = exp(-B|k|^2/(2n^2)) * a_kx (a = 1 at DC and
// Nyquist columns, 2 otherwise: the half-plane counts both halves).
// dJ/ds_fx = -2 sum_k w (2 pi kx/n) Im(conj(S) F_f e^{i phi_f}); same for y.
// Work item (chunk c, frame f): frame f < F accumulates its two gradient
// components over the chunk's elements, f == F accumulates J.  Items of one
// chunk are consecutive on one XCD, so the chunk of S is reused from L2.
__global__ __launch_bounds__(256) void k_sharp_partial(const float2* __restrict__ spectra,
                                                       size_t spec_stride, int nframes, int n,
                                                       const double2* __restrict__ shifts,
                                                       const float2* __restrict__ S,
                                                       const float* __restrict__ wk,
                                                       int nchunks, double2* __restrict__ part) {
  __shared__ double rx[256], ry[256];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int c = L / (nframes + 1), f = L % (nframes + 1);
  const int H = n / 2 + 1;
  const size_t total = (size_t)H * n;
  const size_t e0 = total * c / nchunks, e1 = total * (c + 1) / nchunks;
  double gx = 0, gy = 0;
  const double2 s = f < nframes ? shifts[f] : make_double2(0.0, 0.0);
  const bool skip = f < nframes && s.x != s.x;
  const float2* F = spectra + (size_t)(f < nframes ? f : 0) * spec_stride;
  const double inv_n = 1.0 / n;
  if (!skip) {
    for (size_t e = e0 + threadIdx.x; e < e1; e += 256) {
      const int kx = (int)(e / n), ky = (int)(e % n);
      const int yy = ky < n / 2 ? ky : ky - n;
      const float w = wk[kx] * wk[H + ky];
      const float2 sv = S[e];
      if (f == nframes) {
        gx += (double)w * ((double)sv.x * sv.x + (double)sv.y * sv.y);
      } else {
        double u = (s.x * kx + s.y * yy) * inv_n;
        u -= floor(u);
        float sn, cs;
        sincospif(2.0f * (float)u, &sn, &cs);
        const float2 g = cmul(F[e], make_float2(cs, sn));
        const double im = (double)sv.x * g.y - (double)sv.y * g.x;  // Im(conj(S) G)
        const double t = -2.0 * w * im * (2.0 * M_PI * inv_n);
        gx += t * kx;
        gy += t * yy;
      }
    }
  }
  rx[threadIdx.x] = gx;
  ry[threadIdx.x] = gy;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      rx[threadIdx.x] += rx[threadIdx.x + o];
      ry[threadIdx.x] += ry[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)c * (nframes + 1) + f] = make_double2(rx[0], ry[0]);
}

// Sums the chunk partials in a fixed order: out[f] = (gx, gy), out[F] = (J, 0).
__global__ void k_sharp_final(const double2* __restrict__ part, int nchunks, int nframes,
                              double2* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f > nframes) return;
  double x = 0, y = 0;
  for (int c = 0; c < nchunks; c++) {
    const double2 v = part[(size_t)c * (nframes + 1) + f];
    x += v.x;
    y += v.y;
  }
  out[f] = make_double2(x, y);
}

hipError_t launch_sharpness(const float2* spectra, size_t spec_stride, int nframes, int n,
                            const double2* shifts, const float2* S, const float* wk,
                            double2* part, int nchunks, double2* out, hipStream_t st) {
  const int nb = nchunks * (nframes + 1);
  hipLaunchKernelGGL(k_sharp_partial, dim3(nb), dim3(256), 0, st, spectra, spec_stride, nframes,
                     n, shifts, S, wk, nchunks, part);
  hipLaunchKernelGGL(k_sharp_final, dim3((nframes + 1 + 63) / 64), dim3(64), 0, st, part,
                     nchunks, nframes, out);
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
