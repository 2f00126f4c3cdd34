// mc_sum.hip -- Fourier phase-shift-and-sum of a stack of spectra.
//
// Reference: cuShift (cufunc.cu:372-394) multiplies a frame spectrum by
// e^{i (shx x + shy y')}, shx = sx 2 pi/n, and cuAdd (cufunc.cu:243-252)
// accumulates it into the sum (iteration sums DFAlign.cu:974-989, final sum
// DFAlign.cu:1143-1267).  One pass here reads every frame once and writes
// the sum once: the F shifted spectra are never materialised.  The phase
// (sx kx + sy ky')/n is reduced to [0, 1) in FP64 before the FP32 sincospi,
// so its error does not grow with the shift or the frequency.
#include <hip/hip_runtime.h>

#include <algorithm>

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

// Even/odd half sums of `count` consecutive spectra in one pass (the halves
// of the FSC table): frame f adds to h0 when f is even, else to h1, after its
// phase shift (shifts == nullptr: unshifted; a NaN shift skips the frame).
// Each half accumulates in frame order (DFAlign.cu:808-812 raw halves,
// 1143-1243 aligned halves).  count == 0 writes zeros.
__global__ __launch_bounds__(256) void k_half_sums(const float2* __restrict__ spectra,
                                                   size_t spec_stride, int count, int n,
                                                   const double2* __restrict__ shifts,
                                                   float2* __restrict__ h0,
                                                   float2* __restrict__ h1) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  const int kx = (int)(e / n), ky = (int)(e % n);
  const int yy = ky < n / 2 ? ky : ky - n;
  const double inv_n = 1.0 / n;
  float2 acc[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
  constexpr int U = 8;
  for (int f0 = 0; f0 < count; f0 += U) {
    float2 v[U];
    double2 sh[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      sh[u] = f0 + u >= count ? make_double2(NAN, NAN)
                              : (shifts ? shifts[f0 + u] : make_double2(0.0, 0.0));
      v[u] = sh[u].x == sh[u].x ? spectra[(size_t)(f0 + u) * spec_stride + e]
                                : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const double2 s = sh[u];
      if (s.x != s.x) continue;
      float2 t = v[u];
      if (shifts) {
        double w = (s.x * kx + s.y * yy) * inv_n;
        w -= floor(w);
        float sn, cs;
        sincospif(2.0f * (float)w, &sn, &cs);
        t = cmul(t, make_float2(cs, sn));
      }
      acc[(f0 + u) & 1] = cadd(acc[(f0 + u) & 1], t);
    }
  }
  h0[e] = acc[0];
  h1[e] = acc[1];
}

hipError_t launch_half_sums(const float2* spectra, size_t spec_stride, int count, int n,
                            const double2* shifts, float2* h0, float2* h1, hipStream_t st) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  hipLaunchKernelGGL(k_half_sums, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     spectra, spec_stride, count, n, shifts, h0, h1);
  return hipGetLastError();
}

// ------------------------------------------------- FSC shells (DFAlign.cu:212-297)
// Element (kx, y) of the half spectrum (y = ky for ky < n/2, else ky - n,
// cufunc.cu:309-340) with r = sqrtf(kx^2 + y^2) in float belongs to shell
// int(r / step) unless int(r) <= 0 or r >= n/2+1, on side 0 (parallel to the
// drift direction d) when float(|kx dx + y dy| / r) > float(cos(pi/4)), else
// side 1.  Per (shell, side) the sums of Re(a conj b), |a|^2 and |b|^2 are
// taken in FP64.
//
// The float square root and division are rounded to nearest exactly
// (sqrt_rn, div_rn), so shell membership is the reference's whatever the
// compiler's lowering of sqrtf and '/'.

// Side of (kx, y): FP64 products, sum and quotient each rounded once as in
// the reference's host code (no fused multiply-add).
__device__ inline int fsc_side(int kx, int y, float r, double dx, double dy) {
#pragma clang fp contract(off)
  const double p = (double)kx * dx;
  const double q = (double)y * dy;
  const float angle = (float)(fabs(p + q) / (double)r);
  return angle > (float)0.70710678118654752440 ? 0 : 1;
}

// Block-partial shell sums.  Each wave takes chunks of 64 consecutive
// elements (chunk c goes to wave c mod (waves of the grid)), groups the
// lanes of a chunk by (shell, side) and adds each group's wave-reduced sums
// into its own LDS table in program order; the block's tables are then added
// in wave order.  The result depends only on n and the grid, not on timing.
// padded = 1: row-major [ky][kx] layout (pitch n/2+1), else [kx][ky].
constexpr int FSC_NW = 4;

__global__ __launch_bounds__(FSC_NW * 64) void k_fsc_partial(const float2* __restrict__ a,
                                                            const float2* __restrict__ b, int n,
                                                            int padded, int step, int nbox,
                                                            double dx, double dy,
                                                            double* __restrict__ part) {
  extern __shared__ double fsc_acc[];  // [FSC_NW][nbox][2][3]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int H = n / 2 + 1, total = H * n, nchunk = (total + 63) / 64;
  const int nk = nbox * 6;
  for (int i = threadIdx.x; i < FSC_NW * nk; i += FSC_NW * 64) fsc_acc[i] = 0.0;
  __syncthreads();
  double* my = fsc_acc + w * nk;
  for (int c = blockIdx.x * FSC_NW + w; c < nchunk; c += gridDim.x * FSC_NW) {
    const int e = c * 64 + lane;
    int key = -1;
    double sab = 0.0, saa = 0.0, sbb = 0.0;
    if (e < total) {
      int kx, ky;
      if (padded) {
        ky = e / H;
        kx = e - ky * H;
      } else {
        kx = e / n;
        ky = e - kx * n;
      }
      const int y = ky < n / 2 ? ky : ky - n;
      const int q = kx * kx + y * y;
      if (q > 0) {
        const float r = sqrt_rn((float)q);
        if ((int)r > 0 && r < (float)H) {
          const int id = min((int)div_rn(r, (float)step), nbox - 1);  // r < H keeps id < nbox
          key = 2 * id + fsc_side(kx, y, r, dx, dy);
          const float2 va = a[e], vb = b[e];
          sab = (double)va.x * vb.x + (double)va.y * vb.y;
          saa = (double)va.x * va.x + (double)va.y * va.y;
          sbb = (double)vb.x * vb.x + (double)vb.y * vb.y;
        }
      }
    }
    unsigned long long rem = __ballot(key >= 0);
    while (rem) {
      const int k = __shfl(key, __ffsll((unsigned long long)rem) - 1);
      const bool mine = key == k;
      double t0 = mine ? sab : 0.0, t1 = mine ? saa : 0.0, t2 = mine ? sbb : 0.0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        t0 += __shfl_xor(t0, o);
        t1 += __shfl_xor(t1, o);
        t2 += __shfl_xor(t2, o);
      }
      if (lane == 0) {
        my[3 * k] += t0;
        my[3 * k + 1] += t1;
        my[3 * k + 2] += t2;
      }
      rem &= ~__ballot(mine);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nk; i += FSC_NW * 64) {
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < FSC_NW; v++) s += fsc_acc[v * nk + i];
    part[(size_t)blockIdx.x * nk + i] = s;
  }
}

__global__ __launch_bounds__(256) void k_fsc_final(const double* __restrict__ part, int nblocks,
                                                   int nk, double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nk) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += part[(size_t)b * nk + i];
  out[i] = s;
}

int fsc_bins(int n, int* step) {
  int st = n / 400;  // DFAlign.cu:214 (n < 400 would divide by zero there)
  if (st < 1) st = 1;
  if (step) *step = st;
  return (n / 2 + 1) / st + 1;  // DFAlign.cu:218
}

int fsc_grid(int n) {
  const long long nchunk = ((long long)(n / 2 + 1) * n + 63) / 64;
  return (int)std::max(1LL, std::min(512LL, nchunk / 64));
}

hipError_t launch_fsc_sums(const float2* a, const float2* b, int n, int padded, double dx,
                           double dy, double* part, double* sums, hipStream_t st) {
  int step;
  const int nbox = fsc_bins(n, &step);
  const int nk = nbox * 6, G = fsc_grid(n);
  const size_t lds = sizeof(double) * FSC_NW * nk;
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fsc_partial, dim3(G), dim3(FSC_NW * 64), lds, st, a, b, n, padded, step,
                     nbox, dx, dy, part);
  hipLaunchKernelGGL(k_fsc_final, dim3((nk + 255) / 256), dim3(256), 0, st, part, G, nk, sums);
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
                                                       int nchunks, int jonly,
                                                       double2* __restrict__ part) {
  __shared__ double rx[256], ry[256];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int c = jonly ? L : L / (nframes + 1), f = jonly ? nframes : L % (nframes + 1);
  const int H = n / 2 + 1;
  const size_t total = (size_t)H * n;
  const size_t e0 = total * c / nchunks, e1 = total * (c + 1) / nchunks;
  double gx = 0, gy = 0;
  const double2 s = f < nframes ? shifts[f] : make_double2(0.0, 0.0);
  const bool skip = f < nframes && s.x != s.x;
  const float2* F = spectra + (size_t)(f < nframes ? f : 0) * spec_stride;
  const double inv_n = 1.0 / n;
  if (!skip) {
    // U elements per round, all loads issued first (at valid addresses; the
    // elements past the chunk get weight 0); (kx, ky) advance by the stride
    // instead of being divided out of the element index.
    constexpr int U = 4;
    const int ie1 = (int)e1;
    int e = (int)e0 + threadIdx.x;
    int kx = e / n, ky = e - kx * n;
    const int kadv = 256 / n, yadv = 256 % n;  // (kx, ky) step of e += 256
    for (; e < ie1; e += 256 * U) {
      float2 sv[U], fv[U];
      float w[U];
      int kxs[U], yys[U];
      int ee = e;
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int ec = ee < ie1 ? ee : ie1 - 1;
        sv[u] = S[ec];
        fv[u] = F[ec];
        w[u] = wk[kx] * wk[H + ky] * (ee < ie1 ? 1.f : 0.f);
        kxs[u] = kx;
        yys[u] = ky < n / 2 ? ky : ky - n;
        ee += 256;
        ky += yadv;
        kx += kadv;
        const bool wrap = ky >= n;
        ky = wrap ? ky - n : ky;
        kx = wrap ? kx + 1 : kx;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (f == nframes) {
          gx += (double)w[u] * ((double)sv[u].x * sv[u].x + (double)sv[u].y * sv[u].y);
        } else {
          double uu = (s.x * kxs[u] + s.y * yys[u]) * inv_n;
          uu -= floor(uu);
          float sn, cs;
          sincospif(2.0f * (float)uu, &sn, &cs);
          const float2 g = cmul(fv[u], make_float2(cs, sn));
          const double im = (double)sv[u].x * g.y - (double)sv[u].y * g.x;  // Im(conj(S) G)
          const double t = -2.0 * w[u] * im * (2.0 * M_PI * inv_n);
          gx += t * kxs[u];
          gy += t * yys[u];
        }
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
                            double2* part, int nchunks, double2* out, int jonly,
                            hipStream_t st) {
  const int nb = nchunks * (jonly ? 1 : nframes + 1);
  hipLaunchKernelGGL(k_sharp_partial, dim3(nb), dim3(256), 0, st, spectra, spec_stride, nframes,
                     n, shifts, S, wk, nchunks, jonly, part);
  hipLaunchKernelGGL(k_sharp_final, dim3((nframes + 1 + 63) / 64), dim3(64), 0, st, part,
                     nchunks, nframes, out);
  return hipGetLastError();
}

// ------------------------------------------- elementwise (cufunc.cu:243-263)
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
