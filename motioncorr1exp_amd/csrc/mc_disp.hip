// mc_disp.hip -- quick-look images (dosef_quick/): the 400 x 400 displays
// of a sum spectrum that the reference builds on the host.
//
// Reference: the aligned sum (DFAlign.cu:1281-1284 with ImageOutputThread,
// DFAlign.cu:73-90) is the Fourier crop of the sum spectrum to 400 x 400
// (cufunc.cu:147-161 row pattern), an inverse C2R and an 8-bit mapping; the
// log-modulus images (DFAlign.cu:842-848, 908-911, 1270-1278) take
// log(1 + |X| scale) (cufunc.cu:178-188), lay it out as the full centred
// n x n image (func.cpp:459-489), and bin that image by a forward R2C, the
// same crop, a C2R and / 400^2 (func.cpp:524-555).  The 8-bit mapping clips
// the histogram's 0.5% tails (func.cpp:372-443).
//
// Here the n x n log-modulus image is written by a tiled transpose from the
// resident [kx][ky] spectrum, binned by the stage-1 transform (row FFTs and
// only the kept 201 columns), and the 400-point inverses are direct DFTs
// from LDS (0.1 GFLOP per image; 400 = 2^4 * 5^2 has no plan here).  The
// mapping runs in one workgroup with the reference's float arithmetic.
#include <hip/hip_runtime.h>

#include "mc_device.h"
#include "mc_kernels.h"

namespace mc {

// Output tile (j0..j0+63, i0..i0+63) of the centred image.  Right half
// (i >= n/2): kx = i - n/2, ky = (j + n/2) mod n.  Left half: the point
// mirror, kx = n/2 - i, ky = (n/2 - j) mod n, except row 0, which repeats
// row 1 (ky = n/2 - 1).  A tile reads a 64 x 64 box of the spectrum
// (coalesced along ky) into LDS and writes its rows coalesced along i.
// Requires n % 128 == 0, so that no tile straddles column n/2.
__global__ __launch_bounds__(256) void k_logmod_image(const float2* __restrict__ S, int n,
                                                      float scale, float* __restrict__ img) {
  __shared__ float box[64][65];
  const int hn = n / 2;
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const bool right = i0 >= hn;
  const int kx0 = right ? i0 - hn : hn - i0 - 63;
  const int ky0 = right ? (j0 + hn) % n : ((hn - j0 - 63) % n + n) % n;
  for (int t = threadIdx.x; t < 64 * 64; t += 256) {
    const int bx = t >> 6, by = t & 63;
    const int ky = (ky0 + by) % n;
    const float2 v = S[(size_t)(kx0 + bx) * n + ky];
    box[bx][by] = logf(1.f + sqrtf(v.x * v.x + v.y * v.y) * scale);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 64 * 64; t += 256) {
    const int r = t >> 6, q = t & 63;
    const int j = j0 + r;
    const int bx = right ? q : 63 - q;
    const int by = right ? r : (j == 0 ? 62 : 63 - r);
    img[(size_t)j * n + i0 + q] = box[bx][by];
  }
}

// The same centred image for any even n (tiles would straddle column n/2):
// one thread per output pixel, the spectrum element read directly.
__global__ __launch_bounds__(256) void k_logmod_image_any(const float2* __restrict__ S, int n,
                                                          float scale, float* __restrict__ img) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)n * n) return;
  const int hn = n / 2;
  const int j = (int)(t / n), i = (int)(t % n);
  int kx, ky;
  if (i >= hn) {
    kx = i - hn;
    ky = (j + hn) % n;
  } else {
    kx = hn - i;
    ky = j == 0 ? hn - 1 : ((hn - j) % n + n) % n;
  }
  const float2 v = S[(size_t)kx * n + ky];
  img[t] = logf(1.f + sqrtf(v.x * v.x + v.y * v.y) * scale);
}

hipError_t launch_logmod_image(const float2* spec, int n, float scale, float* img,
                               hipStream_t st) {
  if (n < 2 || n % 2) return hipErrorInvalidValue;
  if (n % 128 == 0)
    hipLaunchKernelGGL(k_logmod_image, dim3(n / 64, n / 64), dim3(256), 0, st, spec, n, scale, img);
  else
    hipLaunchKernelGGL(k_logmod_image_any, dim3((unsigned)(((long long)n * n + 255) / 256)),
                       dim3(256), 0, st, spec, n, scale, img);
  return hipGetLastError();
}

// Column pass of a d-point C2R: Y[kx][y] = sum_t X'[kx][t] e^{+2 pi i t y/d}
// for kx in [0, d/2], where X' is the d-row crop of the [kx][ky] spectrum X
// of column length ld (rows t < d/2 from ky = t, the others from
// ky = ld - d + t).
__global__ __launch_bounds__(256) void k_dft_cols_inv(const float2* __restrict__ X, int ld, int d,
                                                      const float2* __restrict__ tw,
                                                      float2* __restrict__ Y) {
  extern __shared__ float2 dsh[];
  float2* col = dsh;
  float2* w = dsh + d;
  const int kx = blockIdx.x;
  for (int t = threadIdx.x; t < d; t += 256) {
    const int ky = t < d / 2 ? t : ld - d + t;
    col[t] = X[(size_t)kx * ld + ky];
    w[t] = tw[t];
  }
  __syncthreads();
  for (int y = threadIdx.x; y < d; y += 256) {
    float2 acc = make_float2(0.f, 0.f);
    int e = 0;
    for (int t = 0; t < d; t++) {
      acc = cadd(acc, cmul(col[t], w[e]));
      e += y;
      if (e >= d) e -= d;
    }
    Y[(size_t)kx * d + y] = acc;
  }
}

// Row pass: out[y][x] = Re Y[0][y] + (-1)^x Re Y[d/2][y]
//                       + 2 sum_{kx=1}^{d/2-1} Re(Y[kx][y] e^{+2 pi i kx x/d}),
// the C2R that ignores the imaginary parts of the DC and Nyquist columns;
// divided by `div` when div > 0.
__global__ __launch_bounds__(256) void k_dft_rows_c2r(const float2* __restrict__ Y, int d,
                                                      const float2* __restrict__ tw, float div,
                                                      float* __restrict__ out) {
  extern __shared__ float2 dsh[];
  const int h = d / 2;
  float2* row = dsh;
  float2* w = dsh + h + 1;
  const int y = blockIdx.x;
  for (int t = threadIdx.x; t <= h; t += 256) row[t] = Y[(size_t)t * d + y];
  for (int t = threadIdx.x; t < d; t += 256) w[t] = tw[t];
  __syncthreads();
  for (int x = threadIdx.x; x < d; x += 256) {
    float acc = 0.f;
    int e = x;
    for (int kx = 1; kx < h; kx++) {
      acc += row[kx].x * w[e].x - row[kx].y * w[e].y;
      e += x;
      if (e >= d) e -= d;
    }
    float v = row[0].x + ((x & 1) ? -row[h].x : row[h].x) + 2.f * acc;
    if (div > 0.f) v = v / div;
    out[(size_t)y * d + x] = v;
  }
}

hipError_t launch_small_c2r(const float2* X, int ld, int d, float div, float2* tmp, float* out,
                            hipStream_t st) {
  if (d < 2 || d % 2 || d > 2048 || ld < d) return hipErrorInvalidValue;
  const float2* tw = twiddles(d);
  if (!tw) return hipErrorOutOfMemory;
  hipLaunchKernelGGL(k_dft_cols_inv, dim3(d / 2 + 1), dim3(256), sizeof(float2) * 2 * d, st, X,
                     ld, d, tw, tmp);
  hipLaunchKernelGGL(k_dft_rows_c2r, dim3(d), dim3(256), sizeof(float2) * (d / 2 + 1 + d), st,
                     tmp, d, tw, div, out);
  return hipGetLastError();
}

// histMinMax (400 bins, threshold 0.005) and buf2DispShort in one workgroup,
// in the reference's float arithmetic: min/max (order-free), the bin width
// (max - min)/400 and bin indices int((v - min)/width) rounded exactly,
// integer bin counts, the tails cut by whole bins, then
// (v - min') * float(255.0 / (max' - min')) truncated.  An empty range
// (constant image) gives zeros; the reference leaves its buffer unwritten.
constexpr int DISP_NBIN = 400;

__global__ __launch_bounds__(1024) void k_disp_short(const float* __restrict__ buf, int size,
                                                     short* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ int hist[DISP_NBIN];
  __shared__ float wmn[16], wmx[16];
  __shared__ float cut[2];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  float mn = 1e20f, mx = -1e20f;
  for (int i = tid; i < size; i += 1024) {
    const float v = buf[i];
    if (v < mn) mn = v;
    if (v > mx) mx = v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float a = __shfl_xor(mn, o), b = __shfl_xor(mx, o);
    if (a < mn) mn = a;
    if (b > mx) mx = b;
  }
  if (lane == 0) {
    wmn[w] = mn;
    wmx[w] = mx;
  }
  for (int i = tid; i < DISP_NBIN; i += 1024) hist[i] = 0;
  __syncthreads();
  mn = wmn[0];
  mx = wmx[0];
  for (int k = 1; k < 16; k++) {
    if (wmn[k] < mn) mn = wmn[k];
    if (wmx[k] > mx) mx = wmx[k];
  }
  if (!(mn < mx)) {
    for (int i = tid; i < size; i += 1024) out[i] = 0;
    return;
  }
  const float step = div_rn(mx - mn, (float)DISP_NBIN);
  for (int i = tid; i < size; i += 1024) {
    int id = (int)div_rn(buf[i] - mn, step);
    id = id < 0 ? 0 : (id >= DISP_NBIN ? DISP_NBIN - 1 : id);
    atomicAdd(&hist[id], 1);
  }
  __syncthreads();
  if (tid == 0) {
    const int percent = (int)((float)size * 0.005f);
    float lo = mn, hi = mx;
    int count = 0;
    for (int i = 0; i < DISP_NBIN; i++) {
      count += hist[i];
      if (count > percent) {
        lo = lo + (float)i * step;
        break;
      }
    }
    count = 0;
    for (int i = DISP_NBIN - 1; i > 0; i--) {
      count += hist[i];
      if (count > percent) {
        hi = hi - (float)(DISP_NBIN - 1 - i) * step;
        break;
      }
    }
    cut[0] = lo;
    cut[1] = hi;
  }
  __syncthreads();
  const float lo = cut[0], hi = cut[1];
  if (!(lo < hi)) {
    for (int i = tid; i < size; i += 1024) out[i] = 0;
    return;
  }
  const float sc = (float)(255.0 / (double)(hi - lo));
  for (int i = tid; i < size; i += 1024) {
    const float v = buf[i];
    out[i] = v <= lo ? (short)0 : (v >= hi ? (short)255 : (short)((v - lo) * sc));
  }
}

hipError_t launch_disp_short(const float* buf, int size, short* out, hipStream_t st) {
  if (size <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_disp_short, dim3(1), dim3(1024), 0, st, buf, size, out);
  return hipGetLastError();
}

}  // namespace mc
