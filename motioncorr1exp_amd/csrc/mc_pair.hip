// mc_pair.hip -- B-factor-weighted cross-power of frame pairs and the pruned
// inverse that produces only the box x box window of each correlation map.
//
// Reference: cuShiftCC (cufunc.cu:392-410, called with sx = sy = 0 at
// DFAlign.cu:894,1006) forms dfft <- dsum * conj(dfft) * z / sizec with
// z = (-1)^{x+y'} exp(-B (x^2+y'^2)/(2 n^2)) (MkPosList, cufunc.cu:358-390);
// a full n x n C2R follows (DFAlign.cu:897,1008) and only the central box is
// read back (FindShift, cufunc.cu:447-456).  Here the product, the weight
// and the inverse restricted to the box are fused:
//
//  k_pair_cols : one (pair, kx) per workgroup.  The product X[ky] is formed
//    on the fly from the two spectra columns (contiguous in the [kx][ky]
//    layout), decimated ky = a + A b (A = n/M), transformed by A inverse
//    M-point FFTs in LDS, and combined as
//      T[r] = sum_a e^{2 pi i a R_r/n} Y_a[R_r mod M],  R_r = (n-box)/2 + r,
//    which equals the column inverse at the box rows (M >= box).
//  k_pair_rows : one pair per workgroup: C[r][c] = sum_kx a_kx Re(T[kx][r]
//    e^{2 pi i kx R_c/n}) with a_0 = a_{n/2} = 1, a_kx = 2 (the C2R row
//    transform with the DC/Nyquist projection), register-tiled.
#include <hip/hip_runtime.h>

#include "mc_device.h"
#include "mc_kernels.h"

namespace mc {

template <int N, int M, int NT>
__global__ __launch_bounds__(NT) void k_pair_cols(const float2* __restrict__ spectra,
                                                  size_t spec_stride,
                                                  const int2* __restrict__ pairs, int npairs,
                                                  float mb, float inv_sizec, int box,
                                                  float2* __restrict__ tbuf,
                                                  const float2* __restrict__ tw) {
  constexpr int A = N / M;
  constexpr int H = N / 2 + 1;
  constexpr int TS = M + 1;
  __shared__ float2 buf[A * TS];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int p = L % npairs, kx = L / npairs;
  const int2 pr = pairs[p];
  const float2* a = spectra + (size_t)pr.x * spec_stride + (size_t)kx * N;
  const float2* b = spectra + (size_t)pr.y * spec_stride + (size_t)kx * N;
  const int kx2 = kx * kx;
  for (int e = threadIdx.x; e < N; e += NT) {
    const float2 va = a[e], vb = b[e];
    const int yy = e < N / 2 ? e : e - N;
    float z = expf(mb * (float)(kx2 + yy * yy));
    if ((kx + yy) & 1) z = -z;
    float2 x = cmulc(va, vb);
    x = cscale(cscale(x, inv_sizec), z);
    buf[(e % A) * TS + e / A] = x;
  }
  __syncthreads();
  fft_lds<M, NT, 1, A, TS>(buf, tw, N);
  const int ori = (N - box) / 2;
  float2* out = tbuf + ((size_t)p * H + kx) * box;
  for (int r = threadIdx.x; r < box; r += NT) {
    const int R = ori + r;
    const int Rm = R & (M - 1);
    float2 acc = buf[Rm];
#pragma unroll 4
    for (int q = 1; q < A; q++) acc = cadd(acc, cmul(tw[(q * R) & (N - 1)], buf[q * TS + Rm]));
    out[r] = acc;
  }
}

template <int BT>
__global__ __launch_bounds__(256) void k_pair_rows(const float2* __restrict__ tbuf, int n,
                                                   int box, float* __restrict__ boxmaps,
                                                   const float2* __restrict__ tw) {
  constexpr int KC = 16;
  constexpr int W = 16 * BT;
  __shared__ float2 Ts[KC][W];
  __shared__ float2 Es[KC][W];
  const int p = blockIdx.x;
  const int H = n / 2 + 1;
  const int ori = (n - box) / 2;
  const float2* T = tbuf + (size_t)p * H * box;
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  float acc[BT][BT];
#pragma unroll
  for (int i = 0; i < BT; i++)
#pragma unroll
    for (int j = 0; j < BT; j++) acc[i][j] = 0.f;
  for (int k0 = 0; k0 < H; k0 += KC) {
    for (int i = threadIdx.x; i < KC * W; i += 256) {
      const int kk = i / W, c = i % W;
      const int kx = k0 + kk;
      float2 t = make_float2(0.f, 0.f), e = make_float2(0.f, 0.f);
      if (kx < H && c < box) {
        const float al = (kx == 0 || 2 * kx == n) ? 1.f : 2.f;
        t = cscale(T[(size_t)kx * box + c], al);
        e = tw[(int)(((long long)kx * (ori + c)) & (n - 1))];
      }
      Ts[kk][c] = t;
      Es[kk][c] = e;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KC; kk++) {
      float2 tv[BT], ev[BT];
#pragma unroll
      for (int i = 0; i < BT; i++) {
        tv[i] = Ts[kk][tr + 16 * i];
        ev[i] = Es[kk][tc + 16 * i];
      }
#pragma unroll
      for (int i = 0; i < BT; i++)
#pragma unroll
        for (int j = 0; j < BT; j++) acc[i][j] += tv[i].x * ev[j].x - tv[i].y * ev[j].y;
    }
    __syncthreads();
  }
  float* out = boxmaps + (size_t)p * box * box;
#pragma unroll
  for (int i = 0; i < BT; i++)
#pragma unroll
    for (int j = 0; j < BT; j++) {
      const int r = tr + 16 * i, c = tc + 16 * j;
      if (r < box && c < box) out[(size_t)r * box + c] = acc[i][j];
    }
}

template <int N, int M>
static hipError_t pair_cols_t(const float2* spectra, size_t spec_stride, const int2* pairs,
                              int npairs, float mb, float inv_sizec, int box, float2* tbuf,
                              hipStream_t st) {
  constexpr int NT = (N / 16) < 64 ? 64 : ((N / 16) > 256 ? 256 : N / 16);
  const float2* tw = twiddles(N);
  if (!tw) return hipErrorOutOfMemory;
  const long long nb = (long long)npairs * (N / 2 + 1);
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_pair_cols<N, M, NT>), dim3((unsigned)nb), dim3(NT), 0, st, spectra,
                     spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, tw);
  return hipGetLastError();
}

template <int N>
static hipError_t pair_cols_n(const float2* spectra, size_t spec_stride, const int2* pairs,
                              int npairs, float mb, float inv_sizec, int box, float2* tbuf,
                              hipStream_t st) {
  if constexpr (N >= 128) {
    if (box <= 128)
      return pair_cols_t<N, 128>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
  }
  if constexpr (N >= 256) {
    if (box <= 256)
      return pair_cols_t<N, 256>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
  }
  if constexpr (N == 64) {
    if (box <= 64)
      return pair_cols_t<64, 64>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_pair_cols(const float2* spectra, size_t spec_stride, const int2* pairs,
                            int npairs, int n, float bfactor, int box, float2* tbuf,
                            hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  if (box < 1 || box > n || box > 256) return hipErrorInvalidValue;
  // reference MkPosList: float m = -0.5*bfactor/nsam/nsam (cufunc.cu:365)
  const float mb = (float)(-0.5 * bfactor / n / n);
  const float inv_sizec = 1.0f / (float)((n / 2 + 1) * n);
  switch (n) {
    case 64: return pair_cols_n<64>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    case 128: return pair_cols_n<128>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    case 256: return pair_cols_n<256>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    case 512: return pair_cols_n<512>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    case 1024: return pair_cols_n<1024>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    case 2048: return pair_cols_n<2048>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    case 4096: return pair_cols_n<4096>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    case 8192: return pair_cols_n<8192>(spectra, spec_stride, pairs, npairs, mb, inv_sizec, box, tbuf, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pair_rows(const float2* tbuf, int npairs, int n, int box, float* boxmaps,
                            hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  if (box < 1 || box > n || box > 256) return hipErrorInvalidValue;
  const float2* tw = twiddles(n);
  if (!tw) return hipErrorOutOfMemory;
  const int bt = (box + 15) / 16;
  switch (bt) {
#define MC_ROWS_CASE(B) \
  case B: hipLaunchKernelGGL(k_pair_rows<B>, dim3(npairs), dim3(256), 0, st, tbuf, n, box, boxmaps, tw); break;
    MC_ROWS_CASE(1) MC_ROWS_CASE(2) MC_ROWS_CASE(3) MC_ROWS_CASE(4) MC_ROWS_CASE(5)
    MC_ROWS_CASE(6) MC_ROWS_CASE(7) MC_ROWS_CASE(8) MC_ROWS_CASE(9) MC_ROWS_CASE(10)
    MC_ROWS_CASE(11) MC_ROWS_CASE(12) MC_ROWS_CASE(13) MC_ROWS_CASE(14) MC_ROWS_CASE(15)
    MC_ROWS_CASE(16)
#undef MC_ROWS_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mc
