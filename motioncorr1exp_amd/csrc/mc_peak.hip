// mc_peak.hip -- integer peak and 1/32-pixel Fourier-interpolated peak of
// each box map, in FP64.
//
// Reference FindShift (cufunc.cu:445-546):
//  1. first maximum of the box in row-major order, excluding
//     |c - box/2| <= w && |r - box/2| <= w (w = -nps - 1 ... default -1);
//  2. 16 x 16 window at o = clamp(peak - 8, 0, box - 16), real FFT
//     (func.cpp:10-15), cosine mask 0.5 + 0.5 cos(pi r/8) for
//     r = floor(sqrt(kx^2 + ky'^2)) < 8, else 0 (func.cpp:348-370);
//  3. zero-pad to 512 x 512 (rows 0..7 and 504..511, columns 0..8) and
//     C2R (func.cpp:18-23);
//  4. first strict maximum over the 512 x 512 map starting from U[0], with
//     the exclusion scaled by 32; shift = o + p/32 - box/2, cc = U/(n^2 256).
// The 512 x 512 map is evaluated exactly from the 16 x 9 masked spectrum:
//  U[q][p] = Re sum_{kx=0}^{8} a_kx e^{2 pi i kx p/512}
//                sum_{ky} G[ky][kx] e^{2 pi i ky' q/512},
// a_0 = 1, a_kx = 2 (the C2R row weight), ky' = ky (ky < 8) or ky - 16.
#include <hip/hip_runtime.h>

#include "mc_device.h"
#include "mc_kernels.h"

namespace mc {

struct ValIdx {
  double v;
  int i;
};

// max value, ties -> smaller index (row-major first maximum)
MC_DEV ValIdx better(ValIdx a, ValIdx b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

// NT threads (256 or 512): each thread evaluates 512 / NT rows of the map.
template <int NT>
__global__ __launch_bounds__(NT) void k_find_peaks(const float* __restrict__ boxmaps, int box,
                                                   int n, int wnoise, double* __restrict__ out) {
  static_assert(NT == 256 || NT == 512, "threads");
  __shared__ double sv[NT];
  __shared__ int si[NT];
  __shared__ double sub[16][16];
  __shared__ double2 G[16][9];
  __shared__ double2 t512[512];
  __shared__ int soff[2];
  const int p = blockIdx.x;
  const int tid = threadIdx.x;
  const float* bm = boxmaps + (size_t)p * box * box;
  for (int i = tid; i < 512; i += NT) {
    double s, c;
    sincospi((double)i / 256.0, &s, &c);
    t512[i] = make_double2(c, s);
  }
  // 1. integer peak (cufunc.cu:458-473)
  ValIdx best = {-1e9, 0x7fffffff};
  for (int i = tid; i < box * box; i += NT) {
    const int r = i / box, c = i % box;
    if (abs(c - box / 2) <= wnoise && abs(r - box / 2) <= wnoise) continue;
    const double v = (double)bm[i];
    if (v > best.v) best = ValIdx{v, i};
  }
  sv[tid] = best.v;
  si[tid] = best.i;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if (tid < o) {
      ValIdx a{sv[tid], si[tid]}, b{sv[tid + o], si[tid + o]};
      a = better(a, b);
      sv[tid] = a.v;
      si[tid] = a.i;
    }
    __syncthreads();
  }
  if (tid == 0) {
    int ix = 0, iy = 0;
    if (si[0] != 0x7fffffff) {  // no candidate: the reference keeps (0, 0)
      ix = si[0] % box;
      iy = si[0] / box;
    }
    int offx = ix - 8, offy = iy - 8;  // cufunc.cu:479-484
    if (offx + 16 > box) offx = box - 16;
    else if (offx < 0) offx = 0;
    if (offy + 16 > box) offy = box - 16;
    else if (offy < 0) offy = 0;
    soff[0] = offx;
    soff[1] = offy;
  }
  __syncthreads();
  const int offx = soff[0], offy = soff[1];
  if (tid < 256) sub[tid / 16][tid % 16] = (double)bm[(size_t)(offy + tid / 16) * box + offx + tid % 16];
  __syncthreads();
  // 2. 16 x 16 R2C (forward) and cosine mask, times the C2R row weight
  if (tid < 144) {
    const int ky = tid / 9, kx = tid % 9;
    double re = 0, im = 0;
    for (int y = 0; y < 16; y++)
      for (int x = 0; x < 16; x++) {
        const int e = ((kx * x + ky * y) & 15) * 32;  // e^{-2 pi i e/512}
        const double2 w = t512[e];
        re += sub[y][x] * w.x;
        im -= sub[y][x] * w.y;
      }
    const int yy = ky < 8 ? ky : ky - 16;
    const int rr = (int)sqrt((double)(kx * kx + yy * yy));
    double m = 0.0;
    if (rr < 8) m = cospi((double)rr / 8.0) / 2 + 0.5;
    if (kx > 0) m *= 2.0;
    G[ky][kx] = make_double2(re * m, im * m);
  }
  __syncthreads();
  // 3-4. exact evaluation of the 512 x 512 interpolated map, first maximum.
  // Thread t owns the columns c0 + 128 j (c0 = t % 128, j < 4) in the rows
  // q = t / 128 mod 4: the eight twiddles of each column stay in registers, and
  // the row sums Hq[kx] of QC rows at a time are formed in LDS, where every
  // lane reads the same address (a broadcast).  Each point keeps the
  // arithmetic of a per-point evaluation, so the map does not depend on the
  // thread layout; the first maximum in row-major order is the reduction's
  // (value, smallest index).
  static_assert(NT == 512, "four columns and one row class per thread");
  constexpr int QC = 128;
  __shared__ double2 Hs[QC][9];
  const int scale = 32;
  const int wns = scale * wnoise;
  const int ox = (box / 2 - offx) * scale, oy = (box / 2 - offy) * scale;
  constexpr int NC = 4;
  const int c0 = tid % (512 / NC), par = tid / (512 / NC);
  double2 wc[NC][8];
#pragma unroll
  for (int j = 0; j < NC; j++)
#pragma unroll
    for (int kx = 1; kx < 9; kx++) wc[j][kx - 1] = t512[(kx * (c0 + (512 / NC) * j)) & 511];
  ValIdx bb = {-1e300, 0x7fffffff};
  double u0 = 0.0;
  for (int q0 = 0; q0 < 512; q0 += QC) {
    __syncthreads();  // the previous chunk's rows are read
    for (int i = tid; i < QC * 9; i += NT) {
      const int q = q0 + i / 9, kx = i % 9;
      double2 h = make_double2(0.0, 0.0);
      for (int ky = 0; ky < 16; ky++) {
        const int yy = ky < 8 ? ky : ky - 16;
        const double2 w = t512[(yy * q) & 511];
        const double2 g = G[ky][kx];
        h.x += g.x * w.x - g.y * w.y;
        h.y += g.x * w.y + g.y * w.x;
      }
      Hs[i / 9][kx] = h;
    }
    __syncthreads();
    for (int qi = par; qi < QC; qi += NC) {
      const int q = q0 + qi;
      const bool rowex = abs(q - oy) <= wns;
      double2 hq[9];
#pragma unroll
      for (int kx = 0; kx < 9; kx++) hq[kx] = Hs[qi][kx];
#pragma unroll
      for (int j = 0; j < NC; j++) {
        const int pp = c0 + (512 / NC) * j;
        double u = hq[0].x;
#pragma unroll
        for (int kx = 1; kx < 9; kx++) u += hq[kx].x * wc[j][kx - 1].x - hq[kx].y * wc[j][kx - 1].y;
        if (q == 0 && pp == 0) u0 = u;
        if (rowex && abs(pp - ox) <= wns) continue;
        if (u > bb.v) bb = ValIdx{u, q * 512 + pp};
      }
    }
  }
  __syncthreads();
  sv[tid] = bb.v;
  si[tid] = bb.i;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if (tid < o) {
      ValIdx a{sv[tid], si[tid]}, b{sv[tid + o], si[tid + o]};
      a = better(a, b);
      sv[tid] = a.v;
      si[tid] = a.i;
    }
    __syncthreads();
  }
  if (tid == 0) {
    // sequential search starts from U[0] with strict '>' (cufunc.cu:509-525)
    double bv = u0;
    int sxp = 0, syp = 0;
    if (si[0] != 0x7fffffff && sv[0] > u0) {
      bv = sv[0];
      sxp = si[0] % 512;
      syp = si[0] / 512;
    }
    out[3 * p + 0] = offx + sxp / (double)scale - box / 2;
    out[3 * p + 1] = offy + syp / (double)scale - box / 2;
    out[3 * p + 2] = bv / n / n / 16 / 16;
  }
}

hipError_t launch_find_peaks(const float* boxmaps, int npairs, int box, int n, int wnoise,
                             double* out, hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  if (box < 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_find_peaks<512>, dim3(npairs), dim3(512), 0, st, boxmaps, box, n, wnoise,
                     out);
  return hipGetLastError();
}

}  // namespace mc
