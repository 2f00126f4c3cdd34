// mc_peak.hip -- integer peak and 1/32-pixel Fourier-interpolated peak of
// each box map, in FP64.
//
// Reference FindShift (cufunc.cu:396-497):
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

#include <atomic>

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

// Result of pair p from the first strict maximum (v, i) over the map points
// other than U[0]: the sequential search starts from U[0] with a strict '>'
// (cufunc.cu:460-476); shift = o + p/32 - box/2, cc = U/(n^2 256).
MC_DEV void peak_result(int p, double v, int i, double u0, int offx, int offy, int box, int n,
                        double* __restrict__ out) {
  double bv = u0;
  int sxp = 0, syp = 0;
  if (i != 0x7fffffff && v > u0) {
    bv = v;
    sxp = i % 512;
    syp = i / 512;
  }
  out[3 * p + 0] = offx + sxp / 32.0 - box / 2;
  out[3 * p + 1] = offy + syp / 32.0 - box / 2;
  out[3 * p + 2] = bv / n / n / 16 / 16;
}

// One workgroup per (pair, slice of 512/NS map rows), 512 threads.  NS = 1
// writes the result; NS > 1 writes the slice's (max value, index, U[0],
// window origin) to part[(p NS + slice) 5 ...] for k_find_peaks_final.
template <int NT, int NS>
__global__ __launch_bounds__(NT) void k_find_peaks(const float* __restrict__ boxmaps, int box,
                                                   int n, int wnoise, double* __restrict__ out,
                                                   double* __restrict__ part) {
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
  // 1. integer peak (cufunc.cu:409-424)
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
    int offx = ix - 8, offy = iy - 8;  // cufunc.cu:430-435
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
  static_assert((512 / NS) % QC == 0, "whole row chunks per slice");
  const int qb = blockIdx.y * (512 / NS);
  for (int q0 = qb; q0 < qb + 512 / NS; q0 += QC) {
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
    if (NS == 1) {
      peak_result(p, sv[0], si[0], u0, offx, offy, box, n, out);
    } else {
      // U[0] is evaluated by thread 0 of slice 0 (row 0, column 0)
      double* q = part + ((size_t)p * NS + blockIdx.y) * 5;
      q[0] = sv[0];
      q[1] = (double)si[0];
      q[2] = u0;
      q[3] = offx;
      q[4] = offy;
    }
  }
}

// Combines the NS slices of each pair (one thread per pair): first maximum
// in row-major order over the slices, then the result as for one slice.
template <int NS>
__global__ __launch_bounds__(64) void k_find_peaks_final(const double* __restrict__ part,
                                                         int npairs, int box, int n,
                                                         double* __restrict__ out) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  if (p >= npairs) return;
  const double* q = part + (size_t)p * NS * 5;
  ValIdx b = {q[0], (int)q[1]};
#pragma unroll
  for (int s = 1; s < NS; s++) b = better(b, ValIdx{q[5 * s], (int)q[5 * s + 1]});
  peak_result(p, b.v, b.i, q[2], (int)q[3], (int)q[4], box, n, out);
}

// Row slices per pair: one workgroup per pair fills the device from half as
// many pairs as CUs; fewer pairs split their maps over 2 or 4 workgroups.
static int peak_slices(int npairs) {
  static std::atomic<int> cache[64];
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= 64) dev = 0;
  int cus = cache[dev].load(std::memory_order_relaxed);
  if (cus <= 0) {
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cache[dev].store(cus, std::memory_order_relaxed);
  }
  return npairs >= cus / 2 ? 1 : (npairs >= cus / 4 ? 2 : 4);
}

size_t find_peaks_out_elems(int npairs) {
  const int ns = peak_slices(npairs);
  return (size_t)npairs * 3 + (ns > 1 ? (size_t)npairs * ns * 5 : 0);
}

hipError_t launch_find_peaks(const float* boxmaps, int npairs, int box, int n, int wnoise,
                             double* out, hipStream_t st, double* part) {
  if (npairs <= 0) return hipSuccess;
  if (box < 16) return hipErrorInvalidValue;
  if (!part) part = out + (size_t)npairs * 3;
  const unsigned fg = (unsigned)((npairs + 63) / 64);
  switch (peak_slices(npairs)) {
    case 1:
      hipLaunchKernelGGL((k_find_peaks<512, 1>), dim3(npairs), dim3(512), 0, st, boxmaps, box, n,
                         wnoise, out, part);
      break;
    case 2:
      hipLaunchKernelGGL((k_find_peaks<512, 2>), dim3(npairs, 2), dim3(512), 0, st, boxmaps, box,
                         n, wnoise, out, part);
      hipLaunchKernelGGL(k_find_peaks_final<2>, dim3(fg), dim3(64), 0, st, part, npairs, box, n,
                         out);
      break;
    default:
      hipLaunchKernelGGL((k_find_peaks<512, 4>), dim3(npairs, 4), dim3(512), 0, st, boxmaps, box,
                         n, wnoise, out, part);
      hipLaunchKernelGGL(k_find_peaks_final<4>, dim3(fg), dim3(64), 0, st, part, npairs, box, n,
                         out);
      break;
  }
  return hipGetLastError();
}

}  // namespace mc
