// mc_fft.hip -- stage 1 (frames -> spectra) and sum-image inverse transforms.
//
// Spectrum layout in HBM ("column-major half spectrum"): S[kx][ky], kx in
// [0, n/2], ky in [0, n), complex float, ky contiguous.  It holds the same
// values as the cuFFT in-place R2C result the reference keeps on the host
// (reference cufunc.cu:223-227 called at DFAlign.cu:772,797), transposed so
// that every consumer of the alignment path (pair correlation, phase shift
// and sum) streams contiguous ky vectors.
//
// Forward 2-D R2C of an nu x nu canvas in two passes:
//  k_r2c_rows : ROWS image rows per workgroup; each real row is packed into
//               nu/2 complex values, transformed in LDS (Stockham), unpacked
//               to nu/2+1 bins, and only the bins kept after Fourier-crop
//               binning (kx <= n/2) are stored, transposed, to RT[kx][y].
//               Crop offset, element decode (int8/uint8/int16/uint16/f32)
//               and zero fill outside the data region are fused in the load.
//  k_fill     : per-frame mean of the data region (pad geometry only).
//  k_c2c_cols : one column kx per workgroup: adds fill * (spectrum of the
//               fill mask), transforms along y in LDS, keeps the rows of
//               the Fourier crop (reference cufunc.cu:162-176) and writes
//               S[kx][ky] contiguously.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "mc_device.h"
#include "mc_frames.h"
#include "mc_kernels.h"

namespace mc {

// ----------------------------------------------------------- twiddle cache
const float2* twiddles(int n) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, float2*> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, n);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  std::vector<float2> h(n);
  for (int e = 0; e < n; e++) {
    double a = 2.0 * M_PI * (double)e / (double)n;
    h[e] = make_float2((float)cos(a), (float)sin(a));
  }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * n) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  cache[key] = d;
  return d;
}

// ------------------------------------------------------------- row pass
template <int NU, int ROWS, int NT, int DT>
__global__ __launch_bounds__(NT) void k_r2c_rows(R2CGeom g, int f0, int nkx, float2* rt,
                                                 size_t rt_fstride, float* rowsum,
                                                 const float2* __restrict__ gtw) {
  typedef typename Elem<DT>::T T;
  constexpr int N = NU / 2;
  constexpr int PB = 4;
  constexpr int TS = lds_len<PB>(N + 1) | 1;  // odd row stride (float2): rows in distinct banks
  constexpr int LB = tw_lbits(NU);
  __shared__ float2 buf[ROWS * TS];
  __shared__ float2 tlo[1 << LB], thi[NU >> LB];
  const int f = blockIdx.y;
  const int y0 = blockIdx.x * ROWS;
  const char* fbase = g.src + (size_t)(f0 + f) * g.frame_pitch;
  tw_split_load(tlo, thi, gtw, NU, LB, threadIdx.x, NT);
  // packed row z[m] = (x[2m], x[2m+1])
  struct alignas(2 * sizeof(T)) Pair {
    T a, b;
  };
  const bool corr = g.dark || g.gain || g.xs < 1.f || g.ys < 1.f;
  // All loads of the workgroup's rows are issued before the first LDS write
  // (a load-then-store loop waits out the memory latency once per element).
  constexpr int PER = (N + NT - 1) / NT;
  float2 zv[ROWS][PER];
  // the path is uniform; inside it every load is predicated, not branched
  if (corr) {
    // Row by row: every pixel's loads (frame, dark, gain at its stretch
    // source) are issued unconditionally, at the region origin for pixels
    // outside the region, whose loaded values are then replaced by zeros
    // (which the corrections map to 0).
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const int y = y0 + r;
      T raw[PER][2];
      float dk[PER][2], gk[PER][2];
      bool val[PER][2];
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int m = threadIdx.x + k * NT, x = 2 * m;
        const bool ok = (N % NT == 0 || m < N) && y < g.h;
#pragma unroll
        for (int e = 0; e < 2; e++) {
          val[k][e] = ok && x + e < g.w;
          const int2 sp = val[k][e] ? stretch_src(g, g.offx + x + e, g.offy + y)
                                    : make_int2(g.offx, g.offy);
          raw[k][e] = reinterpret_cast<const T*>(fbase + (size_t)sp.y * g.row_pitch)[sp.x];
          const size_t idx = (size_t)sp.y * g.nx + sp.x;
          dk[k][e] = g.dark ? g.dark[idx] : 0.f;
          gk[k][e] = g.gain ? g.gain[idx] : 1.f;
        }
      }
#pragma unroll
      for (int k = 0; k < PER; k++) {
        float c[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const T rv = val[k][e] ? raw[k][e] : (T)0;
          c[e] = correct(g, (float)rv, val[k][e] ? dk[k][e] : 0.f, val[k][e] ? gk[k][e] : 0.f);
        }
        zv[r][k] = make_float2(c[0], c[1]);
      }
    }
  } else if (g.vec && g.w >= 2 && g.h >= 1) {
    // Unconditional pair loads at clamped (always valid) addresses, zeroed
    // outside the region afterwards: no load sits in a branch, so all are in
    // flight together and the element conversions follow the last wait.
    const int xmax = (g.w - 2) & ~1;
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const int y = y0 + r;
      const int yc = y < g.h ? y : g.h - 1;
      const T* row = reinterpret_cast<const T*>(fbase + (size_t)(yc + g.offy) * g.row_pitch) + g.offx;
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int m = threadIdx.x + k * NT, x = 2 * m;
        const bool pv = (N % NT == 0 || m < N) && y < g.h && x + 1 < g.w;
        const Pair v = *reinterpret_cast<const Pair*>(row + (x < xmax ? x : xmax));
        if constexpr (DT == DT_FLOAT32) {
          zv[r][k] = pv ? make_float2((float)v.a, (float)v.b) : make_float2(0.f, 0.f);
        } else {
          // integer counts are finite: a 0/1 factor instead of a select keeps
          // the compiler from moving each load into a branch of its own
          const float mk = pv ? 1.f : 0.f;
          zv[r][k] = make_float2((float)v.a * mk, (float)v.b * mk);
        }
      }
    }
  } else if (g.w >= 1 && g.h >= 1) {
    // element loads at clamped addresses, the loaded values then zeroed
    // outside the region (a select on the load result, not a branch)
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const int y = y0 + r;
      const int yc = y < g.h ? y : g.h - 1;
      const T* row = reinterpret_cast<const T*>(fbase + (size_t)(yc + g.offy) * g.row_pitch) + g.offx;
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int m = threadIdx.x + k * NT, x = 2 * m;
        const bool ok = (N % NT == 0 || m < N) && y < g.h;
        const T a = row[x < g.w ? x : g.w - 1], b = row[x + 1 < g.w ? x + 1 : g.w - 1];
        const bool va = ok && x < g.w, vb = ok && x + 1 < g.w;
        if constexpr (DT == DT_FLOAT32)
          zv[r][k] = make_float2(va ? (float)a : 0.f, vb ? (float)b : 0.f);
        else  // finite counts: 0/1 factors (see the pair path above)
          zv[r][k] = make_float2((float)a * (va ? 1.f : 0.f), (float)b * (vb ? 1.f : 0.f));
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; r++)
#pragma unroll
      for (int k = 0; k < PER; k++) zv[r][k] = make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int r = 0; r < ROWS; r++)
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int m = threadIdx.x + k * NT;
      if (N % NT == 0 || m < N) buf[r * TS + lix<PB>(m)] = zv[r][k];
    }
  if (!corr && g.vec && g.w >= 2 && (g.w & 1) && (g.w - 1) / 2 % NT == (int)threadIdx.x) {
    // odd width: the last column has no pair partner (same thread as above)
    const int x = g.w - 1;
    for (int r = 0; r < ROWS; r++) {
      const int y = y0 + r;
      if (y >= g.h) break;
      const T* row = reinterpret_cast<const T*>(fbase + (size_t)(y + g.offy) * g.row_pitch) + g.offx;
      buf[r * TS + lix<PB>(x / 2)] = make_float2((float)row[x], 0.f);
    }
  }
  __syncthreads();
  const TwSplit tw{tlo, thi, LB, NU};
  fft_lds_t<N, NT, -1, ROWS, TS, TwSplit, PB>(buf, tw, NU);
  // unpack in place, one thread per bin pair (k, N-k), k in [0, N/2]:
  // X[k] = E[k] + e^{-2 pi i k/NU} O[k], E = (Z[k] + conj Z[N-k]) / 2,
  // O = (Z[k] - conj Z[N-k]) / 2i; X[N] from k = 0.
  // A thread takes a bin pair for all ROWS rows, so its two twiddles are
  // looked up once.
  constexpr int NP = N / 2 + 1;
  for (int k = threadIdx.x; k < NP; k += NT) {
    const float2 wk = twv<-1>(tw, k), wm = twv<-1>(tw, N - k);
    const int lk = lix<PB>(k), lm = lix<PB>(k ? N - k : 0);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      float2* Z = buf + r * TS;
      const float2 zk = Z[lk], zm = Z[lm];
      // bin k
      {
        const float2 zc = cconj(zm);
        const float2 E = cscale(cadd(zk, zc), 0.5f);
        const float2 D = csub(zk, zc);
        const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
        const float2 X = cadd(E, cmul(wk, O));
        if (k == 0) {
          // X[N] = E[0] - O[0]
          Z[lix<PB>(N)] = csub(E, O);
        }
        Z[lk] = X;
      }
      // bin N-k (k in (0, N/2))
      if (k > 0 && k < N - k) {
        const float2 zc = cconj(zk);
        const float2 E = cscale(cadd(zm, zc), 0.5f);
        const float2 D = csub(zm, zc);
        const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
        Z[lm] = cadd(E, cmul(wm, O));
      }
    }
  }
  __syncthreads();
  // transposed store, bins in groups of four: RT[kx/4][y][kx%4]; 16
  // consecutive lanes write 4 bins x 4 rows = one 128-byte line.  A thread
  // keeps its bin-in-group and row; its group advances by NT/(4 ROWS).
  static_assert(NT % (4 * ROWS) == 0, "store tiling");
  float2* out = rt + (size_t)f * rt_fstride;
  const int nkg = (nkx + 3) / 4;
  {
    const int kk = threadIdx.x & 3, r = (threadIdx.x >> 2) % ROWS;
    const int y = y0 + r;
    if (y < g.h) {
      for (int kg = (threadIdx.x >> 2) / ROWS; kg < nkg; kg += NT / (4 * ROWS)) {
        const int k = 4 * kg + kk;
        if (k >= nkx) break;
        const float2 X = buf[r * TS + lix<PB>(k)];
        out[((size_t)kg * NU + y) * 4 + kk] = X;
        if (k == 0 && rowsum) rowsum[(size_t)f * NU + y] = X.x;
      }
    }
  }
}

// ------------------------------------------------- per-frame fill value
__global__ __launch_bounds__(256) void k_fill(const float* rowsum, int nu, int h, int w,
                                              float* fill) {
  __shared__ double red[256];
  const int f = blockIdx.x;
  const float* rs = rowsum + (size_t)f * nu;
  // sixteen loads per thread in flight (clamped, unused past h), then added
  // in the order y = tid, tid + 256, ...
  constexpr int U = 16;
  double s = 0;
  for (int y0 = threadIdx.x; y0 < h; y0 += 256 * U) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = rs[min(y0 + 256 * u, h - 1)];
#pragma unroll
    for (int u = 0; u < U; u++)
      if (y0 + 256 * u < h) s += v[u];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) fill[f] = (float)(red[0] / ((double)w * (double)h));
}

hipError_t launch_fill(const float* rowsum, int nu, int h, int w, float* fill, int nframes,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(nframes), dim3(256), 0, st, rowsum, nu, h, w, fill);
  return hipGetLastError();
}

// ---------------------------------------------------------- column pass
// Loads column kx of the row-pass scratch (rows [0, h_data); the canvas rows
// beyond are the fill) into LDS, adding fill x the spectrum P of the fill
// mask in pad mode.  All NU/NT loads are issued before the LDS writes.
template <int NU, int NT, int PB>
MC_DEV void col_load(const float2* src, int h_data, bool fill, float fv, float2 P, int kx,
                     float2* buf) {
  constexpr int PER = (NU + NT - 1) / NT;
  float2 vv[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = threadIdx.x + k * NT;
    vv[k] = (NU % NT == 0 || i < NU) && i < h_data ? src[(size_t)i * 4] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = threadIdx.x + k * NT;
    if (NU % NT != 0 && i >= NU) continue;
    float2 v = vv[k];
    if (i < h_data) {
      if (fill) v = cadd(v, cscale(P, fv));
    } else if (fill && kx == 0) {
      v.x = fv * (float)NU;
    }
    buf[lix<PB>(i)] = v;
  }
}

// PLAN selects the radix plan of the column FFT and PB the LDS index padding
// (one float2 after every 2^PB; 0: none); the autotuner chooses both.
template <int NU, int NT, int PLAN = 0, int PB = 4>
__global__ __launch_bounds__(NT) void k_c2c_cols(const float2* rt, size_t rt_fstride, int nkx,
                                                 int h_data, const float* fill,
                                                 const float2* __restrict__ padP, int n,
                                                 float2* spec, size_t spec_stride,
                                                 const float2* __restrict__ gtw) {
  constexpr int LB = tw_lbits(NU);
  __shared__ float2 buf[lds_len<PB>(NU)];
  __shared__ float2 tlo[1 << LB], thi[NU >> LB];
  tw_split_load(tlo, thi, gtw, NU, LB, threadIdx.x, NT);
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int kx = L % nkx, f = L / nkx;
  const float2* src = rt + (size_t)f * rt_fstride + (size_t)(kx >> 2) * NU * 4 + (kx & 3);
  float fv = 0.f;
  float2 P = make_float2(0.f, 0.f);
  if (fill) {
    fv = fill[f];
    P = padP[kx];
  }
  col_load<NU, NT, PB>(src, h_data, fill != nullptr, fv, P, kx, buf);
  __syncthreads();
  fft_lds_t<NU, NT, -1, 1, NU, TwSplit, PB, PLAN>(buf, TwSplit{tlo, thi, LB, NU}, NU);
  float2* dst = spec + (size_t)f * spec_stride + (size_t)kx * n;
  for (int i = threadIdx.x; i < n; i += NT) {
    const int si = i < n / 2 ? i : NU - n + i;
    dst[i] = buf[lix<PB>(si)];
  }
}

// Column-pass variant v of the autotuner: radix plan v / 2, LDS padding
// (v % 2 ? none : one float2 per 16); plans 1 and 2 are compiled for the
// canvases 1024 to 8192.
template <int NU> constexpr bool cols_plans_compiled() {
  return NU == 1024 || NU == 2048 || NU == 4096 || NU == 8192;
}
template <int NU, int NTC>
static hipError_t launch_cols_plan(int cplan, dim3 grid, hipStream_t st, const float2* rt,
                                   size_t rt_fstride, int nkx, int h, const float* fill,
                                   const float2* padP, int n, float2* spectra,
                                   size_t spec_stride, const float2* tw) {
#define MC_COLS(PL, PBV) hipLaunchKernelGGL((k_c2c_cols<NU, NTC, PL, PBV>), grid, dim3(NTC), 0, st, rt, rt_fstride, nkx, h, fill, padP, n, spectra, spec_stride, tw)
  if constexpr (cols_plans_compiled<NU>()) {
    switch (cplan) {
      case 1: MC_COLS(0, 0); return hipGetLastError();
      case 2: MC_COLS(1, 4); return hipGetLastError();
      case 3: MC_COLS(1, 0); return hipGetLastError();
      case 4: MC_COLS(2, 4); return hipGetLastError();
      case 5: MC_COLS(2, 0); return hipGetLastError();
      default: break;
    }
  }
  MC_COLS(0, 4);
#undef MC_COLS
  return hipGetLastError();
}

// Persistent form of the column pass: each workgroup walks a contiguous
// range of (frame, kx) columns owned by its XCD and keeps the FFT twiddles
// of its threads in registers (no twiddle-table reads in the loop).
template <int NU, int NT>
__global__ __launch_bounds__(NT) void k_c2c_cols_p(const float2* rt, size_t rt_fstride, int nkx,
                                                   int nitems, int h_data, const float* fill,
                                                   const float2* __restrict__ padP, int n,
                                                   float2* spec, size_t spec_stride,
                                                   const float2* __restrict__ gtw) {
  constexpr int PB = 4;
  __shared__ float2 buf[lds_len<PB>(NU)];
  FftTw<NU, NT, -1, 1> ftw;
  ftw.init(TwTable{gtw, NU}, NU);
  const int xcd = blockIdx.x % 8, lid = blockIdx.x / 8, nper = gridDim.x / 8;
  const int i0 = (int)((long long)nitems * xcd / 8), i1 = (int)((long long)nitems * (xcd + 1) / 8);
  for (int L = i0 + lid; L < i1; L += nper) {
    const int kx = L % nkx, f = L / nkx;
    const float2* src = rt + (size_t)f * rt_fstride + (size_t)(kx >> 2) * NU * 4 + (kx & 3);
    float fv = 0.f;
    float2 P = make_float2(0.f, 0.f);
    if (fill) {
      fv = fill[f];
      P = padP[kx];
    }
    col_load<NU, NT, PB>(src, h_data, fill != nullptr, fv, P, kx, buf);
    __syncthreads();
    fft_lds_rw<NU, NT, -1, 1, NU, PB>(buf, ftw);
    float2* dst = spec + (size_t)f * spec_stride + (size_t)kx * n;
    for (int i = threadIdx.x; i < n; i += NT) {
      const int si = i < n / 2 ? i : NU - n + i;
      dst[i] = buf[lix<PB>(si)];
    }
    __syncthreads();  // buf is rewritten by the next column
  }
}

template <int NU, int ROWS, int NT>
static hipError_t launch_rows(const R2CGeom& g, int f0, int nframes, int nkx, float2* rt,
                              size_t rt_fstride, float* rs, const float2* tw, hipStream_t st) {
  dim3 gr((g.h + ROWS - 1) / ROWS, nframes);
  switch (g.dtype) {
#define MC_R2C(D) case D: hipLaunchKernelGGL((k_r2c_rows<NU, ROWS, NT, D>), gr, dim3(NT), 0, st, g, f0, nkx, rt, rt_fstride, rs, tw); break;
    MC_R2C(DT_INT8) MC_R2C(DT_UINT8) MC_R2C(DT_INT16) MC_R2C(DT_UINT16) MC_R2C(DT_FLOAT32)
#undef MC_R2C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

constexpr int rows_nt(int rows, int n) {
  return (rows * n / 16) < 64 ? 64 : ((rows * n / 16) > 256 ? 256 : rows * n / 16);
}

template <int NU>
static hipError_t frames_to_spectra_t(const R2CGeom& g, int f0, int nframes, float2* spectra,
                                      size_t spec_stride, float2* rt, float* rowsum,
                                      float* fill, const float2* padP, int variant,
                                      hipStream_t st) {
  constexpr int N = NU / 2;
  constexpr int NTC = (NU / 16) < 64 ? 64 : ((NU / 16) > 256 ? 256 : NU / 16);
  const float2* tw = twiddles(NU);
  if (!tw) return hipErrorOutOfMemory;
  const int nkx = g.n / 2 + 1;
  const size_t rt_fstride = rt_frame_elems(g.n, NU);
  float* rs = g.pad_mode ? rowsum : nullptr;
  R2CGeom gv = g;
  {
    // pair loads need every (x, x+1) element pair aligned to two elements
    size_t es = 1;
    switch (g.dtype) {
      case DT_INT16: case DT_UINT16: es = 2; break;
      case DT_FLOAT32: es = 4; break;
      default: es = 1;
    }
    const size_t a = 2 * es;
    gv.vec = ((uintptr_t)g.src % a == 0 && g.row_pitch % a == 0 && g.frame_pitch % a == 0 &&
              (g.offx & 1) == 0) ? 1 : 0;
  }
  // bit R2C_COLS_P selects the persistent column pass, bits from
  // R2C_CPLAN_SHIFT the column plan and padding of the per-column pass
  const bool cols_persistent = (variant & R2C_COLS_P) != 0;
  const int cplan = (variant >> R2C_CPLAN_SHIFT) & 7;
  variant &= R2C_COLS_P - 1;
  // row-pass shape: 4 rows x 512 threads on canvases >= 4096 by default
  if (variant == R2C_AUTO) variant = (4 * N / 16 >= 512) ? R2C_R4_T512 : R2C_R4_T256;
  if (variant == R2C_R4_T512 && 4 * N / 16 < 512) variant = R2C_R4_T256;
  hipError_t e = hipSuccess;
  bool done = false;
  if constexpr (4 * N / 16 >= 512) {
    if (variant == R2C_R4_T512) {
      e = launch_rows<NU, 4, 512>(gv, f0, nframes, nkx, rt, rt_fstride, rs, tw, st);
      done = true;
    }
  }
  if (!done) {
    if (variant == R2C_R2_T256)
      e = launch_rows<NU, 2, rows_nt(2, N)>(gv, f0, nframes, nkx, rt, rt_fstride, rs, tw, st);
    else
      e = launch_rows<NU, 4, rows_nt(4, N)>(gv, f0, nframes, nkx, rt, rt_fstride, rs, tw, st);
  }
  if (e != hipSuccess) return e;
  if (g.pad_mode) {
    hipLaunchKernelGGL(k_fill, dim3(nframes), dim3(256), 0, st, rowsum, NU, g.h, g.w, fill);
  }
  if (cols_persistent) {
    const int nitems = nkx * nframes;
    const int grid = resident_grid(k_c2c_cols_p<NU, NTC>, NTC, nitems);
    hipLaunchKernelGGL((k_c2c_cols_p<NU, NTC>), dim3(grid), dim3(NTC), 0, st, rt, rt_fstride, nkx,
                       nitems, g.h, g.pad_mode ? fill : nullptr, padP, g.n, spectra, spec_stride,
                       tw);
  } else {
    return launch_cols_plan<NU, NTC>(cplan, dim3(nkx * nframes), st, rt, rt_fstride, nkx, g.h,
                                     g.pad_mode ? fill : nullptr, padP, g.n, spectra, spec_stride,
                                     tw);
  }
  return hipGetLastError();
}

hipError_t launch_frames_to_spectra(const R2CGeom& g, int f0, int nframes, float2* spectra,
                                    size_t spec_stride, float2* rt, float* rowsum, float* fill,
                                    const float2* padP, int variant, hipStream_t st) {
  switch (g.nu) {
#define MC_F2S(NU) case NU: return frames_to_spectra_t<NU>(g, f0, nframes, spectra, spec_stride, rt, rowsum, fill, padP, variant, st);
    MC_F2S(64) MC_F2S(128) MC_F2S(256) MC_F2S(512) MC_F2S(1024) MC_F2S(2048) MC_F2S(4096) MC_F2S(8192)
    MC_F2S(768) MC_F2S(1536) MC_F2S(3072) MC_F2S(6144)
    MC_F2S(1280) MC_F2S(2560) MC_F2S(5120) MC_F2S(1920) MC_F2S(3840)
#undef MC_F2S
    default:  // any other even size: chirp-z transforms (mc_czt.hip)
      return launch_frames_to_spectra_czt(g, f0, nframes, spectra, spec_stride, rt, rowsum, fill,
                                          padP, st);
  }
}

// =================================================== inverse (sum images)
// Column pass: inverse transform of S[kx][:] along ky, stored in groups of
// four bins, tmp[kx/4][ky][kx%4] (the row-pass scratch layout of the forward
// transform): the four columns of a group run on one XCD (remapped block
// order), so each 128-byte line of tmp is completed in that XCD's L2 rather
// than written to HBM piecewise.
template <int N, int NT>
__global__ __launch_bounds__(NT) void k_cols_inv(const float2* spec, float2* tmp,
                                                 const float2* __restrict__ gtw) {
  // padded LDS indices and split twiddle tables, as in the forward column pass
  constexpr int LB = tw_lbits(N);
  constexpr int PB = 4;
  __shared__ float2 buf[lds_len<PB>(N)];
  __shared__ float2 tlo[1 << LB], thi[N >> LB];
  tw_split_load(tlo, thi, gtw, N, LB, threadIdx.x, NT);
  constexpr int PER = (N + NT - 1) / NT;
  const int kx = xcd_remap(blockIdx.x, gridDim.x);
  const float2* src = spec + (size_t)kx * N;
  float2 vv[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = threadIdx.x + k * NT;
    if (N % NT == 0 || i < N) vv[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = threadIdx.x + k * NT;
    if (N % NT == 0 || i < N) buf[lix<PB>(i)] = vv[k];
  }
  __syncthreads();
  fft_lds_t<N, NT, 1, 1, N, TwSplit, PB>(buf, TwSplit{tlo, thi, LB, N}, N);
  float2* dst = tmp + (size_t)(kx >> 2) * N * 4 + (kx & 3);
  for (int i = threadIdx.x; i < N; i += NT) dst[(size_t)i * 4] = buf[lix<PB>(i)];
}

// Row pass: C2R of ROWS rows of tmp with the DC/Nyquist projection
// (imaginary parts of bins 0 and N/2 ignored), x *= scale.
template <int N, int ROWS, int NT>
__global__ __launch_bounds__(NT) void k_c2r_rows(const float2* tmp, float scale, float* img,
                                                 size_t pitch, const float2* __restrict__ tw) {
  constexpr int NH = N / 2;        // packed complex length
  constexpr int H = N / 2 + 1;     // bins per row
  constexpr int TS = NH + 1;       // LDS row stride
  __shared__ float2 buf[ROWS * TS];
  const int y0 = blockIdx.x * ROWS;
  // tmp[kx/4][y][kx%4]: 16 lanes read 4 bins x 4 rows = one 128-byte line.
  // Every load of the workgroup is issued before its first LDS write; the
  // bins past n/2 in the last group of four exist in tmp and are not kept.
  constexpr int TOT = ROWS * 4 * ((H + 3) / 4);
  constexpr int NL = (TOT + NT - 1) / NT;
  float2 vv[NL];
#pragma unroll
  for (int q = 0; q < NL; q++) {
    const int i = min(threadIdx.x + q * NT, TOT - 1);
    const int kk = i & 3, r = (i >> 2) % ROWS, kg = (i >> 2) / ROWS;
    vv[q] = tmp[((size_t)kg * N + y0 + r) * 4 + kk];
  }
#pragma unroll
  for (int q = 0; q < NL; q++) {
    const int i = threadIdx.x + q * NT;
    const int kk = i & 3, r = (i >> 2) % ROWS, kg = (i >> 2) / ROWS;
    const int k = 4 * kg + kk;
    if (i >= TOT || k >= H) continue;
    float2 v = vv[q];
    if (k == 0 || k == NH) v.y = 0.f;
    buf[r * TS + k] = v;
  }
  __syncthreads();
  // Z[k] = (X[k] + conj X[NH-k]) + i e^{+2 pi i k/N} (X[k] - conj X[NH-k])
  constexpr int NP = NH / 2 + 1;  // pairs (k, NH-k), k in [0, NH/2]
  float2 za[(ROWS * NP + NT - 1) / NT], zb[(ROWS * NP + NT - 1) / NT];
#pragma unroll
  for (int p = 0; p < (ROWS * NP + NT - 1) / NT; p++) {
    const int i = threadIdx.x + p * NT;
    if (i < ROWS * NP) {
      const int r = i / NP, k = i % NP;
      const float2 a = buf[r * TS + k], b = buf[r * TS + NH - k];
      const float2 s1 = cadd(a, cconj(b)), d1 = csub(a, cconj(b));
      za[p] = cadd(s1, mul_si<1>(cmul(tw[k], d1)));
      const float2 s2 = cadd(b, cconj(a)), d2 = csub(b, cconj(a));
      zb[p] = cadd(s2, mul_si<1>(cmul(tw[NH - k], d2)));
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < (ROWS * NP + NT - 1) / NT; p++) {
    const int i = threadIdx.x + p * NT;
    if (i < ROWS * NP) {
      const int r = i / NP, k = i % NP;
      buf[r * TS + k] = za[p];
      if (k != 0 && k != NH - k) buf[r * TS + NH - k] = zb[p];
    }
  }
  __syncthreads();
  fft_lds<NH, NT, 1, ROWS, TS>(buf, tw, N);
  for (int i = threadIdx.x; i < ROWS * NH; i += NT) {
    const int r = i / NH, m = i % NH;
    const float2 z = buf[r * TS + m];
    reinterpret_cast<float2*>(img + (size_t)(y0 + r) * pitch)[m] =
        make_float2(z.x * scale, z.y * scale);
  }
}

template <int N>
static hipError_t spectrum_to_image_t(const float2* spec, float scale, float* img, size_t pitch,
                                      float2* tmp, hipStream_t st) {
  constexpr int NTC = (N / 16) < 64 ? 64 : ((N / 16) > 256 ? 256 : N / 16);
  constexpr int ROWS = 4;
  constexpr int NTR = (ROWS * N / 32) < 64 ? 64 : ((ROWS * N / 32) > 256 ? 256 : ROWS * N / 32);
  const float2* tw = twiddles(N);
  if (!tw) return hipErrorOutOfMemory;
  hipLaunchKernelGGL((k_cols_inv<N, NTC>), dim3(N / 2 + 1), dim3(NTC), 0, st, spec, tmp, tw);
  hipLaunchKernelGGL((k_c2r_rows<N, ROWS, NTR>), dim3(N / ROWS), dim3(NTR), 0, st, tmp, scale,
                     img, pitch, tw);
  return hipGetLastError();
}

hipError_t launch_spectrum_to_image(const float2* spec, int n, float scale, float* img,
                                    size_t img_pitch, float2* tmp, hipStream_t st) {
  switch (n) {
    case 64: return spectrum_to_image_t<64>(spec, scale, img, img_pitch, tmp, st);
    case 128: return spectrum_to_image_t<128>(spec, scale, img, img_pitch, tmp, st);
    case 256: return spectrum_to_image_t<256>(spec, scale, img, img_pitch, tmp, st);
    case 512: return spectrum_to_image_t<512>(spec, scale, img, img_pitch, tmp, st);
    case 1024: return spectrum_to_image_t<1024>(spec, scale, img, img_pitch, tmp, st);
    case 2048: return spectrum_to_image_t<2048>(spec, scale, img, img_pitch, tmp, st);
    case 4096: return spectrum_to_image_t<4096>(spec, scale, img, img_pitch, tmp, st);
    case 8192: return spectrum_to_image_t<8192>(spec, scale, img, img_pitch, tmp, st);
    case 768: return spectrum_to_image_t<768>(spec, scale, img, img_pitch, tmp, st);
    case 1536: return spectrum_to_image_t<1536>(spec, scale, img, img_pitch, tmp, st);
    case 3072: return spectrum_to_image_t<3072>(spec, scale, img, img_pitch, tmp, st);
    case 6144: return spectrum_to_image_t<6144>(spec, scale, img, img_pitch, tmp, st);
    case 1280: return spectrum_to_image_t<1280>(spec, scale, img, img_pitch, tmp, st);
    case 2560: return spectrum_to_image_t<2560>(spec, scale, img, img_pitch, tmp, st);
    case 5120: return spectrum_to_image_t<5120>(spec, scale, img, img_pitch, tmp, st);
    case 1920: return spectrum_to_image_t<1920>(spec, scale, img, img_pitch, tmp, st);
    case 3840: return spectrum_to_image_t<3840>(spec, scale, img, img_pitch, tmp, st);
    default:  // any other even size: chirp-z transforms (mc_czt.hip)
      return launch_spectrum_to_image_czt(spec, n, scale, img, img_pitch, tmp, st);
  }
}

}  // namespace mc
