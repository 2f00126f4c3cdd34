// mc_czt.hip -- transforms of any even length (the reference's default
// square crop, e.g. 3710 = 2*5*7*53 for a 3838 x 3710 K2 frame, or a
// non-integer -bin): Bluestein's chirp-z algorithm on the power-of-two LDS
// Stockham engine of mc_device.h.  The reference plans cuFFT at any size
// (cufunc.cu:202-215, called for nsamUnbin and nsam, DFAlign.cu:596,838).
//
// A length-L DFT with sign S, X_k = sum_m x_m e^{S 2 pi i m k/L}, is
//   X_k = c_k sum_m (x_m c_m) h_{k-m},  c_j = e^{S i pi j^2/L}, h = conj(c),
// a linear convolution, evaluated as a cyclic one of power-of-two length
// M >= 2L - 1: FFT_M, product with the precomputed FFT_M(h)/M, inverse FFT_M.
//
// The pair stage needs only the B box rows y = ori + u of the column
// inverse of length n, so it runs the chirp-z transform restricted to them:
//   Y[ori+u] = e^{i pi (ori+u)^2/n} sum_ky (X[ky] e^{i pi ky^2/n}) g[n-1+u-ky],
//   g_j = e^{-i pi (ori-n+1+j)^2/n},  j < n+B-1,
// a cyclic convolution of length M >= n + B - 1 (4096 for n = 3710, B = 96).
// Every chirp phase pi j^2/L is reduced exactly (j^2 mod 2L in integers)
// and evaluated in double on the host; the tables are float2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "mc_device.h"
#include "mc_frames.h"
#include "mc_kernels.h"

namespace mc {

// ------------------------------------------------------------ host tables
// In-place radix-2 FFT in double, a.size() a power of two, sign S.
static void host_fft(std::vector<std::complex<double>>& a, int S) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; k++) {
        const double t = S * 2.0 * M_PI * (double)k / (double)len;
        const std::complex<double> w(std::cos(t), std::sin(t));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

// e^{S i pi j^2/L} with the phase reduced exactly
static std::complex<double> chirp(long long j, int L, int S) {
  const long long r = ((j % (2LL * L)) * (j % (2LL * L))) % (2LL * L);
  const double t = S * M_PI * (double)r / (double)L;
  return std::complex<double>(std::cos(t), std::sin(t));
}

int czt_len(int need) {
  int M = 64;
  while (M < need) M <<= 1;
  return M;
}

static float2* upload(const std::vector<float2>& h) {
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * h.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

static float2 f2(std::complex<double> z) { return make_float2((float)z.real(), (float)z.imag()); }

// Bluestein tables of a length-L DFT with sign S: chirp[j] = c_j (j < L),
// filt[k] = FFT_M(h)[k]/M (k < M), M = czt_len(2L - 1).
struct BsTab {
  const float2* chirp;
  const float2* filt;
  int M;
};

static const BsTab* bs_tables(int L, int S) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, BsTab> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, L, S);
  auto it = cache.find(key);
  if (it != cache.end()) return &it->second;
  const int M = czt_len(2 * L - 1);
  std::vector<float2> c(L);
  std::vector<std::complex<double>> h(M, 0.0);
  for (int j = 0; j < L; j++) {
    const std::complex<double> cj = chirp(j, L, S);
    c[j] = f2(cj);
    h[j] = std::conj(cj);
    if (j > 0) h[M - j] = std::conj(cj);
  }
  host_fft(h, -1);
  std::vector<float2> F(M);
  for (int k = 0; k < M; k++) F[k] = f2(h[k] / (double)M);
  BsTab t;
  t.chirp = upload(c);
  t.filt = upload(F);
  t.M = M;
  if (!t.chirp || !t.filt) return nullptr;
  return &(cache[key] = t);
}

// Pruned chirp-z tables of the pair stage's column inverse (size n, box B):
// pre[ky] = e^{i pi ky^2/n} (ky < n), filt = FFT_M(g)/M, post[u] =
// e^{i pi (ori+u)^2/n} (u < B), M = czt_len(n + B - 1).
struct CztTab {
  const float2* pre;
  const float2* filt;
  const float2* post;
  int M;
};

static const CztTab* czt_tables(int n, int box) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, CztTab> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, n, box);
  auto it = cache.find(key);
  if (it != cache.end()) return &it->second;
  const int M = czt_len(n + box - 1);
  const int ori = (n - box) / 2;
  std::vector<float2> pre(n), post(box);
  for (int k = 0; k < n; k++) pre[k] = f2(chirp(k, n, 1));
  for (int u = 0; u < box; u++) post[u] = f2(chirp(ori + u, n, 1));
  std::vector<std::complex<double>> g(M, 0.0);
  for (int j = 0; j < n + box - 1; j++) g[j] = chirp(std::llabs((long long)ori - n + 1 + j), n, -1);
  host_fft(g, -1);
  std::vector<float2> F(M);
  for (int k = 0; k < M; k++) F[k] = f2(g[k] / (double)M);
  CztTab t;
  t.pre = upload(pre);
  t.filt = upload(F);
  t.post = upload(post);
  t.M = M;
  if (!t.pre || !t.filt || !t.post) return nullptr;
  return &(cache[key] = t);
}

// Twiddles of the two-level column transform n = Q P:
// tw[k2 Q + a] = e^{-2 pi i ((a k2) mod n)/n}, k2 < P, a < Q.
static const float2* qp_twiddles(int Q, int P) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, const float2*> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, Q, P);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int n = Q * P;
  std::vector<float2> h((size_t)n);
  for (int k2 = 0; k2 < P; k2++)
    for (int a = 0; a < Q; a++) {
      const double t = -2.0 * M_PI * (double)((a * k2) % n) / (double)n;
      h[(size_t)k2 * Q + a] = make_float2((float)std::cos(t), (float)std::sin(t));
    }
  const float2* d = upload(h);
  if (d) cache[key] = d;
  return d;
}

// ------------------------------------------------------------ device part
constexpr int CZT_PB = 4;
// Threads per workgroup: one radix-16 butterfly of stage 0 per thread at
// the largest size (the 16384-point stages spill at 512 threads).
template <int M> constexpr int czt_nt() { return M >= 16384 ? 1024 : (M >= 8192 ? 512 : 256); }
// Minimum waves per SIMD the kernels are built for.  With 512 threads (M =
// 8192) the LDS buffer admits two workgroups per CU, 4 waves per SIMD, which
// needs at most 128 VGPRs: uncapped the kernels took 130-170, which is one
// workgroup per CU.  With 256 threads (M <= 4096) the cap to 128 costs
// spills and measured no faster than 3 workgroups per CU, so none is set.
template <int NT> constexpr int czt_waves() { return NT >= 512 ? 4 : 1; }

// Length-L DFT with sign S of buf[lix(j)], j < L (buf holds lds_len(M)
// entries), in place; entries from L on are overwritten.  Caller
// synchronises before; returns synchronised.
template <int M, int NT, int S>
MC_DEV void bluestein(float2* buf, const TwSplit& tw, const float2* __restrict__ chirp,
                      const float2* __restrict__ filt, int L) {
  constexpr int PB = CZT_PB;
  for (int j = threadIdx.x; j < M; j += NT) {
    const int lj = lix<PB>(j);
    buf[lj] = j < L ? cmul(buf[lj], chirp[j]) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds_t<M, NT, -1, 1, M, TwSplit, PB>(buf, tw, M);
  for (int k = threadIdx.x; k < M; k += NT) {
    const int lk = lix<PB>(k);
    buf[lk] = cmul(buf[lk], filt[k]);
  }
  __syncthreads();
  fft_lds_t<M, NT, 1, 1, M, TwSplit, PB>(buf, tw, M);
  for (int k = threadIdx.x; k < L; k += NT) {
    const int lk = lix<PB>(k);
    buf[lk] = cmul(buf[lk], chirp[k]);
  }
  __syncthreads();
}

// The same with the M-point FFT twiddles of the calling thread held in
// registers (forward ones; the inverse FFT uses their conjugates), for
// persistent kernels whose threads keep their butterflies from item to item.
// The caller applies both chirp products as it writes the input and reads
// the result (two LDS passes and two barriers fewer): buf holds x[j] chirp[j]
// for j < L and zeros up to M, synchronised; on return buf[k] chirp[k] is the
// DFT, k < L.
template <int M, int NT>
MC_DEV void bluestein_rw(float2* buf, const FftTw<M, NT, -1, 1>& ftw,
                         const float2* __restrict__ filt) {
  constexpr int PB = CZT_PB;
  fft_lds_rw<M, NT, -1, 1, M, PB>(buf, ftw);
  for (int k = threadIdx.x; k < M; k += NT) {
    const int lk = lix<PB>(k);
    buf[lk] = cmul(buf[lk], filt[k]);
  }
  __syncthreads();
  fft_lds_rw<M, NT, 1, 1, M, PB>(buf, ftw);
}

// Pixel x of canvas row y of a frame: decoded, corrected, zero outside the
// data region.
template <int DT>
MC_DEV float row_px(const R2CGeom& g, const char* fbase, bool corr, int x, int y) {
  typedef typename Elem<DT>::T T;
  if (x >= g.w) return 0.f;
  const int2 sp = stretch_src(g, g.offx + x, g.offy + y);
  const float v = (float)reinterpret_cast<const T*>(fbase + (size_t)sp.y * g.row_pitch)[sp.x];
  if (!corr) return v;
  const size_t idx = (size_t)sp.y * g.nx + sp.x;
  return correct(g, v, g.dark ? g.dark[idx] : 0.f, g.gain ? g.gain[idx] : 1.f);
}

// Stage-1 row pass of one canvas row y < h of frame f (any even nu): the
// row packed into nu/2 complex values, a Bluestein DFT of length nu/2 (bs),
// unpack to the kept bins kx < nkx, stored as RT[kx/4][y][kx%4] (the layout
// of the power-of-two row pass; the column pass reads it the same way).
// Caller synchronises before; buf is read up to the return.
template <int M, int NT, int DT, class BS>
MC_DEV void row_czt(const R2CGeom& g, int f0, int f, int y, int nkx, float2* rt,
                    size_t rt_fstride, float* rowsum, const float2* __restrict__ twnu,
                    float2* buf, const BS& bs, const float2* __restrict__ chirp = nullptr) {
  constexpr int PB = CZT_PB;
  const int NU = g.nu, N = NU / 2;
  const char* fbase = g.src + (size_t)(f0 + f) * g.frame_pitch;
  const bool corr = g.dark || g.gain || g.xs < 1.f || g.ys < 1.f;
  auto px = [&](int x) -> float { return row_px<DT>(g, fbase, corr, x, y); };
  if (chirp) {  // the first chirp product of the Bluestein DFT, zeros up to M
    for (int m = threadIdx.x; m < M; m += NT)
      buf[lix<PB>(m)] = m < N ? cmul(make_float2(px(2 * m), px(2 * m + 1)), chirp[m])
                              : make_float2(0.f, 0.f);
  } else {
    for (int m = threadIdx.x; m < N; m += NT) buf[lix<PB>(m)] = make_float2(px(2 * m), px(2 * m + 1));
  }
  __syncthreads();
  bs(buf, N);
  // unpack, one thread per bin pair (k, N-k): X[k] = E[k] + e^{-2 pi i k/NU} O[k]
  for (int k = threadIdx.x; k <= N / 2; k += NT) {
    const int lk = lix<PB>(k), lm = lix<PB>(k ? N - k : 0);
    float2 zk = buf[lk], zm = buf[lm];
    if (chirp) {  // the final chirp product of the Bluestein DFT
      zk = cmul(zk, chirp[k]);
      zm = cmul(zm, chirp[k ? N - k : 0]);
    }
    {
      const float2 zc = cconj(zm);
      const float2 E = cscale(cadd(zk, zc), 0.5f);
      const float2 D = csub(zk, zc);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
      const float2 w = cconj(twnu[k]);
      if (k == 0) buf[lix<PB>(N)] = csub(E, O);
      buf[lk] = cadd(E, cmul(w, O));
    }
    if (k > 0 && k < N - k) {
      const float2 zc = cconj(zk);
      const float2 E = cscale(cadd(zm, zc), 0.5f);
      const float2 D = csub(zm, zc);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
      buf[lm] = cadd(E, cmul(cconj(twnu[N - k]), O));
    }
  }
  __syncthreads();
  float2* out = rt + (size_t)f * rt_fstride;
  for (int k = threadIdx.x; k < nkx; k += NT) {
    const float2 X = buf[lix<PB>(k)];
    out[((size_t)(k >> 2) * NU + y) * 4 + (k & 3)] = X;
    if (k == 0 && rowsum) rowsum[(size_t)f * NU + y] = X.x;
  }
}

// One workgroup per row, twiddles from the split LDS tables.
template <int M, int NT, int DT>
__global__ __launch_bounds__(NT, czt_waves<NT>()) void k_r2c_rows_czt(
    R2CGeom g, int f0, int nkx, float2* rt, size_t rt_fstride, float* rowsum,
    const float2* __restrict__ gtw, const float2* __restrict__ chirpv,
    const float2* __restrict__ filt, const float2* __restrict__ twnu) {
  constexpr int PB = CZT_PB;
  constexpr int LB = tw_lbits(M);
  __shared__ float2 buf[lds_len<PB>(M)];
  __shared__ float2 tlo[1 << LB], thi[M >> LB];
  tw_split_load(tlo, thi, gtw, M, LB, threadIdx.x, NT);
  const TwSplit tw{tlo, thi, LB, M};
  row_czt<M, NT, DT>(g, f0, blockIdx.y, blockIdx.x, nkx, rt, rt_fstride, rowsum, twnu, buf,
                     [&](float2* b, int L) { bluestein<M, NT, -1>(b, tw, chirpv, filt, L); });
}

// Persistent form: each workgroup walks a contiguous range of the (frame,
// row) items owned by its XCD with the FFT twiddles in registers.
template <int M, int NT, int DT>
__global__ __launch_bounds__(NT, czt_waves<NT>()) void k_r2c_rows_czt_p(
    R2CGeom g, int f0, int nframes, int nkx, float2* rt, size_t rt_fstride, float* rowsum,
    const float2* __restrict__ gtw, const float2* __restrict__ chirpv,
    const float2* __restrict__ filt, const float2* __restrict__ twnu) {
  constexpr int PB = CZT_PB;
  __shared__ float2 buf[lds_len<PB>(M)];
  FftTw<M, NT, -1, 1> ftw;
  ftw.init(TwTable{gtw, M}, M);
  const int nitems = g.h * nframes;
  const int xcd = blockIdx.x % 8, lid = blockIdx.x / 8, nper = gridDim.x / 8;
  const int i0 = (int)((long long)nitems * xcd / 8), i1 = (int)((long long)nitems * (xcd + 1) / 8);
  for (int L = i0 + lid; L < i1; L += nper) {
    row_czt<M, NT, DT>(g, f0, L / g.h, L % g.h, nkx, rt, rt_fstride, rowsum, twnu, buf,
                       [&](float2* b, int) { bluestein_rw<M, NT>(b, ftw, filt); }, chirpv);
    __syncthreads();  // buf is rewritten by the next row
  }
}

// Stage-1 column pass of column kx (any even nu): rows [0, h) from the
// row-pass scratch (the rest of the canvas is zero, or the fill in pad
// mode, added as fill x the spectrum P of the fill mask), a Bluestein DFT
// of length nu, and the rows of the Fourier crop to n written as S[kx][ky].
template <int M, int NT>
__global__ __launch_bounds__(NT, czt_waves<NT>()) void k_c2c_cols_czt(const float2* rt, size_t rt_fstride, int nkx,
                                                     int NU, int h_data, const float* fill,
                                                     const float2* __restrict__ padP, int n,
                                                     float2* spec, size_t spec_stride,
                                                     const float2* __restrict__ gtw,
                                                     const float2* __restrict__ chirpv,
                                                     const float2* __restrict__ filt) {
  constexpr int PB = CZT_PB;
  constexpr int LB = tw_lbits(M);
  __shared__ float2 buf[lds_len<PB>(M)];
  __shared__ float2 tlo[1 << LB], thi[M >> LB];
  tw_split_load(tlo, thi, gtw, M, LB, threadIdx.x, NT);
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int kx = L % nkx, f = L / nkx;
  const float2* src = rt + (size_t)f * rt_fstride + (size_t)(kx >> 2) * NU * 4 + (kx & 3);
  float fv = 0.f;
  float2 P = make_float2(0.f, 0.f);
  if (fill) {
    fv = fill[f];
    P = padP[kx];
  }
  for (int i = threadIdx.x; i < NU; i += NT) {
    float2 v = make_float2(0.f, 0.f);
    if (i < h_data) {
      v = src[(size_t)i * 4];
      if (fill) v = cadd(v, cscale(P, fv));
    } else if (fill && kx == 0) {
      v.x = fv * (float)NU;
    }
    buf[lix<PB>(i)] = v;
  }
  __syncthreads();
  bluestein<M, NT, -1>(buf, TwSplit{tlo, thi, LB, M}, chirpv, filt, NU);
  float2* dst = spec + (size_t)f * spec_stride + (size_t)kx * n;
  for (int i = threadIdx.x; i < n; i += NT) {
    const int si = i < n / 2 ? i : NU - n + i;
    dst[i] = buf[lix<PB>(si)];
  }
}

// Persistent form of k_c2c_cols_czt: each workgroup walks a contiguous
// range of the (frame, kx) columns owned by its XCD (kx minor, so the four
// columns of one row-pass line group are in flight together), with the FFT
// twiddles in registers.  The LDS buffer allows two workgroups per CU at
// M = 8192, so the registers the twiddles take cost no occupancy there.
template <int M, int NT>
__global__ __launch_bounds__(NT, czt_waves<NT>()) void k_c2c_cols_czt_p(
    const float2* rt, size_t rt_fstride, int nkx, int nitems, int NU, int h_data,
    const float* fill, const float2* __restrict__ padP, int n, float2* spec, size_t spec_stride,
    const float2* __restrict__ gtw, const float2* __restrict__ chirpv,
    const float2* __restrict__ filt) {
  constexpr int PB = CZT_PB;
  __shared__ float2 buf[lds_len<PB>(M)];
  FftTw<M, NT, -1, 1> ftw;
  ftw.init(TwTable{gtw, M}, M);
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
    // with the first chirp product of the Bluestein DFT, zeros up to M
    for (int i = threadIdx.x; i < M; i += NT) {
      float2 v = make_float2(0.f, 0.f);
      if (i < h_data) {
        v = src[(size_t)i * 4];
        if (fill) v = cadd(v, cscale(P, fv));
      } else if (fill && kx == 0 && i < NU) {
        v.x = fv * (float)NU;
      }
      buf[lix<PB>(i)] = i < NU ? cmul(v, chirpv[i]) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    bluestein_rw<M, NT>(buf, ftw, filt);
    float2* dst = spec + (size_t)f * spec_stride + (size_t)kx * n;
    for (int i = threadIdx.x; i < n; i += NT) {
      const int si = i < n / 2 ? i : NU - n + i;
      dst[i] = cmul(buf[lix<PB>(si)], chirpv[si]);  // with the final chirp product
    }
    __syncthreads();  // buf is rewritten by the next column
  }
}

// Stage-1 row pass at nu = 2 Q P (packed length N = Q P, Q with a Stockham
// plan, P odd), for the rows of the 3710 and 4092 crops (N = 1855 = 35 x 53,
// 2046 = 66 x 31): the
// packing, N-point DFT and unpack of row_czt, with the DFT in two levels as
// in k_c2c_cols_qp below.  Lane a < Q reads the packed values z[a + Q b]
// (pixels 2 (a + Q b) and 2 (a + Q b) + 1) straight from the frame and runs
// the P-point direct DFT into LDS row k2, element a; after the twiddles and
// the P Q-point FFTs, Z[k] is row k % P, element k / P (and the unpack puts
// X[N] in the spare element Q of row 0).  The LDS rows are unpadded (odd
// stride; at 3710, 16 KB per row fits 10 workgroups per CU against 9 with
// the padded layout, 0.5 ms less per 40-frame movie for the two passes).
// One row per workgroup of Q lanes rounded up to whole waves;
// the rows are numbered XCD-major, so the four rows that share a line of
// the row-pass scratch are written through one L2.
template <int Q, int P, int R, int NT, int DT>
__global__ __launch_bounds__(NT) void k_r2c_rows_qp(R2CGeom g, int f0, int nframes, int nkx,
                                                    float2* __restrict__ rt, size_t rt_fstride,
                                                    float* __restrict__ rowsum,
                                                    const float2* __restrict__ gtw,
                                                    const float2* __restrict__ twqp,
                                                    const float2* __restrict__ twnu) {
  constexpr int N = Q * P;
  constexpr int NH = N / 2 + 1;
  constexpr int PB = 0;  // unpadded rows (odd stride): more workgroups per CU
  constexpr int TS = lds_len<PB>(Q + 1) | 1;
  constexpr int RS = P * TS;  // one row's LDS block
  static_assert(NT >= R * Q && NT % 64 == 0, "one lane per P-point DFT");
  __shared__ float2 buf[R * RS];
  __shared__ float2 twm[Q];
  const int nitems = g.h * nframes;
  const int L0 = R * ((int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8));
  if (L0 >= nitems) return;
  const int nr = nitems - L0 < R ? nitems - L0 : R;  // rows of this workgroup
  for (int i = threadIdx.x; i < Q; i += NT) twm[i] = gtw[i * P];  // e^{+2 pi i i/Q}
  const bool corr = g.dark || g.gain || g.xs < 1.f || g.ys < 1.f;
  {
    const int r = threadIdx.x / Q, a = threadIdx.x - r * Q;
    if (r < nr) {
      const int L = L0 + r, y = L % g.h, f = L / g.h;
      const char* fbase = g.src + (size_t)(f0 + f) * g.frame_pitch;
      float2 x[P];
      if (!corr && g.vec) {
        // the pixel pairs as single aligned loads (the row is all data)
        typedef typename Elem<DT>::T T;
        struct alignas(2 * sizeof(T)) Pair {
          T a, b;
        };
        const Pair* row = reinterpret_cast<const Pair*>(
            reinterpret_cast<const T*>(fbase + (size_t)(y + g.offy) * g.row_pitch) + g.offx);
#pragma unroll
        for (int b = 0; b < P; b++) {
          const Pair v = row[a + Q * b];
          x[b] = make_float2((float)v.a, (float)v.b);
        }
      } else {
#pragma unroll
        for (int b = 0; b < P; b++) {
          const int m = a + Q * b;
          x[b] = make_float2(row_px<DT>(g, fbase, corr, 2 * m, y),
                             row_px<DT>(g, fbase, corr, 2 * m + 1, y));
        }
      }
      float2* rb = buf + r * RS;
      odd_dft<P, -1>(x, [&](int k, float2 v) { rb[k * TS + lix<PB>(a)] = v; });
    }
  }
  __syncthreads();
  // twiddles e^{-2 pi i a k2/N} (row k2 = 0 has none)
  for (int e = threadIdx.x; e < R * (N - Q); e += NT) {
    const int r = e / (N - Q), j = e - r * (N - Q) + Q;
    const int k2 = j / Q, i = j - k2 * Q;
    float2* p = buf + r * RS + k2 * TS + lix<PB>(i);
    *p = cmul(*p, twqp[j]);
  }
  __syncthreads();
  fft_lds_t<Q, NT, -1, R * P, TS, TwTable, PB>(buf, TwTable{twm, Q}, Q);
  auto at = [&](int k) -> int { return (k % P) * TS + lix<PB>(k / P); };
  // unpack, one thread per bin pair (k, N-k): X[k] = E[k] + e^{-2 pi i k/2N} O[k]
  for (int t = threadIdx.x; t < R * NH; t += NT) {
    const int r = t / NH, k = t - r * NH;
    float2* rb = buf + r * RS;
    const int lk = at(k), lm = at(k ? N - k : 0);
    const float2 zk = rb[lk], zm = rb[lm];
    {
      const float2 zc = cconj(zm);
      const float2 E = cscale(cadd(zk, zc), 0.5f);
      const float2 D = csub(zk, zc);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
      if (k == 0) rb[at(N)] = csub(E, O);
      rb[lk] = cadd(E, cmul(cconj(twnu[k]), O));
    }
    if (k > 0 && k < N - k) {
      const float2 zc = cconj(zk);
      const float2 E = cscale(cadd(zm, zc), 0.5f);
      const float2 D = csub(zm, zc);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
      rb[lm] = cadd(E, cmul(cconj(twnu[N - k]), O));
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nr * nkx; t += NT) {
    const int r = t / nkx, k = t - r * nkx;
    const int L = L0 + r, y = L % g.h, f = L / g.h;
    const float2 X = buf[r * RS + at(k)];
    rt[(size_t)f * rt_fstride + ((size_t)(k >> 2) * (2 * N) + y) * 4 + (k & 3)] = X;
    if (k == 0 && rowsum) rowsum[(size_t)f * (2 * N) + y] = X.x;
  }
}

// Stage-1 column pass at N = Q P without fill rows (Q with a Stockham plan,
// P odd), for the 3710 = 70 x 53 crop of a 3838 x 3710 frame and the 4092 =
// 132 x 31 crop of a K3 frame:
//   X[k2 + P k1] = sum_a e^{-2 pi i a k1/Q} (e^{-2 pi i a k2/N}
//                  sum_b x[a + Q b] e^{-2 pi i b k2/P}).
// Lane a < Q loads x[a + Q b] (b < P) and runs the direct P-point DFT
// (odd_dft) in registers into LDS row k2, element a; a second pass applies
// the twiddles; then P Q-point FFTs run in LDS, and row k2 element k1 is
// X[k2 + P k1]; the rows of the Fourier crop to n are stored.  At 3710
// about 0.4x the arithmetic of the 8192-point chirp-z convolution, in a 31
// KB LDS buffer (unpadded rows of odd stride 71, 5 workgroups per CU).  One
// column per workgroup (a
// persistent loop would hoist the LDS addresses of every FFT stage out of
// the loop and hold them in registers); the columns are numbered XCD-major,
// so the four columns of one row-pass line group share an L2.
template <int Q, int P, int NT, int W>
__global__ __launch_bounds__(NT, W) void k_c2c_cols_qp(const float2* __restrict__ rt,
                                                       size_t rt_fstride, int nkx, int nitems,
                                                       int n, float2* __restrict__ spec,
                                                       size_t spec_stride,
                                                       const float2* __restrict__ gtw,
                                                       const float2* __restrict__ twqp) {
  constexpr int N = Q * P;
  constexpr int PB = 0;  // as in k_r2c_rows_qp
  constexpr int TS = lds_len<PB>(Q) | 1;
  static_assert(NT >= Q && NT % 64 == 0, "one lane per P-point DFT");
  __shared__ float2 buf[P * TS];
  __shared__ float2 twm[Q];
  const int L = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  if (L >= nitems) return;
  for (int i = threadIdx.x; i < Q; i += NT) twm[i] = gtw[i * P];  // e^{+2 pi i i/Q}
  const int a = threadIdx.x;
  const int kx = L % nkx, f = L / nkx;
  const float2* src = rt + (size_t)f * rt_fstride + (size_t)(kx >> 2) * N * 4 + (kx & 3);
  if (a < Q) {
    const auto q = buf_rsrc(src, N * 32);
    float2 x[P];
#pragma unroll
    for (int b = 0; b < P; b++) x[b] = buf_ld2(q, 32 * a, 32 * Q * b);
    odd_dft<P, -1>(x, [&](int k, float2 v) { buf[k * TS + lix<PB>(a)] = v; });
  }
  __syncthreads();
  // twiddles e^{-2 pi i a k2/N} (row k2 = 0 has none)
  for (int e = Q + threadIdx.x; e < N; e += NT) {
    const int k2 = e / Q, i = e - k2 * Q;
    float2* p = buf + k2 * TS + lix<PB>(i);
    *p = cmul(*p, twqp[e]);
  }
  __syncthreads();
  fft_lds_t<Q, NT, -1, P, TS, TwTable, PB>(buf, TwTable{twm, Q}, Q);
  // the rows of the Fourier crop to n (n = N unbinned)
  float2* dst = spec + (size_t)f * spec_stride + (size_t)kx * n;
  for (int o = threadIdx.x; o < n; o += NT) {
    const int si = o < n / 2 ? o : N - n + o;
    dst[o] = buf[(si % P) * TS + lix<PB>(si / P)];
  }
}

// Pair-stage column pass (any even n): per (pair, kx) item, the weighted
// cross power X[ky] = F_a[kx][ky] conj(F_b[kx][ky]) wx[kx] wy[ky], then the
// pruned chirp-z inverse to the box rows, T[p][kx][u].  Persistent: each
// XCD walks a contiguous range of items (kx-major, so the frame columns of
// one kx are read from HBM once and reused from L2 by all pairs).
template <int M, int NT>
__global__ __launch_bounds__(NT, czt_waves<NT>()) void k_pair_cols_czt(const float2* __restrict__ spectra,
                                                      size_t spec_stride,
                                                      const int2* __restrict__ pairs, int npairs,
                                                      const float* __restrict__ wxy, int n,
                                                      int box, float2* __restrict__ tbuf,
                                                      const float2* __restrict__ gtw,
                                                      const float2* __restrict__ pre,
                                                      const float2* __restrict__ filt,
                                                      const float2* __restrict__ post, int kxb,
                                                      int nkx) {
  constexpr int PB = CZT_PB;
  constexpr int LB = tw_lbits(M);
  __shared__ float2 buf[lds_len<PB>(M)];
  __shared__ float2 tlo[1 << LB], thi[M >> LB];
  tw_split_load(tlo, thi, gtw, M, LB, threadIdx.x, NT);
  const TwSplit tw{tlo, thi, LB, M};
  const int H = n / 2 + 1;
  const int xcd = blockIdx.x % 8, lid = blockIdx.x / 8, nper = gridDim.x / 8;
  const long long items = (long long)npairs * nkx;
  const long long i0 = items * xcd / 8, i1 = items * (xcd + 1) / 8;
  __syncthreads();
  for (long long it = i0 + lid; it < i1; it += nper) {
    const int kx = kxb + (int)(it / npairs), p = (int)(it % npairs);
    const int2 pr = pairs[p];
    const float2* sa = spectra + (size_t)pr.x * spec_stride + (size_t)kx * n;
    const float2* sb = spectra + (size_t)pr.y * spec_stride + (size_t)kx * n;
    const float wx = wxy[kx];
    for (int k = threadIdx.x; k < M; k += NT) {
      float2 v = make_float2(0.f, 0.f);
      if (k < n) v = cmul(cscale(cmulc(sa[k], sb[k]), wx * wxy[H + k]), pre[k]);
      buf[lix<PB>(k)] = v;
    }
    __syncthreads();
    fft_lds_t<M, NT, -1, 1, M, TwSplit, PB>(buf, tw, M);
    for (int k = threadIdx.x; k < M; k += NT) {
      const int lk = lix<PB>(k);
      buf[lk] = cmul(buf[lk], filt[k]);
    }
    __syncthreads();
    fft_lds_t<M, NT, 1, 1, M, TwSplit, PB>(buf, tw, M);
    for (int u = threadIdx.x; u < box; u += NT)
      tbuf[((size_t)p * H + kx) * box + u] = cmul(buf[lix<PB>(n - 1 + u)], post[u]);
    __syncthreads();  // buf is rewritten by the next item
  }
}

// Sum-image C2R, column pass (any even n): inverse DFT of S[kx][:] of
// length n, stored as tmp[kx/4][ky][kx%4].
template <int M, int NT>
__global__ __launch_bounds__(NT, czt_waves<NT>()) void k_cols_inv_czt(const float2* spec, int n, float2* tmp,
                                                     const float2* __restrict__ gtw,
                                                     const float2* __restrict__ chirpv,
                                                     const float2* __restrict__ filt) {
  constexpr int PB = CZT_PB;
  constexpr int LB = tw_lbits(M);
  __shared__ float2 buf[lds_len<PB>(M)];
  __shared__ float2 tlo[1 << LB], thi[M >> LB];
  tw_split_load(tlo, thi, gtw, M, LB, threadIdx.x, NT);
  const int kx = xcd_remap(blockIdx.x, gridDim.x);
  const float2* src = spec + (size_t)kx * n;
  for (int i = threadIdx.x; i < n; i += NT) buf[lix<PB>(i)] = src[i];
  __syncthreads();
  bluestein<M, NT, 1>(buf, TwSplit{tlo, thi, LB, M}, chirpv, filt, n);
  float2* dst = tmp + (size_t)(kx >> 2) * n * 4 + (kx & 3);
  for (int i = threadIdx.x; i < n; i += NT) dst[(size_t)i * 4] = buf[lix<PB>(i)];
}

// Sum-image C2R, column pass at N = Q P (the 3710 and 4092 sizes): the
// inverse of k_c2c_cols_qp, X[k2 + P k1] = sum_a e^{2 pi i a k1/Q}
// (e^{2 pi i a k2/N} sum_b x[a + Q b] e^{2 pi i b k2/P}), on column kx of the
// spectrum, stored like k_cols_inv_czt (tmp[kx/4][y][kx%4]).
template <int Q, int P, int NT>
__global__ __launch_bounds__(NT, 2) void k_cols_inv_qp(const float2* __restrict__ spec, int nkx,
                                                       float2* __restrict__ tmp,
                                                       const float2* __restrict__ gtw,
                                                       const float2* __restrict__ twqp) {
  constexpr int N = Q * P;
  constexpr int PB = 0;
  constexpr int TS = lds_len<PB>(Q) | 1;
  static_assert(NT >= Q && NT % 64 == 0, "one lane per P-point DFT");
  __shared__ float2 buf[P * TS];
  __shared__ float2 twm[Q];
  const int kx = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  if (kx >= nkx) return;
  for (int i = threadIdx.x; i < Q; i += NT) twm[i] = gtw[i * P];  // e^{+2 pi i i/Q}
  const int a = threadIdx.x;
  if (a < Q) {
    const auto q = buf_rsrc(spec + (size_t)kx * N, N * 8);
    float2 x[P];
#pragma unroll
    for (int b = 0; b < P; b++) x[b] = buf_ld2(q, 8 * a, 8 * Q * b);
    odd_dft<P, 1>(x, [&](int k, float2 v) { buf[k * TS + lix<PB>(a)] = v; });
  }
  __syncthreads();
  // twiddles e^{+2 pi i a k2/N} (the table holds e^{-2 pi i a k2/N})
  for (int e = Q + threadIdx.x; e < N; e += NT) {
    const int k2 = e / Q, i = e - k2 * Q;
    float2* p = buf + k2 * TS + lix<PB>(i);
    *p = cmulc(*p, twqp[e]);
  }
  __syncthreads();
  fft_lds_t<Q, NT, 1, P, TS, TwTable, PB>(buf, TwTable{twm, Q}, Q);
  float2* dst = tmp + (size_t)(kx >> 2) * N * 4 + (kx & 3);
  for (int o = threadIdx.x; o < N; o += NT) dst[(size_t)o * 4] = buf[(o % P) * TS + lix<PB>(o / P)];
}

// Sum-image C2R, row pass (any even n): the n/2+1 bins of row y with the
// DC/Nyquist projection, packed into n/2 complex values, a Bluestein
// inverse DFT of length n/2, x *= scale.
template <int M, int NT>
__global__ __launch_bounds__(NT, czt_waves<NT>()) void k_c2r_rows_czt(const float2* tmp, int n, float scale,
                                                     float* img, size_t pitch,
                                                     const float2* __restrict__ gtw,
                                                     const float2* __restrict__ chirpv,
                                                     const float2* __restrict__ filt,
                                                     const float2* __restrict__ twn) {
  constexpr int PB = CZT_PB;
  constexpr int LB = tw_lbits(M);
  __shared__ float2 buf[lds_len<PB>(M)];
  __shared__ float2 tlo[1 << LB], thi[M >> LB];
  tw_split_load(tlo, thi, gtw, M, LB, threadIdx.x, NT);
  const int y = blockIdx.x;
  const int NH = n / 2;
  for (int k = threadIdx.x; k <= NH; k += NT) {
    float2 v = tmp[((size_t)(k >> 2) * n + y) * 4 + (k & 3)];
    if (k == 0 || k == NH) v.y = 0.f;
    buf[lix<PB>(k)] = v;
  }
  __syncthreads();
  // Z[k] = (X[k] + conj X[NH-k]) + i e^{+2 pi i k/n} (X[k] - conj X[NH-k]),
  // one thread per pair (k, NH-k)
  for (int k = threadIdx.x; k <= NH / 2; k += NT) {
    const int lk = lix<PB>(k), lm = lix<PB>(NH - k);
    const float2 a = buf[lk], b = buf[lm];
    const float2 za = cadd(cadd(a, cconj(b)), mul_si<1>(cmul(twn[k], csub(a, cconj(b)))));
    const float2 zb = cadd(cadd(b, cconj(a)), mul_si<1>(cmul(twn[NH - k], csub(b, cconj(a)))));
    buf[lk] = za;
    if (k != 0 && k != NH - k) buf[lm] = zb;
  }
  __syncthreads();
  bluestein<M, NT, 1>(buf, TwSplit{tlo, thi, LB, M}, chirpv, filt, NH);
  float2* row = reinterpret_cast<float2*>(img + (size_t)y * pitch);
  for (int m = threadIdx.x; m < NH; m += NT) {
    const float2 z = buf[lix<PB>(m)];
    row[m] = make_float2(z.x * scale, z.y * scale);
  }
}

// ---------------------------------------------------------------- launchers
#define MC_CZT_SIZES(X) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192) X(16384)

template <int M>
static hipError_t rows_czt_t(const R2CGeom& g, int f0, int nframes, int nkx, float2* rt,
                             size_t rt_fstride, float* rs, const BsTab& t, const float2* twnu,
                             hipStream_t st) {
  constexpr int NT = czt_nt<M>();
  const float2* gtw = twiddles(M);
  if (!gtw) return hipErrorOutOfMemory;
  if constexpr (M == 4096 || M == 8192) {
    // persistent, register twiddles (packed rows of 2050-8192 pixels)
    const int nitems = g.h * nframes;
    switch (g.dtype) {
#define MC_RCP(D)                                                                                  \
  case D: {                                                                                        \
    static const int grid = persistent_grid(k_r2c_rows_czt_p<M, NT, D>, NT);                      \
    const int gr = std::min(grid, std::max(8, (nitems + 7) / 8 * 8));                             \
    hipLaunchKernelGGL((k_r2c_rows_czt_p<M, NT, D>), dim3(gr), dim3(NT), 0, st, g, f0, nframes,   \
                       nkx, rt, rt_fstride, rs, gtw, t.chirp, t.filt, twnu);                      \
    break;                                                                                         \
  }
      MC_RCP(DT_INT8) MC_RCP(DT_UINT8) MC_RCP(DT_INT16) MC_RCP(DT_UINT16) MC_RCP(DT_FLOAT32)
#undef MC_RCP
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  dim3 gr(g.h, nframes);
  switch (g.dtype) {
#define MC_RC(D) case D: hipLaunchKernelGGL((k_r2c_rows_czt<M, NT, D>), gr, dim3(NT), 0, st, g, f0, nkx, rt, rt_fstride, rs, gtw, t.chirp, t.filt, twnu); break;
    MC_RC(DT_INT8) MC_RC(DT_UINT8) MC_RC(DT_INT16) MC_RC(DT_UINT16) MC_RC(DT_FLOAT32)
#undef MC_RC
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int M>
static hipError_t cols_czt_t(const R2CGeom& g, int nframes, int nkx, const float2* rt,
                             size_t rt_fstride, const float* fill, const float2* padP,
                             float2* spectra, size_t spec_stride, const BsTab& t, hipStream_t st) {
  constexpr int NT = czt_nt<M>();
  const float2* gtw = twiddles(M);
  if (!gtw) return hipErrorOutOfMemory;
  if constexpr (M == 8192) {
    // persistent, register twiddles (two workgroups per CU: the LDS limit)
    const int nitems = nkx * nframes;
    static const int grid = persistent_grid(k_c2c_cols_czt_p<M, NT>, NT);
    const int gr = std::min(grid, std::max(8, (nitems + 7) / 8 * 8));
    hipLaunchKernelGGL((k_c2c_cols_czt_p<M, NT>), dim3(gr), dim3(NT), 0, st, rt, rt_fstride,
                       nkx, nitems, g.nu, g.h, fill, padP, g.n, spectra, spec_stride, gtw,
                       t.chirp, t.filt);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_c2c_cols_czt<M, NT>), dim3(nkx * nframes), dim3(NT), 0, st, rt,
                     rt_fstride, nkx, g.nu, g.h, fill, padP, g.n, spectra, spec_stride, gtw,
                     t.chirp, t.filt);
  return hipGetLastError();
}

template <int Q, int P>
static hipError_t rows_qp_t(const R2CGeom& g, int f0, int nframes, int nkx, float2* rt,
                            size_t rt_fstride, float* rs, const float2* twnu, hipStream_t st) {
  constexpr int R = 1, NT = (Q + 63) / 64 * 64;
  const float2* gtw = twiddles(Q * P);
  const float2* twqp = qp_twiddles(Q, P);
  if (!gtw || !twqp) return hipErrorOutOfMemory;
  // pixel pairs load as one element when the pairs are aligned
  const size_t es = g.dtype == DT_FLOAT32 ? 4 : (g.dtype == DT_INT16 || g.dtype == DT_UINT16) ? 2 : 1;
  R2CGeom gq = g;
  gq.vec = ((uintptr_t)g.src % (2 * es) == 0 && g.row_pitch % (2 * es) == 0 &&
            g.frame_pitch % (2 * es) == 0 && (g.offx & 1) == 0) ? 1 : 0;
  const int grid = ((g.h * nframes + R - 1) / R + 7) / 8 * 8;
  switch (g.dtype) {
#define MC_RQ(D) case D: hipLaunchKernelGGL((k_r2c_rows_qp<Q, P, R, NT, D>), dim3(grid), dim3(NT), 0, st, gq, f0, nframes, nkx, rt, rt_fstride, rs, gtw, twqp, twnu); break;
    MC_RQ(DT_INT8) MC_RQ(DT_UINT8) MC_RQ(DT_INT16) MC_RQ(DT_UINT16) MC_RQ(DT_FLOAT32)
#undef MC_RQ
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int Q, int P>
static hipError_t cols_qp_t(int nframes, int nkx, const float2* rt, size_t rt_fstride, int n,
                            float2* spectra, size_t spec_stride, hipStream_t st) {
  constexpr int NT = (Q + 63) / 64 * 64, W = 2;
  const float2* gtw = twiddles(Q * P);
  const float2* twqp = qp_twiddles(Q, P);
  if (!gtw || !twqp) return hipErrorOutOfMemory;
  const int grid = (nkx * nframes + 7) / 8 * 8;
  hipLaunchKernelGGL((k_c2c_cols_qp<Q, P, NT, W>), dim3(grid), dim3(NT), 0, st, rt, rt_fstride, nkx,
                     nkx * nframes, n, spectra, spec_stride, gtw, twqp);
  return hipGetLastError();
}

hipError_t launch_frames_to_spectra_czt(const R2CGeom& g, int f0, int nframes, float2* spectra,
                                        size_t spec_stride, float2* rt, float* rowsum,
                                        float* fill, const float2* padP, hipStream_t st) {
  if (g.nu < 64 || g.nu > 8192 || g.nu % 2 || g.n > g.nu || g.n % 2) return hipErrorInvalidValue;
  const int nkx = g.n / 2 + 1;
  const size_t rt_fstride = rt_frame_elems(g.n, g.nu);
  float* rs = g.pad_mode ? rowsum : nullptr;
  R2CGeom gv = g;
  gv.vec = 0;
  const BsTab* tr = bs_tables(g.nu / 2, -1);
  const BsTab* tc = bs_tables(g.nu, -1);
  const float2* twnu = twiddles(g.nu);
  if (!tr || !tc || !twnu) return hipErrorOutOfMemory;
  hipError_t e = hipErrorInvalidValue;
  // square canvases of 3710 (K2 crop) and 4092 (K3 crop) without fill:
  // two-level transforms, 35 x 53 (66 x 31) for the packed rows and 70 x 53
  // (132 x 31) for the columns, any binning
  const int qp = (!g.pad_mode && g.h == g.nu && g.w == g.nu && (g.nu == 3710 || g.nu == 4092))
                     ? g.nu : 0;
  if (qp == 3710) {
    e = rows_qp_t<35, 53>(gv, f0, nframes, nkx, rt, rt_fstride, rs, twnu, st);
  } else if (qp == 4092) {
    e = rows_qp_t<66, 31>(gv, f0, nframes, nkx, rt, rt_fstride, rs, twnu, st);
  } else {
    switch (tr->M) {
#define MC_R(MM) case MM: e = rows_czt_t<MM>(gv, f0, nframes, nkx, rt, rt_fstride, rs, *tr, twnu, st); break;
      MC_CZT_SIZES(MC_R)
#undef MC_R
    }
  }
  if (e != hipSuccess) return e;
  if (g.pad_mode) {
    e = launch_fill(rowsum, g.nu, g.h, g.w, fill, nframes, st);
    if (e != hipSuccess) return e;
  }
  e = hipErrorInvalidValue;
  const float* fl = g.pad_mode ? fill : nullptr;
  if (qp == 3710) {
    e = cols_qp_t<70, 53>(nframes, nkx, rt, rt_fstride, g.n, spectra, spec_stride, st);
  } else if (qp == 4092) {
    e = cols_qp_t<132, 31>(nframes, nkx, rt, rt_fstride, g.n, spectra, spec_stride, st);
  } else {
    switch (tc->M) {
#define MC_C(MM) case MM: e = cols_czt_t<MM>(g, nframes, nkx, rt, rt_fstride, fl, padP, spectra, spec_stride, *tc, st); break;
      MC_CZT_SIZES(MC_C)
#undef MC_C
    }
  }
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

template <int M>
static hipError_t pair_cols_czt_t(const float2* spectra, size_t spec_stride, const int2* pairs,
                                  int npairs, const float* wxy, int n, int box, float2* tbuf,
                                  const CztTab& t, hipStream_t st) {
  constexpr int NT = czt_nt<M>();
  const float2* gtw = twiddles(M);
  if (!gtw) return hipErrorOutOfMemory;
  const int grid = persistent_grid(k_pair_cols_czt<M, NT>, NT);
  const int H = n / 2 + 1;
  hipLaunchKernelGGL((k_pair_cols_czt<M, NT>), dim3(grid), dim3(NT), 0, st, spectra, spec_stride,
                     pairs, npairs, wxy, n, box, tbuf, gtw, t.pre, t.filt, t.post, 0, H);
  return hipGetLastError();
}

hipError_t launch_pair_cols_czt(const float2* spectra, size_t spec_stride, const int2* pairs,
                                int npairs, const float* wxy, int n, int box, float2* tbuf,
                                hipStream_t st) {
  if (n < 64 || n > 8192 || n % 2 || box < 1 || box > n) return hipErrorInvalidValue;
  const CztTab* t = czt_tables(n, box);
  if (!t) return hipErrorOutOfMemory;
  switch (t->M) {
#define MC_P(MM) case MM: return pair_cols_czt_t<MM>(spectra, spec_stride, pairs, npairs, wxy, n, box, tbuf, *t, st);
    MC_CZT_SIZES(MC_P)
#undef MC_P
  }
  return hipErrorInvalidValue;
}

template <int M>
static hipError_t cols_inv_czt_t(const float2* spec, int n, float2* tmp, const BsTab& t,
                                 hipStream_t st) {
  constexpr int NT = czt_nt<M>();
  const float2* gtw = twiddles(M);
  if (!gtw) return hipErrorOutOfMemory;
  hipLaunchKernelGGL((k_cols_inv_czt<M, NT>), dim3(n / 2 + 1), dim3(NT), 0, st, spec, n, tmp, gtw,
                     t.chirp, t.filt);
  return hipGetLastError();
}

template <int M>
static hipError_t c2r_rows_czt_t(const float2* tmp, int n, float scale, float* img, size_t pitch,
                                 const BsTab& t, const float2* twn, hipStream_t st) {
  constexpr int NT = czt_nt<M>();
  const float2* gtw = twiddles(M);
  if (!gtw) return hipErrorOutOfMemory;
  hipLaunchKernelGGL((k_c2r_rows_czt<M, NT>), dim3(n), dim3(NT), 0, st, tmp, n, scale, img, pitch,
                     gtw, t.chirp, t.filt, twn);
  return hipGetLastError();
}

template <int Q, int P>
static hipError_t cols_inv_qp_t(const float2* spec, float2* tmp, hipStream_t st) {
  constexpr int NT = (Q + 63) / 64 * 64;
  const float2* gtw = twiddles(Q * P);
  const float2* twqp = qp_twiddles(Q, P);
  if (!gtw || !twqp) return hipErrorOutOfMemory;
  const int nkx = Q * P / 2 + 1;
  hipLaunchKernelGGL((k_cols_inv_qp<Q, P, NT>), dim3((nkx + 7) / 8 * 8), dim3(NT), 0, st, spec,
                     nkx, tmp, gtw, twqp);
  return hipGetLastError();
}

hipError_t launch_spectrum_to_image_czt(const float2* spec, int n, float scale, float* img,
                                        size_t img_pitch, float2* tmp, hipStream_t st) {
  if (n < 64 || n > 8192 || n % 2) return hipErrorInvalidValue;
  const BsTab* tc = bs_tables(n, 1);
  const BsTab* tr = bs_tables(n / 2, 1);
  const float2* twn = twiddles(n);
  if (!tc || !tr || !twn) return hipErrorOutOfMemory;
  hipError_t e = hipErrorInvalidValue;
  if (n == 3710) {
    e = cols_inv_qp_t<70, 53>(spec, tmp, st);
  } else if (n == 4092) {
    e = cols_inv_qp_t<132, 31>(spec, tmp, st);
  } else {
    switch (tc->M) {
#define MC_CI(MM) case MM: e = cols_inv_czt_t<MM>(spec, n, tmp, *tc, st); break;
      MC_CZT_SIZES(MC_CI)
#undef MC_CI
    }
  }
  if (e != hipSuccess) return e;
  e = hipErrorInvalidValue;
  switch (tr->M) {
#define MC_CR(MM) case MM: e = c2r_rows_czt_t<MM>(tmp, n, scale, img, img_pitch, *tr, twn, st); break;
    MC_CZT_SIZES(MC_CR)
#undef MC_CR
  }
  return e;
}

}  // namespace mc
