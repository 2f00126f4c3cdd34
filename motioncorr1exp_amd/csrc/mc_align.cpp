// mc_align.cpp -- host side of libmc: the C ABI (include/mc_abi.h), the
// resident movie pipeline and the FP64 drift model.
//
// Control flow follows CDFAlign::ThreadFunc_cuAlign (reference
// src-1125/DFAlign.cu:456-1355) with the device work on one HIP stream and
// every spectrum resident in HBM: stage 1 (DFAlign.cu:705-824), pairwise
// shifts (:875-905), least squares with outlier rejection (:914-948),
// iterative align-to-sum (:950-1071) and the final phase-shift-and-sum
// (:1114-1292).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mc_abi.h"
#include "mc_kernels.h"

using mc::R2CGeom;

// ------------------------------------------------------------- errors
static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_err(e_ == hipErrorOutOfMemory ? MC_ERR_NOMEM : MC_ERR_DEVICE, "%s: %s (%s:%d)", \
                     #expr, hipGetErrorString(e_), __FILE__, __LINE__);                     \
  } while (0)

#define MC_TRY(expr)        \
  do {                      \
    int r_ = (expr);        \
    if (r_ != MC_OK) return r_; \
  } while (0)

int mc_internal_set_error(int code, const char* msg) {
  g_err = msg;
  return code;
}

extern "C" const char* mc_last_error(void) { return g_err.c_str(); }
extern "C" const char* mc_version(void) { return "motioncorr1exp_amd 0.1 (gfx950)"; }

// Transform sizes: any even size in [64, 8192], as cuFFT plans any size
// (cufunc.cu:202-215).  Powers of two and 2^k * {3, 5, 15} from 768 have
// Stockham plans; the others run as chirp-z transforms.
static bool fft_size_ok(int n) { return n >= 64 && n <= 8192 && n % 2 == 0; }

// ------------------------------------------------------- device buffers
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) cap = bytes;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

struct KTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

struct mc_context_s {
  int device = 0;
  hipStream_t st = nullptr;
  DevBuf spec, rt, rowsum, fill, padP, pairs_d, tbuf, box, boxpart, peaks, shifts, image, tmp, means,
      sum2img, wk, gpart, gout, jobs, tmax, fsc, disp, plans;
  int graph_captures = 0, graph_replays = 0;
  // autotuned launch shapes (mc_autotune / mc_tune_load)
  struct Tune {
    int nu, n, dtype, box, precision;
    int r2c, chunk, cols, rows;
  };
  std::vector<Tune> tunes;
  // resident spectra
  int n = 0, nframe = 0, nslots = 0, off2 = 0, sum_slot = 0;
  size_t spec_stride = 0;
  size_t spec_off = 0;  // float2 offset of slot 0: the current size group's region in a batch
  int padP_nu = -1, padP_w = -1, padP_nkx = -1;
  // kernel timing
  unsigned kt_mask = 0;  // bit g: kernel group g is timed
  KTimer kt[MC_KT_COUNT];
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  // stage boundaries of a group alignment (created on first use, reused)
  hipEvent_t gev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

  hipEvent_t ev() {
    if (pool_used < pool.size()) return pool[pool_used++];
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    pool.push_back(e);
    pool_used++;
    return e;
  }
};

// Wraps one kernel group with events when timing is enabled.
struct KScope {
  mc_context_s* c;
  int g;
  hipEvent_t a = nullptr;
  KScope(mc_context_s* c_, int g_) : c(c_), g(g_) {
    if ((c->kt_mask >> g) & 1u) {
      a = c->ev();
      if (a) (void)hipEventRecord(a, c->st);
    }
  }
  ~KScope() {
    if (a) {
      hipEvent_t b = c->ev();
      if (b) {
        (void)hipEventRecord(b, c->st);
        c->kt[g].ev.push_back({a, b});
      }
    }
  }
};

extern "C" int mc_context_create(int device, mc_context* ctx) {
  if (!ctx) return set_err(MC_ERR_ARG, "ctx is NULL");
  int nd = 0;
  HIP_TRY(hipGetDeviceCount(&nd));
  if (device < 0 || device >= nd)
    return set_err(MC_ERR_ARG, "GPU ID %d is out of range(%d).", device, nd);
  HIP_TRY(hipSetDevice(device));
  mc_context_s* c = new mc_context_s;
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_err(MC_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  *ctx = c;
  return MC_OK;
}

extern "C" int mc_context_destroy(mc_context ctx) {
  if (!ctx) return MC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->st);
  DevBuf* bufs[] = {&ctx->spec, &ctx->rt, &ctx->rowsum, &ctx->fill, &ctx->padP,
                    &ctx->pairs_d, &ctx->tbuf, &ctx->box, &ctx->boxpart, &ctx->peaks,
                    &ctx->shifts, &ctx->image, &ctx->tmp, &ctx->means, &ctx->sum2img,
                    &ctx->wk, &ctx->gpart, &ctx->gout, &ctx->jobs, &ctx->tmax, &ctx->fsc,
                    &ctx->disp};
  for (DevBuf* b : bufs) b->release();
  for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->gev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->st);
  delete ctx;
  return MC_OK;
}

extern "C" int mc_kernel_timing_groups(mc_context ctx, unsigned mask) {
  if (!ctx) return set_err(MC_ERR_ARG, "ctx is NULL");
  if (mask >> MC_KT_COUNT) return set_err(MC_ERR_ARG, "group mask 0x%x has bits past group %d", mask, MC_KT_COUNT - 1);
  HIP_TRY(hipStreamSynchronize(ctx->st));
  for (auto& k : ctx->kt) k.ev.clear();
  ctx->pool_used = 0;
  ctx->kt_mask = mask;
  return MC_OK;
}

extern "C" int mc_kernel_timing(mc_context ctx, int enable) {
  return mc_kernel_timing_groups(ctx, enable ? (1u << MC_KT_COUNT) - 1 : 0u);
}

extern "C" int mc_kernel_timing_get(mc_context ctx, int group, int* count, double* total_ms) {
  if (!ctx || group < 0 || group >= MC_KT_COUNT) return set_err(MC_ERR_ARG, "bad group");
  HIP_TRY(hipStreamSynchronize(ctx->st));
  double t = 0;
  for (auto& pr : ctx->kt[group].ev) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
    t += ms;
  }
  if (count) *count = (int)ctx->kt[group].ev.size();
  if (total_ms) *total_ms = t;
  return MC_OK;
}

// ============================================================ geometry
// func.cpp:80-103
static int verify_crop_size(int nx, int ny, int offx, int offy, int nsamout, float bin) {
  if (offx < 0 || offx >= nx || offy < 0 || offy >= ny || bin <= 0) return 0;
  if (nsamout <= 0) nsamout = std::min(nx - offx, ny - offy);
  if (nx - offx < nsamout) nsamout = nx - offx;
  if (ny - offy < nsamout) nsamout = ny - offy;
  if (std::fabs(bin - 1.0) < 0.0001) nsamout = nsamout / 2 * 2;
  else if (std::fabs(bin - 2.0) < 0.0001) nsamout = nsamout / 4 * 4;
  else nsamout = nsamout / 2 * 2;
  return nsamout;
}

// DFAlign.cu:515
static int binned_size(int nsu, float bin) { return (int)((float)nsu / bin) / 2 * 2; }

static size_t dtype_size(int dt) {
  switch (dt) {
    case MC_DTYPE_INT8:
    case MC_DTYPE_UINT8: return 1;
    case MC_DTYPE_INT16:
    case MC_DTYPE_UINT16: return 2;
    case MC_DTYPE_FLOAT32: return 4;
    default: return 0;
  }
}

struct Geom {
  int nsu, n, w, h, pad_mode;
  size_t row_pitch, frame_pitch;
};

static int movie_geometry(const mc_movie_desc* m, const mc_params* p, Geom* g) {
  const size_t es = dtype_size(m->dtype);
  if (!es) return set_err(MC_ERR_ARG, "unsupported dtype %d", m->dtype);
  if (m->nx <= 0 || m->ny <= 0 || m->nz <= 0) return set_err(MC_ERR_ARG, "empty stack");
  g->row_pitch = m->row_pitch ? m->row_pitch : (size_t)m->nx * es;
  g->frame_pitch = m->frame_pitch ? m->frame_pitch : g->row_pitch * m->ny;
  if (p->bin <= 0) return set_err(MC_ERR_ARG, "bin must be > 0");
  if (p->pad > 0) {
    if (p->crop_offsetx < 0 || p->crop_offsetx >= m->nx || p->crop_offsety < 0 ||
        p->crop_offsety >= m->ny)
      return set_err(MC_ERR_ARG, "Error: Wrong image Size.");
    g->nsu = p->pad;
    g->w = std::min(m->nx - p->crop_offsetx, p->pad);
    g->h = std::min(m->ny - p->crop_offsety, p->pad);
    g->pad_mode = (g->w < p->pad || g->h < p->pad) ? 1 : 0;
  } else {
    g->nsu = verify_crop_size(m->nx, m->ny, p->crop_offsetx, p->crop_offsety, p->crop_nsam,
                              p->bin);
    if (g->nsu <= 0) return set_err(MC_ERR_ARG, "Error: Wrong image Size.");
    g->w = g->h = g->nsu;
    g->pad_mode = 0;
  }
  g->n = binned_size(g->nsu, p->bin);
  if (!fft_size_ok(g->nsu) || !fft_size_ok(g->n))
    return set_err(MC_ERR_ARG,
                   "transform size %d (binned %d) unsupported: sizes must be powers of two "
                   "in [64, 8192] or 768, 1280, 1536, 1920, 2560, 3072, 3840, 5120, 6144; "
                   "use the pad geometry",
                   g->nsu, g->n);
  return MC_OK;
}

// ========================================================= drift model
// mfunc.h:50-84
extern "C" int mc_pair_list(int nframe, int fod, int* pairs, int cap) {
  int dist = nframe / 2 - fod;
  if (dist <= 0) dist = 1;
  if (dist > nframe / 2) dist = nframe / 2;
  int c = 0;
  for (int i = 0; i < nframe; i++)
    for (int j = i + dist; j < nframe; j++) {
      if (pairs && c < cap) {
        pairs[2 * c] = i;
        pairs[2 * c + 1] = j;
      }
      c++;
    }
  return c;
}

// Least squares of the pair design matrix (A[p][k] = 1 for i <= k < j,
// mfunc.h:76-81) by the FP64 normal equations A^T A s = A^T b, Cholesky.
extern "C" int mc_lsq(const int* pairs, int np, int nframe, const double* b, double* s) {
  const int K = nframe - 1;
  if (K <= 0 || np < K) return set_err(MC_ERR_ARG, "least squares needs >= %d rows", K);
  std::vector<double> N((size_t)K * K, 0.0), r((size_t)2 * K, 0.0);
  // A^T A adds 1 over the block [i, j) x [i, j) of each pair: built from a
  // (K+1)^2 difference array (four corner updates per pair, then 2-D prefix
  // sums), O(np + K^2).  The entries are small integers, exact in any order.
  std::vector<int> D((size_t)(K + 1) * (K + 1), 0);
  for (int q = 0; q < np; q++) {
    const int i = pairs[2 * q], j = pairs[2 * q + 1];
    if (i < 0 || j > K) return set_err(MC_ERR_ARG, "pair %d (%d, %d) out of range", q, i, j);
    if (i >= j) continue;  // an empty increment range adds nothing
    D[(size_t)i * (K + 1) + i] += 1;
    D[(size_t)i * (K + 1) + j] -= 1;
    D[(size_t)j * (K + 1) + i] -= 1;
    D[(size_t)j * (K + 1) + j] += 1;
    for (int k = i; k < j; k++) {
      r[2 * k] += b[2 * q];
      r[2 * k + 1] += b[2 * q + 1];
    }
  }
  for (int k = 0; k < K; k++) {
    int row = 0;
    for (int l = 0; l < K; l++) {
      row += D[(size_t)k * (K + 1) + l];
      const int up = k > 0 ? (int)N[(size_t)(k - 1) * K + l] : 0;
      N[(size_t)k * K + l] = (double)(up + row);
    }
  }
  // Cholesky N = L L^T (lower, in place)
  for (int k = 0; k < K; k++) {
    double d = N[(size_t)k * K + k];
    for (int m = 0; m < k; m++) d -= N[(size_t)k * K + m] * N[(size_t)k * K + m];
    if (!(d > 1e-12)) return set_err(MC_ERR_SOLVE, "singular drift system (column %d)", k);
    d = std::sqrt(d);
    N[(size_t)k * K + k] = d;
    for (int i = k + 1; i < K; i++) {
      double v = N[(size_t)i * K + k];
      for (int m = 0; m < k; m++) v -= N[(size_t)i * K + m] * N[(size_t)k * K + m];
      N[(size_t)i * K + k] = v / d;
    }
  }
  for (int c = 0; c < 2; c++) {
    std::vector<double> y(K);
    for (int i = 0; i < K; i++) {
      double v = r[2 * i + c];
      for (int m = 0; m < i; m++) v -= N[(size_t)i * K + m] * y[m];
      y[i] = v / N[(size_t)i * K + i];
    }
    for (int i = K - 1; i >= 0; i--) {
      double v = y[i];
      for (int m = i + 1; m < K; m++) v -= N[(size_t)m * K + i] * s[2 * m + c];
      s[2 * i + c] = v / N[(size_t)i * K + i];
    }
  }
  return MC_OK;
}

static void residuals(const int* pairs, int np, const double* s, const double* b, double* ki) {
  for (int q = 0; q < np; q++) {
    double x = -b[2 * q], y = -b[2 * q + 1];
    for (int k = pairs[2 * q]; k < pairs[2 * q + 1]; k++) {
      x += s[2 * k];
      y += s[2 * k + 1];
    }
    ki[q] = std::sqrt(x * x + y * y);
  }
}

// rank(A restricted to rows) == nframe-1  <=>  the pair graph is connected
static bool connected(const int* pairs, const int* sel, int nsel, int nframe) {
  std::vector<int> par(nframe);
  for (int i = 0; i < nframe; i++) par[i] = i;
  auto find = [&](int a) {
    while (par[a] != a) a = par[a] = par[par[a]];
    return a;
  };
  int comps = nframe;
  for (int q = 0; q < nsel; q++) {
    int a = find(pairs[2 * sel[q]]), b = find(pairs[2 * sel[q] + 1]);
    if (a != b) {
      par[a] = b;
      comps--;
    }
  }
  return comps == 1;
}

// OD_Threshold (mfunc.h:101-146)
extern "C" int mc_threshold(const int* pairs, int np, int nframe, const double* ki,
                            double kithresh, int* good) {
  const int K = nframe - 1;
  std::vector<int> order(np);
  for (int i = 0; i < np; i++) order[i] = i;
  // ascending ki, ties by lower row index (repeated first-minimum selection)
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ki[a] < ki[b]; });
  int i;
  for (i = np - 1; i >= K; i--)
    if (ki[order[i]] < kithresh) break;
  int ii = i, nsel;
  do {
    nsel = ii + 1;
    for (int q = 0; q < nsel; q++) good[q] = order[q];
    std::sort(good, good + nsel);
    ii++;
    if (connected(pairs, good, nsel, nframe)) break;
  } while (ii < np);
  return nsel;
}

// ========================================================== resident state
static int ensure_padP(mc_context_s* c, int nu, int w, int nkx) {
  if (c->padP_nu == nu && c->padP_w == w && c->padP_nkx == nkx) return MC_OK;
  // P[kx] = sum_{x=w}^{nu-1} e^{-2 pi i kx x/nu}: spectrum of the fill mask row
  std::vector<float2> h(nkx);
  for (int k = 0; k < nkx; k++) {
    double re = 0, im = 0;
    if (k == 0) {
      re = nu - w;
    } else {
      // (e^{-i t w} - e^{-i t nu}) / (1 - e^{-i t}), t = 2 pi k / nu
      const double t = 2.0 * M_PI * k / nu;
      const double ar = std::cos(t * w) - 1.0, ai = -std::sin(t * w);
      const double br = 1.0 - std::cos(t), bi = std::sin(t);
      const double den = br * br + bi * bi;
      re = (ar * br + ai * bi) / den;
      im = (ai * br - ar * bi) / den;
    }
    h[k] = make_float2((float)re, (float)im);
  }
  HIP_TRY(c->padP.ensure(sizeof(float2) * nkx));
  HIP_TRY(hipMemcpyAsync(c->padP.p, h.data(), sizeof(float2) * nkx, hipMemcpyHostToDevice, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  c->padP_nu = nu;
  c->padP_w = w;
  c->padP_nkx = nkx;
  return MC_OK;
}

static int ensure_spectra(mc_context_s* c, int n, int nframe, int two) {
  const size_t stride = (size_t)(n / 2 + 1) * n;
  const int nslots = nframe * (two ? 2 : 1) + 1;
  HIP_TRY(c->spec.ensure(sizeof(float2) * stride * nslots));
  c->n = n;
  c->nframe = nframe;
  c->nslots = nslots;
  c->off2 = two ? nframe : 0;
  c->sum_slot = nslots - 1;
  c->spec_stride = stride;
  c->spec_off = 0;
  return MC_OK;
}

static float2* slot(mc_context_s* c, int s) {
  return c->spec.as<float2>() + c->spec_off + c->spec_stride * s;
}

static const mc_context_s::Tune* stage1_tune(const mc_context_s* c, int nu, int n, int dtype) {
  for (const auto& t : c->tunes)
    if (t.nu == nu && t.n == n && t.dtype == dtype) return &t;
  return nullptr;
}

static const mc_context_s::Tune* pair_tune(const mc_context_s* c, int n, int box, int precision) {
  for (const auto& t : c->tunes)
    if (t.n == n && t.box == box && t.precision == precision) return &t;
  return nullptr;
}

// Stage 1 for frames [first, first+count) of a buffer into slots [s0, ...).
// variant / chunk < 0: the tuned shape for this geometry, else the default.
static int run_stage1(mc_context_s* c, const void* frames, const mc_movie_desc* m,
                      const mc_params* p, const Geom& g, int first, int count, int s0,
                      int variant = -1, int chunk = -1, int section = 1) {
  const int nkx = g.n / 2 + 1;
  const size_t rt_per = mc::rt_frame_elems(g.n, g.nsu);
  const mc_context_s::Tune* tn = stage1_tune(c, g.nsu, g.n, m->dtype);
  if (variant < 0) variant = tn ? tn->r2c : mc::R2C_AUTO;
  if (chunk <= 0 && tn && tn->chunk > 0) chunk = tn->chunk;
  if (chunk <= 0) {
    // Frames per launch pair: the row-pass scratch of a chunk is sized to
    // stay within the 256 MB Infinity Cache between the row and column passes.
    static const size_t chunk_mb = getenv("MC_STAGE1_CHUNK_MB") ? atoi(getenv("MC_STAGE1_CHUNK_MB")) : 192;
    chunk = (int)std::max<size_t>(1, (chunk_mb << 20) / (rt_per * sizeof(float2)));
  }
  chunk = std::min(chunk, count);
  HIP_TRY(c->rt.ensure(sizeof(float2) * rt_per * chunk));
  HIP_TRY(c->rowsum.ensure(sizeof(float) * (size_t)g.nsu * chunk));
  HIP_TRY(c->fill.ensure(sizeof(float) * chunk));
  if (g.pad_mode) MC_TRY(ensure_padP(c, g.nsu, g.w, nkx));
  R2CGeom rg;
  rg.src = static_cast<const char*>(frames) + (size_t)first * g.frame_pitch;
  rg.dtype = m->dtype;
  rg.row_pitch = g.row_pitch;
  rg.frame_pitch = g.frame_pitch;
  rg.offx = p->crop_offsetx;
  rg.offy = p->crop_offsety;
  rg.w = g.w;
  rg.h = g.h;
  rg.nu = g.nsu;
  rg.n = g.n;
  rg.pad_mode = g.pad_mode;
  if (m->corr) {
    rg.dark = section == 1 ? m->corr->dark : m->corr->dark2;
    rg.gain = section == 1 ? m->corr->gain : m->corr->gain2;
    rg.xs = m->corr->xscale;
    rg.ys = m->corr->yscale;
    rg.nx = m->nx;
  }
  for (int f = 0; f < count; f += chunk) {
    const int cnt = std::min(chunk, count - f);
    HIP_TRY(mc::launch_frames_to_spectra(rg, f, cnt, slot(c, s0 + f), c->spec_stride,
                                         c->rt.as<float2>(), c->rowsum.as<float>(),
                                         c->fill.as<float>(), c->padP.as<float2>(), variant,
                                         c->st));
  }
  return MC_OK;
}

static int check_corrections(const mc_movie_desc* m) {
  if (!m->corr) return MC_OK;
  const mc_corrections* k = m->corr;
  if (m->frames2) return set_err(MC_ERR_ARG, "frames2 and device corrections are exclusive");
  if (!(k->xscale > 0.f && k->xscale <= 1.f && k->yscale > 0.f && k->yscale <= 1.f))
    return set_err(MC_ERR_ARG, "stretch factors must be in (0, 1] (1/-scx, 1/-scy)");
  return MC_OK;
}

static int load_frames(mc_context_s* c, const mc_movie_desc* m, const mc_params* p, Geom* g,
                       int* nframe_out) {
  MC_TRY(movie_geometry(m, p, g));
  const int nframe = p->nEnd - p->nStart + 1;
  if (p->nStart < m->first_frame || p->nEnd >= m->first_frame + m->nz || nframe < 2)
    return set_err(MC_ERR_ARG, "frame range [%d, %d] invalid for frames [%d, %d)", p->nStart,
                   p->nEnd, m->first_frame, m->first_frame + m->nz);
  if (!m->frames) return set_err(MC_ERR_ARG, "frames is NULL");
  MC_TRY(check_corrections(m));
  const bool two = m->frames2 != nullptr || (m->corr && m->corr->second);
  MC_TRY(ensure_spectra(c, g->n, nframe, two));
  {
    KScope ks(c, MC_KT_SPECTRA);
    const int first = p->nStart - m->first_frame;
    MC_TRY(run_stage1(c, m->frames, m, p, *g, first, nframe, 0));
    if (m->frames2) MC_TRY(run_stage1(c, m->frames2, m, p, *g, first, nframe, c->off2));
    else if (two) MC_TRY(run_stage1(c, m->frames, m, p, *g, first, nframe, c->off2, -1, -1, 2));
  }
  *nframe_out = nframe;
  return MC_OK;
}

// Pair-stage buffers for np pairs (before any launch, outside graph capture).
static int pair_reserve(mc_context_s* c, int np, int n, int box, int precision) {
  const int H = n / 2 + 1;
  HIP_TRY(c->pairs_d.ensure(sizeof(int) * 2 * (size_t)np));
  HIP_TRY(c->tbuf.ensure(sizeof(float2) * (size_t)np * H * box));
  if (precision == MC_PREC_MFMA_HILO) HIP_TRY(c->tmax.ensure(sizeof(float) * (size_t)np * H));
  HIP_TRY(c->box.ensure(sizeof(float) * (size_t)np * box * box));
  HIP_TRY(c->boxpart.ensure(sizeof(float) * std::max<size_t>(1, mc::pair_rows_part_elems(np, n, box))));
  HIP_TRY(c->peaks.ensure(sizeof(double) * mc::find_peaks_out_elems(np)));
  return MC_OK;
}

// Column pass, row pass and peak search over the np slot pairs in pairs_d
// (one movie's region); results stay on the device (peaks, box maps).
// Capturable.
static int pair_launch(mc_context_s* c, int np, int n, float bfactor, int box, int wnoise,
                       int precision, int cols = -1, int rows = -1) {
  const mc_context_s::Tune* tn = pair_tune(c, n, box, precision);
  if (cols < 0) cols = tn ? tn->cols : mc::PAIR_COLS_AUTO;
  if (rows < 0) rows = tn ? tn->rows : mc::PAIR_ROWS_AUTO;
  {
    KScope ks(c, MC_KT_PAIR_COLS);
    HIP_TRY(mc::launch_pair_cols(slot(c, 0), c->spec_stride, c->pairs_d.as<int2>(), np, n, bfactor,
                                 box, c->tbuf.as<float2>(), c->tmax.as<float>(), precision, cols,
                                 c->st));
  }
  {
    KScope ks(c, MC_KT_PAIR_ROWS);
    HIP_TRY(mc::launch_pair_rows(c->tbuf.as<float2>(), c->tmax.as<float>(), np, n, box,
                                 c->box.as<float>(), c->boxpart.as<float>(), precision, rows,
                                 c->st));
  }
  {
    KScope ks(c, MC_KT_PEAKS);
    HIP_TRY(mc::launch_find_peaks(c->box.as<float>(), np, box, n, wnoise, c->peaks.as<double>(),
                                  c->st));
  }
  return MC_OK;
}

// Pair stage on resident slots: pairs (a, b) are slot indices.
static int pair_stage(mc_context_s* c, const std::vector<int>& slots, float bfactor, int box,
                      int wnoise, int precision, std::vector<double>& res, float* boxmaps_host) {
  const int np = (int)slots.size() / 2;
  const int n = c->n;
  MC_TRY(pair_reserve(c, np, n, box, precision));
  HIP_TRY(hipMemcpyAsync(c->pairs_d.p, slots.data(), sizeof(int) * slots.size(),
                         hipMemcpyHostToDevice, c->st));
  MC_TRY(pair_launch(c, np, n, bfactor, box, wnoise, precision));
  res.resize(3 * (size_t)np);
  HIP_TRY(hipMemcpyAsync(res.data(), c->peaks.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost,
                         c->st));
  if (boxmaps_host)
    HIP_TRY(hipMemcpyAsync(boxmaps_host, c->box.p, sizeof(float) * (size_t)np * box * box,
                           hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  return MC_OK;
}

// Shift-and-sum of slots [base, base+count) with per-frame shifts into `dst`
// (and optionally dst2 for frames <= f2_last), all on device.  The shifts
// are staged in the context's shift buffer; `sh` must stay alive until the
// stream is synchronised.
static int shift_sum(mc_context_s* c, int base, int count, const std::vector<double>& sh,
                     float2* dst, float2* dst2, int f2_last) {
  HIP_TRY(c->shifts.ensure(sizeof(double) * sh.size()));
  HIP_TRY(hipMemcpyAsync(c->shifts.p, sh.data(), sizeof(double) * sh.size(),
                         hipMemcpyHostToDevice, c->st));
  KScope ks(c, MC_KT_SUM);
  HIP_TRY(mc::launch_shift_sum(slot(c, base), c->spec_stride, count, c->n,
                               c->shifts.as<double2>(), dst, dst2, f2_last, c->st));
  return MC_OK;
}

// Sharpness weights: wk[kx] = a_kx exp(-B kx^2/(2n^2)) (a = 1 at kx = 0 and
// n/2, else 2), wk[n/2+1+ky] = exp(-B y'^2/(2n^2)).
static int sharp_weights(mc_context_s* c, int n, float bfactor) {
  static thread_local std::vector<float> h;
  const int H = n / 2 + 1;
  h.assign(H + n, 0.f);
  const double m = -0.5 * (double)bfactor / ((double)n * n);
  for (int kx = 0; kx < H; kx++)
    h[kx] = (float)(((kx == 0 || 2 * kx == n) ? 1.0 : 2.0) * std::exp(m * kx * kx));
  for (int ky = 0; ky < n; ky++) {
    const int yy = ky < n / 2 ? ky : ky - n;
    h[H + ky] = (float)std::exp(m * (double)yy * yy);
  }
  HIP_TRY(c->wk.ensure(sizeof(float) * h.size()));
  HIP_TRY(hipMemcpyAsync(c->wk.p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  return MC_OK;
}

// J(s) and dJ/ds of frames in slots [base, base+F), sum formed in slot
// `sum`; fs: 2F shifts (NaN: frame out).
// J (and with grad, dJ/ds) at the shifts fs.  with_sum = false reuses the
// sum already in the sum slot, formed with the same shifts (still in the
// context's shift buffer).  Without grad only J is computed (same blocks and
// summation order, so the same value).
static int sharpness_eval(mc_context_s* c, int base, int F, int sum, const std::vector<double>& fs,
                          double* J, std::vector<double>* grad, bool with_sum = true) {
  const int nchunks = 256;
  if (with_sum) MC_TRY(shift_sum(c, base, F, fs, slot(c, sum), nullptr, -1));
  HIP_TRY(c->gpart.ensure(sizeof(double2) * (size_t)nchunks * (F + 1)));
  HIP_TRY(c->gout.ensure(sizeof(double2) * (F + 1)));
  {
    KScope ks(c, MC_KT_SUM);
    HIP_TRY(mc::launch_sharpness(slot(c, base), c->spec_stride, F, c->n, c->shifts.as<double2>(),
                                 slot(c, sum), c->wk.as<float>(), c->gpart.as<double2>(), nchunks,
                                 c->gout.as<double2>(), grad ? 0 : 1, c->st));
  }
  std::vector<double> o(2 * (size_t)(F + 1));
  HIP_TRY(hipMemcpyAsync(o.data(), c->gout.p, sizeof(double) * o.size(), hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  if (grad) grad->assign(o.begin(), o.begin() + 2 * F);
  *J = o[2 * F];
  return MC_OK;
}

// Gradient ascent of J over the summed frames' shifts with a trust region:
// s += step * g / max|g|.  A step that raises J is accepted and the next one
// doubled (up to 4x the first step); a step that does not is rejected and
// the step halved.
static int refine_shifts(mc_context_s* c, int base, int F, int sum, std::vector<double>& fs,
                         float bfactor, int steps, float step_px, double* j0, double* j1,
                         int* accepted) {
  MC_TRY(sharp_weights(c, c->n, bfactor));
  double J = 0;
  std::vector<double> g;
  MC_TRY(sharpness_eval(c, base, F, sum, fs, &J, &g));
  *j0 = J;
  const double step0 = step_px > 0 ? step_px : 0.05;
  double step = step0;
  int acc = 0;
  for (int t = 0; t < steps; t++) {
    double gmax = 0;
    for (size_t k = 0; k < g.size(); k++)
      if (!std::isnan(fs[k])) gmax = std::max(gmax, std::fabs(g[k]));
    if (!(gmax > 0)) break;
    std::vector<double> trial(fs);
    for (size_t k = 0; k < g.size(); k++)
      if (!std::isnan(fs[k])) trial[k] += step * g[k] / gmax;
    // the trial's J alone; its gradient only once it is accepted (from the
    // trial's sum, still resident)
    double Jt = 0;
    MC_TRY(sharpness_eval(c, base, F, sum, trial, &Jt, nullptr));
    if (Jt > J) {
      std::vector<double> gt;
      double Jg = 0;
      MC_TRY(sharpness_eval(c, base, F, sum, trial, &Jg, &gt, false));
      fs.swap(trial);
      g.swap(gt);
      J = Jt;
      acc++;
      step = std::min(2.0 * step, 4.0 * step0);
    } else {
      step *= 0.5;
    }
  }
  *j1 = J;
  *accepted = acc;
  return MC_OK;
}

extern "C" int mc_sharpness(mc_context ctx, const double* frame_shift, float bfactor, double* J,
                            double* grad) {
  if (!ctx || !frame_shift) return set_err(MC_ERR_ARG, "NULL argument");
  if (!ctx->nframe) return set_err(MC_ERR_STATE, "no spectra loaded");
  HIP_TRY(hipSetDevice(ctx->device));
  MC_TRY(sharp_weights(ctx, ctx->n, bfactor));
  std::vector<double> fs(frame_shift, frame_shift + 2 * (size_t)ctx->nframe), g;
  double j = 0;
  MC_TRY(sharpness_eval(ctx, 0, ctx->nframe, ctx->sum_slot, fs, &j, &g));
  if (J) *J = j;
  if (grad) memcpy(grad, g.data(), sizeof(double) * g.size());
  return MC_OK;
}

static int to_image(mc_context_s* c, const float2* spec, float* dev_img) {
  const int n = c->n;
  HIP_TRY(c->tmp.ensure(sizeof(float2) * mc::rt_frame_elems(n, n)));
  KScope ks(c, MC_KT_IMAGE);
  // DFAlign.cu:1289-1290: C2R then x 1/size (float 1.0/size)
  HIP_TRY(mc::launch_spectrum_to_image(spec, n, (float)(1.0 / ((double)n * n)), dev_img, n,
                                       c->tmp.as<float2>(), c->st));
  return MC_OK;
}

// ================================================================ level 1
extern "C" int mc_load_frames(mc_context ctx, const mc_movie_desc* m, const mc_params* p,
                              mc_report* rep) {
  if (!ctx || !m || !p) return set_err(MC_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  Geom g;
  int nframe = 0;
  MC_TRY(load_frames(ctx, m, p, &g, &nframe));
  HIP_TRY(hipStreamSynchronize(ctx->st));
  if (rep) {
    memset(rep, 0, sizeof *rep);
    rep->n = g.n;
    rep->nsamUnbin = g.nsu;
    rep->nframe = nframe;
  }
  return MC_OK;
}

extern "C" int mc_set_spectra(mc_context ctx, const float* host_padded, int n, int nframe) {
  if (!ctx || !host_padded || nframe < 1 || !fft_size_ok(n))
    return set_err(MC_ERR_ARG, "bad spectra arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  MC_TRY(ensure_spectra(ctx, n, nframe, 0));
  const size_t per = (size_t)(n / 2 + 1) * n;
  HIP_TRY(ctx->tmp.ensure(sizeof(float2) * per));
  for (int f = 0; f < nframe; f++) {
    HIP_TRY(hipMemcpyAsync(ctx->tmp.p, host_padded + (size_t)f * n * (n + 2),
                           sizeof(float2) * per, hipMemcpyHostToDevice, ctx->st));
    HIP_TRY(mc::launch_padded_to_spec(ctx->tmp.as<float2>(), n, slot(ctx, f), ctx->st));
  }
  HIP_TRY(hipStreamSynchronize(ctx->st));
  return MC_OK;
}

extern "C" int mc_get_spectra(mc_context ctx, float* host_padded) {
  if (!ctx || !host_padded) return set_err(MC_ERR_ARG, "NULL argument");
  if (!ctx->nframe) return set_err(MC_ERR_STATE, "no spectra loaded");
  HIP_TRY(hipSetDevice(ctx->device));
  const int n = ctx->n;
  const size_t per = (size_t)(n / 2 + 1) * n;
  HIP_TRY(ctx->tmp.ensure(sizeof(float2) * per));
  for (int f = 0; f < ctx->nframe; f++) {
    HIP_TRY(mc::launch_spec_to_padded(slot(ctx, f), n, ctx->tmp.as<float2>(), ctx->st));
    HIP_TRY(hipMemcpyAsync(host_padded + (size_t)f * n * (n + 2), ctx->tmp.p,
                           sizeof(float2) * per, hipMemcpyDeviceToHost, ctx->st));
    HIP_TRY(hipStreamSynchronize(ctx->st));
  }
  return MC_OK;
}

extern "C" int mc_pair_shifts(mc_context ctx, const int* pairs, int npairs, float bfactor,
                              int box, int wnoise, double* b, double* cc, float* boxmaps) {
  if (!ctx || !pairs || npairs < 1) return set_err(MC_ERR_ARG, "bad pair arguments");
  if (!ctx->nframe) return set_err(MC_ERR_STATE, "no spectra loaded");
  if (box < 16 || box > ctx->n || box > 256)
    return set_err(MC_ERR_ARG, "peak box %d must be in [16, min(n, 256)]", box);
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<int> s(2 * (size_t)npairs);
  for (int q = 0; q < npairs; q++) {
    if (pairs[2 * q] < 0 || pairs[2 * q] >= ctx->nframe || pairs[2 * q + 1] < 0 ||
        pairs[2 * q + 1] >= ctx->nframe)
      return set_err(MC_ERR_ARG, "pair %d out of range", q);
    s[2 * q] = pairs[2 * q];
    s[2 * q + 1] = pairs[2 * q + 1] + ctx->off2;
  }
  std::vector<double> res;
  MC_TRY(pair_stage(ctx, s, bfactor, box, wnoise, MC_PREC_FP32, res, boxmaps));
  for (int q = 0; q < npairs; q++) {
    if (b) {
      b[2 * q] = res[3 * q];
      b[2 * q + 1] = res[3 * q + 1];
    }
    if (cc) cc[q] = res[3 * q + 2];
  }
  return MC_OK;
}

extern "C" int mc_find_peaks(mc_context ctx, const float* boxmaps, int npairs, int box, int n,
                             int wnoise, double* b, double* cc) {
  if (!ctx || !boxmaps || npairs < 1 || box < 16) return set_err(MC_ERR_ARG, "bad arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(ctx->box.ensure(sizeof(float) * (size_t)npairs * box * box));
  HIP_TRY(ctx->peaks.ensure(sizeof(double) * mc::find_peaks_out_elems(npairs)));
  HIP_TRY(hipMemcpyAsync(ctx->box.p, boxmaps, sizeof(float) * (size_t)npairs * box * box,
                         hipMemcpyHostToDevice, ctx->st));
  HIP_TRY(mc::launch_find_peaks(ctx->box.as<float>(), npairs, box, n, wnoise,
                                ctx->peaks.as<double>(), ctx->st));
  std::vector<double> res(3 * (size_t)npairs);
  HIP_TRY(hipMemcpyAsync(res.data(), ctx->peaks.p, sizeof(double) * res.size(),
                         hipMemcpyDeviceToHost, ctx->st));
  HIP_TRY(hipStreamSynchronize(ctx->st));
  for (int q = 0; q < npairs; q++) {
    if (b) {
      b[2 * q] = res[3 * q];
      b[2 * q + 1] = res[3 * q + 1];
    }
    if (cc) cc[q] = res[3 * q + 2];
  }
  return MC_OK;
}

extern "C" int mc_sum_image(mc_context ctx, const double* frame_shift, float* image) {
  if (!ctx || !frame_shift || !image) return set_err(MC_ERR_ARG, "NULL argument");
  if (!ctx->nframe) return set_err(MC_ERR_STATE, "no spectra loaded");
  HIP_TRY(hipSetDevice(ctx->device));
  const int n = ctx->n;
  std::vector<double> sh(frame_shift, frame_shift + 2 * (size_t)ctx->nframe);
  MC_TRY(shift_sum(ctx, 0, ctx->nframe, sh, slot(ctx, ctx->sum_slot), nullptr, -1));
  HIP_TRY(ctx->image.ensure(sizeof(float) * (size_t)n * n));
  MC_TRY(to_image(ctx, slot(ctx, ctx->sum_slot), ctx->image.as<float>()));
  HIP_TRY(hipMemcpyAsync(image, ctx->image.p, sizeof(float) * (size_t)n * n,
                         hipMemcpyDeviceToHost, ctx->st));
  HIP_TRY(hipStreamSynchronize(ctx->st));
  return MC_OK;
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

// ====================================================== whole movies
// One movie of a call to mc_align_movie / mc_align_batch: its geometry, its
// slot range in the shared spectrum array and the host-side drift state.
struct MovieRun {
  const mc_movie_desc* m = nullptr;
  const mc_params* p = nullptr;
  mc_outputs* out = nullptr;
  mc_report* rep = nullptr;
  Geom g{};
  int F = 0, P = 0, K = 0, box = 0, wnoise = 0;
  int base = 0, off2 = 0, sum = 0, nslot = 0;  // slots [base, base+nslot)
  std::vector<int> pairs;
  std::vector<double> b, s, oldx, newx, fs, sh, sh2;
  int ng = 0, it = 0, converged = 0, rawsum_written = 0, sum2_written = 0, gacc = 0;
  int nss = 0, nes = 0;       // first and last summed frame positions
  double tot_x = 0, tot_y = 0;  // total shift of frame nes (the FSC direction)
  double rmsd = 0, sj0 = 0, sj1 = 0;
  bool iterating() const { return it < p->iter && !converged; }
};

static int prepare_run(MovieRun& r) {
  const mc_movie_desc* m = r.m;
  const mc_params* p = r.p;
  MC_TRY(movie_geometry(m, p, &r.g));
  r.F = p->nEnd - p->nStart + 1;
  if (p->nStart < m->first_frame || p->nEnd >= m->first_frame + m->nz || r.F < 2)
    return set_err(MC_ERR_ARG, "frame range [%d, %d] invalid for frames [%d, %d)", p->nStart,
                   p->nEnd, m->first_frame, m->first_frame + m->nz);
  if (!m->frames) return set_err(MC_ERR_ARG, "frames is NULL");
  r.box = p->CCPeakSearchDim;
  if (r.box < 16 || r.box > r.g.n || r.box > 256)
    return set_err(MC_ERR_ARG, "peak box %d must be in [16, min(n, 256)]", r.box);
  r.wnoise = p->NoisePeakSize - 1;  // DFAlign.cu:900
  if (p->precision != MC_PREC_FP32 && p->precision != MC_PREC_MFMA_HILO)
    return set_err(MC_ERR_ARG, "unknown precision form %d", p->precision);
  if (!mc::pair_form_supported(p->precision, r.g.n, r.box))
    return set_err(MC_ERR_ARG,
                   "precision form %d needs n a multiple of 512 in [512, 8192] and peak box "
                   "<= 96 (n %d, box %d)",
                   p->precision, r.g.n, r.box);
  r.K = r.F - 1;
  r.P = mc_pair_list(r.F, p->FrameDistOffset, nullptr, 0);
  r.pairs.assign(2 * (size_t)r.P, 0);
  mc_pair_list(r.F, p->FrameDistOffset, r.pairs.data(), r.P);
  MC_TRY(check_corrections(m));
  if (r.out && (r.out->disp_raw_fft || r.out->disp_corr_sum || r.out->disp_corr_fft) &&
      r.g.n < MC_DISPDIM)
    return set_err(MC_ERR_ARG, "quick-look images need a transform size >= %d (n %d)",
                   MC_DISPDIM, r.g.n);
  r.nslot = r.F * ((m->frames2 || (m->corr && m->corr->second)) ? 2 : 1) + 1;
  return MC_OK;
}

// Least squares, outlier rejection, least squares (DFAlign.cu:914-948) from
// the pair peaks res (3 doubles per pair), then the start of the
// align-to-sum recursion.
static int solve_drift(MovieRun& r, const double* res) {
  mc_outputs* out = r.out;
  const int P = r.P, K = r.K, F = r.F;
  r.b.assign(2 * (size_t)P, 0.0);
  for (int q = 0; q < P; q++) {
    r.b[2 * q] = res[3 * q];
    r.b[2 * q + 1] = res[3 * q + 1];
    if (out->pair_cc) out->pair_cc[q] = res[3 * q + 2];
  }
  if (out->pairs) memcpy(out->pairs, r.pairs.data(), sizeof(int) * 2 * P);
  if (out->pair_b) memcpy(out->pair_b, r.b.data(), sizeof(double) * 2 * P);
  std::vector<double> ki(P);
  r.s.assign(2 * (size_t)K, 0.0);
  MC_TRY(mc_lsq(r.pairs.data(), P, F, r.b.data(), r.s.data()));
  residuals(r.pairs.data(), P, r.s.data(), r.b.data(), ki.data());
  if (out->shift1) memcpy(out->shift1, r.s.data(), sizeof(double) * 2 * K);
  if (out->ki1) memcpy(out->ki1, ki.data(), sizeof(double) * P);
  std::vector<int> good(P);
  r.ng = mc_threshold(r.pairs.data(), P, F, ki.data(), r.p->kiThresh, good.data());
  std::vector<int> gp(2 * (size_t)r.ng);
  std::vector<double> gb(2 * (size_t)r.ng);
  for (int q = 0; q < r.ng; q++) {
    gp[2 * q] = r.pairs[2 * good[q]];
    gp[2 * q + 1] = r.pairs[2 * good[q] + 1];
    gb[2 * q] = r.b[2 * good[q]];
    gb[2 * q + 1] = r.b[2 * good[q] + 1];
  }
  MC_TRY(mc_lsq(gp.data(), r.ng, F, gb.data(), r.s.data()));
  if (out->good) memcpy(out->good, good.data(), sizeof(int) * r.ng);
  if (out->shift2) memcpy(out->shift2, r.s.data(), sizeof(double) * 2 * K);
  if (out->ki2) residuals(gp.data(), r.ng, r.s.data(), gb.data(), out->ki2);
  // positions of the frames from the increments (DFAlign.cu:955-966)
  r.newx.assign(2 * (size_t)F, 0.0);
  double ax = 0, ay = 0;
  for (int k = 0; k < K; k++) {
    ax += r.s[2 * k];
    ay += r.s[2 * k + 1];
    r.newx[2 * (k + 1)] = ax;
    r.newx[2 * (k + 1) + 1] = ay;
  }
  r.oldx = r.newx;
  return MC_OK;
}

// Shifts of one align-to-sum round: the sum of every frame moved by -oldx
// (DFAlign.cu:974-989).
static void iter_shifts(MovieRun& r) {
  r.sh.assign(2 * (size_t)r.F, 0.0);
  for (int k = 0; k < 2 * r.F; k++) r.sh[k] = -r.oldx[k];
}

// One align-to-sum round from the peaks of (sum, frame k) (DFAlign.cu:990-1071).
static void iter_update(MovieRun& r, const double* res) {
  mc_outputs* out = r.out;
  const int F = r.F, it = r.it;
  for (int k = 0; k < F; k++) {
    r.newx[2 * k] = res[3 * k];
    r.newx[2 * k + 1] = res[3 * k + 1];
    if (out->iter_cc) out->iter_cc[(size_t)it * F + k] = res[3 * k + 2];
  }
  if (out->iter_meas) memcpy(out->iter_meas + (size_t)it * 2 * F, r.newx.data(), sizeof(double) * 2 * F);
  const double rx = r.newx[0], ry = r.newx[1];
  for (int k = 0; k < F; k++) {
    const double dx = r.oldx[2 * k] - r.newx[2 * k], dy = r.oldx[2 * k + 1] - r.newx[2 * k + 1];
    r.oldx[2 * k] = r.newx[2 * k] - rx;
    r.oldx[2 * k + 1] = r.newx[2 * k + 1] - ry;
    r.rmsd += dx * dx + dy * dy;
  }
  r.rmsd /= F;  // accumulates across rounds as the reference does (DFAlign.cu:1036-1041)
  if (out->rmsd) out->rmsd[it] = r.rmsd;
  if (r.rmsd < (double)r.p->iterthres * (double)r.p->iterthres) r.converged = 1;
  r.it++;
}

// Quick-look images of one sum spectrum ([kx][ky], even size n >= 400)
// into host buffers of MC_DISPDIM^2 shorts (DFAlign.cu:73-179,
// 842-848, 908-911, 1270-1285; func.cpp:372-555).  sum_img: crop to 400,
// C2R, 8-bit mapping.  fft_img: centred log-modulus image, its R2C binned
// to 400 by the stage-1 transform, C2R, / 400^2, 8-bit mapping.
static int quicklook(mc_context_s* c, const float2* spec, int n, float fftscale, short* fft_img,
                     short* sum_img) {
  constexpr int D = MC_DISPDIM;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_img = al(sizeof(float) * (size_t)n * n);
  const size_t b_rt = al(sizeof(float2) * mc::rt_frame_elems(D, n));
  const size_t b_sp = al(sizeof(float2) * (size_t)(D / 2 + 1) * D);
  const size_t b_out = al(sizeof(float) * (size_t)D * D);
  const size_t b_sh = al(sizeof(short) * (size_t)D * D);
  HIP_TRY(c->disp.ensure(b_img + b_rt + 2 * b_sp + b_out + b_sh));
  char* base = static_cast<char*>(c->disp.p);
  float* img = reinterpret_cast<float*>(base);
  float2* rt = reinterpret_cast<float2*>(base + b_img);
  float2* dsp = reinterpret_cast<float2*>(base + b_img + b_rt);
  float2* tmp = reinterpret_cast<float2*>(base + b_img + b_rt + b_sp);
  float* outf = reinterpret_cast<float*>(base + b_img + b_rt + 2 * b_sp);
  short* sh = reinterpret_cast<short*>(base + b_img + b_rt + 2 * b_sp + b_out);
  KScope ks(c, MC_KT_IMAGE);
  if (sum_img) {
    HIP_TRY(mc::launch_small_c2r(spec, n, D, 0.f, tmp, outf, c->st));
    HIP_TRY(mc::launch_disp_short(outf, D * D, sh, c->st));
    HIP_TRY(hipMemcpyAsync(sum_img, sh, sizeof(short) * D * D, hipMemcpyDeviceToHost, c->st));
  }
  if (fft_img) {
    HIP_TRY(mc::launch_logmod_image(spec, n, fftscale, img, c->st));
    mc::R2CGeom g{};
    g.src = reinterpret_cast<const char*>(img);
    g.dtype = mc::DT_FLOAT32;
    g.row_pitch = sizeof(float) * (size_t)n;
    g.frame_pitch = sizeof(float) * (size_t)n * n;
    g.w = g.h = g.nu = n;
    g.n = D;
    g.nx = n;
    HIP_TRY(mc::launch_frames_to_spectra(g, 0, 1, dsp, (size_t)(D / 2 + 1) * D, rt, nullptr,
                                         nullptr, nullptr, mc::R2C_AUTO, c->st));
    HIP_TRY(mc::launch_small_c2r(dsp, D, D, (float)(D * D), tmp, outf, c->st));
    HIP_TRY(mc::launch_disp_short(outf, D * D, sh, c->st));
    HIP_TRY(hipMemcpyAsync(fft_img, sh, sizeof(short) * D * D, hipMemcpyDeviceToHost, c->st));
  }
  return MC_OK;
}

static float fft_scale(const mc_params* p) { return p->fftscale > 0.f ? p->fftscale : 0.0001f; }

// Increments after the iteration and the shift of every summed frame
// (DFAlign.cu:1114-1148); frames outside [nss, nes] get NaN (not summed).
static void final_shifts(MovieRun& r, int* nes2_out) {
  const mc_params* p = r.p;
  const int F = r.F, K = r.K;
  for (int k = 0; k < K; k++) {
    r.s[2 * k] = r.oldx[2 * (k + 1)] - r.oldx[2 * k];
    r.s[2 * k + 1] = r.oldx[2 * (k + 1) + 1] - r.oldx[2 * k + 1];
  }
  if (r.out->final_shift) memcpy(r.out->final_shift, r.s.data(), sizeof(double) * 2 * K);
  int nss = p->nStartSum - p->nStart, nes = p->nEndSum - p->nStart, nes2 = p->nEndSum2 - p->nStart;
  if (nss <= 0) nss = 0;
  if (nes <= nss || nes >= F) nes = F - 1;
  if (nes2 <= nss || nes2 >= F) nes2 = F - 1;
  if (nes2 > nes) nes2 = nes;
  double midx = 0, midy = 0;
  if (p->bAlignToMid) {
    const int nzf = r.m->nz_file > 0 ? r.m->nz_file : r.m->nz;
    int ref = nzf / 2 + 1;
    if (ref < p->nStart) ref = p->nStart;
    const int lim = std::min(ref - p->nStart, K);  // guards DFAlign.cu:1138
    for (int j = 0; j < lim; j++) {
      midx += r.s[2 * j];
      midy += r.s[2 * j + 1];
    }
  }
  r.fs.assign(2 * (size_t)F, NAN);
  double tx = 0, ty = 0;
  for (int j = 1; j < nss + 1; j++) {
    tx += r.s[2 * (j - 1)];
    ty += r.s[2 * (j - 1) + 1];
  }
  // first summed frame: shifted only with -atm 1 (DFAlign.cu:1148)
  r.fs[2 * nss] = p->bAlignToMid ? -tx + midx : 0.0;
  r.fs[2 * nss + 1] = p->bAlignToMid ? -ty + midy : 0.0;
  for (int j = nss + 1; j <= nes; j++) {
    tx += r.s[2 * (j - 1)];
    ty += r.s[2 * (j - 1) + 1];
    r.fs[2 * j] = -tx + midx;
    r.fs[2 * j + 1] = -ty + midy;
  }
  r.nss = nss;
  r.nes = nes;
  r.tot_x = tx;
  r.tot_y = ty;
  *nes2_out = nes2;
}

// Shell sums of one pair of half spectra into sums (nbox*6 doubles, device),
// direction normalised as DFAlign.cu:236-237 does.  The caller sizes c->fsc
// to (fsc_grid(n) + 2) * nbox * 6 doubles.
static int fsc_sums(mc_context_s* c, const float2* a, const float2* b, int n, int padded,
                    double dx, double dy, double* sums) {
  const double dn = std::hypot(dx, dy);  // abs(complex<double>)
  if (dn > 0.1) {
    dx /= dn;
    dy /= dn;
  } else {
    dx = 1.0;
    dy = 0.0;
  }
  // c->fsc holds two tables of sums, then the partial sums
  const size_t nk = (size_t)mc::fsc_bins(n, nullptr) * 6;
  KScope ks(c, MC_KT_SUM);
  HIP_TRY(mc::launch_fsc_sums(a, b, n, padded, dx, dy, c->fsc.as<double>() + 2 * nk, sums, c->st));
  return MC_OK;
}

// FSC = sum Re(a conj b) / sqrt(sum|a|^2 sum|b|^2), 0 when the product is
// <= 1e-5 (DFAlign.cu:283-298); column stride `ld`, side 0 then side 1.
static void fsc_ratios(const double* sums, int nbox, double* out, int ld) {
  for (int i = 0; i < nbox; i++)
    for (int side = 0; side < 2; side++) {
      const double* v = sums + (2 * i + side) * 3;
      const double t = v[1] * v[2];
      out[(size_t)i * ld + side] = t > 0.00001 ? v[0] / std::sqrt(t) : 0.0;
    }
}

// FSC table of one movie (DFAlign.cu:808-822, 1143-1266, 1314-1318): raw
// halves are the unshifted frames at alternating positions of
// [nStartSum, nEndSum]; aligned halves are the first summed frame plus every
// second frame after it against the others, with the shifts of the final
// sum (still in the context's shift buffer).  Rows: D_Raw, U_Raw, D_Corr,
// U_Corr.
static int fsc_table(mc_context_s* c, MovieRun& r) {
  const mc_params* p = r.p;
  const int n = c->n;
  const int nbox = mc::fsc_bins(n, nullptr);
  const size_t nk = (size_t)nbox * 6;
  HIP_TRY(c->tmp.ensure(sizeof(float2) * c->spec_stride * 2));
  float2* h0 = c->tmp.as<float2>();
  float2* h1 = h0 + c->spec_stride;
  HIP_TRY(c->fsc.ensure(sizeof(double) * nk * (mc::fsc_grid(n) + 2)));
  double* sums = c->fsc.as<double>();
  const int j0 = std::max(p->nStartSum, p->nStart), j1 = std::min(p->nEndSum, p->nEnd);
  {
    KScope ks(c, MC_KT_SUM);
    HIP_TRY(mc::launch_half_sums(slot(c, r.base + std::max(0, j0 - p->nStart)), c->spec_stride,
                                 std::max(0, j1 - j0 + 1), n, nullptr, h0, h1, c->st));
  }
  MC_TRY(fsc_sums(c, h0, h1, n, 0, r.tot_x, r.tot_y, sums));
  {
    KScope ks(c, MC_KT_SUM);
    HIP_TRY(mc::launch_half_sums(slot(c, r.base + r.nss), c->spec_stride, r.nes - r.nss + 1, n,
                                 c->shifts.as<double2>() + r.nss, h0, h1, c->st));
  }
  MC_TRY(fsc_sums(c, h0, h1, n, 0, r.tot_x, r.tot_y, sums + nk));
  std::vector<double> h(2 * nk);
  HIP_TRY(hipMemcpyAsync(h.data(), sums, sizeof(double) * 2 * nk, hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  fsc_ratios(h.data(), nbox, r.out->fsc, 4);
  fsc_ratios(h.data() + nk, nbox, r.out->fsc + 2, 4);
  return MC_OK;
}

// Final phase-shift-and-sum and images of one movie (DFAlign.cu:1143-1292).
static int final_sum(mc_context_s* c, MovieRun& r) {
  mc_outputs* out = r.out;
  const mc_params* p = r.p;
  const int n = c->n;
  int nes2 = 0;
  final_shifts(r, &nes2);
  if (out->frame_shift) memcpy(out->frame_shift, r.fs.data(), sizeof(double) * 2 * r.F);
  if (p->grad_steps > 0) {
    // adjoint sharpness refinement of the summed frames' shifts
    MC_TRY(refine_shifts(c, r.base, r.F, r.sum, r.fs, p->bfactor, p->grad_steps, p->grad_step_px,
                         &r.sj0, &r.sj1, &r.gacc));
    if (out->frame_shift) memcpy(out->frame_shift, r.fs.data(), sizeof(double) * 2 * r.F);
  }
  const bool want2 = out->sum2_image && p->nEndSum2 > 0 && p->nEndSum2 < p->nEndSum;
  float2* s2 = nullptr;
  if (want2) {
    // after the image scratch of to_image, which runs before s2 is copied out
    const size_t off = std::max(c->spec_stride, mc::rt_frame_elems(n, n));
    HIP_TRY(c->tmp.ensure(sizeof(float2) * (off + c->spec_stride)));
    s2 = c->tmp.as<float2>() + off;
  }
  MC_TRY(shift_sum(c, r.base, r.F, r.fs, slot(c, r.sum), s2, nes2));
  HIP_TRY(c->image.ensure(sizeof(float) * (size_t)n * n));
  float* img = out->sum_image_dev ? static_cast<float*>(out->sum_image_dev) : c->image.as<float>();
  MC_TRY(to_image(c, slot(c, r.sum), img));
  if (out->sum_image)
    HIP_TRY(hipMemcpyAsync(out->sum_image, img, sizeof(float) * (size_t)n * n,
                           hipMemcpyDeviceToHost, c->st));
  if (out->disp_corr_fft || out->disp_corr_sum)  // DFAlign.cu:1270-1285
    MC_TRY(quicklook(c, slot(c, r.sum), n, fft_scale(p), out->disp_corr_fft, out->disp_corr_sum));
  if (want2) {
    // the second sum shares tmp: copy it into a spectrum slot first
    HIP_TRY(hipMemcpyAsync(slot(c, r.sum), s2, sizeof(float2) * c->spec_stride,
                           hipMemcpyDeviceToDevice, c->st));
    HIP_TRY(c->sum2img.ensure(sizeof(float) * (size_t)n * n));
    MC_TRY(to_image(c, slot(c, r.sum), c->sum2img.as<float>()));
    HIP_TRY(hipMemcpyAsync(out->sum2_image, c->sum2img.p, sizeof(float) * (size_t)n * n,
                           hipMemcpyDeviceToHost, c->st));
    r.sum2_written = 1;
  }
  if (out->fsc) MC_TRY(fsc_table(c, r));
  // the host copies above read c->image / sum2img before the next movie's
  // launches on the same stream overwrite them
  HIP_TRY(hipStreamSynchronize(c->st));
  return MC_OK;
}

// Movies of one transform size within an alignment group: their slots form
// one region of the spectrum buffer with a uniform stride.
struct SizeGroup {
  int n = 0;
  size_t stride = 0, off = 0;  // float2 per slot; region start in spec
  int nslots = 0;
  std::vector<MovieRun*> rs;
};

static void use_group(mc_context_s* c, const SizeGroup& g) {
  c->n = g.n;
  c->spec_stride = g.stride;
  c->spec_off = g.off;
}

// The pair stage of several movies (each of its own size group): span i has
// np pairs of slot indices of its group's region, stored in pairs_d from pair
// po, with its T columns from tbuf element to.  Built and uploaded outside
// any graph capture; pairs_launch below only launches.
struct PairSpan {
  int gi = 0, np = 0;
  size_t po = 0, to = 0;
  int cols = mc::PAIR_COLS_AUTO, body = -1;
};
struct PairPlan {
  std::vector<PairSpan> sp;
  size_t np = 0;
  long long items[2] = {0, 0};  // plan-kernel items of each register class
  int nplans[2] = {0, 0};       // class-0 table, then the class-1 table
  float bfactor = 0.f;
  int box = 0, wnoise = 0, precision = 0;
};

// Reserves the pair buffers and uploads pairs (host slot pairs, span by span)
// and the column-pass plan table.  With more than one span the spans that
// the plan kernel covers run as one launch sequence (mc::PairColsPlan).
static int pairs_prepare(mc_context_s* c, const std::vector<SizeGroup>& gs, PairPlan& L,
                         const std::vector<int>& slots) {
  size_t np = 0, tel = 0, part = 0, pscr = 0;
  std::vector<size_t> gnp(gs.size(), 0);
  for (PairSpan& q : L.sp) {
    const int n = gs[q.gi].n;
    q.po = np;
    q.to = tel;
    np += q.np;
    tel += (size_t)q.np * (n / 2 + 1) * L.box;
    gnp[q.gi] += q.np;
    const mc_context_s::Tune* tn = pair_tune(c, n, L.box, L.precision);
    q.cols = tn ? tn->cols : mc::PAIR_COLS_AUTO;
    // several movies: the default-variant ones share one plan table (a
    // single movie runs the same plan kernel with its plan passed by value)
    q.body = (L.precision == MC_PREC_FP32 && L.sp.size() > 1)
                 ? mc::pair_cols_plan_body(n, L.box, q.cols) : -1;
  }
  for (size_t g = 0; g < gs.size(); g++) {
    part = std::max(part, mc::pair_rows_part_elems((int)gnp[g], gs[g].n, L.box));
    pscr = std::max(pscr, mc::find_peaks_out_elems((int)gnp[g]) - 3 * gnp[g]);
  }
  if (slots.size() != 2 * np) return set_err(MC_ERR_ARG, "pair spans cover %zu of %zu pairs", np, slots.size() / 2);
  L.np = np;
  HIP_TRY(c->pairs_d.ensure(sizeof(int) * 2 * std::max<size_t>(1, np)));
  HIP_TRY(c->tbuf.ensure(sizeof(float2) * std::max<size_t>(1, tel)));
  if (L.precision == MC_PREC_MFMA_HILO) HIP_TRY(c->tmax.ensure(sizeof(float) * std::max<size_t>(1, tel / L.box)));
  HIP_TRY(c->box.ensure(sizeof(float) * std::max<size_t>(1, np) * L.box * L.box));
  HIP_TRY(c->boxpart.ensure(sizeof(float) * std::max<size_t>(1, part)));
  HIP_TRY(c->peaks.ensure(sizeof(double) * (3 * np + pscr + 1)));
  HIP_TRY(hipMemcpyAsync(c->pairs_d.p, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice,
                         c->st));
  std::vector<mc::PairColsPlan> plans[2];
  for (int k = 0; k < 2; k++) L.items[k] = 0;
  for (const PairSpan& q : L.sp) {
    if (q.body < 0 || q.np == 0) continue;
    const SizeGroup& g = gs[q.gi];
    const int k = mc::pair_cols_plan_class(q.body);
    mc::PairColsPlan pl;
    pl.spectra = c->spec.as<float2>() + g.off;
    pl.pairs = c->pairs_d.as<int2>() + q.po;
    pl.wxy = mc::pair_weights(g.n, L.bfactor);
    pl.tw = mc::twiddles(g.n);
    if (!pl.wxy || !pl.tw) return set_err(MC_ERR_DEVICE, "pair tables for n %d", g.n);
    pl.tbuf = c->tbuf.as<float2>() + q.to;
    pl.spec_stride = g.stride;
    pl.item0 = L.items[k];
    pl.npairs = q.np;
    pl.box = L.box;
    pl.body = q.body;
    plans[k].push_back(pl);
    L.items[k] += (long long)q.np * (g.n / 2 + 1);
  }
  for (int k = 0; k < 2; k++) L.nplans[k] = (int)plans[k].size();
  plans[0].insert(plans[0].end(), plans[1].begin(), plans[1].end());
  if (!plans[0].empty()) {
    HIP_TRY(c->plans.ensure(sizeof(mc::PairColsPlan) * plans[0].size()));
    HIP_TRY(hipMemcpyAsync(c->plans.p, plans[0].data(), sizeof(mc::PairColsPlan) * plans[0].size(),
                           hipMemcpyHostToDevice, c->st));
    // the host vector goes out of scope: the copy must have read it
    HIP_TRY(hipStreamSynchronize(c->st));
  }
  return MC_OK;
}

// Column pass (plan kernel and per-span launches), then the row pass and the
// peak search per size group.  Results: peaks (3 doubles per pair, span
// order) and box maps.  Capturable.
static int pairs_launch(mc_context_s* c, const std::vector<SizeGroup>& gs, const PairPlan& L) {
  const int box = L.box;
  {
    KScope ks(c, MC_KT_PAIR_COLS);
    for (int k = 0; k < 2; k++)
      if (L.nplans[k] > 0)
        HIP_TRY(mc::launch_pair_cols_plans(c->plans.as<mc::PairColsPlan>() + (k ? L.nplans[0] : 0),
                                           L.nplans[k], L.items[k], k, c->st));
    for (const PairSpan& q : L.sp) {
      if (q.body >= 0 || q.np == 0) continue;
      const SizeGroup& g = gs[q.gi];
      float* tmax = L.precision == MC_PREC_MFMA_HILO ? c->tmax.as<float>() + q.to / box : nullptr;
      HIP_TRY(mc::launch_pair_cols(c->spec.as<float2>() + g.off, g.stride,
                                   c->pairs_d.as<int2>() + q.po, q.np, g.n, L.bfactor, box,
                                   c->tbuf.as<float2>() + q.to, tmax, L.precision, q.cols, c->st));
    }
  }
  // consecutive spans of one size group share a row-pass and a peak launch
  size_t i = 0;
  while (i < L.sp.size()) {
    size_t j = i;
    int np = 0;
    while (j < L.sp.size() && L.sp[j].gi == L.sp[i].gi) np += L.sp[j++].np;
    const PairSpan& q = L.sp[i];
    const int n = gs[q.gi].n;
    i = j;
    if (np == 0) continue;
    const mc_context_s::Tune* tn = pair_tune(c, n, box, L.precision);
    const int rows = tn ? tn->rows : mc::PAIR_ROWS_AUTO;
    float* maps = c->box.as<float>() + q.po * box * box;
    {
      KScope ks(c, MC_KT_PAIR_ROWS);
      float* tmax = L.precision == MC_PREC_MFMA_HILO ? c->tmax.as<float>() + q.to / box : nullptr;
      HIP_TRY(mc::launch_pair_rows(c->tbuf.as<float2>() + q.to, tmax, np, n, box, maps,
                                   c->boxpart.as<float>(), L.precision, rows, c->st));
    }
    {
      KScope ks(c, MC_KT_PEAKS);
      HIP_TRY(mc::launch_find_peaks(maps, np, box, n, L.wnoise, c->peaks.as<double>() + 3 * q.po,
                                    c->st, c->peaks.as<double>() + 3 * L.np));
    }
  }
  return MC_OK;
}

// Graph of one align-to-sum round of a set of movies: the batched sums, the
// pair column/row passes and the peak search, replayed while the set of
// iterating movies is unchanged.  Only kernel launches are captured: the
// shifts are uploaded and the peaks read back around each replay.
struct RoundGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<int> key;
  ~RoundGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

// Align-to-sum rounds run as graph replays unless MC_NO_GRAPH=1 (eager
// launches of the same kernels; read at every call).
static bool graphs_enabled() {
  const char* e = getenv("MC_NO_GRAPH");
  return !(e && atoi(e));
}

// One round: per size group the batched shift-and-sum of its iterating
// movies (jobs [j0, j0 + nj) of the job table), then the pair stage.
struct RoundJobs {
  int gi, j0, nj;
};
static int round_launch(mc_context_s* c, const std::vector<SizeGroup>& gs,
                        const std::vector<RoundJobs>& rj, const PairPlan& L) {
  {
    KScope ks(c, MC_KT_SUM);
    for (const RoundJobs& q : rj) {
      const SizeGroup& g = gs[q.gi];
      HIP_TRY(mc::launch_shift_sum_multi(c->spec.as<float2>() + g.off, g.stride, g.n,
                                         c->jobs.as<int4>() + q.j0, q.nj, c->shifts.as<double2>(),
                                         c->st));
    }
  }
  return pairs_launch(c, gs, L);
}

static int round_graph(mc_context_s* c, RoundGraph& gr, const std::vector<int>& key,
                       const std::vector<SizeGroup>& gs, const std::vector<RoundJobs>& rj,
                       const PairPlan& L) {
  if (gr.exec && gr.key == key) return MC_OK;
  if (gr.exec) {
    (void)hipGraphExecDestroy(gr.exec);
    gr.exec = nullptr;
  }
  const unsigned kt = c->kt_mask;
  c->kt_mask = 0;  // no event records inside the graph
  HIP_TRY(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  const int rc = round_launch(c, gs, rj, L);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->st, &graph);
  c->kt_mask = kt;
  if (rc != MC_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HIP_TRY(ec);
  const hipError_t ei = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_TRY(ei);
  gr.key = key;
  c->graph_captures++;
  return MC_OK;
}

struct GroupTimes {
  float spectra = 0, pairs = 0, solve = 0, iter = 0, sum = 0, total = 0;
};

// Aligns movies with the same peak box, B-factor, noise window and form,
// of any transform sizes, whose spectra fit together in device memory
// (callers order them by size): stage 1 per movie into its size group's
// region, then one pair stage over every pair of every movie (the column
// pass as one plan-kernel launch sequence across sizes when there are
// several movies), the host drift model per movie, align-to-sum rounds over
// the movies still iterating (one graph replay per round), and the final sum
// per movie.
static int align_group(mc_context_s* c, std::vector<MovieRun*>& rs, bool use_graph,
                       GroupTimes* tm) {
  const int box = rs[0]->box, wnoise = rs[0]->wnoise;
  const float bfactor = rs[0]->p->bfactor;
  const int precision = rs[0]->p->precision;
  std::vector<SizeGroup> gs;
  std::vector<int> gi_of(rs.size());
  size_t total = 0;
  int ftot = 0, nmax = 0;
  for (size_t i = 0; i < rs.size(); i++) {
    MovieRun* r = rs[i];
    size_t k = 0;
    while (k < gs.size() && gs[k].n != r->g.n) k++;
    if (k == gs.size()) {
      gs.emplace_back();
      gs.back().n = r->g.n;
      gs.back().stride = (size_t)(r->g.n / 2 + 1) * r->g.n;
    }
    gi_of[i] = (int)k;
    SizeGroup& g = gs[k];
    r->base = g.nslots;
    r->off2 = (r->m->frames2 || (r->m->corr && r->m->corr->second)) ? r->F : 0;
    r->sum = r->base + r->nslot - 1;
    g.nslots += r->nslot;
    g.rs.push_back(r);
    ftot += r->F;
    nmax = std::max(nmax, r->g.n);
  }
  for (SizeGroup& g : gs) {
    g.off = total;
    total += g.stride * g.nslots;
  }
  HIP_TRY(c->spec.ensure(sizeof(float2) * total));
  c->nframe = 0;  // resident single-movie state is rebuilt below
  hipEvent_t* ev = c->gev;
  for (int i = 0; i < 6; i++)
    if (!ev[i]) HIP_TRY(hipEventCreate(&ev[i]));
  HIP_TRY(hipEventRecord(ev[0], c->st));

  // 1. spectra (DFAlign.cu:705-824), frame means and unaligned sums
  HIP_TRY(c->means.ensure(sizeof(double) * ftot));
  {
    KScope ks(c, MC_KT_SPECTRA);
    for (size_t i = 0; i < rs.size(); i++) {
      MovieRun* r = rs[i];
      use_group(c, gs[gi_of[i]]);
      const int first = r->p->nStart - r->m->first_frame;
      MC_TRY(run_stage1(c, r->m->frames, r->m, r->p, r->g, first, r->F, r->base));
      if (r->m->frames2)
        MC_TRY(run_stage1(c, r->m->frames2, r->m, r->p, r->g, first, r->F, r->base + r->off2));
      else if (r->off2)  // section-2 corrections of the same frames
        MC_TRY(run_stage1(c, r->m->frames, r->m, r->p, r->g, first, r->F, r->base + r->off2, -1,
                          -1, 2));
    }
  }
  {
    int mo = 0;
    for (size_t i = 0; i < rs.size(); i++) {
      MovieRun* r = rs[i];
      use_group(c, gs[gi_of[i]]);
      const int n = r->g.n;
      double* md = c->means.as<double>() + mo;
      HIP_TRY(mc::launch_dc_gather(slot(c, r->base), c->spec_stride, r->F, 1.0 / ((double)n * n), md,
                                   c->st));
      if (r->out->frame_mean)
        HIP_TRY(hipMemcpyAsync(r->out->frame_mean, md, sizeof(double) * r->F, hipMemcpyDeviceToHost,
                               c->st));
      mo += r->F;
    }
  }
  HIP_TRY(c->image.ensure(sizeof(float) * (size_t)nmax * nmax));
  for (size_t i = 0; i < rs.size(); i++) {
    MovieRun* r = rs[i];
    if (!r->out->rawsum_image && !r->out->disp_raw_fft) continue;
    use_group(c, gs[gi_of[i]]);
    const int n = r->g.n;
    // DFAlign.cu:808-823,853-857: unaligned sum of frames [nStartSum, nEndSum]
    r->sh.assign(2 * (size_t)r->F, NAN);
    for (int f = 0; f < r->F; f++) {
      const int j = f + r->p->nStart;
      if (j >= r->p->nStartSum && j <= r->p->nEndSum) r->sh[2 * f] = r->sh[2 * f + 1] = 0.0;
    }
    MC_TRY(shift_sum(c, r->base, r->F, r->sh, slot(c, r->sum), nullptr, -1));
    if (r->out->disp_raw_fft)  // DFAlign.cu:842-848, 908-911
      MC_TRY(quicklook(c, slot(c, r->sum), n, fft_scale(r->p), r->out->disp_raw_fft, nullptr));
    if (r->out->rawsum_image) {
      MC_TRY(to_image(c, slot(c, r->sum), c->image.as<float>()));
      HIP_TRY(hipMemcpyAsync(r->out->rawsum_image, c->image.p, sizeof(float) * (size_t)n * n,
                             hipMemcpyDeviceToHost, c->st));
      r->rawsum_written = 1;
    }
    HIP_TRY(hipStreamSynchronize(c->st));
  }
  HIP_TRY(hipEventRecord(ev[1], c->st));

  // 2. pairwise shifts of every movie (DFAlign.cu:875-905)
  PairPlan L;
  L.bfactor = bfactor;
  L.box = box;
  L.wnoise = wnoise;
  L.precision = precision;
  std::vector<int> slots;
  for (size_t i = 0; i < rs.size(); i++) {
    MovieRun* r = rs[i];
    PairSpan q;
    q.gi = gi_of[i];
    q.np = r->P;
    L.sp.push_back(q);
    for (int k = 0; k < r->P; k++) {
      slots.push_back(r->base + r->pairs[2 * k]);
      slots.push_back(r->base + r->off2 + r->pairs[2 * k + 1]);
    }
  }
  MC_TRY(pairs_prepare(c, gs, L, slots));
  MC_TRY(pairs_launch(c, gs, L));
  std::vector<double> res(3 * L.np);
  if (L.np)
    HIP_TRY(hipMemcpyAsync(res.data(), c->peaks.p, sizeof(double) * 3 * L.np, hipMemcpyDeviceToHost,
                           c->st));
  HIP_TRY(hipEventRecord(ev[2], c->st));
  {
    size_t po = 0;
    for (MovieRun* r : rs) {
      if (r->out->boxmaps)
        HIP_TRY(hipMemcpyAsync(r->out->boxmaps, c->box.as<float>() + po * box * box,
                               sizeof(float) * (size_t)r->P * box * box, hipMemcpyDeviceToHost, c->st));
      po += r->P;
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    po = 0;
    for (MovieRun* r : rs) {
      // 3. least squares, outlier rejection, least squares (DFAlign.cu:914-948)
      MC_TRY(solve_drift(*r, res.data() + 3 * po));
      po += r->P;
    }
  }
  HIP_TRY(hipEventRecord(ev[3], c->st));

  // 4. iterative alignment to the sum (DFAlign.cu:950-1071), batched
  RoundGraph gr;
  for (;;) {
    // iterating movies, size group by size group
    std::vector<int> key, jobs, islots;
    std::vector<RoundJobs> rj;
    std::vector<double> shifts;
    std::vector<MovieRun*> act;
    PairPlan R;
    R.bfactor = bfactor;
    R.box = box;
    R.wnoise = wnoise;
    R.precision = precision;
    for (size_t g = 0; g < gs.size(); g++) {
      RoundJobs q{(int)g, (int)(jobs.size() / 4), 0};
      for (MovieRun* r : gs[g].rs) {
        if (!r->iterating()) continue;
        act.push_back(r);
        key.insert(key.end(), {(int)g, r->base});
        iter_shifts(*r);
        jobs.insert(jobs.end(), {r->base, r->F, (int)(shifts.size() / 2), r->sum});
        shifts.insert(shifts.end(), r->sh.begin(), r->sh.end());
        for (int k = 0; k < r->F; k++) {
          islots.push_back(r->sum);  // GPUShiftCC(dfft=frame, dsum=sum): sum * conj(frame)
          islots.push_back(r->base + k);
        }
        PairSpan s;
        s.gi = (int)g;
        s.np = r->F;
        R.sp.push_back(s);
        q.nj++;
      }
      if (q.nj) rj.push_back(q);
    }
    if (act.empty()) break;
    const bool fresh = !(gr.exec && gr.key == key);
    if (fresh) {
      MC_TRY(pairs_prepare(c, gs, R, islots));
      HIP_TRY(c->jobs.ensure(sizeof(int) * jobs.size()));
      HIP_TRY(hipMemcpyAsync(c->jobs.p, jobs.data(), sizeof(int) * jobs.size(),
                             hipMemcpyHostToDevice, c->st));
      HIP_TRY(c->shifts.ensure(sizeof(double) * shifts.size()));
    } else {
      R.np = islots.size() / 2;
    }
    HIP_TRY(hipMemcpyAsync(c->shifts.p, shifts.data(), sizeof(double) * shifts.size(),
                           hipMemcpyHostToDevice, c->st));
    if (use_graph) {
      MC_TRY(round_graph(c, gr, key, gs, rj, R));
      KScope ks(c, MC_KT_ROUNDS);
      HIP_TRY(hipGraphLaunch(gr.exec, c->st));
      c->graph_replays++;
    } else {
      gr.key = key;  // buffers uploaded for this set
      MC_TRY(round_launch(c, gs, rj, R));
    }
    const size_t np = islots.size() / 2;
    res.resize(3 * np);
    HIP_TRY(hipMemcpyAsync(res.data(), c->peaks.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost,
                           c->st));
    size_t po = 0;
    for (MovieRun* r : act) {
      if (r->out->iter_boxmaps)
        HIP_TRY(hipMemcpyAsync(r->out->iter_boxmaps + (size_t)r->it * r->F * box * box,
                               c->box.as<float>() + po * box * box,
                               sizeof(float) * (size_t)r->F * box * box, hipMemcpyDeviceToHost, c->st));
      po += r->F;
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    po = 0;
    for (MovieRun* r : act) {
      iter_update(*r, res.data() + 3 * po);
      po += r->F;
    }
  }
  HIP_TRY(hipEventRecord(ev[4], c->st));

  // 5. final phase-shift-and-sum per movie (DFAlign.cu:1114-1292)
  for (size_t i = 0; i < rs.size(); i++) {
    use_group(c, gs[gi_of[i]]);
    MC_TRY(final_sum(c, *rs[i]));
  }
  HIP_TRY(hipEventRecord(ev[5], c->st));
  HIP_TRY(hipEventSynchronize(ev[5]));
  tm->spectra = elapsed(ev[0], ev[1]);
  tm->pairs = elapsed(ev[1], ev[2]);
  tm->solve = elapsed(ev[2], ev[3]);
  tm->iter = elapsed(ev[3], ev[4]);
  tm->sum = elapsed(ev[4], ev[5]);
  tm->total = elapsed(ev[0], ev[5]);
  if (rs.size() == 1) {
    // the movie stays resident for the stage entry points
    c->nframe = rs[0]->F;
    c->nslots = rs[0]->nslot;
    c->off2 = rs[0]->off2;
    c->sum_slot = rs[0]->sum;
  }
  return MC_OK;
}

static void fill_report(const MovieRun& r, const GroupTimes& tm) {
  mc_report* rep = r.rep;
  if (!rep) return;
  memset(rep, 0, sizeof *rep);
  rep->n = r.g.n;
  rep->nsamUnbin = r.g.nsu;
  rep->nframe = r.F;
  rep->npairs = r.P;
  rep->ngood = r.ng;
  rep->niter = r.it;
  rep->converged = r.converged;
  rep->rmsd = r.rmsd;
  rep->sum2_written = r.sum2_written;
  rep->rawsum_written = r.rawsum_written;
  rep->sharp_j0 = r.sj0;
  rep->sharp_j1 = r.sj1;
  rep->grad_accepted = r.gacc;
  rep->ms_spectra = tm.spectra;
  rep->ms_pairs = tm.pairs;
  rep->ms_solve = tm.solve;
  rep->ms_iter = tm.iter;
  rep->ms_sum = tm.sum;
  rep->ms_total = tm.total;
}

extern "C" int mc_disp_short(mc_context ctx, const float* buf, int size, short* out) {
  if (!ctx || !buf || !out || size <= 0) return set_err(MC_ERR_ARG, "bad argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(ctx->disp.ensure(sizeof(short) * (size_t)size));
  short* d = ctx->disp.as<short>();
  HIP_TRY(mc::launch_disp_short(buf, size, d, ctx->st));
  HIP_TRY(hipMemcpyAsync(out, d, sizeof(short) * (size_t)size, hipMemcpyDeviceToHost, ctx->st));
  HIP_TRY(hipStreamSynchronize(ctx->st));
  return MC_OK;
}

extern "C" int mc_fsc_bins(int n) {
  if (n < 4 || n > 8192 || n % 2) return set_err(MC_ERR_ARG, "FSC size %d must be even, in [4, 8192]", n);
  return mc::fsc_bins(n, nullptr);
}

extern "C" int mc_fsc_table(mc_context ctx, const void* a, const void* b, int n, double dx,
                            double dy, double* out) {
  if (!ctx || !a || !b || !out) return set_err(MC_ERR_ARG, "NULL argument");
  const int nbox = mc_fsc_bins(n);
  if (nbox < 0) return nbox;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t nk = (size_t)nbox * 6;
  HIP_TRY(ctx->fsc.ensure(sizeof(double) * nk * (mc::fsc_grid(n) + 2)));
  MC_TRY(fsc_sums(ctx, static_cast<const float2*>(a), static_cast<const float2*>(b), n, 1, dx, dy,
                  ctx->fsc.as<double>()));
  std::vector<double> h(nk);
  HIP_TRY(hipMemcpyAsync(h.data(), ctx->fsc.p, sizeof(double) * nk, hipMemcpyDeviceToHost,
                         ctx->st));
  HIP_TRY(hipStreamSynchronize(ctx->st));
  fsc_ratios(h.data(), nbox, out, 2);
  return MC_OK;
}

extern "C" int mc_align_movie(mc_context ctx, const mc_movie_desc* m, const mc_params* p,
                              mc_outputs* out, mc_report* rep) {
  if (!ctx || !m || !p) return set_err(MC_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  mc_outputs none;
  memset(&none, 0, sizeof none);
  MovieRun r;
  r.m = m;
  r.p = p;
  r.out = out ? out : &none;
  r.rep = rep;
  MC_TRY(prepare_run(r));
  std::vector<MovieRun*> rs{&r};
  GroupTimes tm;
  MC_TRY(align_group(ctx, rs, graphs_enabled(), &tm));
  fill_report(r, tm);
  return MC_OK;
}

// Device memory a sub-batch may hold for resident spectra and its pair
// stage: MC_BATCH_MEM_MB, or 60% of the free memory at the call.
static size_t batch_budget(mc_context_s* c) {
  if (getenv("MC_BATCH_MEM_MB")) return (size_t)atoll(getenv("MC_BATCH_MEM_MB")) << 20;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)8 << 30;
  return fr / 10 * 6 + c->spec.cap + c->tbuf.cap + c->box.cap;
}

extern "C" int mc_align_batch(mc_context ctx, const mc_batch_item* items, int count,
                              mc_batch_report* brep) {
  if (!ctx || !items || count < 1) return set_err(MC_ERR_ARG, "bad batch arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<mc_outputs> none(count);
  memset(none.data(), 0, sizeof(mc_outputs) * count);
  std::vector<MovieRun> runs(count);
  for (int i = 0; i < count; i++) {
    if (!items[i].movie || !items[i].params) return set_err(MC_ERR_ARG, "batch item %d: NULL", i);
    runs[i].m = items[i].movie;
    runs[i].p = items[i].params;
    runs[i].out = items[i].out ? items[i].out : &none[i];
    runs[i].rep = items[i].rep;
    const int rc = prepare_run(runs[i]);
    if (rc != MC_OK) {
      const std::string msg = g_err;
      return set_err(rc, "batch item %d: %s", i, msg.c_str());
    }
  }
  // groups: same box, B-factor, precision form and noise window (any
  // transform sizes), larger transforms first; each group split into
  // sub-batches whose spectra fit the budget
  std::vector<int> order(count);
  for (int i = 0; i < count; i++) order[i] = i;
  auto group_less = [&](int a, int b) {
    const MovieRun &x = runs[a], &y = runs[b];
    if (x.box != y.box) return x.box < y.box;
    if (x.p->bfactor != y.p->bfactor) return x.p->bfactor < y.p->bfactor;
    if (x.p->precision != y.p->precision) return x.p->precision < y.p->precision;
    return x.wnoise < y.wnoise;
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (group_less(a, b) || group_less(b, a)) return group_less(a, b);
    return runs[a].g.n > runs[b].g.n;
  });
  const size_t budget = batch_budget(ctx);
  const int cap0 = ctx->graph_captures, rep0 = ctx->graph_replays;
  int groups = 0, subs = 0;
  double ms = 0;
  size_t i = 0;
  while (i < order.size()) {
    size_t j = i + 1;
    while (j < order.size() && !group_less(order[i], order[j]) && !group_less(order[j], order[i])) j++;
    groups++;
    size_t k = i;
    while (k < j) {
      std::vector<MovieRun*> rs;
      size_t bytes = 0;
      while (k < j) {
        const MovieRun& r = runs[order[k]];
        // resident spectra, the pair stage's T columns and box maps
        const size_t H = (size_t)(r.g.n / 2 + 1);
        const size_t need = sizeof(float2) * H * r.g.n * r.nslot +
                            sizeof(float2) * H * r.box * r.P + sizeof(float) * r.box * r.box * r.P;
        if (!rs.empty() && bytes + need > budget) break;
        rs.push_back(&runs[order[k]]);
        bytes += need;
        k++;
      }
      GroupTimes tm;
      MC_TRY(align_group(ctx, rs, graphs_enabled(), &tm));
      for (MovieRun* r : rs) fill_report(*r, tm);
      ms += tm.total;
      subs++;
    }
    i = j;
  }
  if (count > 1) ctx->nframe = 0;  // no single resident movie after a batch
  if (brep) {
    brep->groups = groups;
    brep->sub_batches = subs;
    brep->graph_captures = ctx->graph_captures - cap0;
    brep->graph_replays = ctx->graph_replays - rep0;
    brep->ms_total = ms;
  }
  return MC_OK;
}

// ============================================================ autotuner
// Launch shapes are chosen by timing the candidates on the caller's movie:
// the row-pass shape and frames per stage-1 launch (all frames of the movie
// through stage 1), then the pair column and row kernels (all pairs of the
// movie).  Each candidate runs once untimed, then `reps` times between HIP
// events on the context's stream; the minimum is kept.
static const int kChunks[4] = {1, 2, 4, 8};

static int device_arch(int dev, char* arch, size_t len, int* cus) {
  hipDeviceProp_t pr;
  HIP_TRY(hipGetDeviceProperties(&pr, dev));
  snprintf(arch, len, "%s", pr.gcnArchName);
  char* colon = strchr(arch, ':');
  if (colon) *colon = 0;
  *cus = pr.multiProcessorCount;
  return MC_OK;
}

template <class F>
static int time_min(mc_context_s* c, int reps, float* ms, F&& run) {
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  int rc = run();  // untimed
  float best = 1e30f;
  for (int r = 0; rc == MC_OK && r < reps; r++) {
    if (hipEventRecord(a, c->st) != hipSuccess) rc = set_err(MC_ERR_DEVICE, "event record");
    if (rc == MC_OK) rc = run();
    if (rc == MC_OK && hipEventRecord(b, c->st) != hipSuccess) rc = set_err(MC_ERR_DEVICE, "event record");
    if (rc == MC_OK && hipEventSynchronize(b) != hipSuccess) rc = set_err(MC_ERR_DEVICE, "event sync");
    if (rc == MC_OK) best = std::min(best, elapsed(a, b));
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = best;
  return rc;
}

static void set_tune(mc_context_s* c, const mc_context_s::Tune& t) {
  for (auto& u : c->tunes)
    if (u.nu == t.nu && u.n == t.n && u.dtype == t.dtype && u.box == t.box &&
        u.precision == t.precision) {
      u = t;
      return;
    }
  c->tunes.push_back(t);
}

extern "C" int mc_autotune(mc_context ctx, const mc_movie_desc* m, const mc_params* p,
                           const char* cache_path, mc_tune_report* rep) {
  if (!ctx || !m || !p) return set_err(MC_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  mc_context_s* c = ctx;
  MovieRun r;
  mc_outputs none;
  memset(&none, 0, sizeof none);
  r.m = m;
  r.p = p;
  r.out = &none;
  MC_TRY(prepare_run(r));
  const int n = r.g.n, F = r.F, box = r.box;
  const size_t stride = (size_t)(n / 2 + 1) * n;
  HIP_TRY(c->spec.ensure(sizeof(float2) * stride * (F + 1)));
  c->n = n;
  c->spec_stride = stride;
  c->spec_off = 0;
  c->nframe = 0;
  const int reps = 3;
  mc_tune_report tr;
  memset(&tr, 0, sizeof tr);
  // ---- stage 1
  const int first = p->nStart - m->first_frame;
  float best = 1e30f;
  int bv = mc::R2C_AUTO, bc = 0;
  for (int cp = 0; cp < 2; cp++)
    for (int v = mc::R2C_R4_T256; v < mc::R2C_VARIANTS; v++) {
      if (v == mc::R2C_R4_T512 && r.g.nsu < 4096) continue;
      for (int k = 0; k < 4; k++) {
        const int ch = kChunks[k];
        if (ch > F && k > 0) continue;
        const int vv = v | (cp ? mc::R2C_COLS_P : 0);
        float ms = 0;
        MC_TRY(time_min(c, reps, &ms, [&] {
          return run_stage1(c, m->frames, m, p, r.g, first, F, 0, vv, ch);
        }));
        tr.ms_stage1[cp][v][k] = ms;
        if (ms < best) {
          best = ms;
          bv = vv;
          bc = ch;
        }
      }
    }
  // ---- radix plan and LDS padding of the per-column pass (canvases with
  // plan variants), with the chosen row shape and chunk
  tr.cols_plan = 0;
  if (!(bv & mc::R2C_COLS_P) && (r.g.nsu == 1024 || r.g.nsu == 2048 || r.g.nsu == 4096 ||
                                 r.g.nsu == 8192)) {
    for (int cpl = 0; cpl < mc::R2C_CPLANS; cpl++) {
      const int vv = (bv & (mc::R2C_COLS_P - 1)) | (cpl << mc::R2C_CPLAN_SHIFT);
      float ms = 0;
      MC_TRY(time_min(c, reps, &ms, [&] {
        return run_stage1(c, m->frames, m, p, r.g, first, F, 0, vv, bc);
      }));
      tr.ms_cols_plan[cpl] = ms;
      if (cpl == 0 || ms < best) {
        best = ms;
        bv = vv;
        tr.cols_plan = cpl;
      }
    }
  }
  // ---- pair stage (spectra of the last stage-1 run are resident)
  std::vector<int> slots(r.pairs);
  MC_TRY(pair_reserve(c, r.P, n, box, p->precision));
  HIP_TRY(hipMemcpyAsync(c->pairs_d.p, slots.data(), sizeof(int) * slots.size(),
                         hipMemcpyHostToDevice, c->st));
  int bcol = mc::PAIR_COLS_AUTO, brow = mc::PAIR_ROWS_AUTO;
  if (p->precision == MC_PREC_FP32) {
    best = 1e30f;
    for (int v = 0; v < mc::PAIR_COLS_VARIANTS; v++) {
      if (v != mc::PAIR_COLS_GENERIC && !(n >= 512 && box <= 128)) continue;
      float ms = 0;
      MC_TRY(time_min(c, reps, &ms, [&] {
        const hipError_t e = mc::launch_pair_cols(slot(c, 0), c->spec_stride,
                                                  c->pairs_d.as<int2>(), r.P, n, p->bfactor, box,
                                                  c->tbuf.as<float2>(), c->tmax.as<float>(),
                                                  p->precision, v, c->st);
        return e == hipSuccess ? MC_OK : set_err(MC_ERR_DEVICE, "pair columns: %s", hipGetErrorString(e));
      }));
      tr.ms_pair_cols[v] = ms;
      if (ms < best) {
        best = ms;
        bcol = v;
      }
    }
    // The default variant is the plan kernel that mc_align_batch shares across
    // movies: it is kept unless another one is more than 2% faster.
    if (tr.ms_pair_cols[mc::PAIR_COLS_AUTO] > 0 && tr.ms_pair_cols[mc::PAIR_COLS_AUTO] <= 1.02f * best)
      bcol = mc::PAIR_COLS_AUTO;
    best = 1e30f;
    for (int v = 0; v < mc::PAIR_ROWS_VARIANTS; v++) {
      if (v != mc::PAIR_ROWS_VALU && box > 96) continue;
      float ms = 0;
      MC_TRY(time_min(c, reps, &ms, [&] {
        const hipError_t e = mc::launch_pair_rows(
            c->tbuf.as<float2>(), c->tmax.as<float>(), r.P, n, box, c->box.as<float>(),
            c->boxpart.as<float>(), p->precision, v, c->st);
        return e == hipSuccess ? MC_OK : set_err(MC_ERR_DEVICE, "pair rows: %s", hipGetErrorString(e));
      }));
      tr.ms_pair_rows[v] = ms;
      if (ms < best) {
        best = ms;
        brow = v;
      }
    }
  }
  HIP_TRY(hipStreamSynchronize(c->st));
  mc_context_s::Tune t{r.g.nsu, n, m->dtype, box, p->precision, bv, bc, bcol, brow};
  set_tune(c, t);
  tr.r2c_variant = bv;
  tr.chunk_frames = bc;
  tr.pair_cols_variant = bcol;
  tr.pair_rows_variant = brow;
  if (rep) *rep = tr;
  if (cache_path && *cache_path) {
    char arch[64];
    int cus = 0;
    MC_TRY(device_arch(c->device, arch, sizeof arch, &cus));
    FILE* fp = fopen(cache_path, "a");
    if (!fp) return set_err(MC_ERR_IO, "cannot append to %s", cache_path);
    fprintf(fp,
            "{\"arch\": \"%s\", \"cus\": %d, \"nu\": %d, \"n\": %d, \"dtype\": %d, \"box\": %d, "
            "\"precision\": %d, \"r2c\": %d, \"chunk\": %d, \"cols\": %d, \"rows\": %d}\n",
            arch, cus, t.nu, t.n, t.dtype, t.box, t.precision, t.r2c, t.chunk, t.cols, t.rows);
    fclose(fp);
  }
  return MC_OK;
}

static bool json_int(const char* line, const char* key, int* v) {
  char pat[40];
  snprintf(pat, sizeof pat, "\"%s\":", key);
  const char* q = strstr(line, pat);
  return q && sscanf(q + strlen(pat), " %d", v) == 1;
}

extern "C" int mc_tune_load(mc_context ctx, const char* cache_path, int* loaded) {
  if (!ctx || !cache_path) return set_err(MC_ERR_ARG, "NULL argument");
  char arch[64];
  int cus = 0;
  MC_TRY(device_arch(ctx->device, arch, sizeof arch, &cus));
  FILE* fp = fopen(cache_path, "r");
  if (!fp) return set_err(MC_ERR_IO, "cannot open %s", cache_path);
  char line[512], want[96];
  snprintf(want, sizeof want, "\"arch\": \"%s\"", arch);
  int cnt = 0;
  while (fgets(line, sizeof line, fp)) {
    mc_context_s::Tune t;
    int lc = 0;
    if (!strstr(line, want) || !json_int(line, "cus", &lc) || lc != cus) continue;
    if (json_int(line, "nu", &t.nu) && json_int(line, "n", &t.n) &&
        json_int(line, "dtype", &t.dtype) && json_int(line, "box", &t.box) &&
        json_int(line, "precision", &t.precision) && json_int(line, "r2c", &t.r2c) &&
        json_int(line, "chunk", &t.chunk) && json_int(line, "cols", &t.cols) &&
        json_int(line, "rows", &t.rows) && t.r2c >= 0 &&
        (t.r2c & (mc::R2C_COLS_P - 1)) < mc::R2C_VARIANTS &&
        (t.r2c >> mc::R2C_CPLAN_SHIFT) < mc::R2C_CPLANS &&
        t.cols >= 0 && t.cols < mc::PAIR_COLS_VARIANTS && t.rows >= 0 &&
        t.rows < mc::PAIR_ROWS_VARIANTS && t.chunk >= 0) {
      set_tune(ctx, t);  // later lines win
      cnt++;
    }
  }
  fclose(fp);
  if (loaded) *loaded = cnt;
  return MC_OK;
}
