// mc_level0.cpp -- entry points of the cufunc.h seam (reference
// src-1125/cufunc.h:16-49, cufunc.cu:30-497) on the padded in-place
// spectrum layout.  All device work runs on the null stream of the current
// device and is synchronous with respect to the host where the reference
// calls were (cudaMemcpy), asynchronous where they were kernel launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mc_abi.h"
#include "mc_kernels.h"

// Records the text returned by mc_last_error() (defined in mc_align.cpp).
int mc_internal_set_error(int code, const char* msg);

static int l0_err(int code, const char* what, hipError_t e) {
  char buf[384];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  return mc_internal_set_error(code, buf);
}

#define L0_TRY(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return l0_err(e_ == hipErrorOutOfMemory ? MC_ERR_NOMEM : MC_ERR_DEVICE, #expr, e_); \
  } while (0)

// Any even size in [64, 8192]: Stockham plans for powers of two and
// 2^k * {3, 5, 15}, chirp-z transforms otherwise.
static bool size_ok(int n) { return n >= 64 && n <= 8192 && n % 2 == 0; }

struct mc_plan_s {
  int n;
  int inverse;
  float2* spec;  // [kx][ky] working copy, (n/2+1)*n
  float2* rt;    // row-pass scratch, (n/2+1)*n
  float* rowsum;
  float* fill;
};

extern "C" int mc_init_gpu(int gpu) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    return mc_internal_set_error(MC_ERR_DEVICE, "no GPU device");
  }
  if (gpu < 0 || gpu >= n) {
    char buf[128];
    snprintf(buf, sizeof buf, "GPU ID %d is out of range(%d). Abort.", gpu, n);
    return mc_internal_set_error(MC_ERR_ARG, buf);
  }
  L0_TRY(hipSetDevice(gpu));
  return MC_OK;
}

extern "C" int mc_reset_gpu(void) {
  L0_TRY(hipDeviceReset());
  return MC_OK;
}

extern "C" int mc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int mc_mem_check(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  L0_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return MC_OK;
}

extern "C" int mc_mem_alloc(void** buf, size_t bytes) {
  L0_TRY(hipMalloc(buf, bytes));
  return MC_OK;
}
extern "C" int mc_mem_zero(void* buf, size_t bytes) {
  L0_TRY(hipMemset(buf, 0, bytes));
  return MC_OK;
}
extern "C" int mc_mem_free(void* buf) {
  L0_TRY(hipFree(buf));
  return MC_OK;
}
extern "C" int mc_mem_h2d(void* dst, const void* src, size_t bytes) {
  L0_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return MC_OK;
}
extern "C" int mc_mem_d2h(void* dst, const void* src, size_t bytes) {
  L0_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return MC_OK;
}
extern "C" int mc_mem_d2d(void* dst, const void* src, size_t bytes) {
  L0_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return MC_OK;
}

// cufunc.cu:147-176: rows [0, dst_n/2) and the last dst_n/2 rows.
static int bin_copy(float* dst, const float* src, int dst_n, int src_n, hipMemcpyKind kind) {
  if (dst_n < 2 || src_n < dst_n) {
    return mc_internal_set_error(MC_ERR_ARG, "bin sizes");
  }
  const size_t w = sizeof(float) * (dst_n + 2);
  const size_t dp = sizeof(float) * (dst_n + 2), sp = sizeof(float) * (src_n + 2);
  const int h = dst_n / 2;
  L0_TRY(hipMemcpy2D(dst, dp, src, sp, w, h, kind));
  L0_TRY(hipMemcpy2D(dst + (size_t)(dst_n - h) * (dst_n + 2), dp,
                     src + (size_t)(src_n - h) * (src_n + 2), sp, w, h, kind));
  return MC_OK;
}
extern "C" int mc_mem_bin_d2d(float* dst, const float* src, int dst_n, int src_n) {
  return bin_copy(dst, src, dst_n, src_n, hipMemcpyDeviceToDevice);
}
extern "C" int mc_mem_bin_d2h(float* dst, const float* src, int dst_n, int src_n) {
  return bin_copy(dst, src, dst_n, src_n, hipMemcpyDeviceToHost);
}

static int make_plan(mc_plan* plan, int n, int inverse) {
  if (!plan) return mc_internal_set_error(MC_ERR_ARG, "plan is NULL");
  if (!size_ok(n)) return mc_internal_set_error(MC_ERR_ARG, "FFT size must be even, in [64, 8192]");
  mc_plan_s* p = new mc_plan_s();
  p->n = n;
  p->inverse = inverse;
  const size_t per = (size_t)(n / 2 + 1) * n;
  hipError_t e = hipMalloc(&p->spec, sizeof(float2) * per);
  if (e == hipSuccess) e = hipMalloc(&p->rt, sizeof(float2) * std::max(per, mc::rt_frame_elems(n, n)));
  if (e == hipSuccess) e = hipMalloc(&p->rowsum, sizeof(float) * n);
  if (e == hipSuccess) e = hipMalloc(&p->fill, sizeof(float));
  if (e != hipSuccess) {
    (void)hipFree(p->spec);
    (void)hipFree(p->rt);
    (void)hipFree(p->rowsum);
    delete p;
    return l0_err(MC_ERR_NOMEM, "plan buffers", e);
  }
  if (!mc::twiddles(n)) {
    delete p;
    return mc_internal_set_error(MC_ERR_NOMEM, "twiddle table");
  }
  *plan = p;
  return MC_OK;
}
extern "C" int mc_fft_plan(mc_plan* plan, int n) { return make_plan(plan, n, 0); }
extern "C" int mc_ifft_plan(mc_plan* plan, int n) { return make_plan(plan, n, 1); }

extern "C" int mc_fft_destroy(mc_plan plan) {
  if (!plan) return MC_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(plan->spec);
  (void)hipFree(plan->rt);
  (void)hipFree(plan->rowsum);
  (void)hipFree(plan->fill);
  delete plan;
  return MC_OK;
}

// GPUFFT2d: in-place R2C of an n x n image stored with pitch n+2.
extern "C" int mc_fft2d(float* buf, mc_plan plan) {
  if (!buf || !plan) return mc_internal_set_error(MC_ERR_ARG, "NULL argument");
  const int n = plan->n;
  mc::R2CGeom g;
  g.src = reinterpret_cast<const char*>(buf);
  g.dtype = mc::DT_FLOAT32;
  g.row_pitch = sizeof(float) * (n + 2);
  g.frame_pitch = g.row_pitch * n;
  g.offx = g.offy = 0;
  g.w = g.h = n;
  g.nu = n;
  g.n = n;
  g.pad_mode = 0;
  L0_TRY(mc::launch_frames_to_spectra(g, 0, 1, plan->spec, (size_t)(n / 2 + 1) * n, plan->rt,
                                      plan->rowsum, plan->fill, nullptr, mc::R2C_AUTO,
                                      nullptr));
  L0_TRY(mc::launch_spec_to_padded(plan->spec, n, reinterpret_cast<float2*>(buf), nullptr));
  return MC_OK;
}

// GPUIFFT2d: in-place C2R (unnormalised) of the padded layout.
extern "C" int mc_ifft2d(float* buf, mc_plan plan) {
  if (!buf || !plan) return mc_internal_set_error(MC_ERR_ARG, "NULL argument");
  const int n = plan->n;
  L0_TRY(mc::launch_padded_to_spec(reinterpret_cast<const float2*>(buf), n, plan->spec, nullptr));
  L0_TRY(mc::launch_spectrum_to_image(plan->spec, n, 1.0f, buf, n + 2, plan->rt, nullptr));
  return MC_OK;
}

extern "C" int mc_sync(void) {
  L0_TRY(hipDeviceSynchronize());
  L0_TRY(hipGetLastError());
  return MC_OK;
}

extern "C" int mc_add(float* dst, const float* src, size_t nfloats) {
  L0_TRY(mc::launch_add(dst, src, nfloats, nullptr));
  return MC_OK;
}
extern "C" int mc_multiply_num(float* dst, float num, size_t nfloats) {
  L0_TRY(mc::launch_scale(dst, num, nfloats, nullptr));
  return MC_OK;
}
extern "C" int mc_fft_log_modulus(float* dmod, const float* dfft, int n, float scale) {
  L0_TRY(mc::launch_log_modulus(dmod, reinterpret_cast<const float2*>(dfft), n, scale, nullptr));
  return MC_OK;
}

// MkPosList(MASK*, nsam, bfactor) (cufunc.cu:309-341), host table.
extern "C" int mc_make_pos_list(mc_mask* list, int n, float bfactor) {
  if (!list || n < 2) return mc_internal_set_error(MC_ERR_ARG, "bad pos list arguments");
  const int hb = n / 2 + 1, h = n / 2;
  const float m = (float)(-0.5 * bfactor / n / n);
  int count = 0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < hb; i++) {
      const int y = j < h ? j : j - n;
      const int r2 = i * i + y * y;
      const float z = std::exp(m * (float)r2);
      list[count].x = i;
      list[count].y = y;
      list[count].z = ((i + y) % 2 == 0) ? z : -z;
      count++;
    }
  return MC_OK;
}

static_assert(sizeof(mc_mask) == sizeof(mc::MaskElem), "MASK layout");

// GPUShiftCC / GPUShift (cufunc.cu:362-394): shx = sx*2.0*PI/nsam in float
extern "C" int mc_shift_cc(float* dfft, const float* dsum, const mc_mask* dposlist, float sx,
                           float sy, int n) {
  const float shx = (float)(sx * 2.0 * 3.141592653589793 / n);
  const float shy = (float)(sy * 2.0 * 3.141592653589793 / n);
  L0_TRY(mc::launch_mask_shift_cc(reinterpret_cast<float2*>(dfft),
                                  reinterpret_cast<const float2*>(dsum),
                                  reinterpret_cast<const mc::MaskElem*>(dposlist), shx, shy, n,
                                  nullptr));
  return MC_OK;
}
extern "C" int mc_shift(float* dfft, const mc_mask* dposlist, float sx, float sy, int n) {
  const float shx = (float)(sx * 2.0 * 3.141592653589793 / n);
  const float shy = (float)(sy * 2.0 * 3.141592653589793 / n);
  L0_TRY(mc::launch_mask_shift(reinterpret_cast<float2*>(dfft),
                               reinterpret_cast<const mc::MaskElem*>(dposlist), shx, shy, n,
                               nullptr));
  return MC_OK;
}

// FindShift (cufunc.cu:396-497): box read-out to the host buffer plus the
// FP64 peak search on the device.
extern "C" int mc_find_shift(const float* dsrc, int n, float* hboxmap, int box, float* sx,
                             float* sy, float* cc, int wnoise) {
  if (!dsrc || box < 16 || box > n) return mc_internal_set_error(MC_ERR_ARG, "bad box");
  const int ori = (n - box) / 2;
  float* dbox = nullptr;
  double* dres = nullptr;
  L0_TRY(hipMalloc(&dbox, sizeof(float) * (size_t)box * box));
  hipError_t e = hipMalloc(&dres, sizeof(double) * mc::find_peaks_out_elems(1));
  if (e == hipSuccess)
    e = hipMemcpy2D(dbox, sizeof(float) * box, dsrc + (size_t)ori * (n + 2) + ori,
                    sizeof(float) * (n + 2), sizeof(float) * box, box, hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = mc::launch_find_peaks(dbox, 1, box, n, wnoise, dres, nullptr);
  double r[3] = {0, 0, 0};
  if (e == hipSuccess) e = hipMemcpy(r, dres, sizeof r, hipMemcpyDeviceToHost);
  if (e == hipSuccess && hboxmap)
    e = hipMemcpy(hboxmap, dbox, sizeof(float) * (size_t)box * box, hipMemcpyDeviceToHost);
  (void)hipFree(dbox);
  (void)hipFree(dres);
  if (e != hipSuccess) return l0_err(MC_ERR_DEVICE, "FindShift", e);
  if (sx) *sx = (float)r[0];
  if (sy) *sy = (float)r[1];
  if (cc) *cc = (float)r[2];
  return MC_OK;
}
