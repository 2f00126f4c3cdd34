// mc_kernels.h -- host-side launch wrappers of the HIP kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mc {

// Grid of a persistent kernel on the current device: CUs x resident
// workgroups per CU (occupancy query), a multiple of 8 so that every XCD
// gets the same share (workgroup b runs on XCD b % 8), at least 8.
template <class K> inline int persistent_grid(K kernel, int nt) {
  int dev = 0, cus = 256, per = 1;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, nt, 0) != hipSuccess || per < 1)
    per = 1;
  const int g = (cus * per) / 8 * 8;
  return g < 8 ? 8 : g;
}
// The same, with at most one workgroup per item (rounded up to 8).
template <class K> inline int resident_grid(K kernel, int nt, int nitems) {
  const int g = persistent_grid(kernel, nt), need = (nitems + 7) / 8 * 8;
  return need < 8 ? 8 : (g < need ? g : need);
}

// Element type codes of frame data (match MC_DTYPE_* in include/mc_abi.h).
enum { DT_INT8 = 0, DT_INT16 = 1, DT_FLOAT32 = 2, DT_UINT16 = 6, DT_UINT8 = 100 };

// Geometry of one stage-1 launch (frames -> spectra), see mc_fft.hip.
struct R2CGeom {
  const char* src;      // device frames
  int dtype;
  size_t row_pitch;     // bytes
  size_t frame_pitch;   // bytes
  int offx, offy;       // region origin in the frame
  int w, h;             // data region (<= nu); the rest of the canvas is fill
  int nu;               // canvas (unbinned) size, even
  int n;                // binned size (<= nu), even
  int pad_mode;         // 1: canvas outside the region holds the frame mean
  int vec = 0;          // set by the launcher: element pairs are 2-element aligned
  // corrections of the raw frame (frame pixel (X, Y), see mc_corrections):
  const float* dark = nullptr;  // nx*ny or nullptr
  const float* gain = nullptr;  // nx*ny or nullptr
  float xs = 1.f, ys = 1.f;     // nearest-neighbour stretch when either < 1
  int nx = 0;                   // frame width (reference row pitch)
};

// Row-pass scratch of one frame (float2 elements): the kept bins kx in
// [0, n/2] are stored in groups of four, RT[kx/4][y][kx%4], so that four
// bins of four consecutive rows fill one 128-byte line.
inline size_t rt_frame_elems(int n, int nu) { return (size_t)((n / 2 + 1 + 3) / 4) * 4 * nu; }

// Twiddle table e^{+2 pi i e/n}, e in [0, n), computed in double on the host.
const float2* twiddles(int n);

// Stage 1: spectra[f] (layout [kx][ky], kx in [0, n/2], ky in [0, n)) of
// nframes frames starting at frame index f0.  rt: scratch of
// nframes*rt_frame_elems(n, nu) float2; rowsum/fill: scratch floats
// (nframes*nu, nframes).
// Row-pass shapes (rows per workgroup x threads); R2C_AUTO picks 4 x 512 on
// canvases >= 4096 and 4 x min(256, 2nu/16) below.
// R2C_COLS_P (or-ed in) selects the persistent column pass with register
// twiddles instead of one workgroup per column.
// Bits R2C_CPLAN_SHIFT.. (v < R2C_CPLANS) select the radix plan (v / 2: 0
// balanced, 1 reversed order, 2 radix-8 chain) and the LDS padding (v % 2:
// 0 one float2 per 16, 1 none) of the per-column pass; plans 1 and 2 exist
// for canvases 1024 to 8192 (others run plan 0).
enum { R2C_AUTO = 0, R2C_R4_T256 = 1, R2C_R2_T256 = 2, R2C_R4_T512 = 3, R2C_VARIANTS = 4,
       R2C_COLS_P = 8, R2C_CPLAN_SHIFT = 4, R2C_CPLANS = 6 };
hipError_t launch_frames_to_spectra(const R2CGeom& g, int f0, int nframes, float2* spectra,
                                    size_t spec_stride, float2* rt, float* rowsum,
                                    float* fill, const float2* padP, int variant, hipStream_t st);

// Per-frame mean of the data region from the row sums (pad geometry).
hipError_t launch_fill(const float* rowsum, int nu, int h, int w, float* fill, int nframes,
                       hipStream_t st);

// Transforms of sizes without a Stockham plan (any even size in [64, 8192]):
// Bluestein / chirp-z on power-of-two transforms (mc_czt.hip).  The stage-1
// and sum-image entry points above dispatch to them by size; the pair
// column pass computes only the box rows of each column inverse.
int czt_len(int need);
hipError_t launch_frames_to_spectra_czt(const R2CGeom& g, int f0, int nframes, float2* spectra,
                                        size_t spec_stride, float2* rt, float* rowsum,
                                        float* fill, const float2* padP, hipStream_t st);
hipError_t launch_spectrum_to_image_czt(const float2* spec, int n, float scale, float* img,
                                        size_t img_pitch, float2* tmp, hipStream_t st);
hipError_t launch_pair_cols_czt(const float2* spectra, size_t spec_stride, const int2* pairs,
                                int npairs, const float* wxy, int n, int box, float2* tbuf,
                                hipStream_t st);

// Inverse 2-D transform of one spectrum ([kx][ky] layout, size n) to an
// n x n real image scaled by `scale` (C2R with the DC/Nyquist projection).
// tmp: scratch of rt_frame_elems(n, n) float2.
hipError_t launch_spectrum_to_image(const float2* spec, int n, float scale, float* img,
                                    size_t img_pitch, float2* tmp, hipStream_t st);

// Pair stage: for each pair p, T[p][kx][r] (column pass) then the box map
// C[p][r][c] (row pass).  pairs: device int2 (a, b) indices into `spectra`
// (frame slots of stride spec_stride); operand a is kept, b conjugated.
// precision 0: FP32 (vector ALUs; row pass on FP32 MFMA); 1: the first DFT
// stage of the column pass and the row pass as FP16 hi/lo MFMA products
// (tmax: npairs*(n/2+1) floats of column maxima, written by the column pass
// and read by the row pass).
// Plain-form variants (the hi/lo form has one kernel per pass):
//   column pass: AUTO = the M = 128 persistent kernel (n >= 512, box <= 128),
//   M128_PF = the same with next-item register prefetch (fewer resident
//   workgroups), GENERIC = the M-generic persistent kernel, M128_W4 = M = 128
//   compiled for 4 waves per SIMD (128 VGPRs); row pass: AUTO = FP32 MFMA
//   streaming each wave's kx range from memory (box <= 96; from n = 2048
//   the kx range is also split over workgroups, a count set by n alone, whose
//   partial maps go to `part`, pair_rows_part_elems(npairs, n, box) floats,
//   and are added in order), VALU = register-tiled vector ALU kernel, MFMA8 / MFMA4 = FP32
//   MFMA staged through LDS with 8 / 4 waves per (pair, 32-row block).
enum { PAIR_COLS_AUTO = 0, PAIR_COLS_M128_PF = 1, PAIR_COLS_GENERIC = 2, PAIR_COLS_M128_W4 = 3,
       PAIR_COLS_VARIANTS = 4 };
enum { PAIR_ROWS_AUTO = 0, PAIR_ROWS_VALU = 1, PAIR_ROWS_MFMA8 = 2, PAIR_ROWS_MFMA4 = 3,
       PAIR_ROWS_VARIANTS = 4 };
bool pair_form_supported(int precision, int n, int box);
hipError_t launch_pair_cols(const float2* spectra, size_t spec_stride, const int2* pairs,
                            int npairs, int n, float bfactor, int box, float2* tbuf,
                            float* tmax, int precision, int variant, hipStream_t st);
// One movie of a batched column pass: its (pair, kx) items are the global
// indices [item0, item0 + npairs (n/2+1)), kx major; plans are listed in
// increasing item0.
struct PairColsPlan {
  const float2* spectra;  // the movie's slot 0
  const int2* pairs;      // movie-local slot pairs
  const float* wxy;       // pair_weights(n, B-factor)
  const float2* tw;       // twiddles(n)
  float2* tbuf;           // T columns of the movie's pairs, [pair][kx][box]
  unsigned long long spec_stride;
  long long item0;
  int npairs, box;
  int body;               // pair_cols_plan_body(n, box, variant)
};
// The plan-kernel body launch_pair_cols(n, box, variant) runs at precision 0:
// the default variant (PAIR_COLS_AUTO) of the M = 128 decimation at n in
// [512, 3840] and box <= 128; -1 for other variants and sizes, which keep
// their own kernels (the default at 4096 is the 4-wave M = 128 kernel).
int pair_cols_plan_body(int n, int box, int variant);
// Register class of a body: 0 runs at 4 workgroups per CU, 1 at 2.  A plan
// table holds the movies of one class; each class is its own launch sequence.
int pair_cols_plan_class(int body);
const float* pair_weights(int n, float bfactor);
hipError_t launch_pair_cols_plans(const PairColsPlan* plans, int nplans, long long items, int cls,
                                  hipStream_t st);
size_t pair_rows_part_elems(int npairs, int n, int box);
hipError_t launch_pair_rows(const float2* tbuf, const float* tmax, int npairs, int n, int box,
                            float* boxmaps, float* part, int precision, int variant,
                            hipStream_t st);

// Peaks: integer peak + FP64 16x16 -> 512x512 sub-pixel search per box map.
// out: 3 doubles per pair (sx, sy, cc), followed by per-slice scratch when
// the maps of few pairs are split over several workgroups:
// find_peaks_out_elems(npairs) doubles in all.  With `part` the scratch
// (find_peaks_out_elems(npairs) - 3 npairs doubles) is there instead.
size_t find_peaks_out_elems(int npairs);
hipError_t launch_find_peaks(const float* boxmaps, int npairs, int box, int n, int wnoise,
                             double* out, hipStream_t st, double* part = nullptr);

// Phase-shift-and-sum: sum[kx][ky] = sum_f include_f * S_f[kx][ky] *
// exp(i 2 pi (sx_f kx + sy_f ky')/n).  shifts: device double2 per frame;
// frames with a NaN x shift are skipped.  If sum2 != nullptr, frames
// f <= f2_last also accumulate into sum2.
hipError_t launch_shift_sum(const float2* spectra, size_t spec_stride, int nframes, int n,
                            const double2* shifts, float2* sum, float2* sum2, int f2_last,
                            hipStream_t st);

// Batched sums: job j = int4(first slot, frame count, first shift index,
// destination slot) in slot units of `spectra`; one launch for all jobs.
hipError_t launch_shift_sum_multi(float2* spectra, size_t spec_stride, int n, const int4* jobs,
                                  int njobs, const double2* shifts, hipStream_t st);

// Even/odd half sums of `count` consecutive spectra: frame f goes to h0
// (f even) or h1 (f odd) after its phase shift (shifts == nullptr: none;
// NaN: frame skipped).
hipError_t launch_half_sums(const float2* spectra, size_t spec_stride, int count, int n,
                            const double2* shifts, float2* h0, float2* h1, hipStream_t st);

// FSC shells (DFAlign.cu:212-297): shell width and count for size n, the grid
// of the partial-sum pass, and the shell sums of the half spectra a, b
// (padded = 1: row-major [ky][kx], else the resident [kx][ky] layout) split
// by the unit direction (dx, dy).  part: fsc_grid(n) * nbox * 6 doubles;
// sums: nbox * 6 doubles, [shell][side][Re(a conj b), |a|^2, |b|^2].
int fsc_bins(int n, int* step);
int fsc_grid(int n);
hipError_t launch_fsc_sums(const float2* a, const float2* b, int n, int padded, double dx,
                           double dy, double* part, double* sums, hipStream_t st);

// Quick-look images (mc_disp.hip).  Centred n x n log-modulus image
// log(1 + |X| scale) of a [kx][ky] spectrum (n % 128 == 0); d-point C2R of
// the d-row crop of a [kx][ky] spectrum of column length ld (divided by div
// when div > 0; tmp: (d/2+1)*d float2); the 8-bit display mapping of `size`
// floats (one workgroup).
hipError_t launch_logmod_image(const float2* spec, int n, float scale, float* img,
                               hipStream_t st);
hipError_t launch_small_c2r(const float2* X, int ld, int d, float div, float2* tmp, float* out,
                            hipStream_t st);
hipError_t launch_disp_short(const float* buf, int size, short* out, hipStream_t st);

// Sharpness J of the shifted sum S (already formed by launch_shift_sum with
// the same shifts) and dJ/ds_f: out[f] = (dJ/dsx_f, dJ/dsy_f), out[F] = (J, 0).
// wk: [n/2+1] column weights (incl. the half-plane factor) then [n] row
// weights, w(kx, ky) = wk[kx] * wk[n/2+1+ky].  part: nchunks*(F+1) double2.
// jonly = 1 computes only J (out[F]; out[0..F) are then not meaningful), with
// the same blocks and summation order as the full pass.
hipError_t launch_sharpness(const float2* spectra, size_t spec_stride, int nframes, int n,
                            const double2* shifts, const float2* S, const float* wk,
                            double2* part, int nchunks, double2* out, int jonly,
                            hipStream_t st);

// Layout conversion between the resident [kx][ky] spectrum and the padded
// row-major [ky][kx] layout (n rows of n/2+1 complex).
hipError_t launch_spec_to_padded(const float2* spec, int n, float2* padded, hipStream_t st);
hipError_t launch_padded_to_spec(const float2* padded, int n, float2* spec, hipStream_t st);

// cufunc.h-compatible kernels on the padded layout with a MASK table
// (x, y, z) per element (reference cufunc.cu:343-394, 178-188).
struct MaskElem {
  int x, y;
  float z;
};
hipError_t launch_mask_shift_cc(float2* dfft, const float2* dsum, const MaskElem* mask,
                                float shx, float shy, int n, hipStream_t st);
hipError_t launch_mask_shift(float2* dfft, const MaskElem* mask, float shx, float shy, int n,
                             hipStream_t st);
hipError_t launch_log_modulus(float* dmod, const float2* dfft, int n, float scale,
                              hipStream_t st);
// First-element gather: out[f] = spectra[f*stride].x * scale (frame means).
hipError_t launch_dc_gather(const float2* spectra, size_t stride, int nframes, double scale,
                            double* out, hipStream_t st);

// Elementwise helpers for the cufunc.h-compatible layer.
hipError_t launch_add(float* dst, const float* src, size_t n, hipStream_t st);
hipError_t launch_scale(float* dst, float s, size_t n, hipStream_t st);

}  // namespace mc
