/*
 * mc_abi.h -- C ABI of the MI355X drift-correction library (libmc.so).
 *
 * Plain C: POD types, device pointers as void*, byte counts as size_t, int
 * status returns (MC_OK = 0).  No HIP, CUDA or cuFFT types appear here.
 *
 * Level 0 replaces the device seam of the reference, src-1125/cufunc.h:16-49
 * (the functions the CDFAlign worker calls, DFAlign.cu:456-1355), one entry
 * point per function with the same argument meaning and order.  Buffers use
 * the same in-place padded spectrum layout as the reference (n rows of n+2
 * floats = n x (n/2+1) complex, row-major).
 *
 * Level 1 is the resident movie pipeline: all spectra stay in HBM, the pair
 * stage computes only the peak-search box of every correlation map, and one
 * call returns the per-frame shifts and the aligned sum.
 */
#ifndef MC_ABI_H
#define MC_ABI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MC_OK 0
#define MC_ERR_ARG -1      /* invalid argument or unsupported size        */
#define MC_ERR_DEVICE -2   /* HIP runtime error (message: mc_last_error) */
#define MC_ERR_NOMEM -3    /* device allocation failed                    */
#define MC_ERR_SOLVE -4    /* least-squares system singular               */
#define MC_ERR_STATE -5    /* call out of order (e.g. no spectra loaded)  */
#define MC_ERR_IO -6       /* tuning cache file could not be read/written */

/* Element types of frame data.  Codes 0, 1, 2, 6 are MRC modes
 * (src-1125/mrcheader.h:8-14); mode 0 is read as signed bytes as the
 * reference does (mrc.cpp:338-348).  100 is unsigned 8-bit (8-bit TIFF). */
#define MC_DTYPE_INT8 0
#define MC_DTYPE_INT16 1
#define MC_DTYPE_FLOAT32 2
#define MC_DTYPE_UINT16 6
#define MC_DTYPE_UINT8 100

/* Text of the last error on the calling thread. */
const char *mc_last_error(void);
/* Library version string. */
const char *mc_version(void);

/* ======================================================================
 * Level 0: cufunc.h seam (src-1125/cufunc.h:16-49)
 * ==================================================================== */
typedef struct mc_plan_s *mc_plan;

/* initGPU (cufunc.cu:30-65): select device; MC_ERR_ARG if out of range. */
int mc_init_gpu(int gpu);
/* ResetGPU (cufunc.cu:67-71). */
int mc_reset_gpu(void);
/* getGPUList (cufunc.cu:80-103): number of devices. */
int mc_device_count(void);
/* GPUMemCheck (cufunc.cu:105-109). */
int mc_mem_check(size_t *free_bytes, size_t *total_bytes);
/* GPUMemAlloc/Zero/Free/H2D/D2H/D2D (cufunc.cu:111-145); size_t counts. */
int mc_mem_alloc(void **buf, size_t bytes);
int mc_mem_zero(void *buf, size_t bytes);
int mc_mem_free(void *buf);
int mc_mem_h2d(void *dst, const void *src, size_t bytes);
int mc_mem_d2h(void *dst, const void *src, size_t bytes);
int mc_mem_d2d(void *dst, const void *src, size_t bytes);
/* GPUMemBinD2D / GPUMemBinD2H (cufunc.cu:147-176): Fourier crop of a padded
 * spectrum of size src_n to dst_n (rows 0..dst_n/2-1 and the last dst_n/2
 * rows, dst_n+2 floats each). */
int mc_mem_bin_d2d(float *dst, const float *src, int dst_n, int src_n);
int mc_mem_bin_d2h(float *dst_host, const float *src, int dst_n, int src_n);
/* GPUFFTPlan / GPUIFFTPlan (cufunc.cu:202-215): n x n in-place plans.
 * Supported n: any even size from 64 to 8192 (Stockham plans for powers of
 * two and 2^k * 3, 5 or 15; chirp-z transforms for the others, except the
 * forward transforms at 3710 and 4092, which run in two levels: 35 x 53 and
 * 70 x 53, 66 x 31 and 132 x 31). */
int mc_fft_plan(mc_plan *plan, int n);
int mc_ifft_plan(mc_plan *plan, int n);
/* GPUFFT2d / GPUIFFT2d (cufunc.cu:223-233): unnormalised, in place. */
int mc_fft2d(float *buf, mc_plan plan);
int mc_ifft2d(float *buf, mc_plan plan);
/* GPUFFTDestroy (cufunc.cu:217-221). */
int mc_fft_destroy(mc_plan plan);
/* GPUSync (cufunc.cu:235-241). */
int mc_sync(void);
/* GPUAdd / GPUMultiplyNum (cufunc.cu:243-263): counts in floats. */
int mc_add(float *dst, const float *src, size_t nfloats);
int mc_multiply_num(float *dst, float num, size_t nfloats);
/* GPUFFTLogModulus (cufunc.cu:178-188): dmod[i] = log(1+|F_i|*scale). */
int mc_fft_log_modulus(float *dmod, const float *dfft, int n, float scale);
/* MASK of cufunc.h:9-14; MkPosList(MASK*, n, bfactor) (cufunc.cu:309-341)
 * fills a host table of n*(n/2+1) entries. */
typedef struct {
  int x;
  int y;
  float z;
} mc_mask;
int mc_make_pos_list(mc_mask *list, int n, float bfactor);
/* GPUShiftCC (cufunc.cu:343-370): dfft <- dsum*conj(dfft*e^{i th})*z/sizec. */
int mc_shift_cc(float *dfft, const float *dsum, const mc_mask *dposlist, float sx, float sy,
                int n);
/* GPUShift (cufunc.cu:372-394): dfft <- dfft * e^{i th}, th = (2 pi/n)(sx x + sy y'). */
int mc_shift(float *dfft, const mc_mask *dposlist, float sx, float sy, int n);
/* FindShift (cufunc.cu:396-497): box read-out of a C2R output (pitch n+2),
 * integer and 1/32-pixel peak (FP64).  hboxmap (host, box*box floats) gets
 * the box; returns 0 and writes cc. */
int mc_find_shift(const float *dsrc, int n, float *hboxmap, int box, float *sx, float *sy,
                  float *cc, int wnoise);
/* No counterpart in the reference.  mc_find_shift for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments and results, but the box
 * copy and the peak record live in a fixed device array of the library
 * (MC_NOALLOC_MAX_BOX^2 floats + 256 bytes, one per GPU, held until the
 * process ends) and the call makes no allocation.  box > MC_NOALLOC_MAX_BOX
 * returns MC_ERR_ARG.  All calls on one GPU share the array: calls that
 * overlap on one GPU (several host threads) race. */
#define MC_NOALLOC_MAX_BOX 256
int mc_find_shift_noalloc(const float *dsrc, int n, float *hboxmap, int box, float *sx,
                          float *sy, float *cc, int wnoise);

/* ======================================================================
 * Level 1: resident movie pipeline
 * ==================================================================== */
typedef struct mc_context_s *mc_context;

/* A context owns a HIP stream and the device buffers of one movie at a
 * time (grown as needed, kept across calls). */
int mc_context_create(int device, mc_context *ctx);
int mc_context_destroy(mc_context ctx);

/* Frames of one movie in device memory. */
typedef struct {
  const void *frames;  /* nz frames; element type dtype                    */
  const void *frames2; /* second dark/gain-corrected copy used as the
                          conjugated operand of the pair stage (the hbuf2
                          stack, DFAlign.cu:724-749,889-890) or NULL      */
  int dtype;           /* MC_DTYPE_*                                      */
  int nx, ny, nz;      /* frame width, height, count in the buffer        */
  size_t row_pitch;    /* bytes between rows (0: nx * element size)       */
  size_t frame_pitch;  /* bytes between frames (0: ny * row_pitch)        */
  int nz_file;         /* frames in the stack file; RefFrame = nz_file/2+1
                          (DFAlign.cu:1133); 0: nz                        */
  int first_frame;     /* absolute index of frames[0]; frame indices in
                          mc_params are absolute                          */
  const struct mc_corrections_s *corr; /* dark/gain/stretch applied on the
                          device while the frames are loaded, or NULL     */
} mc_movie_desc;

/* Dark/gain correction and stretch of the raw frames, applied on the device
 * in the load of the row FFT (reference DFAlign.cu:614-766, done there on
 * the host).  Per frame pixel (X, Y) of the nx x ny frame:
 *   v1 = (raw - dark) * gain    (a missing reference is skipped; the float
 *   v2 = (raw - dark2) * gain2   operations are the reference's, in order)
 * then, with a stretch (xscale < 1 or yscale < 1), pixel (X, Y), X, Y >= 1,
 * takes the value at (int(X*xscale + 0.5), int(Y*yscale + 0.5)) (nearest
 * neighbour; row 0 and column 0 keep their own values).  xscale, yscale are
 * 1/-scx and 1/-scy.  v1 feeds every spectrum; with `second`, v2 feeds a
 * second spectrum set used as the conjugated operand of the pair stage (the
 * hbuf2 stack, DFAlign.cu:889-890; with single-section references the
 * reference sets dark2 = 0 and gain2 = 1).  References are device buffers
 * of nx*ny floats, row pitch nx.  frames2 must be NULL when corr is used. */
typedef struct mc_corrections_s {
  const float *dark, *gain;    /* section 1 (NULL: none)                  */
  const float *dark2, *gain2;  /* section 2 (NULL: none)                  */
  float xscale, yscale;        /* stretch factors in (0, 1]; 1: none      */
  int second;                  /* build the section-2 spectra             */
} mc_corrections;

/* Alignment parameters: APARA (DFAlign.h:16-63) and CDFAlign::m_iter /
 * m_iterthres (DFAlign.h:75-76), plus the new geometry and form selectors.
 * Frame indices are absolute (0-based) and already clamped as VerifyPara
 * does (dosefgpu_driftcorr.cpp:60-118). */
typedef struct {
  int crop_offsetx, crop_offsety, crop_nsam; /* -crx -cry -crd            */
  float bin;                                 /* -bin (1..2)               */
  int nStart, nEnd;                          /* -nst -ned                 */
  int nStartSum, nEndSum, nEndSum2;          /* -nss -nes -2es            */
  float bfactor;                             /* -bft                      */
  int CCPeakSearchDim;                       /* -pbx                      */
  int FrameDistOffset;                       /* -fod                      */
  int NoisePeakSize;                         /* -nps                      */
  float kiThresh;                            /* -kit                      */
  int bAlignToMid;                           /* -atm                      */
  int iter;                                  /* -itr                      */
  float iterthres;                           /* -ith                      */
  int pad;        /* new: 0 = reference square crop; N = pad the region
                     [crx,nx) x [cry,ny) into an N x N canvas filled with
                     the region mean                                       */
  int precision;  /* MC_PREC_*                                          */
  int grad_steps; /* new: adjoint sharpness refinement steps (0 = off)  */
  float grad_step_px; /* new: trust-region step of the refinement (px)  */
  float fftscale; /* APARA fftscale of the quick-look log-modulus images
                     (0 = 0.0001, the reference default)                  */
} mc_params;

#define MC_PREC_FP32 0      /* FP32 on the vector ALUs (pair row pass: FP32 MFMA) */
#define MC_PREC_MFMA_HILO 1 /* pair stage DFTs as FP16 hi/lo MFMA products with
                               FP32 accumulation (n in [512, 8192], box <= 96);
                               FFTs of stage 1, sums and the FP64 peak fit as
                               in MC_PREC_FP32 */

/* Optional results (host pointers; NULL to skip).  Sizes use
 * F = nEnd-nStart+1, K = F-1, P = pair count, B = CCPeakSearchDim. */
typedef struct {
  int *pairs;          /* [2P] frame positions (i, j) of each pair          */
  double *pair_b;      /* [2P] measured shift of each pair                   */
  double *pair_cc;     /* [P]                                                */
  float *boxmaps;      /* [P*B*B] box of each correlation map                */
  double *shift1;      /* [2K] first least-squares increments               */
  double *ki1;         /* [P] residuals                                      */
  int *good;           /* [P] rows kept by outlier rejection (ascending)    */
  double *shift2;      /* [2K] increments after outlier rejection           */
  double *ki2;         /* [P] residuals of the kept rows                     */
  double *iter_meas;   /* [iter*2F] per-round measured positions            */
  double *iter_cc;     /* [iter*F] per-round CC of each frame vs the sum      */
  float *iter_boxmaps; /* [iter*F*B*B] per-round box of each frame vs sum    */
  double *rmsd;        /* [iter] accumulated RMSD recursion                  */
  double *final_shift; /* [2K] increments after iteration                   */
  double *frame_shift; /* [2F] shift applied to each frame in the sum (NaN:
                          not summed)                                        */
  double *frame_mean;  /* [F] DC/n^2 of each spectrum (the per-frame log)   */
  float *sum_image;    /* [n*n] aligned sum (host)                           */
  float *sum2_image;   /* [n*n] second sum (-2es) (host)                     */
  float *rawsum_image; /* [n*n] unaligned sum (host)                         */
  void *sum_image_dev; /* [n*n] aligned sum left in device memory            */
  double *fsc;         /* [mc_fsc_bins(n)*4] FSC table (-fsc): per shell
                          D_Raw, U_Raw, D_Corr, U_Corr (DFAlign.cu:1314-1318) */
  /* Quick-look images (-dsp 1, dosef_quick/; DFAlign.cu:73-179, 842-848,
   * 908-911, 1270-1285; func.cpp:372-555): MC_DISPDIM^2 8-bit values each,
   * written as MRC mode 1.  Need n >= MC_DISPDIM.                          */
  short *disp_raw_fft;  /* _RawFFT: log-modulus display of the unaligned sum */
  short *disp_corr_sum; /* _CorrSum: the aligned sum binned to 400 x 400     */
  short *disp_corr_fft; /* _CorrFFT: log-modulus display of the aligned sum  */
} mc_outputs;

#define MC_DISPDIM 400 /* DFAlign.h:13 */

typedef struct {
  int n, nsamUnbin, nframe, npairs, ngood, niter, converged;
  int sum2_written, rawsum_written;
  double rmsd;
  float ms_total, ms_spectra, ms_pairs, ms_peaks, ms_solve, ms_iter, ms_sum;
  double sharp_j0, sharp_j1; /* sharpness before / after refinement       */
  int grad_accepted;         /* accepted refinement steps                 */
} mc_report;

/* Whole movie: DFAlign.cu:456-1355 numerics (without file I/O and logs). */
int mc_align_movie(mc_context ctx, const mc_movie_desc *movie, const mc_params *p,
                   mc_outputs *out, mc_report *rep);

/* Many movies (no reference counterpart: the reference aligns one stack per
 * process run, DFAlign.cu:423-453).  Movies with the same peak box,
 * B-factor, precision form and noise window form a group, whatever their
 * transform sizes; a group is split into sub-batches whose resident spectra
 * and pair-stage buffers fit in device memory (MC_BATCH_MEM_MB, default 60%
 * of the free memory).  In a sub-batch, stage 1 runs per movie; the pair
 * column pass of every movie on the default column kernel is one persistent
 * launch sequence across sizes (a plan table, one record per movie), the
 * others keep per-movie launches; the row pass and the peak search run once
 * per transform size; the drift model runs per movie on the host; each
 * align-to-sum round of the movies still iterating (all sizes) is one HIP
 * graph, captured once and replayed while that set of movies is unchanged;
 * the final sums run per movie.  Every movie gets bit for bit the results
 * mc_align_movie would give it: alone, a movie runs the same kernels (the
 * default column kernel with its one plan), and the row pass's summation
 * order depends on n only.  The ms_* fields of its report are those of its
 * sub-batch.  The movies' frame buffers must all be resident. */
typedef struct {
  const mc_movie_desc *movie;
  const mc_params *params;
  mc_outputs *out; /* may be NULL */
  mc_report *rep;  /* may be NULL */
} mc_batch_item;

typedef struct {
  int groups, sub_batches;            /* launch groups and memory splits  */
  int graph_captures, graph_replays;  /* align-to-sum round graphs        */
  double ms_total;                    /* device time of all sub-batches   */
} mc_batch_report;

int mc_align_batch(mc_context ctx, const mc_batch_item *items, int count,
                   mc_batch_report *brep);

/* Autotuner (no reference counterpart).  Times the launch-shape candidates
 * on this movie and keeps the fastest in the context for every later call
 * with the same canvas size, binned size and element type (stage 1) and the
 * same binned size, peak box and precision form (pair stage):
 *   stage 1: row-pass shape (rows per workgroup x threads: 4x256, 2x256,
 *            4x512 for canvases >= 4096) x frames per launch (1, 2, 4, 8)
 *            x column pass (one workgroup per column, or persistent with
 *            register twiddles), then, for the per-column pass on canvases
 *            1024 to 8192, the radix plan of its FFT (balanced, reversed
 *            order, radix-8 chain) x its LDS index padding (one float2 per
 *            16, or none);
 *   pair stage (FP32 form): column kernel (M=128, M=128 with next-item
 *            prefetch, M-generic, M=128 at 4 waves per SIMD) and row kernel
 *            (FP32 MFMA with 4 or 8 waves per pair, vector ALU).
 * Each candidate runs once untimed and 3 times timed (minimum kept); the
 * context's resident spectra are overwritten, so the stage entry points need
 * mc_load_frames again afterwards.  With
 * cache_path, one JSON line per tuned geometry is appended
 * ({"arch", "cus", "nu", "n", "dtype", "box", "precision", "r2c", "chunk",
 * "cols", "rows"}); mc_tune_load reads the lines of the same device type
 * back (later lines win).  Without tuning, the defaults are 4x512 (4x256
 * below 4096), 192 MB of row-pass scratch per launch, the M=128 column
 * pass (the batched plan kernel up to n = 3840, the 4-wave kernel at 4096,
 * the plain one above) and the streaming FP32 MFMA row kernel. */
typedef struct {
  int r2c_variant;        /* 1: 4x256, 2: 2x256, 3: 4x512; +8: persistent
                             column pass with register twiddles           */
  int chunk_frames;       /* frames per stage-1 launch                      */
  int pair_cols_variant;  /* 0: M=128, 1: M=128 + prefetch, 2: M-generic,
                             3: M=128 at 4 waves per SIMD                 */
  int pair_rows_variant;  /* 0: FP32 MFMA streaming (box <= 96), 1: vector
                             ALU, 2: FP32 MFMA through LDS (8 waves),
                             3: FP32 MFMA through LDS (4 waves)             */
  float ms_stage1[2][4][4]; /* [column pass: per column, persistent]
                               [row shape][chunk 1, 2, 4, 8]; 0 = not run  */
  float ms_pair_cols[4];  /* 0 = not run                                    */
  float ms_pair_rows[4];
  int cols_plan;          /* per-column pass of stage 1: radix plan cols_plan/2
                             (0 balanced, 1 reversed order, 2 radix-8 chain)
                             and LDS padding cols_plan%2 (0 one float2 per 16,
                             1 none); r2c_variant holds it in bits 4-6      */
  float ms_cols_plan[6];  /* 0 = not run (canvases 1024 to 8192 only)      */
} mc_tune_report;

int mc_autotune(mc_context ctx, const mc_movie_desc *movie, const mc_params *p,
                const char *cache_path, mc_tune_report *rep);
int mc_tune_load(mc_context ctx, const char *cache_path, int *loaded);

/* FSC table (DFAlign.cu:212-326, CDFAlign::PlotFSC).  Shell count for size
 * n: (n/2+1)/step + 1 with step = n/400, at least 1 (the reference divides
 * by zero for n < 400).  Returns the count or an error code (n odd, < 4 or
 * > 8192). */
int mc_fsc_bins(int n);
/* FSC of two device half spectra a, b in the reference's layout (n rows of
 * n/2+1 complex, i.e. the hFSCRaw/hFSCCorr arrays of DFAlign.cu:821-822,
 * 1265-1266) in shells of width step, split into the elements parallel
 * (out[2i]) and perpendicular (out[2i+1]) to the direction (dx, dy) (the
 * total shift; normalised when its length exceeds 0.1, else (1, 0)).  Sums
 * are taken in FP64 (the reference sums in float).  out: host,
 * 2*mc_fsc_bins(n) doubles. */
int mc_fsc_table(mc_context ctx, const void *a, const void *b, int n, double dx, double dy,
                 double *out);

/* 8-bit display of `size` device floats into host shorts: histMinMax with
 * 400 bins and a 0.5% threshold, then buf2DispShort (func.cpp:372-443), in
 * the reference's float arithmetic.  A constant buffer gives zeros (the
 * reference leaves its output unwritten). */
int mc_disp_short(mc_context ctx, const float *buf, int size, short *out);

/* Stage entry points on the context's resident state ------------------- */
/* Stage 1 only: frames -> resident spectra (sizes written to rep). */
int mc_load_frames(mc_context ctx, const mc_movie_desc *movie, const mc_params *p,
                   mc_report *rep);
/* Replace the resident spectra with host spectra in the padded layout
 * (nframe x n x (n+2) floats). */
int mc_set_spectra(mc_context ctx, const float *host_padded, int n, int nframe);
/* Copy the resident spectra out in the padded layout. */
int mc_get_spectra(mc_context ctx, float *host_padded);
/* Pair stage on resident spectra: pairs[2P] index resident frames (operand
 * i kept, operand j conjugated); writes b (2P), cc (P) and optionally the
 * box maps (P*box*box). */
int mc_pair_shifts(mc_context ctx, const int *pairs, int npairs, float bfactor, int box,
                   int wnoise, double *b, double *cc, float *boxmaps);
/* Peak search on given box maps (host, P*box*box floats). */
int mc_find_peaks(mc_context ctx, const float *boxmaps, int npairs, int box, int n,
                  int wnoise, double *b, double *cc);
/* Phase-shift-and-sum of the resident spectra with per-frame shifts
 * (2F doubles; NaN x = frame skipped) and inverse to an n x n image scaled
 * by 1/n^2 (host). */
int mc_sum_image(mc_context ctx, const double *frame_shift, float *image);

/* Sharpness of the phase-shift-and-sum (no reference counterpart):
 * J = sum_k w(k) |sum_f F_f(k) e^{i 2 pi k.s_f/n}|^2 over the half plane
 * (w = exp(-B|k|^2/(2n^2)), doubled off the DC/Nyquist columns), and its
 * adjoint dJ/ds_f (grad: 2F doubles) on the resident spectra.  Frames with a
 * NaN x shift are left out of the sum and get a zero gradient.  With
 * grad_steps > 0, mc_align_movie maximises J over the summed frames' shifts
 * before the final sum (trust-region gradient ascent: first step
 * grad_step_px, default 0.05 px, doubled after an accepted step up to 4x the
 * first, halved after a rejected one); the peak-search shifts are the
 * start. */
int mc_sharpness(mc_context ctx, const double *frame_shift, float bfactor, double *J,
                 double *grad);

/* Host FP64 drift model (mfunc.h:50-146, DFAlign.cu:914-948). */
int mc_pair_list(int nframe, int fod, int *pairs, int cap);
int mc_lsq(const int *pairs, int npairs, int nframe, const double *b, double *s);
int mc_threshold(const int *pairs, int npairs, int nframe, const double *ki, double kithresh,
                 int *good);

/* Kernel timing over a region (HIP events on the context's stream around
 * every launch of the named kernel group).  mc_kernel_timing times every
 * group (enable != 0) or none; mc_kernel_timing_groups times only the groups
 * whose bit (1u << group) is set in mask (0: none), so that a timed region
 * carries only the events it needs.  Either call clears earlier records. */
#define MC_KT_SPECTRA 0
#define MC_KT_PAIR_COLS 1
#define MC_KT_PAIR_ROWS 2
#define MC_KT_PEAKS 3
#define MC_KT_SUM 4
#define MC_KT_IMAGE 5
#define MC_KT_ROUNDS 6 /* graph replays of whole align-to-sum rounds */
#define MC_KT_COUNT 7
int mc_kernel_timing(mc_context ctx, int enable);
int mc_kernel_timing_groups(mc_context ctx, unsigned mask);
int mc_kernel_timing_get(mc_context ctx, int group, int *count, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
