/*
 * mc_oracle.h -- CPU restatement of the dosefgpu_driftcorr alignment path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * host driver, the CLI) includes, links or calls this code.  It is the
 * checker: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it to verify and to time the reference algorithm on the host.
 *
 * The same source is compiled twice:
 *   liboracle_f64.so  (MC_REAL=double) -- the FP64 oracle every GPU result is
 *                                         compared against;
 *   liboracle_f32.so  (MC_REAL=float)  -- the reference algorithm in float,
 *                                         timed as the host-CPU baseline.
 *
 * Reference files followed (paths relative to the reference tree):
 *   src-1125/DFAlign.cu:456-1355   control flow, clamps, quirks
 *   src-1125/cufunc.cu:162-176     Fourier-crop binning
 *   src-1125/cufunc.cu:309-394     weight table, cross-power, phase shift
 *   src-1125/cufunc.cu:396-497     box read-out, integer peak, sub-pixel peak
 *   src-1125/func.cpp:80-103,157-173,193-201,348-370
 *   src-1125/mfunc.h:50-146       pair list, design matrix, outlier rejection
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#ifndef MC_REAL
#define MC_REAL double
#endif
typedef MC_REAL or_real;

/* ---- FFT (unnormalised; forward = e^{-2 pi i}) ------------------------- */
/* 1-D complex transform of n interleaved (re,im) values, in place.
 * sign = -1 forward, +1 inverse.  Any n >= 1. */
int or_fft_c2c(or_real *data, int n, int sign);
/* 2-D real-to-complex transform of an n x n image held in the in-place padded
 * layout (n rows of n+2 reals).  Result: n rows x (n/2+1) complex.
 * Follows cufftExecR2C in place (cufunc.cu:223-227). */
int or_r2c2d(or_real *buf, int n);
/* 2-D complex-to-real inverse of the same layout, unnormalised.  Column
 * transforms are complex; the row transform takes the real part of
 * sum_{kx=0}^{n/2} a_kx Y[kx] e^{+2 pi i kx c/n} with a_0 = a_{n/2} = 1 and
 * a_kx = 2 otherwise (imaginary parts of the DC and Nyquist bins of each row
 * are ignored).  Follows cufftExecC2R in place (cufunc.cu:229-233). */
int or_c2r2d(or_real *buf, int n);

/* ---- geometry ---------------------------------------------------------- */
/* func.cpp:80-103 */
int or_verify_crop_size(int nx, int ny, int offsetx, int offsety, int nsamout,
                        float bin);
/* DFAlign.cu:515 */
int or_binned_size(int nsamUnbin, float bin);
/* func.cpp:157-173: copy a square crop into the padded layout (pitch n+2). */
int or_crop2fft(const float *in, int nx, int ny, or_real *out, int offsetx,
                int offsety, int nsam);
/* Padding geometry (no reference counterpart): the region [offsetx, nx) x
 * [offsety, ny) is placed at the canvas origin of an npad x npad canvas; the
 * rest of the canvas holds the mean of the region.  Pitch npad+2. */
int or_pad2fft(const float *in, int nx, int ny, or_real *out, int offsetx,
               int offsety, int npad);
/* cufunc.cu:162-176: Fourier-crop a spectrum of size src_n to dst_n. */
int or_bin_crop(or_real *dst, const or_real *src, int dst_n, int src_n);

/* ---- spectral kernels -------------------------------------------------- */
/* cufunc.cu:309-341: weight z(x,y') of the MASK table for element id of the
 * n x (n/2+1) spectrum (row-major, row pitch n/2+1 complex). */
or_real or_weight(int n, float bfactor, int kx, int ky);
/* cufunc.cu:343-370 with sx = sy = 0: dfft <- dsum * conj(dfft) * z / sizec */
int or_cross_power(or_real *dfft, const or_real *dsum, int n, float bfactor);
/* cufunc.cu:372-394: F <- F * exp(i (2 pi/n)(sx*x + sy*y')) */
int or_phase_shift(or_real *dfft, int n, double sx, double sy);

/* cufunc.cu:396-497: box read-out of the C2R output (pitch n+2), integer
 * peak, 16x16 -> 512x512 Fourier interpolation, 1/32-pixel peak.
 * boxmap (box*box) receives the box.  Returns cc. */
double or_find_shift(const or_real *cmap, int n, or_real *boxmap, int box,
                     double *sx, double *sy, int wnoise);
/* Same search applied to a box map that is already extracted (the pruned
 * inverse below produces it directly). */
double or_find_shift_box(const or_real *boxmap, int n, int box, double *sx,
                         double *sy, int wnoise);
/* Pruned inverse: box x box window of C2R(X) around the map centre, where X
 * is the n x (n/2+1) cross-power spectrum; element [r][c] equals
 * C2R(X)[(n-box)/2 + r][(n-box)/2 + c].  Exact-arithmetic equal to the full
 * inverse followed by the read-out of cufunc.cu:404-407. */
int or_pruned_box(const or_real *X, int n, int box, or_real *boxmap);

/* ---- drift model (mfunc.h) --------------------------------------------- */
/* mfunc.h:50-84.  Returns the pair count; fills pairs (2*count ints). */
int or_pair_list(int nframe, int fod, int *pairs, int cap);
/* Least-squares A s = b for the 0/1 design matrix of the pair list (FP64,
 * Householder QR).  b, s: interleaved (x,y).  Returns 0 on success. */
int or_lsq(const int *pairs, int npairs, int nframe, const double *b,
           double *s);
/* ki_p = |A s - b|_p  (DFAlign.cu:916) */
void or_residual(const int *pairs, int npairs, const double *s,
                 const double *b, double *ki);
/* mfunc.h:101-146: outlier rejection; returns the number of kept rows and
 * writes their indices (ascending) to good. */
int or_threshold(const int *pairs, int npairs, int nframe, const double *ki,
                 double kithresh, int *good);

/* ---- whole movie (DFAlign.cu:456-1355, numerics only) ------------------ */
typedef struct {
  int nx, ny, nz;                 /* stack geometry (nz = frames in file)  */
  int crop_offsetx, crop_offsety; /* -crx -cry                              */
  int crop_nsam;                  /* -crd                                   */
  float bin;                      /* -bin                                   */
  int pad;                        /* 0: reference square crop; N: pad to N */
  int nStart, nEnd;               /* -nst -ned (already clamped)            */
  int nStartSum, nEndSum, nEndSum2;
  float bfactor;                  /* -bft */
  int box;                        /* -pbx */
  int fod;                        /* -fod */
  int nps;                        /* -nps (wNoise = nps - 1)                */
  float kit;                      /* -kit */
  int atm;                        /* -atm */
  int iter;                       /* -itr */
  float iterthres;                /* -ith */
  int full_c2r;                   /* 1: full n x n inverse per pair as the
                                     reference does; 0: pruned box inverse */
  int nthreads;                   /* OpenMP threads (0 = default)           */
  float fftscale;                 /* APARA fftscale (0: 0.0001, the default) */
} or_params;

typedef struct {
  int n, nsamUnbin, nframe, npairs, ncols, ngood, niter;
  /* Optional outputs (NULL to skip); sizes in brackets. */
  int *pairs;            /* [2*npairs]                                     */
  double *pair_b;        /* [2*npairs] measured (sx, sy) per pair           */
  double *pair_cc;       /* [npairs]                                        */
  or_real *boxmaps;      /* [npairs*box*box]                                */
  double *shift1;        /* [2*ncols] first LSQ solution                    */
  double *ki1;           /* [npairs]                                        */
  int *good;             /* [npairs] kept rows                              */
  double *shift2;        /* [2*ncols] after outlier rejection               */
  double *ki2;           /* [npairs]                                        */
  double *iter_meas;     /* [iter*2*nframe] per-round measured positions   */
  double *rmsd;          /* [iter]                                          */
  double *final_shift;   /* [2*ncols] increments after iteration           */
  double *frame_shift;   /* [2*nframe] shift applied to each frame in the
                            final sum (x, y); NaN for frames not summed     */
  or_real *sum_image;    /* [n*n]                                           */
  or_real *sum2_image;   /* [n*n] second sum (-2es), if produced            */
  or_real *rawsum_image; /* [n*n]                                           */
  or_real *spectra;      /* [nframe*n*(n+2)] frame spectra (padded layout)  */
  int sum2_written, rawsum_written;
  double *fsc;           /* [or_fsc_bins(n)*4] FSC table: per shell D_Raw,
                            U_Raw, D_Corr, U_Corr (DFAlign.cu:1314-1318)   */
  short *disp_raw_fft;   /* [400*400] dosef_quick/ _RawFFT (n >= 400)       */
  short *disp_corr_sum;  /* [400*400] _CorrSum                              */
  short *disp_corr_fft;  /* [400*400] _CorrFFT                              */
  int disp_written;      /* bit 0 raw FFT, bit 1 sum, bit 2 corrected FFT   */
} or_result;

/* frames: nz x ny x nx float (decoded, dark/gain corrected = hbuf source);
 * frames2: the second corrected copy (hbuf2 source, DFAlign.cu:724-749),
 * or NULL when it equals frames.  Returns 0 on success. */
int or_align_movie(const float *frames, const float *frames2,
                   const or_params *p, or_result *r);

/* ---- FSC table (DFAlign.cu:212-326) ------------------------------------ */
/* Shell width step = n/400 (at least 1: the reference divides by zero for
 * n < 400) and shell count (n/2+1)/step + 1 (DFAlign.cu:214-218). */
int or_fsc_bins(int n, int *step);
/* FSC of two half spectra a, b (padded layout) in shells of width step,
 * split into elements parallel (|k.d|/|k| > cos(pi/4), out[2i]) and
 * perpendicular (out[2i+1]) to the direction d = (dx, dy) (normalised when
 * |d| > 0.1, else (1, 0)).  Element (kx, y) with r = sqrtf(kx^2+y^2) in
 * float goes to shell int(r/step); DC and r >= n/2+1 are skipped.  Sums of
 * Re(a conj b), |a|^2, |b|^2 per shell in FP64 (the reference sums in
 * float); FSC = sum Re / sqrt(sum|a|^2 sum|b|^2), 0 when that product is
 * <= 1e-5.  DFAlign.cu:212-297. */
int or_fsc(const or_real *a, const or_real *b, int n, double dx, double dy,
           double *out);

/* ---- quick-look images (dosef_quick/) ------------------------------------ */
#define OR_DISPDIM 400 /* DFAlign.h:13 */
/* Log-modulus log(1 + |X| scale) of the half spectrum spec (padded layout;
 * cufunc.cu:178-188) laid out as the full n x n image (pitch n) with DC at
 * row n/2, column n/2 (func.cpp:459-489 FFTModulusToDispBuf). */
int or_logmod_image(const or_real *spec, int n, float scale, or_real *img);
/* func.cpp:524-555 (BinFFTDispBufToChar up to the display mapping): R2C of
 * the log-modulus image, Fourier crop to d x d (the row pattern of
 * cufunc.cu:147-161), C2R, / d^2.  out: d x d.  n >= d. */
int or_fft_display(const or_real *spec, int n, float scale, int d, or_real *out);
/* DFAlign.cu:1281-1284 with ImageOutputThread (DFAlign.cu:73-90): Fourier
 * crop of the sum spectrum to d x d, C2R, no scaling.  n >= d. */
int or_sum_display(const or_real *spec, int n, int d, or_real *out);
/* func.cpp:372-411 histMinMax (400 bins, threshold 0.005) and
 * func.cpp:441-455 buf2DispShort, in the reference's float arithmetic.
 * Returns 1 (out untouched) when the clipped range is empty, else 0. */
int or_disp_short(const float *buf, int size, short *out);

/* Sharpness objective of the phase-shift-and-sum and its gradient with
 * respect to per-frame shifts (no reference counterpart):
 *   J(s) = sum_k w(k) |sum_f F_f(k) e^{i 2 pi k.s_f / n}|^2,
 * w(k) = exp(-B |k|^2/(2 n^2)) * h(kx), h = 1 for kx in {0, n/2}, else 2.
 * spectra: nframe padded-layout spectra; s: 2*nframe (x,y); grad: 2*nframe. */
double or_sharpness(const or_real *spectra, int nframe, int n, float bfactor,
                    const double *s, double *grad);

#ifdef __cplusplus
}
#endif
#endif
