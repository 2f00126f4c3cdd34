/*
 * mc_oracle.c -- CPU restatement of the dosefgpu_driftcorr alignment path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the cpu_baseline leg of bench.py load this library; the product never does.
 * Each function names the reference lines it restates (paths relative to the
 * reference tree).  Compiled as FP64 (the oracle) and as FP32 (the timed
 * host baseline) from this one file; see mc_oracle.h.
 */
#include "mc_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef or_real R;
typedef struct { R re, im; } cpx;

#define OR_PI 3.14159265358979323846

/* ======================================================================
 * FFT: Stockham auto-sort, mixed radix (4, 2, 3, 5 and generic primes).
 * Twiddles are cos/sin(2 pi e / n) evaluated in double and rounded to R.
 * ==================================================================== */
typedef struct {
  int n;
  int nst;
  int radix[64];
  R *wr, *wi; /* wr[e] = cos(2 pi e/n), wi[e] = sin(2 pi e/n) */
} fplan;

static void plan_init(fplan *P, int n) {
  P->n = n;
  P->nst = 0;
  int m = n;
  while (m % 4 == 0) { P->radix[P->nst++] = 4; m /= 4; }
  while (m % 2 == 0) { P->radix[P->nst++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { P->radix[P->nst++] = f; m /= f; }
  }
  P->wr = (R *)malloc(sizeof(R) * (size_t)n);
  P->wi = (R *)malloc(sizeof(R) * (size_t)n);
  for (int e = 0; e < n; e++) {
    /* exact reduction: angle = 2 pi e/n with e in [0, n) */
    double a = 2.0 * OR_PI * (double)e / (double)n;
    P->wr[e] = (R)cos(a);
    P->wi[e] = (R)sin(a);
  }
}

static void plan_free(fplan *P) {
  free(P->wr);
  free(P->wi);
}

/* W^e with W = exp(sign * 2 pi i / n) */
static inline cpx tw(const fplan *P, long e, int sign) {
  int n = P->n;
  e %= n;
  if (e < 0) e += n;
  cpx w;
  w.re = P->wr[e];
  w.im = sign < 0 ? -P->wi[e] : P->wi[e];
  return w;
}

static inline cpx cmul(cpx a, cpx b) {
  cpx c;
  c.re = a.re * b.re - a.im * b.im;
  c.im = a.re * b.im + a.im * b.re;
  return c;
}

/* One Stockham stage: x -> y, radix Rr, Ns = length of finished
 * sub-transforms.  v and t are scratch of at least Rr entries. */
static void stage(const cpx *x, cpx *y, const fplan *P, int Rr, int Ns,
                  int sign, cpx *v, cpx *t) {
  const int n = P->n;
  const int NR = n / Rr;
  const long tstep = n / (Ns * Rr); /* twiddle W_{Ns*Rr}^{k r} = W_n^{k r tstep} */
  for (int j = 0; j < NR; j++) {
    const int k = j % Ns;
    for (int r = 0; r < Rr; r++) v[r] = x[j + r * NR];
    if (k) {
      for (int r = 1; r < Rr; r++) v[r] = cmul(v[r], tw(P, (long)k * r * tstep, sign));
    }
    if (Rr == 2) {
      cpx a = v[0], b = v[1];
      v[0].re = a.re + b.re; v[0].im = a.im + b.im;
      v[1].re = a.re - b.re; v[1].im = a.im - b.im;
    } else if (Rr == 4) {
      cpx t0 = {v[0].re + v[2].re, v[0].im + v[2].im};
      cpx t1 = {v[0].re - v[2].re, v[0].im - v[2].im};
      cpx t2 = {v[1].re + v[3].re, v[1].im + v[3].im};
      cpx t3 = {v[1].re - v[3].re, v[1].im - v[3].im};
      /* w t3 with w = sign * i */
      cpx wt3 = {-(R)sign * t3.im, (R)sign * t3.re};
      v[0].re = t0.re + t2.re; v[0].im = t0.im + t2.im;
      v[2].re = t0.re - t2.re; v[2].im = t0.im - t2.im;
      v[1].re = t1.re + wt3.re; v[1].im = t1.im + wt3.im;
      v[3].re = t1.re - wt3.re; v[3].im = t1.im - wt3.im;
    } else {
      const long rstep = n / Rr; /* W_Rr^{q r} = W_n^{q r rstep} */
      for (int q = 0; q < Rr; q++) {
        cpx acc = v[0];
        for (int r = 1; r < Rr; r++) {
          cpx w = tw(P, ((long)q * r) % Rr * rstep, sign);
          cpx pr = cmul(v[r], w);
          acc.re += pr.re;
          acc.im += pr.im;
        }
        t[q] = acc;
      }
      for (int q = 0; q < Rr; q++) v[q] = t[q];
    }
    const int idx = (j / Ns) * Ns * Rr + k;
    for (int r = 0; r < Rr; r++) y[idx + r * Ns] = v[r];
  }
}

/* In-place 1-D transform using caller scratch (n complex + 2*maxradix). */
static void fft1(const fplan *P, cpx *data, cpx *scratch, int sign) {
  const int n = P->n;
  if (n <= 1) return;
  cpx *v = scratch + n;
  int maxr = 1;
  for (int s = 0; s < P->nst; s++)
    if (P->radix[s] > maxr) maxr = P->radix[s];
  cpx *t = v + maxr;
  cpx *x = data, *y = scratch;
  int Ns = 1;
  for (int s = 0; s < P->nst; s++) {
    stage(x, y, P, P->radix[s], Ns, sign, v, t);
    Ns *= P->radix[s];
    cpx *tmp = x; x = y; y = tmp;
  }
  if (x != data) memcpy(data, x, sizeof(cpx) * (size_t)n);
}

static size_t scratch_len(const fplan *P) {
  int maxr = 1;
  for (int s = 0; s < P->nst; s++)
    if (P->radix[s] > maxr) maxr = P->radix[s];
  return (size_t)P->n + 2 * (size_t)maxr;
}

int or_fft_c2c(R *data, int n, int sign) {
  if (n < 1) return -1;
  fplan P;
  plan_init(&P, n);
  cpx *s = (cpx *)malloc(sizeof(cpx) * scratch_len(&P));
  fft1(&P, (cpx *)data, s, sign);
  free(s);
  plan_free(&P);
  return 0;
}

/* cufunc.cu:223-227 (cufftExecR2C in place, padded pitch n+2) */
int or_r2c2d(R *buf, int n) {
  if (n < 2 || (n & 1)) return -1;
  const int h = n / 2 + 1;
  const int pitch = n + 2;
  fplan P;
  plan_init(&P, n);
  const size_t sl = scratch_len(&P);
#pragma omp parallel
  {
    cpx *row = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    cpx *s = (cpx *)malloc(sizeof(cpx) * sl);
#pragma omp for schedule(static)
    for (int y = 0; y < n; y++) {
      R *rp = buf + (size_t)y * pitch;
      for (int x = 0; x < n; x++) { row[x].re = rp[x]; row[x].im = 0; }
      fft1(&P, row, s, -1);
      cpx *out = (cpx *)rp;
      for (int k = 0; k < h; k++) out[k] = row[k];
    }
#pragma omp for schedule(static)
    for (int kx = 0; kx < h; kx++) {
      cpx *c = (cpx *)buf;
      for (int y = 0; y < n; y++) row[y] = c[(size_t)y * h + kx];
      fft1(&P, row, s, -1);
      for (int y = 0; y < n; y++) c[(size_t)y * h + kx] = row[y];
    }
    free(row);
    free(s);
  }
  plan_free(&P);
  return 0;
}

/* cufunc.cu:229-233 (cufftExecC2R in place); DC/Nyquist imaginary parts of
 * each row are ignored after the column pass (see header). */
int or_c2r2d(R *buf, int n) {
  if (n < 2 || (n & 1)) return -1;
  const int h = n / 2 + 1;
  const int pitch = n + 2;
  fplan P;
  plan_init(&P, n);
  const size_t sl = scratch_len(&P);
#pragma omp parallel
  {
    cpx *row = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    cpx *s = (cpx *)malloc(sizeof(cpx) * sl);
#pragma omp for schedule(static)
    for (int kx = 0; kx < h; kx++) {
      cpx *c = (cpx *)buf;
      for (int y = 0; y < n; y++) row[y] = c[(size_t)y * h + kx];
      fft1(&P, row, s, +1);
      for (int y = 0; y < n; y++) c[(size_t)y * h + kx] = row[y];
    }
#pragma omp for schedule(static)
    for (int y = 0; y < n; y++) {
      R *rp = buf + (size_t)y * pitch;
      const cpx *in = (const cpx *)rp;
      for (int k = 0; k < n; k++) { row[k].re = 0; row[k].im = 0; }
      for (int k = 0; k < h; k++) {
        R a = (k == 0 || k == n / 2) ? (R)1 : (R)2;
        row[k].re = a * in[k].re;
        row[k].im = a * in[k].im;
      }
      fft1(&P, row, s, +1);
      for (int x = 0; x < n; x++) rp[x] = row[x].re;
      rp[n] = 0;
      rp[n + 1] = 0;
    }
    free(row);
    free(s);
  }
  plan_free(&P);
  return 0;
}

/* ======================================================================
 * Geometry
 * ==================================================================== */
/* func.cpp:80-103 */
int or_verify_crop_size(int nx, int ny, int offsetx, int offsety, int nsamout,
                        float bin) {
  if (offsetx < 0 || offsetx >= nx || offsety < 0 || offsety >= ny || bin <= 0)
    return 0;
  if (nsamout <= 0) nsamout = (nx - offsetx) < (ny - offsety) ? (nx - offsetx) : (ny - offsety);
  if ((nx - offsetx) < nsamout) nsamout = nx - offsetx;
  if ((ny - offsety) < nsamout) nsamout = ny - offsety;
  if (fabs(bin - 1.0) < 0.0001)
    nsamout = nsamout / 2 * 2;
  else if (fabs(bin - 2.0) < 0.0001)
    nsamout = nsamout / 4 * 4;
  else
    nsamout = nsamout / 2 * 2;
  return nsamout;
}

/* DFAlign.cu:515 */
int or_binned_size(int nsamUnbin, float bin) {
  return (int)((float)nsamUnbin / bin) / 2 * 2;
}

/* func.cpp:157-173 */
int or_crop2fft(const float *in, int nx, int ny, R *out, int offsetx,
                int offsety, int nsam) {
  (void)ny;
  for (int i = 0; i < nsam; i++) {
    const float *src = in + (size_t)(i + offsety) * nx + offsetx;
    R *dst = out + (size_t)i * (nsam + 2);
    for (int x = 0; x < nsam; x++) dst[x] = (R)src[x];
    dst[nsam] = 0;
    dst[nsam + 1] = 0;
  }
  return nsam;
}

/* Padding geometry (new; no reference counterpart). */
int or_pad2fft(const float *in, int nx, int ny, R *out, int offsetx,
               int offsety, int npad) {
  int w = nx - offsetx, hgt = ny - offsety;
  if (w > npad) w = npad;
  if (hgt > npad) hgt = npad;
  double sum = 0;
  for (int y = 0; y < hgt; y++)
    for (int x = 0; x < w; x++) sum += in[(size_t)(y + offsety) * nx + x + offsetx];
  const R fill = (R)(sum / ((double)w * (double)hgt));
  for (int y = 0; y < npad; y++) {
    R *dst = out + (size_t)y * (npad + 2);
    for (int x = 0; x < npad; x++)
      dst[x] = (y < hgt && x < w) ? (R)in[(size_t)(y + offsety) * nx + x + offsetx] : fill;
    dst[npad] = 0;
    dst[npad + 1] = 0;
  }
  return npad;
}

/* cufunc.cu:162-176: rows 0..n/2-1 and the last n/2 rows, n+2 floats each */
int or_bin_crop(R *dst, const R *src, int dst_n, int src_n) {
  const size_t rowb = sizeof(R) * (size_t)(dst_n + 2);
  for (int i = 0; i < dst_n / 2; i++) {
    memcpy(dst + (size_t)i * (dst_n + 2), src + (size_t)i * (src_n + 2), rowb);
    memcpy(dst + (size_t)(dst_n - 1 - i) * (dst_n + 2),
           src + (size_t)(src_n - 1 - i) * (src_n + 2), rowb);
  }
  return 0;
}

/* ======================================================================
 * Spectral kernels
 * ==================================================================== */
/* cufunc.cu:309-341.  Element (kx, ky) of the n x (n/2+1) spectrum with
 * y' = ky (ky < n/2) or ky - n; z = (-1)^{x+y'} exp(-B (x^2+y'^2)/(2 n^2)).
 * The FP32 build evaluates exactly as the reference does (float m, expf). */
R or_weight(int n, float bfactor, int kx, int ky) {
  int yy = ky < n / 2 ? ky : ky - n;
  int r2 = kx * kx + yy * yy;
#if defined(MC_ORACLE_F32)
  float m = (float)(-0.5 * bfactor / n / n);
  float z = expf(m * (float)r2);
#else
  double z = exp(-0.5 * (double)bfactor * (double)r2 / ((double)n * (double)n));
#endif
  return ((kx + yy) % 2 == 0) ? (R)z : -(R)z;
}

/* cufunc.cu:343-370, sx = sy = 0 (DFAlign.cu:892-894,1004-1006) */
int or_cross_power(R *dfft, const R *dsum, int n, float bfactor) {
  const int h = n / 2 + 1;
#if !defined(MC_ORACLE_F32)
  const double sizec = (double)h * (double)n;
#endif
#pragma omp parallel for schedule(static)
  for (int ky = 0; ky < n; ky++) {
    cpx *f = (cpx *)(dfft + (size_t)ky * (n + 2));
    const cpx *s = (const cpx *)(dsum + (size_t)ky * (n + 2));
    for (int kx = 0; kx < h; kx++) {
      R z = or_weight(n, bfactor, kx, ky);
      cpx a = s[kx], b = f[kx];
      cpx v = {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; /* a*conj(b) */
#if defined(MC_ORACLE_F32)
      v.re /= (R)(h * n);
      v.re *= z;
      v.im /= (R)(h * n);
      v.im *= z;
#else
      v.re = v.re / (R)sizec * z;
      v.im = v.im / (R)sizec * z;
#endif
      f[kx] = v;
    }
  }
  return 0;
}

/* cufunc.cu:372-394 */
int or_phase_shift(R *dfft, int n, double sx, double sy) {
  const int h = n / 2 + 1;
#if defined(MC_ORACLE_F32)
  const float shx = (float)((float)sx * 2.0 * OR_PI / n);
  const float shy = (float)((float)sy * 2.0 * OR_PI / n);
#else
  const double shx = sx * 2.0 * OR_PI / n;
  const double shy = sy * 2.0 * OR_PI / n;
#endif
#pragma omp parallel for schedule(static)
  for (int ky = 0; ky < n; ky++) {
    int yy = ky < n / 2 ? ky : ky - n;
    cpx *f = (cpx *)(dfft + (size_t)ky * (n + 2));
    for (int kx = 0; kx < h; kx++) {
#if defined(MC_ORACLE_F32)
      float a = shx * kx + shy * yy;
      cpx ph = {cosf(a), sinf(a)};
#else
      double a = shx * kx + shy * yy;
      cpx ph = {cos(a), sin(a)};
#endif
      f[kx] = cmul(f[kx], ph);
    }
  }
  return 0;
}

/* func.cpp:348-370 applied to a 16 x 9 complex half spectrum */
static void cosmask(cpx *pf, int ns) {
  const int hn = ns / 2, hnb = hn + 1;
  const double step = OR_PI / hn;
  for (int j = 0; j < hn; j++)
    for (int i = 0; i < hnb; i++) {
      int id = j * hnb + i;
      int r = (int)sqrt((double)(i * i + j * j));
      if (r < hn) {
        R m = (R)(cos(r * step) / 2 + 0.5);
        pf[id].re *= m;
        pf[id].im *= m;
      } else {
        pf[id].re = 0;
        pf[id].im = 0;
      }
      id = (j + hn) * hnb + i;
      int jj = j - hn;
      r = (int)sqrt((double)(i * i + jj * jj));
      if (r < hn) {
        R m = (R)(cos(r * step) / 2 + 0.5);
        pf[id].re *= m;
        pf[id].im *= m;
      } else {
        pf[id].re = 0;
        pf[id].im = 0;
      }
    }
}

/* cufunc.cu:409-496 applied to an extracted box map */
double or_find_shift_box(const R *boxmap, int n, int box, double *psx,
                         double *psy, int wnoise) {
  double best = -1e9;
  int sx = 0, sy = 0;
  /* cufunc.cu:411-424: first maximum in row-major order, noise excluded */
  for (int j = 0; j < box; j++)
    for (int i = 0; i < box; i++) {
      if (abs(i - box / 2) <= wnoise && abs(j - box / 2) <= wnoise) continue;
      if (boxmap[j * box + i] > best) {
        best = boxmap[j * box + i];
        sx = i;
        sy = j;
      }
    }
  /* cufunc.cu:427-437 */
  const int sub = 16;
  int offx = sx - sub / 2, offy = sy - sub / 2;
  if (offx + sub > box) offx = box - sub;
  else if (offx < 0) offx = 0;
  if (offy + sub > box) offy = box - sub;
  else if (offy < 0) offy = 0;
  R subm[16 * 18];
  for (int y = 0; y < sub; y++) {
    for (int x = 0; x < sub; x++) subm[y * 18 + x] = boxmap[(y + offy) * box + x + offx];
    subm[y * 18 + 16] = 0;
    subm[y * 18 + 17] = 0;
  }
  or_r2c2d(subm, sub);
  cosmask((cpx *)subm, sub);
  /* cufunc.cu:441-455: zero-pad to 512 x 512 and inverse */
  const int scale = 32, pad = sub * scale;
  const int wns = scale * wnoise;
  R *pm = (R *)calloc((size_t)(pad + 2) * pad, sizeof(R));
  for (int i = 0; i < sub / 2; i++) {
    memcpy(pm + (size_t)i * (pad + 2), subm + i * 18, sizeof(R) * 18);
    memcpy(pm + (size_t)(pad - 1 - i) * (pad + 2), subm + (sub - 1 - i) * 18,
           sizeof(R) * 18);
  }
  or_c2r2d(pm, pad);
  /* cufunc.cu:456-476 */
  const int ox = (box / 2 - offx) * scale, oy = (box / 2 - offy) * scale;
  int sxp = 0, syp = 0;
  R bestp = pm[0];
  for (int j = 0; j < pad; j++) {
    const R *rp = pm + (size_t)j * (pad + 2);
    for (int i = 0; i < pad; i++) {
      if (abs(i - ox) <= wns && abs(j - oy) <= wns) continue;
      if (rp[i] > bestp) {
        bestp = rp[i];
        sxp = i;
        syp = j;
      }
    }
  }
  free(pm);
  /* cufunc.cu:478-496 */
  *psx = offx + sxp / (double)scale - box / 2;
  *psy = offy + syp / (double)scale - box / 2;
  return (double)bestp / n / n / sub / sub;
}

/* cufunc.cu:396-497 */
double or_find_shift(const R *cmap, int n, R *boxmap, int box, double *psx,
                     double *psy, int wnoise) {
  const int ori = (n - box) / 2;
  for (int i = 0; i < box; i++)
    memcpy(boxmap + (size_t)i * box, cmap + (size_t)(i + ori) * (n + 2) + ori,
           sizeof(R) * box);
  return or_find_shift_box(boxmap, n, box, psx, psy, wnoise);
}

/* Pruned inverse (exact-arithmetic equal to full C2R + box read-out):
 * every column is inverse-transformed, but only the box rows are kept and
 * only those rows get the row C2R; then only the box columns are read. */
int or_pruned_box(const R *X, int n, int box, R *boxmap) {
  const int h = n / 2 + 1;
  const int ori = (n - box) / 2;
  fplan P;
  plan_init(&P, n);
  const size_t sl = scratch_len(&P);
  cpx *T = (cpx *)malloc(sizeof(cpx) * (size_t)box * h);
  const cpx *Xc = (const cpx *)X;
#pragma omp parallel
  {
    cpx *col = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    cpx *s = (cpx *)malloc(sizeof(cpx) * sl);
#pragma omp for schedule(static)
    for (int kx = 0; kx < h; kx++) {
      for (int ky = 0; ky < n; ky++) col[ky] = Xc[(size_t)ky * h + kx];
      fft1(&P, col, s, +1);
      for (int r = 0; r < box; r++) T[(size_t)r * h + kx] = col[ori + r];
    }
#pragma omp for schedule(static)
    for (int r = 0; r < box; r++) {
      for (int k = 0; k < n; k++) { col[k].re = 0; col[k].im = 0; }
      for (int k = 0; k < h; k++) {
        R a = (k == 0 || k == n / 2) ? (R)1 : (R)2;
        col[k].re = a * T[(size_t)r * h + k].re;
        col[k].im = a * T[(size_t)r * h + k].im;
      }
      fft1(&P, col, s, +1);
      for (int c = 0; c < box; c++) boxmap[(size_t)r * box + c] = col[ori + c].re;
    }
    free(col);
    free(s);
  }
  free(T);
  plan_free(&P);
  return 0;
}

/* ======================================================================
 * Drift model (mfunc.h)
 * ==================================================================== */
/* mfunc.h:50-84 */
int or_pair_list(int nframe, int fod, int *pairs, int cap) {
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

/* Least squares for the pair design matrix A[p][k] = 1 for i <= k < j
 * (mfunc.h:76-81), solved by Householder QR in FP64 (replaces SP++
 * lsSolver at DFAlign.cu:915,936). */
int or_lsq(const int *pairs, int np, int nframe, const double *b, double *s) {
  const int K = nframe - 1;
  if (K <= 0) return -1;
  if (np < K) return -2;
  double *A = (double *)calloc((size_t)np * K, sizeof(double));
  double *B = (double *)malloc(sizeof(double) * (size_t)np * 2);
  for (int p = 0; p < np; p++) {
    for (int k = pairs[2 * p]; k < pairs[2 * p + 1]; k++) A[(size_t)p * K + k] = 1.0;
    B[2 * p] = b[2 * p];
    B[2 * p + 1] = b[2 * p + 1];
  }
  int rc = 0;
  for (int c = 0; c < K; c++) {
    double nrm = 0;
    for (int p = c; p < np; p++) nrm += A[(size_t)p * K + c] * A[(size_t)p * K + c];
    nrm = sqrt(nrm);
    if (nrm == 0) { rc = -3; break; }
    double alpha = A[(size_t)c * K + c] > 0 ? -nrm : nrm;
    /* v = a - alpha e1, stored in place */
    A[(size_t)c * K + c] -= alpha;
    double vnorm2 = 0;
    for (int p = c; p < np; p++) vnorm2 += A[(size_t)p * K + c] * A[(size_t)p * K + c];
    if (vnorm2 > 0) {
      for (int cc = c + 1; cc < K; cc++) {
        double d = 0;
        for (int p = c; p < np; p++) d += A[(size_t)p * K + c] * A[(size_t)p * K + cc];
        d = 2 * d / vnorm2;
        for (int p = c; p < np; p++) A[(size_t)p * K + cc] -= d * A[(size_t)p * K + c];
      }
      for (int q = 0; q < 2; q++) {
        double d = 0;
        for (int p = c; p < np; p++) d += A[(size_t)p * K + c] * B[2 * p + q];
        d = 2 * d / vnorm2;
        for (int p = c; p < np; p++) B[2 * p + q] -= d * A[(size_t)p * K + c];
      }
    }
    A[(size_t)c * K + c] = alpha; /* R diagonal; below-diagonal part unused */
  }
  if (rc == 0) {
    for (int q = 0; q < 2; q++)
      for (int c = K - 1; c >= 0; c--) {
        double acc = B[2 * c + q];
        for (int cc = c + 1; cc < K; cc++) acc -= A[(size_t)c * K + cc] * s[2 * cc + q];
        s[2 * c + q] = acc / A[(size_t)c * K + c];
      }
  }
  free(A);
  free(B);
  return rc;
}

/* DFAlign.cu:916: ki = abs(A*shift - b) (complex modulus) */
void or_residual(const int *pairs, int np, const double *s, const double *b,
                 double *ki) {
  for (int p = 0; p < np; p++) {
    double x = 0, y = 0;
    for (int k = pairs[2 * p]; k < pairs[2 * p + 1]; k++) {
      x += s[2 * k];
      y += s[2 * k + 1];
    }
    x -= b[2 * p];
    y -= b[2 * p + 1];
    ki[p] = sqrt(x * x + y * y);
  }
}

static int uf_find(int *par, int a) {
  while (par[a] != a) {
    par[a] = par[par[a]];
    a = par[a];
  }
  return a;
}

/* Rank of the selected design rows == nframe-1 iff the pair graph on the
 * frames is connected (row (i,j) = e_j - e_i in position coordinates).
 * Replaces CSVD::rank (mfunc.h:139-140). */
static int rows_full_rank(const int *pairs, const int *sel, int nsel,
                          int nframe) {
  int *par = (int *)malloc(sizeof(int) * (size_t)nframe);
  for (int i = 0; i < nframe; i++) par[i] = i;
  int comps = nframe;
  for (int q = 0; q < nsel; q++) {
    int a = uf_find(par, pairs[2 * sel[q]]), b = uf_find(par, pairs[2 * sel[q] + 1]);
    if (a != b) {
      par[a] = b;
      comps--;
    }
  }
  free(par);
  return comps == 1;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* mfunc.h:86-146 */
int or_threshold(const int *pairs, int np, int nframe, const double *ki,
                 double kithresh, int *good) {
  const int K = nframe - 1;
  int *order = (int *)malloc(sizeof(int) * (size_t)np);
  double *tki = (double *)malloc(sizeof(double) * (size_t)np);
  memcpy(tki, ki, sizeof(double) * (size_t)np);
  /* mfunc.h:111-116: repeated first-minimum selection */
  for (int i = 0; i < np; i++) {
    int pos = 0;
    double m = tki[0];
    for (int q = 1; q < np; q++)
      if (tki[q] < m) {
        m = tki[q];
        pos = q;
      }
    order[i] = pos;
    tki[pos] = 1e20;
  }
  /* mfunc.h:119-123 */
  int i;
  for (i = np - 1; i >= K; i--)
    if (ki[order[i]] < kithresh) break;
  int ii = i;
  int nsel;
  do {
    nsel = ii + 1;
    for (int q = 0; q < nsel; q++) good[q] = order[q];
    qsort(good, (size_t)nsel, sizeof(int), cmp_int);
    int full = rows_full_rank(pairs, good, nsel, nframe);
    ii++;
    if (full) break;
  } while (ii < np);
  free(order);
  free(tki);
  return nsel;
}

/* ======================================================================
 * Whole movie (DFAlign.cu:456-1355; numerics only)
 * ==================================================================== */
static void c2r_to_image(R *spec, int n, R *img) {
  /* DFAlign.cu:1288-1292: C2R, scale 1/n^2 (float 1.0/size), unpad */
  or_c2r2d(spec, n);
  const R sc = (R)(1.0 / ((double)n * (double)n));
  for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++) img[(size_t)y * n + x] = spec[(size_t)y * (n + 2) + x] * sc;
}

static int disp_image(const R *spec, int n, float scale, int fft, short *out);

static void add_spec(R *dst, const R *src, size_t len) {
  for (size_t i = 0; i < len; i++) dst[i] += src[i];
}

/* Frame -> spectrum: DFAlign.cu:768-779 / 792-805 */
static int frame_spectrum(const float *frame, const or_params *p, int nsamUnbin,
                          int n, R *work, R *out) {
  if (p->pad > 0)
    or_pad2fft(frame, p->nx, p->ny, work, p->crop_offsetx, p->crop_offsety, nsamUnbin);
  else
    or_crop2fft(frame, p->nx, p->ny, work, p->crop_offsetx, p->crop_offsety, nsamUnbin);
  or_r2c2d(work, nsamUnbin);
  if (p->bin > 1.0001f)
    or_bin_crop(out, work, n, nsamUnbin);
  else
    memcpy(out, work, sizeof(R) * (size_t)n * (n + 2));
  return 0;
}

/* Cross-correlate spectra a (first operand) and b (conjugated); write the
 * measured (sx, sy) and return cc.  DFAlign.cu:889-900 / 1000-1010. */
static double correlate(const R *a, const R *b, const or_params *p, int n,
                        R *work, R *boxmap, double *sx, double *sy) {
  const size_t sb = (size_t)n * (n + 2);
  memcpy(work, b, sizeof(R) * sb);
  or_cross_power(work, a, n, p->bfactor);
  if (p->full_c2r) {
    or_c2r2d(work, n);
    return or_find_shift(work, n, boxmap, p->box, sx, sy, p->nps - 1);
  }
  or_pruned_box(work, n, p->box, boxmap);
  return or_find_shift_box(boxmap, n, p->box, sx, sy, p->nps - 1);
}

int or_align_movie(const float *frames, const float *frames2,
                   const or_params *p, or_result *r) {
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  int nsamUnbin;
  if (p->pad > 0) {
    nsamUnbin = p->pad;
  } else {
    nsamUnbin = or_verify_crop_size(p->nx, p->ny, p->crop_offsetx, p->crop_offsety,
                                    p->crop_nsam, p->bin);
  }
  if (nsamUnbin <= 0) return -1;
  const int n = p->pad > 0 ? or_binned_size(nsamUnbin, p->bin)
                           : or_binned_size(nsamUnbin, p->bin);
  const int nframe = p->nEnd - p->nStart + 1;
  if (nframe < 2) return -2;
  if (p->box < 16 || p->box > n) return -3;
  const size_t sb = (size_t)n * (n + 2);
  const size_t sbu = (size_t)nsamUnbin * (nsamUnbin + 2);
  r->n = n;
  r->nsamUnbin = nsamUnbin;
  r->nframe = nframe;

  /* 1. spectra of all frames (hbuf from frames, hbuf2 from frames2) */
  R *hbuf = (R *)malloc(sizeof(R) * sb * nframe);
  R *hbuf2 = frames2 ? (R *)malloc(sizeof(R) * sb * nframe) : hbuf;
  R *work = (R *)malloc(sizeof(R) * (sbu > sb ? sbu : sb));
  const size_t fsz = (size_t)p->nx * p->ny;
  R *rs0 = (R *)calloc(sb, sizeof(R)), *rs1 = (R *)calloc(sb, sizeof(R));
  int even = 1;
  for (int j = p->nStart; j <= p->nEnd; j++) {
    R *dst = hbuf + (size_t)(j - p->nStart) * sb;
    if (frames2)
      frame_spectrum(frames2 + (size_t)j * fsz, p, nsamUnbin, n, work,
                     hbuf2 + (size_t)(j - p->nStart) * sb);
    frame_spectrum(frames + (size_t)j * fsz, p, nsamUnbin, n, work, dst);
    /* DFAlign.cu:808-813 raw even/odd sums */
    if (j >= p->nStartSum && j <= p->nEndSum) {
      add_spec(even ? rs0 : rs1, dst, sb);
      even = !even;
    }
  }
  if (r->spectra) memcpy(r->spectra, hbuf, sizeof(R) * sb * nframe);
  /* DFAlign.cu:821-822: the raw halves are kept for the FSC table */
  R *fsc_raw0 = NULL, *fsc_raw1 = NULL;
  if (r->fsc) {
    fsc_raw0 = (R *)malloc(sizeof(R) * sb);
    memcpy(fsc_raw0, rs0, sizeof(R) * sb);
    fsc_raw1 = rs1;
    rs1 = NULL;
  }
  const float fftscale = p->fftscale > 0 ? p->fftscale : 0.0001f;
  r->disp_written = 0;
  if (r->rawsum_image || (r->disp_raw_fft && n >= OR_DISPDIM)) {
    add_spec(rs0, fsc_raw1 ? fsc_raw1 : rs1, sb); /* DFAlign.cu:823 */
    /* DFAlign.cu:842-848, 908-911: log-modulus display of the raw sum */
    if (r->disp_raw_fft && n >= OR_DISPDIM && disp_image(rs0, n, fftscale, 1, r->disp_raw_fft))
      r->disp_written |= 1;
    if (r->rawsum_image) {
      c2r_to_image(rs0, n, r->rawsum_image);
      r->rawsum_written = 1;
    }
  }
  free(rs0);
  free(rs1);

  /* 2. pairwise shifts (DFAlign.cu:875-905) */
  const int np = or_pair_list(nframe, p->fod, NULL, 0);
  const int K = nframe - 1;
  r->npairs = np;
  r->ncols = K;
  int *pairs = (int *)malloc(sizeof(int) * 2 * (size_t)np);
  or_pair_list(nframe, p->fod, pairs, np);
  double *b = (double *)malloc(sizeof(double) * 2 * (size_t)np);
  const int box = p->box;
  R *boxes = (R *)malloc(sizeof(R) * (size_t)box * box * np);
  for (int q = 0; q < np; q++) {
    double sx, sy;
    double cc = correlate(hbuf + (size_t)pairs[2 * q] * sb,
                          hbuf2 + (size_t)pairs[2 * q + 1] * sb, p, n, work,
                          boxes + (size_t)q * box * box, &sx, &sy);
    /* DFAlign.cu:900-901: FindShift returns float sx, sy, cc */
    b[2 * q] = (double)(or_real)sx;
    b[2 * q + 1] = (double)(or_real)sy;
    if (r->pair_cc) r->pair_cc[q] = cc;
  }
  if (r->pairs) memcpy(r->pairs, pairs, sizeof(int) * 2 * (size_t)np);
  if (r->pair_b) memcpy(r->pair_b, b, sizeof(double) * 2 * (size_t)np);
  if (r->boxmaps) memcpy(r->boxmaps, boxes, sizeof(R) * (size_t)box * box * np);

  /* 3. LSQ, outlier rejection, LSQ (DFAlign.cu:914-948) */
  double *s = (double *)malloc(sizeof(double) * 2 * (size_t)K);
  double *ki = (double *)malloc(sizeof(double) * (size_t)np);
  int rc = or_lsq(pairs, np, nframe, b, s);
  if (rc) goto fail;
  or_residual(pairs, np, s, b, ki);
  if (r->shift1) memcpy(r->shift1, s, sizeof(double) * 2 * K);
  if (r->ki1) memcpy(r->ki1, ki, sizeof(double) * np);
  int *good = (int *)malloc(sizeof(int) * (size_t)np);
  int ng = or_threshold(pairs, np, nframe, ki, p->kit, good);
  int *gp = (int *)malloc(sizeof(int) * 2 * (size_t)ng);
  double *gb = (double *)malloc(sizeof(double) * 2 * (size_t)ng);
  for (int q = 0; q < ng; q++) {
    gp[2 * q] = pairs[2 * good[q]];
    gp[2 * q + 1] = pairs[2 * good[q] + 1];
    gb[2 * q] = b[2 * good[q]];
    gb[2 * q + 1] = b[2 * good[q] + 1];
  }
  rc = or_lsq(gp, ng, nframe, gb, s);
  r->ngood = ng;
  if (r->good) memcpy(r->good, good, sizeof(int) * ng);
  if (r->shift2) memcpy(r->shift2, s, sizeof(double) * 2 * K);
  if (r->ki2) or_residual(gp, ng, s, gb, r->ki2);
  free(gp);
  free(gb);
  free(good);
  if (rc) goto fail;

  /* 4. iterative alignment (DFAlign.cu:950-1071) */
  {
    double *oldx = (double *)malloc(sizeof(double) * 2 * nframe);
    double *newx = (double *)malloc(sizeof(double) * 2 * nframe);
    double ax = 0, ay = 0;
    newx[0] = 0;
    newx[1] = 0;
    for (int k = 0; k < K; k++) {
      newx[2 * (k + 1)] = ax + s[2 * k];
      newx[2 * (k + 1) + 1] = ay + s[2 * k + 1];
      ax += s[2 * k];
      ay += s[2 * k + 1];
    }
    memcpy(oldx, newx, sizeof(double) * 2 * nframe);
    double rmsd = 0;
    int it, converged = 0;
    R *sum = (R *)malloc(sizeof(R) * sb);
    R *tmp = (R *)malloc(sizeof(R) * sb);
    for (it = 0; it < p->iter && !converged; it++) {
      memcpy(sum, hbuf, sizeof(R) * sb);
      or_phase_shift(sum, n, -oldx[0], -oldx[1]);
      for (int k = 1; k < nframe; k++) {
        memcpy(tmp, hbuf + (size_t)k * sb, sizeof(R) * sb);
        or_phase_shift(tmp, n, -oldx[2 * k], -oldx[2 * k + 1]);
        add_spec(sum, tmp, sb);
      }
      for (int k = 0; k < nframe; k++) {
        double sx, sy;
        correlate(sum, hbuf + (size_t)k * sb, p, n, work, boxes, &sx, &sy);
        newx[2 * k] = (double)(or_real)sx;
        newx[2 * k + 1] = (double)(or_real)sy;
      }
      if (r->iter_meas) memcpy(r->iter_meas + (size_t)it * 2 * nframe, newx, sizeof(double) * 2 * nframe);
      double rx = newx[0], ry = newx[1];
      for (int k = 0; k < nframe; k++) {
        double dx = oldx[2 * k] - newx[2 * k], dy = oldx[2 * k + 1] - newx[2 * k + 1];
        oldx[2 * k] = newx[2 * k] - rx;
        oldx[2 * k + 1] = newx[2 * k + 1] - ry;
        rmsd += dx * dx + dy * dy;
      }
      rmsd /= nframe;
      if (r->rmsd) r->rmsd[it] = rmsd;
      if (rmsd < (double)p->iterthres * (double)p->iterthres) converged = 1;
    }
    r->niter = it;
    for (int k = 0; k < K; k++) {
      s[2 * k] = oldx[2 * (k + 1)] - oldx[2 * k];
      s[2 * k + 1] = oldx[2 * (k + 1) + 1] - oldx[2 * k + 1];
    }
    if (r->final_shift) memcpy(r->final_shift, s, sizeof(double) * 2 * K);
    free(oldx);
    free(newx);
    free(tmp);

    /* 5. final sum (DFAlign.cu:1114-1292) */
    int nss = p->nStartSum - p->nStart, nes = p->nEndSum - p->nStart,
        nes2 = p->nEndSum2 - p->nStart;
    if (nss <= 0) nss = 0;
    if (nes <= nss || nes >= nframe) nes = nframe - 1;
    if (nes2 <= nss || nes2 >= nframe) nes2 = nframe - 1;
    if (nes2 > nes) nes2 = nes;
    double midx = 0, midy = 0;
    if (p->atm) {
      int ref = p->nz / 2 + 1;
      if (ref < p->nStart) ref = p->nStart;
      int lim = ref - p->nStart;
      if (lim > K) lim = K; /* guards the out-of-range read of DFAlign.cu:1138 */
      for (int j = 0; j < lim; j++) {
        midx += s[2 * j];
        midy += s[2 * j + 1];
      }
    }
    if (r->frame_shift)
      for (int k = 0; k < 2 * nframe; k++) r->frame_shift[k] = NAN;
    double tx = 0, ty = 0;
    for (int j = 1; j < nss + 1; j++) {
      tx += s[2 * (j - 1)];
      ty += s[2 * (j - 1) + 1];
    }
    R *dsum = (R *)calloc(sb, sizeof(R));
    R *dsumcorr = sum;
    memcpy(dsumcorr, hbuf + (size_t)nss * sb, sizeof(R) * sb);
    if (p->atm) or_phase_shift(dsumcorr, n, -tx + midx, -ty + midy);
    if (r->frame_shift) {
      r->frame_shift[2 * nss] = p->atm ? (-tx + midx) : 0.0;
      r->frame_shift[2 * nss + 1] = p->atm ? (-ty + midy) : 0.0;
    }
    int fsceven = 0;
    R *fr = (R *)malloc(sizeof(R) * sb);
    for (int j = nss + 1; j <= nes; j++) {
      tx += s[2 * (j - 1)];
      ty += s[2 * (j - 1) + 1];
      memcpy(fr, hbuf + (size_t)j * sb, sizeof(R) * sb);
      or_phase_shift(fr, n, -tx + midx, -ty + midy);
      if (r->frame_shift) {
        r->frame_shift[2 * j] = -tx + midx;
        r->frame_shift[2 * j + 1] = -ty + midy;
      }
      add_spec(fsceven ? dsumcorr : dsum, fr, sb);
      fsceven = !fsceven;
      if (j == nes2 && p->nEndSum2 > 0 && p->nEndSum2 < p->nEndSum && r->sum2_image) {
        /* DFAlign.cu:1198-1229 second sum */
        memcpy(fr, dsumcorr, sizeof(R) * sb);
        add_spec(fr, dsum, sb);
        c2r_to_image(fr, n, r->sum2_image);
        r->sum2_written = 1;
      }
    }
    if (r->fsc) {
      /* DFAlign.cu:1265-1266,1314-1318: Corr0 = dsumcorr, Corr1 = dsum,
       * direction = total shift of the last summed frame */
      const int nb = or_fsc_bins(n, NULL);
      double *t = (double *)malloc(sizeof(double) * 4 * (size_t)nb);
      or_fsc(fsc_raw0, fsc_raw1, n, tx, ty, t);
      or_fsc(dsumcorr, dsum, n, tx, ty, t + 2 * (size_t)nb);
      for (int i = 0; i < nb; i++) {
        r->fsc[4 * i] = t[2 * i];
        r->fsc[4 * i + 1] = t[2 * i + 1];
        r->fsc[4 * i + 2] = t[2 * nb + 2 * i];
        r->fsc[4 * i + 3] = t[2 * nb + 2 * i + 1];
      }
      free(t);
    }
    add_spec(dsumcorr, dsum, sb);
    /* DFAlign.cu:1270-1285 quick-look images of the aligned sum */
    if (r->disp_corr_fft && n >= OR_DISPDIM && disp_image(dsumcorr, n, fftscale, 1, r->disp_corr_fft))
      r->disp_written |= 4;
    if (r->disp_corr_sum && n >= OR_DISPDIM && disp_image(dsumcorr, n, fftscale, 0, r->disp_corr_sum))
      r->disp_written |= 2;
    if (r->sum_image) c2r_to_image(dsumcorr, n, r->sum_image);
    free(fr);
    free(dsum);
    free(sum);
  }
  free(s);
  free(ki);
  free(b);
  free(boxes);
  free(pairs);
  free(work);
  free(fsc_raw0);
  free(fsc_raw1);
  if (hbuf2 != hbuf) free(hbuf2);
  free(hbuf);
  return 0;
fail:
  free(fsc_raw0);
  free(fsc_raw1);
  free(s);
  free(ki);
  free(b);
  free(boxes);
  free(pairs);
  free(work);
  if (hbuf2 != hbuf) free(hbuf2);
  free(hbuf);
  return -10;
}

/* DFAlign.cu:214-218 */
int or_fsc_bins(int n, int *step) {
  int st = n / 400;
  if (st < 1) st = 1;
  if (step) *step = st;
  return (n / 2 + 1) / st + 1;
}

/* DFAlign.cu:212-297.  The element order of the reference loop is the
 * padded layout's (row ky, column kx); MkPosList (cufunc.cu:309-340) gives
 * row ky the frequency y = ky for ky < n/2, else ky - n. */
int or_fsc(const R *a, const R *b, int n, double dx, double dy, double *out) {
  int step;
  const int nbox = or_fsc_bins(n, &step);
  const int nsamc = n / 2 + 1;
  const double dn = hypot(dx, dy); /* abs(complex<double>) */
  if (dn > 0.1) {
    dx /= dn;
    dy /= dn;
  } else {
    dx = 1.0;
    dy = 0.0;
  }
  const float edge = (float)cos(OR_PI / 4);
  double *acc = (double *)calloc((size_t)nbox * 6, sizeof(double));
  if (!acc) return -1;
  for (int ky = 0; ky < n; ky++) {
    const int y = ky < n / 2 ? ky : ky - n;
    const cpx *A = (const cpx *)(a + (size_t)ky * (n + 2));
    const cpx *B = (const cpx *)(b + (size_t)ky * (n + 2));
    for (int x = 0; x < nsamc; x++) {
      const float r = sqrtf((float)(x * x + y * y));
      if ((int)r <= 0 || r >= (float)nsamc) continue;
      const int id = (int)(r / (float)step);
      const float angle = (float)(fabs(x * dx + y * dy) / (double)r);
      double *s = acc + ((size_t)id * 2 + (angle > edge ? 0 : 1)) * 3;
      s[0] += (double)A[x].re * B[x].re + (double)A[x].im * B[x].im;
      s[1] += (double)A[x].re * A[x].re + (double)A[x].im * A[x].im;
      s[2] += (double)B[x].re * B[x].re + (double)B[x].im * B[x].im;
    }
  }
  for (int i = 0; i < 2 * nbox; i++) {
    const double *s = acc + (size_t)i * 3;
    const double t = s[1] * s[2];
    out[i] = t > 0.00001 ? s[0] / sqrt(t) : 0.0;
  }
  free(acc);
  return 0;
}

/* cufunc.cu:178-188 and func.cpp:459-489 */
int or_logmod_image(const R *spec, int n, float scale, R *img) {
  if (n < 4 || (n & 1)) return -1;
  const int hn = n / 2;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      int kx, ky;
      if (i >= hn) { /* right half: rows fftshifted */
        kx = i - hn;
        ky = (j + hn) % n;
      } else if (j >= 1) { /* left half: point mirror of the right half */
        kx = hn - i;
        ky = (hn - j + n) % n;
      } else { /* row 0 copies row 1 */
        kx = hn - i;
        ky = hn - 1;
      }
      const cpx v = ((const cpx *)(spec + (size_t)ky * (n + 2)))[kx];
      img[(size_t)j * n + i] = (R)log(1.0 + sqrt((double)v.re * v.re + (double)v.im * v.im) * scale);
    }
  return 0;
}

/* func.cpp:524-555 up to buf2DispShort */
int or_fft_display(const R *spec, int n, float scale, int d, R *out) {
  if (n < d || d < 2 || (d & 1)) return -1;
  R *buf = (R *)calloc((size_t)n * (n + 2), sizeof(R));
  R *img = (R *)malloc(sizeof(R) * (size_t)n * n);
  R *c = (R *)calloc((size_t)d * (d + 2), sizeof(R));
  or_logmod_image(spec, n, scale, img);
  for (int j = 0; j < n; j++) memcpy(buf + (size_t)j * (n + 2), img + (size_t)j * n, sizeof(R) * n);
  or_r2c2d(buf, n);
  or_bin_crop(c, buf, d, n);
  or_c2r2d(c, d);
  const R dd = (R)((double)d * d);
  for (int y = 0; y < d; y++)
    for (int x = 0; x < d; x++) out[(size_t)y * d + x] = c[(size_t)y * (d + 2) + x] / dd;
  free(buf);
  free(img);
  free(c);
  return 0;
}

/* DFAlign.cu:1281-1284, 73-90 */
int or_sum_display(const R *spec, int n, int d, R *out) {
  if (n < d || d < 2 || (d & 1)) return -1;
  R *c = (R *)calloc((size_t)d * (d + 2), sizeof(R));
  or_bin_crop(c, spec, d, n);
  or_c2r2d(c, d);
  for (int y = 0; y < d; y++)
    for (int x = 0; x < d; x++) out[(size_t)y * d + x] = c[(size_t)y * (d + 2) + x];
  free(c);
  return 0;
}

/* func.cpp:372-411 (histMinMax, threshold 0.005) and 429-443 (buf2DispShort) */
int or_disp_short(const float *buf, int size, short *out) {
  const int nbin = 400;
  const float threshold = 0.005f;
  float mn = 1e20f, mx = -1e20f;
  for (int i = 0; i < size; i++) {
    if (buf[i] < mn) mn = buf[i];
    if (buf[i] > mx) mx = buf[i];
  }
  if (mn >= mx) return 1;
  int *hist = (int *)calloc(nbin, sizeof(int));
  const float step = (mx - mn) / nbin;
  for (int i = 0; i < size; i++) {
    int id = (int)((buf[i] - mn) / step);
    if (id >= nbin) id = nbin - 1;
    hist[id]++;
  }
  const int percent = (int)(size * threshold);
  int count = 0;
  for (int i = 0; i < nbin; i++) {
    count += hist[i];
    if (count > percent) {
      mn += i * step;
      break;
    }
  }
  count = 0;
  for (int i = nbin - 1; i > 0; i--) {
    count += hist[i];
    if (count > percent) {
      mx -= (nbin - 1 - i) * step;
      break;
    }
  }
  free(hist);
  if (mn >= mx) return 1;
  const float sc = 255.0 / (mx - mn);
  for (int i = 0; i < size; i++) {
    if (buf[i] <= mn)
      out[i] = 0;
    else if (buf[i] >= mx)
      out[i] = 255;
    else
      out[i] = (short)((buf[i] - mn) * sc);
  }
  return 0;
}

/* One quick-look image: the float display buffer, then the 8-bit mapping. */
static int disp_image(const R *spec, int n, float scale, int fft, short *out) {
  const int d = OR_DISPDIM;
  R *f = (R *)malloc(sizeof(R) * (size_t)d * d);
  float *ff = (float *)malloc(sizeof(float) * (size_t)d * d);
  if (fft)
    or_fft_display(spec, n, scale, d, f);
  else
    or_sum_display(spec, n, d, f);
  for (size_t i = 0; i < (size_t)d * d; i++) ff[i] = (float)f[i];
  const int rc = or_disp_short(ff, d * d, out);
  free(f);
  free(ff);
  return rc == 0;
}

/* Sharpness objective and gradient (no reference counterpart). */
double or_sharpness(const R *spectra, int nframe, int n, float bfactor,
                    const double *s, double *grad) {
  const int h = n / 2 + 1;
  const size_t sb = (size_t)n * (n + 2);
  double J = 0;
  double *gl = (double *)calloc(2 * (size_t)nframe, sizeof(double));
  /* rows in parallel; each thread sums its rows, the partials are added in
   * thread order */
#pragma omp parallel
  {
    double Jt = 0;
    double *gt = (double *)calloc(2 * (size_t)nframe, sizeof(double));
#pragma omp for schedule(static)
    for (int ky = 0; ky < n; ky++) {
      int yy = ky < n / 2 ? ky : ky - n;
      for (int kx = 0; kx < h; kx++) {
        double w = exp(-0.5 * (double)bfactor * (double)(kx * kx + yy * yy) /
                       ((double)n * (double)n)) *
                   ((kx == 0 || kx == n / 2) ? 1.0 : 2.0);
        double Sr = 0, Si = 0;
        for (int f = 0; f < nframe; f++) {
          const cpx *F = (const cpx *)(spectra + (size_t)f * sb + (size_t)ky * (n + 2));
          double a = 2.0 * OR_PI * (kx * s[2 * f] + yy * s[2 * f + 1]) / n;
          double c = cos(a), sn = sin(a);
          Sr += F[kx].re * c - F[kx].im * sn;
          Si += F[kx].re * sn + F[kx].im * c;
        }
        Jt += w * (Sr * Sr + Si * Si);
        if (grad) {
          for (int f = 0; f < nframe; f++) {
            const cpx *F = (const cpx *)(spectra + (size_t)f * sb + (size_t)ky * (n + 2));
            double a = 2.0 * OR_PI * (kx * s[2 * f] + yy * s[2 * f + 1]) / n;
            double c = cos(a), sn = sin(a);
            double Gr = F[kx].re * c - F[kx].im * sn, Gi = F[kx].re * sn + F[kx].im * c;
            double im = Sr * Gi - Si * Gr; /* Im(conj(S) G) */
            gt[2 * f] += -2.0 * w * (2.0 * OR_PI * kx / n) * im;
            gt[2 * f + 1] += -2.0 * w * (2.0 * OR_PI * yy / n) * im;
          }
        }
      }
    }
#pragma omp for ordered schedule(static, 1)
    for (int t = 0; t < omp_get_num_threads(); t++) {
#pragma omp ordered
      {
        J += Jt;
        for (int f = 0; f < 2 * nframe; f++) gl[f] += gt[f];
      }
    }
    free(gt);
  }
  if (grad)
    for (int f = 0; f < 2 * nframe; f++) grad[f] = gl[f];
  free(gl);
  return J;
}
