// mc_device.h -- device-side building blocks shared by the HIP kernels:
// complex helpers, register DFT butterflies, the LDS Stockham FFT engine and
// XCD-aware block remapping for gfx950 (CDNA4, wave64, 160 KB LDS per WG).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MC_DEV __device__ __forceinline__

namespace mc {

// ---------------------------------------------------------------- complex
MC_DEV float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
MC_DEV float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
MC_DEV float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
MC_DEV float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
MC_DEV float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
MC_DEV float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (S i), S = +1 or -1
template <int S> MC_DEV float2 mul_si(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ------------------------------------------------------- buffer loads
// A raw buffer resource over `bytes` bytes at p (uniform).  A load through it
// takes a per-lane 32-bit byte offset and a scalar byte offset, so the base is
// added by the address unit instead of by 64-bit vector adds per load, and an
// offset at or past `bytes` reads zero.  Word 3 (0x00020000) is the gfx9
// family's 32-bit data format with no swizzle, as raw loads need.
MC_DEV __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}
// (The 64-bit result is reinterpreted whole: taking its two 32-bit elements
// one by one made hipcc emit a 32-bit load and reuse the first word.)
MC_DEV float2 buf_ld2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v v = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  return make_float2(v.x, v.y);
}

// twiddle e^{S * 2 pi i e / N} from a table t[e] = e^{+2 pi i e / N}
template <int S> MC_DEV float2 twid(const float2* t, int e) {
  float2 w = t[e];
  return S > 0 ? w : make_float2(w.x, -w.y);
}

// ------------------------------------------------------ register butterflies
// In-place DFT of R registers, X[k] = sum_r v[r] e^{S 2 pi i r k / R}.
template <int S> MC_DEV void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S> MC_DEV void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_si<S>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}
template <int S> MC_DEV void dft4(float2* v) { dft4<S>(v[0], v[1], v[2], v[3]); }

template <int S> MC_DEV void dft8(float2* v) {
  const float c = 0.70710678118654752440f;
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<S>(e0, e1, e2, e3);
  dft4<S>(o0, o1, o2, o3);
  // W8^1 = (c, S c), W8^2 = S i, W8^3 = (-c, S c)
  o1 = cmul(o1, make_float2(c, S * c));
  o2 = mul_si<S>(o2);
  o3 = cmul(o3, make_float2(-c, S * c));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

template <int S> MC_DEV void dft16(float2* v) {
  // 16 = 4 x 4: X[k1 + 4 k2] = sum_n2 W4^{n2 k2} W16^{n2 k1} sum_n1 v[n2 + 4 n1] W4^{n1 k1}
  const float c1 = 0.92387953251128675613f;  // cos(pi/8)
  const float s1 = 0.38268343236508977173f;  // sin(pi/8)
  const float c2 = 0.70710678118654752440f;
  float2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) {
    a[n2][0] = v[n2];
    a[n2][1] = v[n2 + 4];
    a[n2][2] = v[n2 + 8];
    a[n2][3] = v[n2 + 12];
    dft4<S>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
  }
  // twiddles W16^{n2 k1}, exponent e = n2*k1 in {0..9}
  // W16^1 = (c1, S s1), ^2 = (c2, S c2), ^3 = (s1, S c1), ^4 = S i,
  // ^6 = (-c2, S c2), ^9 = (-c1, -S s1)
  a[1][1] = cmul(a[1][1], make_float2(c1, S * s1));
  a[1][2] = cmul(a[1][2], make_float2(c2, S * c2));
  a[1][3] = cmul(a[1][3], make_float2(s1, S * c1));
  a[2][1] = cmul(a[2][1], make_float2(c2, S * c2));
  a[2][2] = mul_si<S>(a[2][2]);
  a[2][3] = cmul(a[2][3], make_float2(-c2, S * c2));
  a[3][1] = cmul(a[3][1], make_float2(s1, S * c1));
  a[3][2] = cmul(a[3][2], make_float2(-c2, S * c2));
  a[3][3] = cmul(a[3][3], make_float2(-c1, -S * s1));
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    float2 b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
    dft4<S>(b0, b1, b2, b3);
    v[k1] = b0;
    v[k1 + 4] = b1;
    v[k1 + 8] = b2;
    v[k1 + 12] = b3;
  }
}

// X[k] = sum_n v[n] W3^{S n k}: W3 = -1/2 + S i sqrt(3)/2.
template <int S> MC_DEV void dft3(float2* v) {
  const float h = 0.86602540378443864676f;  // sqrt(3)/2
  const float2 a = v[0], b = v[1], c = v[2];
  const float2 s = cadd(b, c), d = csub(b, c);
  const float2 m = make_float2(a.x - 0.5f * s.x, a.y - 0.5f * s.y);
  const float2 r = make_float2(-S * h * d.y, S * h * d.x);  // S i h (b - c)
  v[0] = cadd(a, s);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

// X[k] = sum_n v[n] W5^{S n k} (Winograd-style: 2 real rotations shared).
template <int S> MC_DEV void dft5(float2* v) {
  const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
  const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
  const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
  const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
  const float2 a = v[0];
  const float2 p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
  const float2 p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
  const float2 r1 = make_float2(a.x + c1 * p1.x + c2 * p2.x, a.y + c1 * p1.y + c2 * p2.y);
  const float2 r2 = make_float2(a.x + c2 * p1.x + c1 * p2.x, a.y + c2 * p1.y + c1 * p2.y);
  // S i (s1 m1 + s2 m2) and S i (s2 m1 - s1 m2)
  const float2 q1 = make_float2(s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y);
  const float2 q2 = make_float2(s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y);
  const float2 i1 = make_float2(-S * q1.y, S * q1.x), i2 = make_float2(-S * q2.y, S * q2.x);
  v[0] = cadd(a, cadd(p1, p2));
  v[1] = cadd(r1, i1);
  v[4] = csub(r1, i1);
  v[2] = cadd(r2, i2);
  v[3] = csub(r2, i2);
}

// X[k] = sum_n v[n] W7^{S n k}: the pairs v[m] +- v[7-m] (m = 1, 2, 3) with
// the cosines and sines of 2 pi m k / 7.
template <int S> MC_DEV void dft7(float2* v) {
  const float c1 = 0.62348980185873353053f;   // cos(2 pi/7)
  const float c2 = -0.22252093395631440429f;  // cos(4 pi/7)
  const float c3 = -0.90096886790241912624f;  // cos(6 pi/7)
  const float s1 = 0.78183148246802980871f;   // sin(2 pi/7)
  const float s2 = 0.97492791218182360702f;   // sin(4 pi/7)
  const float s3 = 0.43388373911755812048f;   // sin(6 pi/7)
  const float2 a = v[0];
  const float2 p1 = cadd(v[1], v[6]), m1 = csub(v[1], v[6]);
  const float2 p2 = cadd(v[2], v[5]), m2 = csub(v[2], v[5]);
  const float2 p3 = cadd(v[3], v[4]), m3 = csub(v[3], v[4]);
  const float2 r1 = make_float2(a.x + c1 * p1.x + c2 * p2.x + c3 * p3.x, a.y + c1 * p1.y + c2 * p2.y + c3 * p3.y);
  const float2 r2 = make_float2(a.x + c2 * p1.x + c3 * p2.x + c1 * p3.x, a.y + c2 * p1.y + c3 * p2.y + c1 * p3.y);
  const float2 r3 = make_float2(a.x + c3 * p1.x + c1 * p2.x + c2 * p3.x, a.y + c3 * p1.y + c1 * p2.y + c2 * p3.y);
  const float2 q1 = make_float2(s1 * m1.x + s2 * m2.x + s3 * m3.x, s1 * m1.y + s2 * m2.y + s3 * m3.y);
  const float2 q2 = make_float2(s2 * m1.x - s3 * m2.x - s1 * m3.x, s2 * m1.y - s3 * m2.y - s1 * m3.y);
  const float2 q3 = make_float2(s3 * m1.x - s1 * m2.x + s2 * m3.x, s3 * m1.y - s1 * m2.y + s2 * m3.y);
  // S i q
  const float2 i1 = make_float2(-S * q1.y, S * q1.x), i2 = make_float2(-S * q2.y, S * q2.x),
               i3 = make_float2(-S * q3.y, S * q3.x);
  v[0] = cadd(a, cadd(p1, cadd(p2, p3)));
  v[1] = cadd(r1, i1);
  v[6] = csub(r1, i1);
  v[2] = cadd(r2, i2);
  v[5] = csub(r2, i2);
  v[3] = cadd(r3, i3);
  v[4] = csub(r3, i3);
}

// X[k] = sum_n v[n] W11^{S n k}: the pairs v[m] +- v[11-m] (m = 1..5) with
// the cosines and sines of 2 pi j / 11, j = m k mod 11 (cos(2 pi (11-j)/11)
// = cos(2 pi j/11), sin(2 pi (11-j)/11) = -sin(2 pi j/11)).
template <int S> MC_DEV void dft11(float2* v) {
  const float C[6] = {1.f, 0.84125353283118116886f, 0.41541501300188642553f,
                      -0.14231483827328514044f, -0.65486073394528506406f,
                      -0.95949297361449738989f};
  const float SN[6] = {0.f, 0.54064081745559758211f, 0.90963199535451837141f,
                       0.98982144188093273238f, 0.75574957435425828377f,
                       0.28173255684142969771f};
  float2 p[6], m[6];
  const float2 a = v[0];
  float2 x0 = a;
#pragma unroll
  for (int j = 1; j <= 5; j++) {
    p[j] = cadd(v[j], v[11 - j]);
    m[j] = csub(v[j], v[11 - j]);
    x0 = cadd(x0, p[j]);
  }
#pragma unroll
  for (int k = 1; k <= 5; k++) {
    float2 r = a, q = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 1; j <= 5; j++) {
      const int e = (j * k) % 11;
      const float c = C[e <= 5 ? e : 11 - e];
      const float sn = e <= 5 ? SN[e] : -SN[11 - e];
      r = make_float2(r.x + c * p[j].x, r.y + c * p[j].y);
      q = make_float2(q.x + sn * m[j].x, q.y + sn * m[j].y);
    }
    const float2 iq = make_float2(-S * q.y, S * q.x);  // S i q
    v[k] = cadd(r, iq);
    v[11 - k] = csub(r, iq);
  }
  v[0] = x0;
}

template <int R, int S> MC_DEV void dft(float2* v) {
  if constexpr (R == 2) dft2<S>(v);
  else if constexpr (R == 3) dft3<S>(v);
  else if constexpr (R == 4) dft4<S>(v);
  else if constexpr (R == 5) dft5<S>(v);
  else if constexpr (R == 7) dft7<S>(v);
  else if constexpr (R == 11) dft11<S>(v);
  else if constexpr (R == 8) dft8<S>(v);
  else if constexpr (R == 16) dft16<S>(v);
  else static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 7 || R == 8 || R == 11 || R == 16, "radix");
}

// ------------------------------------------- direct DFT of an odd length
// cos and sin of 2 pi m/P evaluated at compile time: Taylor series in
// double on the angle reduced to [-pi, pi] (40 terms: the truncation error
// is far below one double ulp), rounded once to float.
__host__ __device__ constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 40; i++) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
__host__ __device__ constexpr double cx_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 40; i++) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
template <int P> struct OddTrig {
  float c[P], s[P];
  __host__ __device__ constexpr OddTrig() : c{}, s{} {
    for (int m = 0; m < P; m++) {
      const int mm = 2 * m < P ? m : m - P;
      const double t = 2.0 * 3.14159265358979323846 * mm / P;
      c[m] = (float)cx_cos(t);
      s[m] = (float)cx_sin(t);
    }
  }
};

// DFT of odd length P in registers, X_k = sum_n x_n e^{S 2 pi i n k/P}, by
// pairing the inputs n and P - n (s_n = x_n + x_{P-n}, d_n = x_n - x_{P-n})
// and the outputs k and P - k:
//   X_k, X_{P-k} = x_0 + sum_n s_n cos(2 pi n k/P) -/+ i S' sum_n d_n sin(...)
// that is (P - 1)^2 real FMAs instead of the 4 P^2 of the plain sum.  Every
// lane uses the same coefficient at the same time, so with the loops
// unrolled the coefficients are literal operands.  x is overwritten;
// out(k, X_k) is called once for every k.
template <int P, int S, class OUT>
MC_DEV void odd_dft(float2 (&x)[P], OUT&& out) {
  static_assert(P % 2 == 1 && P >= 3, "odd length");
  constexpr int H = (P - 1) / 2;
  constexpr OddTrig<P> T{};
  const float2 x0 = x[0];
  float2 sum = x0;
#pragma unroll
  for (int n = 1; n <= H; n++) {
    const float2 a = x[n], b = x[P - n];
    x[n] = cadd(a, b);
    x[P - n] = csub(a, b);
    sum = cadd(sum, x[n]);
  }
  out(0, sum);
#pragma unroll
  for (int k = 1; k <= H; k++) {
    float2 A = x0, B = make_float2(0.f, 0.f);
#pragma unroll
    for (int n = 1; n <= H; n++) {
      const float c = T.c[(n * k) % P], s = T.s[(n * k) % P];
      A.x = fmaf(x[n].x, c, A.x);
      A.y = fmaf(x[n].y, c, A.y);
      B.x = fmaf(x[P - n].x, s, B.x);
      B.y = fmaf(x[P - n].y, s, B.y);
    }
    // A -/+ i B: S = -1 gives X_k = A - i B, X_{P-k} = A + i B
    const float2 m = make_float2(A.x + B.y, A.y - B.x), p = make_float2(A.x - B.y, A.y + B.x);
    out(k, S < 0 ? m : p);
    out(P - k, S < 0 ? p : m);
    // one output pair at a time: interleaving them keeps 4 accumulators per
    // pair live and spills
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ---------------------------------------------------- Stockham stage plans
// Lengths N = 2^k * m, m a product of distinct factors 3, 5, 7 and 11 (k = 0
// included: no power-of-two stage).
// Power-of-two plans are balanced; the odd factors add a radix-3, radix-5,
// radix-7 and/or radix-11 stage after the plan of 2^k.
__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr bool is_pow2c(int n) { return n > 0 && (n & (n - 1)) == 0; }
__host__ __device__ constexpr int pow2_stage_count(int N) {
  return N <= 16 ? 1 : (N == 32 || N == 64 || N == 128 || N == 256) ? 2 : (N <= 4096 ? 3 : 4);
}
__host__ __device__ constexpr int pow2_stage_radix(int N, int s) {
  // 16:[16] 32:[8,4] 64:[8,8] 128:[16,8] 256:[16,16] 512:[8,8,8]
  // 1024:[16,8,8] 2048:[16,16,8] 4096:[16,16,16] 8192:[16,16,8,4]
  // 16384:[16,16,16,4]
  return N <= 16 ? N
       : N == 32 ? (s == 0 ? 8 : 4)
       : N == 64 ? 8
       : N == 128 ? (s == 0 ? 16 : 8)
       : N == 256 ? 16
       : N == 512 ? 8
       : N == 1024 ? (s == 0 ? 16 : 8)
       : N == 2048 ? (s < 2 ? 16 : 8)
       : N == 4096 ? 16
       : N == 8192 ? (s < 2 ? 16 : (s == 2 ? 8 : 4))
       : (s < 3 ? 16 : 4);
}
// Radix plans of the power-of-two part (the autotuner chooses among them):
//   plan 0: the balanced plans above;
//   plan 1: the same radices in the reverse order (e.g. 2048: [8,16,16]);
//   plan 2: a radix-8 chain, the remainder as a last radix-16 or radix-4
//           stage (4096: [8,8,8,8], 2048: [8,8,8,4], 1024: [8,8,16]); up to
//           8192.
__host__ __device__ constexpr int r8_stage_count(int N) {
  return N <= 16 ? 1 : (ilog2c(N) % 3 == 2 ? ilog2c(N) / 3 + 1 : ilog2c(N) / 3);
}
__host__ __device__ constexpr int r8_stage_radix(int N, int s) {
  return N <= 16 ? N
       : (ilog2c(N) % 3 == 1 && s == ilog2c(N) / 3 - 1) ? 16
       : (ilog2c(N) % 3 == 2 && s == ilog2c(N) / 3) ? 4
       : 8;
}
__host__ __device__ constexpr int plan_p2_count(int N, int P) {
  return N == 1 ? 0 : P == 2 ? r8_stage_count(N) : pow2_stage_count(N);
}
__host__ __device__ constexpr int plan_p2_radix(int N, int s, int P) {
  return P == 2 ? r8_stage_radix(N, s)
       : P == 1 ? pow2_stage_radix(N, pow2_stage_count(N) - 1 - s)
       : pow2_stage_radix(N, s);
}
__host__ __device__ constexpr int odd_part(int N) { return N % 2 ? N : odd_part(N / 2); }
// Stages of a length N = 2^k * m under plan P: the power-of-two stages, then
// one stage per odd factor of m, in the order 3, 5, 7, 11.
__host__ __device__ constexpr int odd_factor(int k) {
  return k == 0 ? 3 : (k == 1 ? 5 : (k == 2 ? 7 : 11));
}
__host__ __device__ constexpr int odd_radix_from(int m, int i, int k) {
  return k > 3 ? 11
       : m % odd_factor(k) != 0 ? odd_radix_from(m, i, k + 1)
       : i == 0 ? odd_factor(k)
                : odd_radix_from(m, i - 1, k + 1);
}
// the i-th of the factors 3, 5, 7, 11 that divide m
__host__ __device__ constexpr int odd_radix(int m, int i) { return odd_radix_from(m, i, 0); }
__host__ __device__ constexpr int stage_count(int N, int P = 0) {
  return plan_p2_count(N / odd_part(N), P) + (odd_part(N) % 3 == 0 ? 1 : 0) +
         (odd_part(N) % 5 == 0 ? 1 : 0) + (odd_part(N) % 7 == 0 ? 1 : 0) +
         (odd_part(N) % 11 == 0 ? 1 : 0);
}
__host__ __device__ constexpr int stage_radix(int N, int s, int P = 0) {
  return s < plan_p2_count(N / odd_part(N), P)
             ? plan_p2_radix(N / odd_part(N), s, P)
             : odd_radix(odd_part(N), s - plan_p2_count(N / odd_part(N), P));
}
static_assert(stage_radix(70, 0) == 2 && stage_radix(70, 1) == 5 && stage_radix(70, 2) == 7 &&
                  stage_radix(3840, 2) == 3 && stage_radix(3840, 3) == 5 &&
                  stage_count(132) == 3 && stage_radix(132, 0) == 4 &&
                  stage_radix(132, 1) == 3 && stage_radix(132, 2) == 11 &&
                  stage_count(35) == 2 && stage_radix(35, 0) == 5 && stage_radix(35, 1) == 7,
              "stage plans");
// Table index modulo the table size (any size; a power of two masks).
MC_DEV int tw_mod(int e, int ntw) { return (ntw & (ntw - 1)) == 0 ? (e & (ntw - 1)) : (e % ntw); }
__host__ __device__ constexpr int stage_ns(int N, int s, int P = 0) {
  return s == 0 ? 1 : stage_ns(N, s - 1, P) * stage_radix(N, s - 1, P);
}

// Twiddle sources: e^{+2 pi i e/ntw} for e in [0, ntw).
struct TwTable {  // full table (global or LDS)
  const float2* t;
  int ntw;
  MC_DEV float2 operator()(int e) const { return t[tw_mod(e, ntw)]; }
};
// Two-level table in LDS: w(e) = lo[e & (L-1)] * hi[e >> log2 L].  With
// correctly rounded entries the product is within ~1.5 ulp.
struct TwSplit {
  const float2* lo;
  const float2* hi;
  int lbits, ntw;
  MC_DEV float2 operator()(int e) const {
    e = tw_mod(e, ntw);
    return cmul(lo[e & ((1 << lbits) - 1)], hi[e >> lbits]);
  }
};
// Fills lo (2^lbits) and hi (ntw >> lbits) from a global full table.
MC_DEV void tw_split_load(float2* lo, float2* hi, const float2* __restrict__ g, int ntw,
                          int lbits, int tid, int nt) {
  const int L = 1 << lbits;
  for (int i = tid; i < L; i += nt) lo[i] = g[i];
  for (int i = tid; i < (ntw >> lbits); i += nt) hi[i] = g[i << lbits];
}
template <int S, class TW> MC_DEV float2 twv(const TW& tw, int e) {
  float2 w = tw(e);
  return S > 0 ? w : make_float2(w.x, -w.y);
}

// One Stockham butterfly of stage (R, Ns) for transform length N at index j
// (0 <= j < N/R): gather v[r] = x[j + r N/R], twiddle by W_{Ns R}^{(j%Ns) r},
// DFT_R.  tw: e^{+2 pi i e/ntw}, ntw a multiple of N.
template <int N, int R, int Ns, int S, class TW>
MC_DEV void stockham_bfly(float2* v, int j, const TW& tw, int ntw) {
  if constexpr (Ns > 1) {
    const int k = j % Ns;
    const int step = ntw / (Ns * R);
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], twv<S>(tw, k * r * step));
  }
  dft<R, S>(v);
}
template <int R, int Ns> MC_DEV int stockham_out(int j) {
  return (j / Ns) * Ns * R + (j % Ns);
}

// LDS index padding: one float2 of padding after every 2^PB elements, so
// that the stride-R writes of Stockham stage 0 spread over all banks.
template <int PB> MC_DEV int lix(int i) { return PB ? i + (i >> PB) : i; }
template <int PB> __host__ __device__ constexpr int lds_len(int n) {
  return PB ? n + ((n - 1) >> PB) + 1 : n;
}
// lix(i0 + r*STRIDE) for a compile-time r, given li0 = lix(i0).  When STRIDE
// is a multiple of 2^PB, adding r*STRIDE leaves the low PB bits of i0 alone,
// so the padded index is li0 plus a constant (an LDS instruction's immediate
// offset instead of a shift-and-add per element).  With STRIDE == 1 the same
// holds when i0 is a multiple of R and R divides 2^PB (r < R).
template <int PB, int STRIDE, int R = 0> MC_DEV int lix_at(int li0, int i0, int r) {
  if constexpr (PB == 0)
    return i0 + r * STRIDE;
  else if constexpr (STRIDE % (1 << PB) == 0)
    return li0 + r * (STRIDE + (STRIDE >> PB));
  else if constexpr (STRIDE == 1 && R > 0 && (1 << PB) % R == 0)
    return li0 + r;
  else
    return lix<PB>(i0 + r * STRIDE);
}

// Full in-LDS FFT of NB transforms of length N (buf[b*TS + lix(i)]), NT
// threads.  Data and result in natural order.  Caller synchronises before.
template <int N, int NT, int S, int NB, int STAGE, int TS, class TW, int PB = 0, int PLAN = 0>
MC_DEV void fft_lds_stage(float2* buf, const TW& tw, int ntw) {
  constexpr int R = stage_radix(N, STAGE, PLAN);
  constexpr int Ns = stage_ns(N, STAGE, PLAN);
  constexpr int NBF = NB * (N / R);
  constexpr int PER = (NBF + NT - 1) / NT;
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int q = tid + p * NT;
    if (NBF % NT == 0 || q < NBF) {
      const int b = q / (N / R), j = q % (N / R);
      float2* src = buf + b * TS;
      const int lj = lix<PB>(j);
#pragma unroll
      for (int r = 0; r < R; r++) v[p][r] = src[lix_at<PB, N / R>(lj, j, r)];
      stockham_bfly<N, R, Ns, S>(v[p], j, tw, ntw);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int q = tid + p * NT;
    if (NBF % NT == 0 || q < NBF) {
      const int b = q / (N / R), j = q % (N / R);
      const int o = stockham_out<R, Ns>(j);
      float2* dst = buf + b * TS;
      const int lo = lix<PB>(o);
#pragma unroll
      for (int r = 0; r < R; r++) dst[lix_at<PB, Ns, R>(lo, o, r)] = v[p][r];
    }
  }
  __syncthreads();
}

template <int N, int NT, int S, int NB, int TS, class TW, int PB = 0, int PLAN = 0>
MC_DEV void fft_lds_t(float2* buf, const TW& tw, int ntw) {
  static_assert(stage_count(N, PLAN) <= 4, "at most four stages");
  fft_lds_stage<N, NT, S, NB, 0, TS, TW, PB, PLAN>(buf, tw, ntw);
  if constexpr (stage_count(N, PLAN) > 1) fft_lds_stage<N, NT, S, NB, 1, TS, TW, PB, PLAN>(buf, tw, ntw);
  if constexpr (stage_count(N, PLAN) > 2) fft_lds_stage<N, NT, S, NB, 2, TS, TW, PB, PLAN>(buf, tw, ntw);
  if constexpr (stage_count(N, PLAN) > 3) fft_lds_stage<N, NT, S, NB, 3, TS, TW, PB, PLAN>(buf, tw, ntw);
}
template <int N, int NT, int S, int NB, int TS = N>
MC_DEV void fft_lds(float2* buf, const float2* tw, int ntw) {
  fft_lds_t<N, NT, S, NB, TS>(buf, TwTable{tw, ntw}, ntw);
}

// ------------------------------------- register-resident stage twiddles
// For kernels whose threads keep the same butterflies from item to item
// (persistent loops), the twiddles of stage STAGE are computed once per
// thread and held in registers: w[p][r] = W_{Ns R}^{(j%Ns) r} for the
// butterfly j of slot p.  Stage 0 (Ns = 1) has none.  A stage beyond the
// plan is an empty placeholder.
template <int N, int NT, int S, int NB, int STAGE, bool USED = (STAGE < stage_count(N))>
struct StageTw {
  static constexpr int R = 1, PER = 1;
  float2 w[1][1];
  template <class TW> MC_DEV void init(const TW&, int) {}
};
template <int N, int NT, int S, int NB, int STAGE>
struct StageTw<N, NT, S, NB, STAGE, true> {
  static constexpr int R = stage_radix(N, STAGE);
  static constexpr int Ns = stage_ns(N, STAGE);
  static constexpr int NBF = NB * (N / R);
  static constexpr int PER = (NBF + NT - 1) / NT;
  float2 w[PER][R];
  template <class TW> MC_DEV void init(const TW& tw, int ntw) {
#pragma unroll
    for (int p = 0; p < PER; p++) {
      const int j = (threadIdx.x + p * NT) % (N / R);
      const int k = j % Ns;
      const int step = ntw / (Ns * R);
      w[p][0] = make_float2(1.f, 0.f);
#pragma unroll
      for (int r = 1; r < R; r++) w[p][r] = twv<S>(tw, k * r * step);
    }
  }
};

template <int N, int NT, int S, int NB, int STAGE, int TS, int PB, class WT, bool CONJ = false>
MC_DEV void fft_lds_stage_rw(float2* buf, const WT& tw) {
  constexpr int R = stage_radix(N, STAGE);
  constexpr int Ns = stage_ns(N, STAGE);
  constexpr int NBF = NB * (N / R);
  constexpr int PER = (NBF + NT - 1) / NT;
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int q = tid + p * NT;
    if (NBF % NT == 0 || q < NBF) {
      const int b = q / (N / R), j = q % (N / R);
      float2* src = buf + b * TS;
      const int lj = lix<PB>(j);
#pragma unroll
      for (int r = 0; r < R; r++) v[p][r] = src[lix_at<PB, N / R>(lj, j, r)];
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++)
          v[p][r] = CONJ ? cmulc(v[p][r], tw.w[p][r]) : cmul(v[p][r], tw.w[p][r]);
      }
      dft<R, S>(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int q = tid + p * NT;
    if (NBF % NT == 0 || q < NBF) {
      const int b = q / (N / R), j = q % (N / R);
      const int o = stockham_out<R, Ns>(j);
      float2* dst = buf + b * TS;
      const int lo = lix<PB>(o);
#pragma unroll
      for (int r = 0; r < R; r++) dst[lix_at<PB, Ns, R>(lo, o, r)] = v[p][r];
    }
  }
  __syncthreads();
}

// All twiddles of an N-point in-LDS FFT (NB transforms, NT threads) held in
// registers.
template <int N, int NT, int S, int NB>
struct FftTw {
  StageTw<N, NT, S, NB, 0> s0;  // placeholder-sized: stage 0 has no twiddles
  StageTw<N, NT, S, NB, 1> s1;
  StageTw<N, NT, S, NB, 2> s2;
  StageTw<N, NT, S, NB, 3> s3;
  template <class TW> MC_DEV void init(const TW& tw, int ntw) {
    s1.init(tw, ntw);
    s2.init(tw, ntw);
    s3.init(tw, ntw);
  }
};

// Sign S with twiddles held for sign SW: when they differ, the held
// twiddles are conjugated on use (the inverse of the same plan from the
// forward twiddles).
template <int N, int NT, int S, int NB, int TS, int PB, int SW>
MC_DEV void fft_lds_rw(float2* buf, const FftTw<N, NT, SW, NB>& tw) {
  constexpr bool CJ = S != SW;
  using W1 = StageTw<N, NT, SW, NB, 1>;
  using W2 = StageTw<N, NT, SW, NB, 2>;
  using W3 = StageTw<N, NT, SW, NB, 3>;
  fft_lds_stage_rw<N, NT, S, NB, 0, TS, PB, StageTw<N, NT, SW, NB, 0>, CJ>(buf, tw.s0);
  if constexpr (stage_count(N) > 1) fft_lds_stage_rw<N, NT, S, NB, 1, TS, PB, W1, CJ>(buf, tw.s1);
  if constexpr (stage_count(N) > 2) fft_lds_stage_rw<N, NT, S, NB, 2, TS, PB, W2, CJ>(buf, tw.s2);
  if constexpr (stage_count(N) > 3) fft_lds_stage_rw<N, NT, S, NB, 3, TS, PB, W3, CJ>(buf, tw.s3);
}

// -------------------------------------------- float results rounded exactly
// Host code of the reference computes some float square roots and quotients
// whose rounding decides a discrete outcome (a shell index, a histogram
// bin).  These return the correctly rounded float result whatever the
// compiler's lowering of sqrtf and '/': the first guess is corrected against
// the midpoints to its neighbours, whose squares (products with b) are exact
// in FP64 (25-bit by 24-bit significands); such square roots and quotients
// never fall on a midpoint.  Operands are positive and finite.
MC_DEV float f32_next(float x, int d) { return __int_as_float(__float_as_int(x) + d); }

MC_DEV float sqrt_rn(float q) {  // q > 0
  float r = sqrtf(q);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double hi = 0.5 * ((double)r + (double)f32_next(r, 1));
    const double lo = 0.5 * ((double)r + (double)f32_next(r, -1));
    if ((double)q > hi * hi)
      r = f32_next(r, 1);
    else if ((double)q < lo * lo)
      r = f32_next(r, -1);
  }
  return r;
}

MC_DEV float div_rn(float a, float b) {  // a >= 0, b > 0
  if (a == 0.f) return 0.f;
  float d = (float)((double)a / (double)b);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double hi = 0.5 * ((double)d + (double)f32_next(d, 1));
    const double lo = 0.5 * ((double)d + (double)f32_next(d, -1));
    if ((double)a > hi * (double)b)
      d = f32_next(d, 1);
    else if ((double)a < lo * (double)b)
      d = f32_next(d, -1);
  }
  return d;
}

// --------------------------------------------------------- XCD remapping
// gfx950 dispatches consecutive workgroups round-robin over 8 XCDs, each
// with its own L2.  Remap so that a contiguous range of logical block ids
// runs on one XCD in order: neighbouring logical blocks then share L2.
MC_DEV int xcd_remap(int b, int nblocks) {
  const int nx = 8;
  const int per = nblocks / nx;  // blocks per XCD (full rounds)
  const int full = per * nx;
  if (b >= full) return b;       // tail: identity
  return (b % nx) * per + b / nx;
}

}  // namespace mc
