// mc_pair.hip -- B-factor-weighted cross-power of frame pairs and the pruned
// inverse that produces only the box x box window of each correlation map.
//
// Reference: cuShiftCC (cufunc.cu:343-361, called with sx = sy = 0 at
// DFAlign.cu:894,1006) forms dfft <- dsum * conj(dfft) * z / sizec with
// z = (-1)^{x+y'} exp(-B (x^2+y'^2)/(2 n^2)) (MkPosList, cufunc.cu:309-341);
// a full n x n C2R follows (DFAlign.cu:897,1008) and only the central box is
// read back (FindShift, cufunc.cu:398-407).  Here the product, the weight
// and the inverse restricted to the box are fused:
//
//  k_pair_cols : one (pair, kx) per workgroup.  The product X[ky] is formed
//    on the fly from the two spectra columns (contiguous in the [kx][ky]
//    layout), decimated ky = a + A b (A = n/M), transformed by A inverse
//    M-point FFTs in LDS, and combined as
//      T[r] = sum_a e^{2 pi i a R_r/n} Y_a[R_r mod M],  R_r = (n-box)/2 + r,
//    which equals the column inverse at the box rows (M >= box).
//  k_pair_rows : one pair per workgroup: C[r][c] = sum_kx a_kx Re(T[kx][r]
//    e^{2 pi i kx R_c/n}) with a_0 = a_{n/2} = 1, a_kx = 2 (the C2R row
//    transform with the DC/Nyquist projection): a real GEMM with K =
//    2(n/2+1), on the matrix cores (k_pair_rows_mfma, box <= 96) or
//    register-tiled on the vector ALUs.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "mc_device.h"
#include "mc_kernels.h"

namespace mc {

// MFMA operand and accumulator vectors
typedef float v16f __attribute__((ext_vector_type(16)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

// Persistent: the grid is a multiple of 8 workgroups; workgroup b runs on
// XCD b % 8 and walks a contiguous range of work items (kx major, pair
// minor) owned by that XCD, so the frame columns of one kx are fetched from
// HBM once into that XCD's L2 and reused by all pairs.  Per-workgroup
// constants (M-point twiddles in LDS; ky weights, combine twiddles in
// registers) are loaded once.
//
// The per-item arithmetic of each column pass lives in a body struct
// (ColGen, ColM128) shared by the single-movie kernels and the batched plan
// kernel, so a pair's T column has the same FP32 operations in both.
//
// ColGen: generic decimated pass, A = N/M inverse M-point FFTs in LDS
// (fft_lds_t; M any Stockham length), combine by lane groups of SP lanes
// whose lane h takes the terms a = h, h + SP, ... < A, summed through LDS.
// F0 (for a small first radix R0, such as M = 70 = 2 5 7, or M = 128 under
// stage plan PL = 1, 8 16): the first Stockham stage runs in registers on
// the loaded products.  Its butterfly (a, j) takes the elements e + r N/R0
// (e = a + A j < N/R0, r < R0), and writes row a at j R0 + r, so the
// products skip one LDS round trip and two barriers.
template <int N, int M, int NT, int SP, bool F0 = false, int PL = 0>
struct ColGen {
  static constexpr int NN = N;
  static constexpr int A = N / M;
  static constexpr int H = N / 2 + 1;
  static constexpr int PB = 4;
  static constexpr int TS = lds_len<PB>(M) | 1;
  static constexpr int R0 = stage_radix(M, 0, PL);
  static constexpr int E0 = N / R0;              // first-stage butterflies (all rows)
  static constexpr int EP0 = (E0 + NT - 1) / NT;  // ... per thread
  static constexpr int EPT = F0 ? EP0 * R0 : (N + NT - 1) / NT;  // elements per thread
  static constexpr int AS = (A + SP - 1) / SP;   // combine terms per thread (at most)
  static constexpr int LDS = A * TS + M + SP * 256;  // float2 elements
  static_assert(A * M == N, "decimation");
  static_assert(!F0 || (M % R0 == 0 && stage_count(M, PL) >= 2), "first stage");
  static_assert(F0 || PL == 0, "plan");
  float2* buf;
  float2* twm;
  float2* part;  // [SP][256]
  float wy[EPT];
  float2 ctw[AS];
  int cr, ch, Rm, box, tid;
  bool cact;

  // Loads the workgroup constants; ends with a barrier.
  MC_DEV void init(float2* lds, const float* __restrict__ wxy, const float2* __restrict__ tw,
                   int box_, int tid_) {
    tid = tid_;
    box = box_;
    buf = lds;
    twm = lds + A * TS;
    part = twm + M;
    for (int i = tid; i < M; i += NT) twm[i] = tw[i * (N / M)];
    // ky weights (-1)^{y'} exp(m y'^2) / sizec for this thread's elements
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int e = F0 ? tid + (i / R0) * NT + (i % R0) * E0 : tid + i * NT;
      wy[i] = (F0 ? tid + (i / R0) * NT < E0 : e < N) ? wxy[H + e] : 0.f;
    }
    // combine: thread -> (output r, term block h)
    const int ori = (N - box) / 2;
    cr = tid % box;
    ch = tid / box;
    cact = tid < box * SP;
    const int R = ori + cr;
#pragma unroll
    for (int i = 0; i < AS; i++) {
      const int a = i * SP + ch;
      ctw[i] = cact && a < A ? tw[(a * R) % N] : make_float2(0.f, 0.f);
    }
    Rm = R % M;
    __syncthreads();
  }
  // Item (pair columns a, b at bin kx with column weight wx) -> tout[0, box).
  // buf and part are rewritten only after the next item's first barrier.
  MC_DEV void item(const float2* a, const float2* b, const float* __restrict__, float wx,
                   float2* tout) {
    const auto qa = buf_rsrc(a, N * 8), qb = buf_rsrc(b, N * 8);
    float2 va[EPT], vb[EPT];
    if constexpr (F0) {
#pragma unroll
      for (int p = 0; p < EP0; p++) {
        if (E0 % NT == 0 || tid + p * NT < E0) {
#pragma unroll
          for (int r = 0; r < R0; r++) {
            va[p * R0 + r] = buf_ld2(qa, 8 * tid, 8 * (p * NT + r * E0));
            vb[p * R0 + r] = buf_ld2(qb, 8 * tid, 8 * (p * NT + r * E0));
          }
        }
      }
      // products (column weight wx applied to the combined outputs), then
      // the first radix-R0 stage (Ns = 1: no twiddles) into the LDS rows
#pragma unroll
      for (int p = 0; p < EP0; p++) {
        const int e = tid + p * NT;
        if (E0 % NT == 0 || e < E0) {
          float2 v[R0];
#pragma unroll
          for (int r = 0; r < R0; r++) v[r] = cscale(cmulc(va[p * R0 + r], vb[p * R0 + r]), wy[p * R0 + r]);
          dft<R0, 1>(v);
          float2* row = buf + (e % A) * TS;
          const int j = e / A;
#pragma unroll
          for (int r = 0; r < R0; r++) row[lix<PB>(j * R0 + r)] = v[r];
        }
      }
      __syncthreads();
      const TwTable twt{twm, M};
      fft_lds_stage<M, NT, 1, A, 1, TS, TwTable, PB, PL>(buf, twt, M);
      if constexpr (stage_count(M, PL) > 2) fft_lds_stage<M, NT, 1, A, 2, TS, TwTable, PB, PL>(buf, twt, M);
      if constexpr (stage_count(M, PL) > 3) fft_lds_stage<M, NT, 1, A, 3, TS, TwTable, PB, PL>(buf, twt, M);
    } else {
#pragma unroll
      for (int i = 0; i < EPT; i++) {
        const int e = tid + i * NT;
        if (e < N) {
          va[i] = buf_ld2(qa, 8 * tid, 8 * NT * i);
          vb[i] = buf_ld2(qb, 8 * tid, 8 * NT * i);
        }
      }
      // the column weight wx is applied to the combined outputs
#pragma unroll
      for (int i = 0; i < EPT; i++) {
        const int e = tid + i * NT;
        if (e < N) buf[(e % A) * TS + lix<PB>(e / A)] = cscale(cmulc(va[i], vb[i]), wy[i]);
      }
      __syncthreads();
      fft_lds_t<M, NT, 1, A, TS, TwTable, PB>(buf, TwTable{twm, M}, M);
    }
    if (cact) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int i = 0; i < AS; i++)
        if (A % SP == 0 || i * SP + ch < A)
          acc = cadd(acc, cmul(ctw[i], buf[(i * SP + ch) * TS + lix<PB>(Rm)]));
      part[ch * 256 + cr] = acc;
    }
    __syncthreads();
    if (tid < box) {
      float2 acc = part[tid];
#pragma unroll
      for (int h = 1; h < SP; h++) acc = cadd(acc, part[h * 256 + tid]);
      tout[tid] = cscale(acc, wx);
    }
  }
};

template <int N, int M, int NT, int SP, bool F0 = false, int PL = 0>
__global__ __launch_bounds__(NT) void k_pair_cols(const float2* __restrict__ spectra,
                                                  size_t spec_stride,
                                                  const int2* __restrict__ pairs, int npairs,
                                                  const float* __restrict__ wxy, int box,
                                                  float2* __restrict__ tbuf,
                                                  const float2* __restrict__ tw) {
  typedef ColGen<N, M, NT, SP, F0, PL> B;
  __shared__ float2 lds[B::LDS];
  B b;
  b.init(lds, wxy, tw, box, threadIdx.x);
  const int xcd = blockIdx.x % 8, lid = blockIdx.x / 8, nper = gridDim.x / 8;
  const long long items = (long long)npairs * B::H;
  const long long i0 = items * xcd / 8, i1 = items * (xcd + 1) / 8;
  for (long long it = i0 + lid; it < i1; it += nper) {
    const int kx = (int)(it / npairs), p = (int)(it % npairs);
    const int2 pr = pairs[p];
    b.item(spectra + (size_t)pr.x * spec_stride + (size_t)kx * N,
           spectra + (size_t)pr.y * spec_stride + (size_t)kx * N, wxy, wxy[kx],
           tbuf + ((size_t)p * B::H + kx) * box);
  }
}

// Specialised column pass for M = 128 (box <= 128), n >= 512, NT = 256:
//  stage 0 (radix 16 over b): the 16 inputs X[a + A(j + 8r)] of butterfly
//    (a, j) are loaded straight from the two spectra (coalesced over a),
//    multiplied and weighted in registers, transformed and written to LDS;
//  stage 1 (radix 8, Ns = 16): reads and writes the same 8 LDS slots per
//    butterfly, so it runs in place without a barrier; its 7 twiddles depend
//    only on tid % 16 and stay in registers;
//  combine: SP adjacent lanes share one output row and reduce with xor
//    shuffles (no LDS round trip).
// WYG (for the 128-VGPR cap of 4 workgroups per CU): the stage-0 weights are
// re-read per item (L1-resident, coalesced) instead of held in 16 registers,
// and the stage-1 twiddles come from a 128-entry LDS table, which frees 16
// more VGPRs.  The values, and so the results, are the same either way.
template <int N, int SP, bool WYG>
struct ColM128 {
  static constexpr int NN = N;
  static constexpr int NT = 256;
  static constexpr int M = 128;
  static constexpr int A = N / M;
  static constexpr int H = N / 2 + 1;
  static constexpr int PB = 4;
  static constexpr int TS = lds_len<PB>(M) | 1;
  static constexpr int NB0 = A * 8;  // stage-0 butterflies
  static constexpr int P0 = (NB0 + NT - 1) / NT;
  static constexpr int NB1 = A * 16;  // stage-1 butterflies
  static constexpr int P1 = (NB1 + NT - 1) / NT;
  static constexpr int AS = A / SP;
  static constexpr int LDS = A * TS + (WYG ? M : 0);  // float2 elements
  static_assert(NT % 16 == 0 && A % SP == 0, "shape");
  float2* buf;
  float2* tw128;
  float wy[P0][16];
  float2 t1[8];
  float2 ctw[AS];
  int cr, ch, Rm, tid;
  bool cact;

  // Loads the workgroup constants (the LDS twiddle table is complete after
  // the first barrier of the first item).
  MC_DEV void init(float2* lds, const float* __restrict__ wxy, const float2* __restrict__ tw,
                   int box, int tid_) {
    tid = tid_;
    buf = lds;
    tw128 = lds + A * TS;
    // stage-0 element weights (fixed per thread)
#pragma unroll
    for (int p = 0; p < P0; p++) {
      const int q = tid + p * NT;
      const int a = q % A, j = q / A;
#pragma unroll
      for (int r = 0; r < 16; r++) wy[p][r] = (!WYG && q < NB0) ? wxy[H + a + A * (j + 8 * r)] : 0.f;
    }
    // stage-1 twiddles W_128^{k r}, k = tid % 16 (inverse sign)
    if constexpr (WYG) {
      for (int i = tid; i < M; i += NT) tw128[i] = tw[i * (N / M)];
    } else {
#pragma unroll
      for (int r = 0; r < 8; r++) t1[r] = tw[((tid % 16) * r * (N / M)) % N];
    }
    // combine: lane group (SP lanes) per output row
    const int ori = (N - box) / 2;
    cr = tid / SP;
    ch = tid % SP;
    cact = cr < box;
    const int R = ori + (cact ? cr : 0);
    Rm = R & (M - 1);
#pragma unroll
    for (int i = 0; i < AS; i++) ctw[i] = tw[((i * SP + ch) * R) % N];
  }
  // Raw stage-0 inputs of one item: the two spectrum columns at bin kx.
  // (As buffer loads, with the column base in scalar registers, the loop
  // issues 28 fewer VALU instructions per item at the same speed, but the
  // pass fetched 13% more bytes past L2, so these stay global loads.)
  MC_DEV void fetch(float2 (&ra)[P0][16], float2 (&rb)[P0][16], const float2* sa,
                    const float2* sb) const {
#pragma unroll
    for (int pp = 0; pp < P0; pp++) {
      const int q = tid + pp * NT;
      if (NB0 % NT == 0 || q < NB0) {
        const int a = q % A, j = q / A;
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int e = a + A * (j + 8 * r);
          ra[pp][r] = sa[e];
          rb[pp][r] = sb[e];
        }
      }
    }
  }
  // Stage 0 from the fetched registers into LDS (no barrier).  The column
  // weight is left to finish(): the transform is linear, so scaling its
  // box outputs by wx replaces 16 products per lane.
  MC_DEV void stage0(const float2 (&ra)[P0][16], const float2 (&rb)[P0][16],
                     const float* __restrict__ wxy) {
#pragma unroll
    for (int pp = 0; pp < P0; pp++) {
      const int q = tid + pp * NT;
      if (NB0 % NT == 0 || q < NB0) {
        const int a = q % A, j = q / A;
        float2 v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const float w = WYG ? wxy[H + a + A * (j + 8 * r)] : wy[pp][r];
          v[r] = cscale(cmulc(ra[pp][r], rb[pp][r]), w);
        }
        dft16<1>(v);
#pragma unroll
        for (int r = 0; r < 16; r++) buf[a * TS + lix_at<PB, 1, 16>(lix<PB>(16 * j), 16 * j, r)] = v[r];
      }
    }
  }
  // Stage 1 and the combine into tout[0, box), scaled by the column weight
  // wx; ends with a barrier (buf is rewritten by the next item's stage 0).
  MC_DEV void finish(float2* tout, float wx) {
    __syncthreads();
    // ---- stage 1 in place (radix 8, Ns = 16)
#pragma unroll
    for (int pp = 0; pp < P1; pp++) {
      const int q = tid + pp * NT;
      if (NB1 % NT == 0 || q < NB1) {
        const int a = q / 16, j = q % 16;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = buf[a * TS + lix_at<PB, 16>(lix<PB>(j), j, r)];
#pragma unroll
        for (int r = 1; r < 8; r++)
          v[r] = cmul(v[r], WYG ? tw128[((tid % 16) * r) % M] : t1[r]);
        dft8<1>(v);
#pragma unroll
        for (int r = 0; r < 8; r++) buf[a * TS + lix_at<PB, 16>(lix<PB>(j), j, r)] = v[r];
      }
    }
    __syncthreads();
    // ---- combine T[R] = sum_a e^{2 pi i a R/N} Y_a[R mod M]
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < AS; i++) acc = cadd(acc, cmul(ctw[i], buf[(i * SP + ch) * TS + lix<PB>(Rm)]));
#pragma unroll
    for (int o = 1; o < SP; o <<= 1) {
      acc.x += __shfl_xor(acc.x, o);
      acc.y += __shfl_xor(acc.y, o);
    }
    if (cact && ch == 0) tout[cr] = cscale(acc, wx);
    __syncthreads();
  }
  // One whole item (no prefetch).
  MC_DEV void item(const float2* a, const float2* b, const float* __restrict__ wxy, float wx,
                   float2* tout) {
    float2 ra[P0][16], rb[P0][16];
    fetch(ra, rb, a, b);
    stage0(ra, rb, wxy);
    finish(tout, wx);
  }
};

// A launch covers the bins kx in [kxb, kxb + nkx) of every pair.  With PF
// (software pipeline) the raw spectrum values of the next item are loaded
// into registers while the current item goes through stage 1 and the
// combine, so the global-load latency overlaps LDS work (at the cost of 64
// more VGPRs, i.e. one workgroup less per CU).
template <int N, int SP, bool PF, int MINW>
__global__ __launch_bounds__(256, (N <= 4096 ? MINW : 1)) void k_pair_cols_m128(const float2* __restrict__ spectra,
                                                        size_t spec_stride,
                                                        const int2* __restrict__ pairs,
                                                        int npairs,
                                                        const float* __restrict__ wxy, int box,
                                                        float2* __restrict__ tbuf,
                                                        const float2* __restrict__ tw,
                                                        int kxb, int nkx) {
  typedef ColM128<N, SP, (MINW >= 4)> B;
  __shared__ float2 lds[B::LDS];
  B b;
  b.init(lds, wxy, tw, box, threadIdx.x);
  const int xcd = blockIdx.x % 8, lid = blockIdx.x / 8, nper = gridDim.x / 8;
  const long long items = (long long)npairs * nkx;
  const long long i0 = items * xcd / 8, i1 = items * (xcd + 1) / 8;
  float2 ra[B::P0][16], rb[B::P0][16];
  auto fetch = [&](long long it) {
    const int kx = kxb + (int)(it / npairs), p = (int)(it % npairs);
    const int2 pr = pairs[p];
    b.fetch(ra, rb, spectra + (size_t)pr.x * spec_stride + (size_t)kx * N,
            spectra + (size_t)pr.y * spec_stride + (size_t)kx * N);
  };
  if (PF && i0 + lid < i1) fetch(i0 + lid);
  for (long long it = i0 + lid; it < i1; it += nper) {
    const int kx = kxb + (int)(it / npairs), p = (int)(it % npairs);
    if (!PF) fetch(it);
    b.stage0(ra, rb, wxy);
    if (PF && it + nper < i1) fetch(it + nper);
    b.finish(tbuf + ((size_t)p * B::H + kx) * box, wxy[kx]);
  }
}

// ------------------------------------------- batched (plan) column pass
// One persistent launch over the column-pass items of one movie or of many
// movies of different transform sizes (the default column pass of the M = 128
// decimation).  Movie m's items are the global indices [item0_m, item0_m +
// npairs_m (n_m/2+1)), kx major, pair minor; a plan table lists the movies in
// increasing item0 (one movie: the plan is passed by value).  A launch covers
// the global range [g0, g1), split over the XCDs as in the single-movie
// kernels; a workgroup runs consecutive items of one movie with that movie's
// body (ColM128 at its n) and sets the body up again when its next item
// belongs to the next movie.  Bodies are keyed n | SP << 14.  A movie alone
// and a movie in a batch run the same machine code for its items, so its T
// columns are bitwise the same (two builds of one body may contract
// multiply-adds differently, so this is a property of the code, not of the
// source).
// The items [it, stop) (step nper) of one movie.  The movie's pointers are
// parameters with __restrict__ so that, as in the single-movie kernels, its
// uniform reads (pair slots, column weight) are scalar loads.
template <class B>
MC_DEV long long plan_items(const float2* __restrict__ spectra, size_t spec_stride,
                            const int2* __restrict__ pairs, int np, const float* __restrict__ wxy,
                            const float2* __restrict__ tw, float2* __restrict__ tbuf, long long it,
                            long long li0, long long stop, int nper, int box, float2* lds,
                            int tid) {
  constexpr int N = B::NN;
  constexpr int H = N / 2 + 1;
  B b;
  __syncthreads();  // the previous body's LDS reads are done
  b.init(lds, wxy, tw, box, tid);
  // a movie has fewer than 2^31 items: kx and pair advance incrementally
  const int li = (int)li0;
  int kx = li / np, p = li - kx * np;
  const int dkx = nper / np, dp = nper - dkx * np;
  for (; it < stop; it += nper) {
    const int2 pr = pairs[p];
    b.item(spectra + (size_t)pr.x * spec_stride + (size_t)kx * N,
           spectra + (size_t)pr.y * spec_stride + (size_t)kx * N, wxy, wxy[kx],
           tbuf + ((size_t)p * H + kx) * box);
    kx += dkx;
    p += dp;
    if (p >= np) {
      p -= np;
      kx++;
    }
  }
  return it;
}

constexpr int plan_key(int n, int sp) { return n | (sp << 14); }
constexpr int kPlanLds = ColM128<3840, 2, true>::LDS;  // the largest body arena (float2)
// The LDS arena of the plan kernels, shared by all bodies.
__shared__ float2 plan_lds[kPlanLds];

// Runs the items [it, stop) of plan pl with body B.  The thread index comes
// in through an opaque move made here, once per plan: the values a body
// derives from it are then computed per plan rather than hoisted to the
// kernel entry, where those of every body would be live at once.  The box
// comes from the plan record for the same reason.
template <class B>
MC_DEV long long plan_run(const PairColsPlan& pl, long long it, long long stop, int nper) {
  int tid = threadIdx.x;
  asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"(tid));
  return plan_items<B>(pl.spectra, pl.spec_stride, pl.pairs, pl.npairs, pl.wxy, pl.tw, pl.tbuf, it,
                       it - pl.item0, stop, nper, pl.box, plan_lds, tid);
}

// Two register classes: ColM128 at n <= 1536 fits the 128-VGPR cap of 4
// workgroups per CU and runs in k_pair_cols_plan<4>; at 1920..3840 it needs
// 150-180 VGPRs and runs in k_pair_cols_plan<2>.  At n = 4096 the body needs
// 120-125 VGPRs alone and spills inside the plan kernel, so movies of that
// size take k_pair_cols_m128<4096, SP, false, 4> (single and batched alike).
constexpr bool plan_wide(int n) { return n >= 1920 && n < 4096; }

template <int MINW>
__global__ __launch_bounds__(256, MINW) void k_pair_cols_plan(PairColsPlan one,
                                                              const PairColsPlan* __restrict__ plans,
                                                              int nplans, long long g0,
                                                              long long g1) {
  const int xcd = blockIdx.x % 8, lid = blockIdx.x / 8, nper = gridDim.x / 8;
  const long long tot = g1 - g0;
  const long long x1 = g0 + tot * (xcd + 1) / 8;
  long long it = g0 + tot * xcd / 8 + lid;
  if (it >= x1) return;
  if (!plans) nplans = 1;
  auto plan = [&](int i) -> PairColsPlan { return plans ? plans[i] : one; };
  // the movie of the first item (binary search), then forward as `it` grows
  int lo = 0, hi = nplans - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (plan(mid).item0 <= it) lo = mid;
    else hi = mid - 1;
  }
  int pi = lo;
  while (it < x1) {
    while (pi + 1 < nplans && plan(pi + 1).item0 <= it) pi++;
    const PairColsPlan pl = plan(pi);
    const long long end = pi + 1 < nplans ? plan(pi + 1).item0 : x1;
    const long long stop = end < x1 ? end : x1;
    switch (pl.body) {
// A key of the other class is never in this class's table; it is skipped.
#define MC_PLAN_M(NN, SPV)                                                                   \
  case plan_key(NN, SPV):                                                             \
    if constexpr (plan_wide(NN) == (MINW == 2))                                              \
      it = plan_run<ColM128<NN, SPV, true>>(pl, it, stop, nper);                            \
    else                                                                                     \
      it += (stop - it + nper - 1) / nper * nper;                                            \
    break;
      MC_PLAN_M(512, 4) MC_PLAN_M(512, 2) MC_PLAN_M(768, 2) MC_PLAN_M(1024, 4) MC_PLAN_M(1024, 2)
      MC_PLAN_M(1280, 2) MC_PLAN_M(1536, 4) MC_PLAN_M(1536, 2) MC_PLAN_M(1920, 1)
      MC_PLAN_M(2048, 4) MC_PLAN_M(2048, 2) MC_PLAN_M(2560, 4) MC_PLAN_M(2560, 2)
      MC_PLAN_M(3072, 4) MC_PLAN_M(3072, 2) MC_PLAN_M(3840, 2)
#undef MC_PLAN_M
      default:  // not reached: the host admits only the keys above
        it += (stop - it + nper - 1) / nper * nper;
        break;
    }
    // `it` is now the workgroup's first item at or past `stop`
  }
}

// ----------------------------------------------- MFMA FP16 hi/lo form
// Same decimated pruned inverse as k_pair_cols_m128, with its first DFT
// stage (16 points, stride 8A in ky) run on the matrix cores.  With
// ky = c + 8A b1 (c in [0, 8A), b1 in [0, 16)) the stage is the complex
// product Z[k1][c] = sum_b1 e^{2 pi i b1 k1/16} X[c + 8A b1], i.e. the real
// product [Zr; Zi] = [[C, -S], [S, C]] [Xr; Xi] with K = 32, on 32-column
// blocks of c.  Operands are split into FP16 pairs x = hi + lo (hi =
// fp16(x), lo = fp16(x - hi)) and multiplied as hi*hi + hi*lo + lo*hi with
// FP32 accumulation (v_mfma_f32_32x32x8f16): relative error about 2^-21 per
// product.  Each column of X is scaled by a power of two that puts its
// largest component in [2^13, 2^14) before the split and unscaled after,
// so the FP16 range never clips.  Stage 2 (radix 8 with twiddles) and the
// combine run on the vector ALUs in FP32 as in the plain form.
// Lane layout (32x32x8 f16): lane l feeds A[l % 32][4 (l / 32) + t] and
// B[4 (l / 32) + t][l % 32], t = 0..3; accumulator v of lane l is
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
MC_DEV void split_hilo(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

template <int N, int SP>
__global__ __launch_bounds__(256) void k_pair_cols_mfma(const float2* __restrict__ spectra,
                                                        size_t spec_stride,
                                                        const int2* __restrict__ pairs,
                                                        int npairs,
                                                        const float* __restrict__ wxy, int box,
                                                        float2* __restrict__ tbuf,
                                                        float* __restrict__ tmax,
                                                        const float2* __restrict__ tw,
                                                        int kxb, int nkx) {
  constexpr int NT = 256;
  constexpr int M = 128;
  constexpr int A = N / M;
  constexpr int H = N / 2 + 1;
  constexpr int PB = 4;
  constexpr int TS = lds_len<PB>(M) | 1;
  constexpr int CB = 8 * A / 32;       // 32-column blocks of c (block cb -> wave cb % 4)
  constexpr int CBW = (CB + 3) / 4;    // blocks per wave (at most)
  constexpr int NB1 = A * 16;
  constexpr int P1 = (NB1 + NT - 1) / NT;
  constexpr int AS = A / SP;
  static_assert(CB >= 1 && (8 * A) % 32 == 0 && A % SP == 0, "shape");
  __shared__ float2 buf[A * TS];
  __shared__ float red[NT / 64];
  const int tid = threadIdx.x;
  const int wv = tid / 64, l = tid % 64, li = l % 32, lh = l / 32;
  // DFT-16 matrix fragments (rows (k1, re|im), columns (b1, re|im)), hi/lo
  v4h dh[4], dl[4];
  {
    const int i = li;
    const int k1 = i & 15;
    const bool im_row = i >= 16;
#pragma unroll
    for (int ks = 0; ks < 4; ks++)
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int kg = 8 * ks + 4 * lh + t;
        const int b1 = kg & 15;
        const bool im_col = kg >= 16;
        const float2 e = tw[((b1 * k1) & 15) * (N / 16)];  // (cos, sin)(2 pi b1 k1/16)
        float v;
        if (!im_row) v = im_col ? -e.y : e.x;
        else v = im_col ? e.x : e.y;
        _Float16 h, lo;
        split_hilo(v, h, lo);
        dh[ks][t] = h;
        dl[ks][t] = lo;
      }
  }
  // per-lane ky weights: column c = 32 (wv + 4 u) + li, rows b1 = 4 lh + t
  // and 8 + 4 lh + t
  float wy[CBW][8];
#pragma unroll
  for (int u = 0; u < CBW; u++) {
    const int cb = wv + 4 * u;
    const int c = 32 * (cb < CB ? cb : 0) + li;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int b1 = (t < 4 ? 0 : 8) + 4 * lh + (t & 3);
      wy[u][t] = wxy[H + c + 8 * A * b1];
    }
  }
  float2 t1[8];
#pragma unroll
  for (int r = 0; r < 8; r++) t1[r] = tw[((tid % 16) * r * (N / M)) % N];
  const int ori = (N - box) / 2;
  const int cr = tid / SP, ch = tid % SP;
  const bool cact = cr < box;
  const int R = ori + (cact ? cr : 0);
  const int Rm = R & (M - 1);
  float2 ctw[AS];
#pragma unroll
  for (int i = 0; i < AS; i++) ctw[i] = tw[((i * SP + ch) * R) % N];
  const int xcd = blockIdx.x % 8, lid = blockIdx.x / 8, nper = gridDim.x / 8;
  const long long items = (long long)npairs * nkx;
  const long long i0 = items * xcd / 8, i1 = items * (xcd + 1) / 8;
  for (long long it = i0 + lid; it < i1; it += nper) {
    const int kx = kxb + (int)(it / npairs), p = (int)(it % npairs);
    const int2 pr = pairs[p];
    const float2* sa = spectra + (size_t)pr.x * spec_stride + (size_t)kx * N;
    const float2* sb = spectra + (size_t)pr.y * spec_stride + (size_t)kx * N;
    const float wx = wxy[kx];
#pragma unroll
    for (int u = 0; u < CBW; u++) {
      const int cb = wv + 4 * u;
      if (cb >= CB) continue;  // uniform per wave
      const int c = 32 * cb + li;
      float2 x[8];
      float m = 0.f;
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const int b1 = (t < 4 ? 0 : 8) + 4 * lh + (t & 3);
        const int e = c + 8 * A * b1;
        x[t] = cscale(cmulc(sa[e], sb[e]), wx * wy[u][t]);
        m = fmaxf(m, fmaxf(fabsf(x[t].x), fabsf(x[t].y)));
      }
      m = fmaxf(m, __shfl_xor(m, 32));
      int ex = 0;
      (void)frexpf(m, &ex);
      const int sh = m > 0.f ? 14 - ex : 0;
      v4h bh[4], bl[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        _Float16 h, lo;
        split_hilo(ldexpf(x[t].x, sh), h, lo);      // ks 0: Re, b1 = 4 lh + t
        bh[0][t] = h; bl[0][t] = lo;
        split_hilo(ldexpf(x[t + 4].x, sh), h, lo);  // ks 1: Re, b1 = 8 + 4 lh + t
        bh[1][t] = h; bl[1][t] = lo;
        split_hilo(ldexpf(x[t].y, sh), h, lo);      // ks 2: Im
        bh[2][t] = h; bl[2][t] = lo;
        split_hilo(ldexpf(x[t + 4].y, sh), h, lo);  // ks 3: Im
        bh[3][t] = h; bl[3][t] = lo;
      }
      v16f acc;
#pragma unroll
      for (int v = 0; v < 16; v++) acc[v] = 0.f;
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(dh[ks], bl[ks], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(dl[ks], bh[ks], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(dh[ks], bh[ks], acc, 0, 0, 0);
      }
      // rows 8 (v/4) + 4 lh + v%4: Re of k1 for v < 8, Im of the same k1 at v + 8
      const int a = c % A, j = c / A;
#pragma unroll
      for (int v = 0; v < 8; v++) {
        const int k1 = 8 * (v / 4) + 4 * lh + (v % 4);
        buf[a * TS + lix<PB>(16 * j + k1)] = make_float2(ldexpf(acc[v], -sh), ldexpf(acc[v + 8], -sh));
      }
    }
    __syncthreads();
    // ---- stage 1 in place (radix 8, Ns = 16), FP32
#pragma unroll
    for (int pp = 0; pp < P1; pp++) {
      const int q = tid + pp * NT;
      if (NB1 % NT == 0 || q < NB1) {
        const int a = q / 16, j = q % 16;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = buf[a * TS + lix<PB>(j + 16 * r)];
#pragma unroll
        for (int r = 1; r < 8; r++) v[r] = cmul(v[r], t1[r]);
        dft8<1>(v);
#pragma unroll
        for (int r = 0; r < 8; r++) buf[a * TS + lix<PB>(j + 16 * r)] = v[r];
      }
    }
    __syncthreads();
    // ---- combine T[R] = sum_a e^{2 pi i a R/N} Y_a[R mod M]
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < AS; i++) acc = cadd(acc, cmul(ctw[i], buf[(i * SP + ch) * TS + lix<PB>(Rm)]));
#pragma unroll
    for (int o = 1; o < SP; o <<= 1) {
      acc.x += __shfl_xor(acc.x, o);
      acc.y += __shfl_xor(acc.y, o);
    }
    float mx = (cact && ch == 0) ? fmaxf(fabsf(acc.x), fabsf(acc.y)) : 0.f;
    if (cact && ch == 0) tbuf[((size_t)p * H + kx) * box + cr] = acc;
    // column maximum of T for the scale of the row pass
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (l == 0) red[wv] = mx;
    __syncthreads();  // also: buf is rewritten by the next item's stage 0
    if (tid == 0) tmax[(size_t)p * H + kx] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  }
}

template <int BT>
__global__ __launch_bounds__(256) void k_pair_rows(const float2* __restrict__ tbuf, int n,
                                                   int box, float* __restrict__ boxmaps,
                                                   const float2* __restrict__ tw) {
  constexpr int KC = 16;
  constexpr int W = 16 * BT;
  __shared__ float2 Ts[KC][W];
  __shared__ float2 Es[KC][W];
  const int p = blockIdx.x;
  const int H = n / 2 + 1;
  const int ori = (n - box) / 2;
  const float2* T = tbuf + (size_t)p * H * box;
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  float acc[BT][BT];
#pragma unroll
  for (int i = 0; i < BT; i++)
#pragma unroll
    for (int j = 0; j < BT; j++) acc[i][j] = 0.f;
  for (int k0 = 0; k0 < H; k0 += KC) {
    for (int i = threadIdx.x; i < KC * W; i += 256) {
      const int kk = i / W, c = i % W;
      const int kx = k0 + kk;
      float2 t = make_float2(0.f, 0.f), e = make_float2(0.f, 0.f);
      if (kx < H && c < box) {
        const float al = (kx == 0 || 2 * kx == n) ? 1.f : 2.f;
        t = cscale(T[(size_t)kx * box + c], al);
        e = tw[(int)(((long long)kx * (ori + c)) % n)];
      }
      Ts[kk][c] = t;
      Es[kk][c] = e;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KC; kk++) {
      float2 tv[BT], ev[BT];
#pragma unroll
      for (int i = 0; i < BT; i++) {
        tv[i] = Ts[kk][tr + 16 * i];
        ev[i] = Es[kk][tc + 16 * i];
      }
#pragma unroll
      for (int i = 0; i < BT; i++)
#pragma unroll
        for (int j = 0; j < BT; j++) acc[i][j] += tv[i].x * ev[j].x - tv[i].y * ev[j].y;
    }
    __syncthreads();
  }
  float* out = boxmaps + (size_t)p * box * box;
#pragma unroll
  for (int i = 0; i < BT; i++)
#pragma unroll
    for (int j = 0; j < BT; j++) {
      const int r = tr + 16 * i, c = tc + 16 * j;
      if (r < box && c < box) out[(size_t)r * box + c] = acc[i][j];
    }
}

// The same row pass on the matrix cores: per pair the box map is the real
// product C = A B with A[r][(kx, 0|1)] = a_kx (Re, Im) T[kx][r] and
// B[(kx, 0|1)][c] = (cos, -sin)(2 pi kx R_c/n), K = 2(n/2+1).  One
// v_mfma_f32_32x32x2f32 consumes one kx for a 32 x 32 block (FP32 operands
// and accumulation).  Chunks of KC bins of T and of the twiddles are staged
// in LDS (double-buffered: the next chunk is fetched into registers while
// the current one is multiplied); each of the four waves takes a quarter of
// every chunk for all NB x NB blocks, and the four partial maps are summed
// in LDS in wave order.  Lane layout (32x32x2 f32): lane l feeds
// A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator v of lane l is
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
template <int NB, int NT>
__global__ __launch_bounds__(NT) void k_pair_rows_mfma(const float2* __restrict__ tbuf, int n,
                                                        int box, float* __restrict__ boxmaps,
                                                        const float2* __restrict__ etab) {
  constexpr int W = 32 * NB;
  constexpr int KC = 32;
  constexpr int NW = NT / 64;  // waves; each takes KC/NW bins of every chunk
  static_assert(KC % NW == 0, "chunk split");
  constexpr int PT = (KC * 32 + NT - 1) / NT;  // T elements (this row block) per thread
  constexpr int PE = (KC * W + NT - 1) / NT;   // twiddles per thread
  __shared__ float2 Ts[2][KC][32];
  __shared__ float2 Es[2][KC][W];
  const int p = blockIdx.x;
  const int rb = blockIdx.y;  // 32-row block of the box map
  const int H = n / 2 + 1;
  const float2* T = tbuf + (size_t)p * H * box;
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64;
  const int li = l % 32, lh = l / 32;
  v16f acc[NB];
#pragma unroll
  for (int j = 0; j < NB; j++)
#pragma unroll
    for (int v = 0; v < 16; v++) acc[j][v] = 0.f;
  float2 pt[PT], pe[PE];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int i = threadIdx.x + NT * q;
      const int kk = i / 32, r = 32 * rb + i % 32, kx = k0 + kk;
      float2 t = make_float2(0.f, 0.f);
      if (i < KC * 32 && kx < H && r < box) {
        const float al = (kx == 0 || 2 * kx == n) ? 1.f : 2.f;
        t = cscale(T[(size_t)kx * box + r], al);
      }
      pt[q] = t;
    }
#pragma unroll
    for (int q = 0; q < PE; q++) {
      const int i = threadIdx.x + NT * q;
      const int kk = i / W, c = i % W, kx = k0 + kk;
      float2 e = make_float2(0.f, 0.f);
      if (i < KC * W && kx < H && c < box) e = etab[(size_t)kx * box + c];
      pe[q] = e;
    }
  };
  auto stash = [&](int b) {
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int i = threadIdx.x + NT * q;
      if (i < KC * 32) Ts[b][i / 32][i % 32] = pt[q];
    }
#pragma unroll
    for (int q = 0; q < PE; q++) {
      const int i = threadIdx.x + NT * q;
      if (i < KC * W) Es[b][i / W][i % W] = pe[q];
    }
  };
  const int nch = (H + KC - 1) / KC;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const int b = ch & 1;
    if (ch + 1 < nch) fetch((ch + 1) * KC);
#pragma unroll 2
    for (int kk = wv * (KC / NW); kk < (wv + 1) * (KC / NW); kk++) {
      const float2 tv = Ts[b][kk][li];
      const float a = lh ? tv.y : tv.x;
      float bb[NB];
#pragma unroll
      for (int j = 0; j < NB; j++) {
        const float2 ev = Es[b][kk][32 * j + li];
        bb[j] = lh ? -ev.y : ev.x;
      }
#pragma unroll
      for (int j = 0; j < NB; j++)
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bb[j], acc[j], 0, 0, 0);
    }
    if (ch + 1 < nch) stash(b ^ 1);
    __syncthreads();
  }
  // ordered sum of the NW partial row blocks in LDS (reuses the buffers)
  float* red = reinterpret_cast<float*>(&Es[0][0][0]);
  static_assert(sizeof(Es) >= sizeof(float) * 32 * W, "reduction buffer");
  for (int w = 0; w < NW; w++) {
    if (wv == w) {
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
          const int r = 8 * (v / 4) + 4 * lh + (v % 4), c = 32 * j + li;
          if (w == 0) red[r * W + c] = acc[j][v];
          else red[r * W + c] += acc[j][v];
        }
    }
    __syncthreads();
  }
  float* out = boxmaps + (size_t)p * box * box;
  const int rows = min(32, box - 32 * rb);
  for (int e = threadIdx.x; e < rows * box; e += NT)
    out[(size_t)(32 * rb + e / box) * box + e % box] = red[(e / box) * W + e % box];
}

// Streaming form of the same GEMM (box <= 96), with no LDS staging in the
// main loop.  The K range of a pair (bins kx in [0, n/2]) is split over ks
// workgroups and over the NW waves of each; a wave walks its contiguous kx
// range, each lane loading just the float it feeds (A[l % 32][l / 32] =
// (Re | Im) T[kx][r]; B[l / 32][l % 32] from a table that holds
// a_kx (cos, -sin)(2 pi kx R_c/n), zero past the box), D bins ahead in
// registers, and keeps all NB x NB accumulator blocks.  Rows of A past the
// box repeat the last box row; their outputs are never stored.  The NW
// partial maps are summed in LDS in wave order; with ks > 1 workgroup s
// writes partial map s of its pair and k_sum_parts adds them in order.
template <int NB, int NW, int D>
__global__ __launch_bounds__(NW * 64, (8 / NW > 1 ? 8 / NW : 1)) void k_pair_rows_stream(const float* __restrict__ tf, int n,
                                                              int box, int ks,
                                                              float* __restrict__ out,
                                                              const float* __restrict__ ef) {
  constexpr int W = 32 * NB;
  const int p = blockIdx.x, s = blockIdx.y;
  const int H = n / 2 + 1;
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64, li = l % 32, lh = l / 32;
  const int nsub = ks * NW, sub = s * NW + wv;
  const int k0 = (int)((long long)H * sub / nsub), k1 = (int)((long long)H * (sub + 1) / nsub);
  const float* T = tf + (size_t)p * H * box * 2;
  int roff[NB];
#pragma unroll
  for (int i = 0; i < NB; i++) roff[i] = 2 * min(32 * i + li, box - 1) + lh;
  const int eoff = 2 * li + lh;
  v16f acc[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; i++)
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[i][j][v] = 0.f;
  float a[D][NB], b[D][NB];
  // operands of bin min(kx, n/2) into slot d (bins past the wave's range
  // are loaded but never multiplied)
  auto load = [&](int d, int kx) {
    const int kc = min(kx, H - 1);
    const float* tr = T + (size_t)kc * box * 2;
    const float* er = ef + (size_t)kc * W * 2;
#pragma unroll
    for (int i = 0; i < NB; i++) a[d][i] = tr[roff[i]];
#pragma unroll
    for (int j = 0; j < NB; j++) b[d][j] = er[eoff + 64 * j];
  };
  auto mul = [&](int d) {
#pragma unroll
    for (int i = 0; i < NB; i++)
#pragma unroll
      for (int j = 0; j < NB; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[d][i], b[d][j], acc[i][j], 0, 0, 0);
  };
#pragma unroll
  for (int d = 0; d < D; d++) load(d, k0 + d);
  const int nfull = (k1 - k0) / D, rem = (k1 - k0) % D;
  for (int it = 0; it < nfull; it++) {
    const int kn = k0 + (it + 1) * D;
#pragma unroll
    for (int d = 0; d < D; d++) {
      mul(d);
      load(d, kn + d);
      // keeps each step's loads in that step, D bins ahead of their use: the
      // memory clobber stops them sinking into the next iteration, the
      // scheduling barrier stops the MFMAs of later steps moving above them
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int d = 0; d < D; d++)
    if (d < rem) mul(d);
  // ordered sum of the NW wave maps
  __shared__ float red[W * W];
  for (int w = 0; w < NW; w++) {
    if (wv == w) {
#pragma unroll
      for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
          for (int v = 0; v < 16; v++) {
            const int r = 32 * i + 8 * (v / 4) + 4 * lh + (v % 4), c = 32 * j + li;
            if (w == 0) red[r * W + c] = acc[i][j][v];
            else red[r * W + c] += acc[i][j][v];
          }
    }
    __syncthreads();
  }
  float* o = out + ((size_t)p * ks + s) * box * box;
  for (int e = threadIdx.x; e < box * box; e += NW * 64) o[e] = red[(e / box) * W + e % box];
}

// out[p][e] = sum_{s < ks} part[p][s][e] (e < m), added in order of s.
__global__ __launch_bounds__(256) void k_sum_parts(const float* __restrict__ part, int ks, int m,
                                                   long long total, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const long long p = i / m, e = i % m;
  const float* q = part + p * ks * m + e;
  float v = q[0];
  for (int s = 1; s < ks; s++) v += q[(size_t)s * m];
  out[i] = v;
}

// Row pass of the MFMA FP16 hi/lo form: the GEMM of k_pair_rows_mfma with
// v_mfma_f32_32x32x8f16 (four kx per instruction: K index 2 kk + (Re|Im)).
// T of a pair is scaled by the power of two that puts max|T| (the column
// maxima tmax written by k_pair_cols_mfma) in [2^13, 2^14); T and the
// twiddles are split into FP16 hi/lo pairs when staged in LDS and
// multiplied as hi*hi + hi*lo + lo*hi.
template <int NB>
__global__ __launch_bounds__(256) void k_pair_rows_mfma16(const float2* __restrict__ tbuf,
                                                          const float* __restrict__ tmax, int n,
                                                          int box, float* __restrict__ boxmaps,
                                                          const float2* __restrict__ etab) {
  constexpr int W = 32 * NB;
  constexpr int KC = 32;
  constexpr int PER = (KC * W + 255) / 256;
  __shared__ h2 S4[4][2][KC][W];  // T hi, T lo, E hi, E lo; double-buffered
  __shared__ float red[4];
  auto Th = S4[0];
  auto Tl = S4[1];
  auto Eh = S4[2];
  auto El = S4[3];
  const int p = blockIdx.x;
  const int H = n / 2 + 1;
  const float2* T = tbuf + (size_t)p * H * box;
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64;
  const int li = l % 32, lh = l / 32;
  // pair scale
  float m = 0.f;
  for (int k = threadIdx.x; k < H; k += 256) m = fmaxf(m, tmax[(size_t)p * H + k]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (l == 0) red[wv] = m;
  __syncthreads();
  m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  int ex = 0;
  (void)frexpf(m, &ex);
  const int sh = m > 0.f ? 14 - ex : 0;
  v16f acc[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; i++)
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[i][j][v] = 0.f;
  float2 pt[PER], pe[PER];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int i = threadIdx.x + 256 * q;
      const int kk = i / W, c = i % W, kx = k0 + kk;
      float2 t = make_float2(0.f, 0.f), e = make_float2(0.f, 0.f);
      if (i < KC * W && kx < H && c < box) {
        const float al = (kx == 0 || 2 * kx == n) ? 1.f : 2.f;
        t = T[(size_t)kx * box + c];
        t = make_float2(ldexpf(al * t.x, sh), ldexpf(al * t.y, sh));
        e = etab[(size_t)kx * box + c];
        e.y = -e.y;
      }
      pt[q] = t;
      pe[q] = e;
    }
  };
  auto stash = [&](int b) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int i = threadIdx.x + 256 * q;
      if (i < KC * W) {
        _Float16 a0, a1, b0, b1;
        split_hilo(pt[q].x, a0, a1);
        split_hilo(pt[q].y, b0, b1);
        Th[b][i / W][i % W] = h2{a0, b0};
        Tl[b][i / W][i % W] = h2{a1, b1};
        split_hilo(pe[q].x, a0, a1);
        split_hilo(pe[q].y, b0, b1);
        Eh[b][i / W][i % W] = h2{a0, b0};
        El[b][i / W][i % W] = h2{a1, b1};
      }
    }
  };
  const int nch = (H + KC - 1) / KC;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const int b = ch & 1;
    if (ch + 1 < nch) fetch((ch + 1) * KC);
#pragma unroll
    for (int ks = 0; ks < KC / 16; ks++) {
      const int kk = wv * (KC / 4) + 4 * ks + 2 * lh;  // this lane's two bins
      v4h ah[NB], al_[NB], bh[NB], bl[NB];
#pragma unroll
      for (int i = 0; i < NB; i++) {
        const h2 x0 = Th[b][kk][32 * i + li], x1 = Th[b][kk + 1][32 * i + li];
        const h2 y0 = Tl[b][kk][32 * i + li], y1 = Tl[b][kk + 1][32 * i + li];
        ah[i] = v4h{x0[0], x0[1], x1[0], x1[1]};
        al_[i] = v4h{y0[0], y0[1], y1[0], y1[1]};
        const h2 e0 = Eh[b][kk][32 * i + li], e1 = Eh[b][kk + 1][32 * i + li];
        const h2 f0 = El[b][kk][32 * i + li], f1 = El[b][kk + 1][32 * i + li];
        bh[i] = v4h{e0[0], e0[1], e1[0], e1[1]};
        bl[i] = v4h{f0[0], f0[1], f1[0], f1[1]};
      }
#pragma unroll
      for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x8f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x8f16(al_[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x8f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    if (ch + 1 < nch) stash(b ^ 1);
    __syncthreads();
  }
  float* redm = reinterpret_cast<float*>(&S4[0][0][0][0]);
  static_assert(sizeof(S4) >= sizeof(float) * W * W, "reduction buffer");
  for (int w = 0; w < 4; w++) {
    if (wv == w) {
#pragma unroll
      for (int i = 0; i < NB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
          for (int v = 0; v < 16; v++) {
            const int r = 32 * i + 8 * (v / 4) + 4 * lh + (v % 4), c = 32 * j + li;
            if (w == 0) redm[r * W + c] = acc[i][j][v];
            else redm[r * W + c] += acc[i][j][v];
          }
    }
    __syncthreads();
  }
  float* out = boxmaps + (size_t)p * box * box;
  for (int e = threadIdx.x; e < box * box; e += 256)
    out[e] = ldexpf(redm[(e / box) * W + e % box], -sh);
}

// Weight table: wxy[kx] = (-1)^kx exp(m kx^2), kx in [0, n/2]; wxy[n/2+1+ky]
// = (-1)^{y'} exp(m y'^2) / sizec, so z/sizec = wxy[kx] * wxy[n/2+1+ky] with
// m = -0.5 B/n^2 as in MkPosList (cufunc.cu:309-341).
static const float* weights(int n, float bfactor) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, float>, float*> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, n, bfactor);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int H = n / 2 + 1;
  const double m = -0.5 * (double)bfactor / n / n;
  const double sizec = (double)H * n;
  std::vector<float> h(H + n);
  for (int kx = 0; kx < H; kx++) h[kx] = (float)((kx & 1 ? -1.0 : 1.0) * std::exp(m * kx * kx));
  for (int ky = 0; ky < n; ky++) {
    const int yy = ky < n / 2 ? ky : ky - n;
    h[H + ky] = (float)(((yy & 1) ? -1.0 : 1.0) * std::exp(m * (double)yy * yy) / sizec);
  }
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * h.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  cache[key] = d;
  return d;
}

// Row-pass twiddles of a box: etab[kx * box + c] = e^{2 pi i kx R_c/n},
// R_c = (n - box)/2 + c, kx in [0, n/2] (double on the host, the same values
// as the n-point table), so the MFMA row kernels stream them like T.
static const float2* row_twiddles(int n, int box) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, float2*> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, n, box);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int H = n / 2 + 1, ori = (n - box) / 2;
  std::vector<float2> h((size_t)H * box);
  for (int kx = 0; kx < H; kx++)
    for (int c = 0; c < box; c++) {
      const long long e = ((long long)kx * (ori + c)) % n;
      const double a = 2.0 * M_PI * (double)e / (double)n;
      h[(size_t)kx * box + c] = make_float2((float)cos(a), (float)sin(a));
    }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * h.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  cache[key] = d;
  return d;
}

// B operand table of k_pair_rows_stream: per bin kx, 32 ceil(box/32)
// columns of (a_kx cos t, -a_kx sin t), t = 2 pi (kx R_c mod n)/n, with
// a_0 = a_{n/2} = 1, a_kx = 2 otherwise, and zeros past the box.
static const float* row_twiddles_stream(int n, int box) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, float*> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, n, box);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int H = n / 2 + 1, ori = (n - box) / 2, W = 32 * ((box + 31) / 32);
  std::vector<float> h((size_t)H * W * 2, 0.f);
  for (int kx = 0; kx < H; kx++) {
    const float al = (kx == 0 || 2 * kx == n) ? 1.f : 2.f;
    for (int c = 0; c < box; c++) {
      const long long e = ((long long)kx * (ori + c)) % n;
      const double a = 2.0 * M_PI * (double)e / (double)n;
      h[((size_t)kx * W + c) * 2] = al * (float)cos(a);
      h[((size_t)kx * W + c) * 2 + 1] = -al * (float)sin(a);
    }
  }
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * h.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  cache[key] = d;
  return d;
}

// Waves per workgroup of the streaming row pass; two workgroups share a CU.
constexpr int ROWS_NW = 4;

// K splits of the streaming row pass: ks = clamp((n/2+1)/512, 1, 8), so that
// a wave walks about 128 bins.  The split fixes the order in which the FP32
// partial products of a box map are added, so it depends on the transform
// size alone: a pair gives the same map whatever the number of pairs in the
// launch (one movie or a batch, a full pair set or an align-to-sum round)
// and whatever the device's CU count.
static int pair_rows_splits(int n) {
  return std::max(1, std::min(8, (n / 2 + 1) / 512));
}

size_t pair_rows_part_elems(int npairs, int n, int box) {
  const int ks = pair_rows_splits(n);
  return ks > 1 ? (size_t)npairs * ks * box * box : 0;
}


template <int N, int M, int SP, bool F0 = false, int PL = 0>
static hipError_t pair_cols_t(const float2* spectra, size_t spec_stride, const int2* pairs,
                              int npairs, const float* wxy, int box, float2* tbuf,
                              hipStream_t st) {
  constexpr int NT = 256;
  const float2* tw = twiddles(N);
  if (!tw) return hipErrorOutOfMemory;
  static const int grid = persistent_grid(k_pair_cols<N, M, NT, SP, F0, PL>, NT);
  hipLaunchKernelGGL((k_pair_cols<N, M, NT, SP, F0, PL>), dim3(grid), dim3(NT), 0, st, spectra,
                     spec_stride, pairs, npairs, wxy, box, tbuf, tw);
  return hipGetLastError();
}

// kx segments of a persistent column-pass launch: about 128 items per
// workgroup each, at most 8.  Within one launch the workgroups of an XCD
// drift apart as they go, and the frame columns of the kx values in flight
// outgrow its L2; a new launch lines them up again (3838x3710x40: 3.09 ms in
// one launch, 2.93 in four).
static int col_segments(int npairs, int H, int grid) {
  const long long per_wg = (long long)npairs * H / std::max(1, grid);
  return (int)std::max(1LL, std::min(8LL, (per_wg + 64) / 128));
}

template <int N, int SP>
static hipError_t pair_cols_m128(const float2* spectra, size_t spec_stride, const int2* pairs,
                                 int npairs, const float* wxy, int box, float2* tbuf, int mode,
                                 hipStream_t st) {
  const float2* tw = twiddles(N);
  if (!tw) return hipErrorOutOfMemory;
  static const int g1 = persistent_grid(k_pair_cols_m128<N, SP, true, 1>, 256);
  static const int g0 = persistent_grid(k_pair_cols_m128<N, SP, false, 1>, 256);
  static const int g4 = persistent_grid(k_pair_cols_m128<N, SP, false, 4>, 256);
  const int H = N / 2 + 1;
  const int S = col_segments(npairs, H, mode == 1 ? g1 : (mode == 2 ? g4 : g0));
  for (int sgi = 0; sgi < S; sgi++) {
    const int kxb = H * sgi / S, nkx = H * (sgi + 1) / S - kxb;
    if (mode == 1)
      hipLaunchKernelGGL((k_pair_cols_m128<N, SP, true, 1>), dim3(g1), dim3(256), 0, st, spectra,
                         spec_stride, pairs, npairs, wxy, box, tbuf, tw, kxb, nkx);
    else if (mode == 2)
      hipLaunchKernelGGL((k_pair_cols_m128<N, SP, false, 4>), dim3(g4), dim3(256), 0, st,
                         spectra, spec_stride, pairs, npairs, wxy, box, tbuf, tw, kxb, nkx);
    else
      hipLaunchKernelGGL((k_pair_cols_m128<N, SP, false, 1>), dim3(g0), dim3(256), 0, st,
                         spectra, spec_stride, pairs, npairs, wxy, box, tbuf, tw, kxb, nkx);
  }
  return hipGetLastError();
}

template <int N, int SP>
static hipError_t pair_cols_mfma(const float2* spectra, size_t spec_stride, const int2* pairs,
                                 int npairs, const float* wxy, int box, float2* tbuf,
                                 float* tmax, hipStream_t st) {
  const float2* tw = twiddles(N);
  if (!tw) return hipErrorOutOfMemory;
  static const int g = persistent_grid(k_pair_cols_mfma<N, SP>, 256);
  const int H = N / 2 + 1;
  const int S = col_segments(npairs, H, g);
  for (int sgi = 0; sgi < S; sgi++) {
    const int kxb = H * sgi / S, nkx = H * (sgi + 1) / S - kxb;
    hipLaunchKernelGGL((k_pair_cols_mfma<N, SP>), dim3(g), dim3(256), 0, st, spectra, spec_stride,
                       pairs, npairs, wxy, box, tbuf, tmax, tw, kxb, nkx);
  }
  return hipGetLastError();
}

template <int N>
static hipError_t pair_cols_n(const float2* spectra, size_t spec_stride, const int2* pairs,
                              int npairs, const float* wxy, int box, float2* tbuf, int variant,
                              hipStream_t st) {
  if constexpr (N >= 512) {
    if (variant != PAIR_COLS_GENERIC) {
      // the default at 4096 (no plan body there) is the 4-workgroup kernel
      const bool w4 = variant == PAIR_COLS_M128_W4 || (variant == PAIR_COLS_AUTO && N == 4096);
      const int mode = variant == PAIR_COLS_M128_PF ? 1 : (w4 ? 2 : 0);
      if constexpr ((N / 128) % 4 == 0) {
        if (box <= 64)
          return pair_cols_m128<N, 4>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, mode, st);
      }
      if constexpr ((N / 128) % 2 == 0) {
        if (box <= 128)
          return pair_cols_m128<N, 2>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, mode, st);
      } else {
        if (box <= 128)
          return pair_cols_m128<N, 1>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, mode, st);
      }
    }
  }
  // the generic pass at M = 128: first stage (radix 8 of plan 1, 8 x 16) in
  // registers on the loaded products
  if constexpr (N % 512 == 0) {
    if (box <= 64)
      return pair_cols_t<N, 128, 4, true, 1>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, st);
  }
  if constexpr (N % 256 == 0) {
    if (box <= 128)
      return pair_cols_t<N, 128, 2, true, 1>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, st);
  }
  if constexpr (N % 128 == 0) {
    if (box <= 128)
      return pair_cols_t<N, 128, 1, true, 1>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, st);
  }
  if constexpr (N % 256 == 0) {
    if (box <= 256)
      return pair_cols_t<N, 256, 1>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, st);
  }
  if constexpr (N == 64) {
    if (box <= 64)
      return pair_cols_t<64, 64, 1>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, st);
  }
  return hipErrorInvalidValue;
}

// Sizes of the decimated column pass (n = 128 A, and 64); other even sizes
// take the chirp-z column pass (mc_czt.hip).
static bool pair_cols_decimated(int n) {
  switch (n) {
    case 64: case 128: case 256: case 512: case 1024: case 2048: case 4096: case 8192:
    case 768: case 1536: case 3072: case 6144: case 1280: case 2560: case 5120: case 1920:
    case 3840:
      return true;
    default:
      return false;
  }
}

bool pair_form_supported(int precision, int n, int box) {
  if (precision == 0) return box >= 1 && box <= n && box <= 256 && n >= 64 && n <= 8192 && n % 2 == 0;
  // the hi/lo column kernel needs 32-column blocks of 8 n/128 columns
  if (precision == 1) return n >= 512 && n <= 8192 && n % 512 == 0 && box >= 1 && box <= 96;
  return false;
}

static hipError_t plan_launches(const PairColsPlan& one, const PairColsPlan* plans, int nplans,
                                long long items, int cls, hipStream_t st);

hipError_t launch_pair_cols(const float2* spectra, size_t spec_stride, const int2* pairs,
                            int npairs, int n, float bfactor, int box, float2* tbuf,
                            float* tmax, int precision, int variant, hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  if (!pair_form_supported(precision, n, box)) return hipErrorInvalidValue;
  const float* wxy = weights(n, bfactor);
  if (!wxy) return hipErrorOutOfMemory;
  if (precision == 1) {
    if (!tmax) return hipErrorInvalidValue;
    switch (n) {
#define MC_COLS_MFMA(NN) \
  case NN: return box <= 64 ? pair_cols_mfma<NN, 4>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, tmax, st) \
                            : pair_cols_mfma<NN, 2>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, tmax, st);
      MC_COLS_MFMA(512) MC_COLS_MFMA(1024) MC_COLS_MFMA(2048) MC_COLS_MFMA(4096) MC_COLS_MFMA(8192)
      MC_COLS_MFMA(1536) MC_COLS_MFMA(3072) MC_COLS_MFMA(6144) MC_COLS_MFMA(2560)
      MC_COLS_MFMA(5120)
#undef MC_COLS_MFMA
      default: return hipErrorInvalidValue;
    }
  }
  if (n == 3710 && box <= 128)
    // the default square crop of a 3838 x 3710 K2 frame: n = 70 * 53,
    // decimated into 53 inverse 70-point FFTs (radix 2, 5, 7), the radix-2
    // stage in registers on the loaded products
    return pair_cols_t<3710, 70, 2, true>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, st);
  if (n == 4092 && box <= 128)
    // the default square crop of a 5760 x 4092 K3 frame: n = 132 * 31,
    // decimated into 31 inverse 132-point FFTs (radix 4, 3, 11)
    return pair_cols_t<4092, 132, 2, true>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, st);
  if (!pair_cols_decimated(n))
    return launch_pair_cols_czt(spectra, spec_stride, pairs, npairs, wxy, n, box, tbuf, st);
  const int body = pair_cols_plan_body(n, box, variant);
  if (body >= 0) {  // the default: the plan kernel with this one movie
    PairColsPlan one{};
    one.spectra = spectra;
    one.pairs = pairs;
    one.wxy = wxy;
    one.tw = twiddles(n);
    if (!one.tw) return hipErrorOutOfMemory;
    one.tbuf = tbuf;
    one.spec_stride = spec_stride;
    one.item0 = 0;
    one.npairs = npairs;
    one.box = box;
    one.body = body;
    return plan_launches(one, nullptr, 1, (long long)npairs * (n / 2 + 1),
                         pair_cols_plan_class(body), st);
  }
  switch (n) {
    case 64: return pair_cols_n<64>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 128: return pair_cols_n<128>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 256: return pair_cols_n<256>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 512: return pair_cols_n<512>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 1024: return pair_cols_n<1024>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 2048: return pair_cols_n<2048>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 4096: return pair_cols_n<4096>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 8192: return pair_cols_n<8192>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 768: return pair_cols_n<768>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 1536: return pair_cols_n<1536>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 3072: return pair_cols_n<3072>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 6144: return pair_cols_n<6144>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 1280: return pair_cols_n<1280>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 2560: return pair_cols_n<2560>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 5120: return pair_cols_n<5120>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 1920: return pair_cols_n<1920>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    case 3840: return pair_cols_n<3840>(spectra, spec_stride, pairs, npairs, wxy, box, tbuf, variant, st);
    default: return hipErrorInvalidValue;
  }
}

int pair_cols_plan_body(int n, int box, int variant) {
  if (variant != PAIR_COLS_AUTO) return -1;
  if (box < 1 || box > 128 || n < 512 || n >= 4096 || n % 128 != 0 || !pair_cols_decimated(n))
    return -1;
  const int A = n / 128;
  return plan_key(n, (A % 4 == 0 && box <= 64) ? 4 : (A % 2 == 0 ? 2 : 1));
}

int pair_cols_plan_class(int body) { return plan_wide(body & 0x3fff) ? 1 : 0; }

const float* pair_weights(int n, float bfactor) { return weights(n, bfactor); }

// Launches of about 128 items per workgroup (see col_segments) over a plan
// table (plans != nullptr) or the one plan `one`.
static hipError_t plan_launches(const PairColsPlan& one, const PairColsPlan* plans, int nplans,
                                long long items, int cls, hipStream_t st) {
  if (items <= 0) return hipSuccess;
  static const int g4 = persistent_grid(k_pair_cols_plan<4>, 256);
  static const int g2 = persistent_grid(k_pair_cols_plan<2>, 256);
  const int grid = cls ? g2 : g4;
  const long long per_wg = items / grid;
  const long long S = std::max(1LL, (per_wg + 64) / 128);
  for (long long s = 0; s < S; s++) {
    const long long a = items * s / S, b = items * (s + 1) / S;
    if (cls)
      hipLaunchKernelGGL(k_pair_cols_plan<2>, dim3(grid), dim3(256), 0, st, one, plans, nplans, a, b);
    else
      hipLaunchKernelGGL(k_pair_cols_plan<4>, dim3(grid), dim3(256), 0, st, one, plans, nplans, a, b);
  }
  return hipGetLastError();
}

hipError_t launch_pair_cols_plans(const PairColsPlan* plans, int nplans, long long items, int cls,
                                  hipStream_t st) {
  if (nplans <= 0) return hipSuccess;
  PairColsPlan none{};
  return plan_launches(none, plans, nplans, items, cls, st);
}

hipError_t launch_pair_rows(const float2* tbuf, const float* tmax, int npairs, int n, int box,
                            float* boxmaps, float* part, int precision, int variant,
                            hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  if (!pair_form_supported(precision, n, box)) return hipErrorInvalidValue;
  const float2* tw = twiddles(n);
  if (!tw) return hipErrorOutOfMemory;
  if (precision == 0 && variant == PAIR_ROWS_AUTO && box <= 96) {
    const float* ef = row_twiddles_stream(n, box);
    if (!ef) return hipErrorOutOfMemory;
    const int ks = pair_rows_splits(n);
    if (ks > 1 && !part) return hipErrorInvalidValue;
    float* dst = ks > 1 ? part : boxmaps;
    const float* tf = reinterpret_cast<const float*>(tbuf);
    constexpr int NT = ROWS_NW * 64;
    switch ((box + 31) / 32) {
      case 1: hipLaunchKernelGGL((k_pair_rows_stream<1, ROWS_NW, 16>), dim3(npairs, ks), dim3(NT), 0, st, tf, n, box, ks, dst, ef); break;
      case 2: hipLaunchKernelGGL((k_pair_rows_stream<2, ROWS_NW, 12>), dim3(npairs, ks), dim3(NT), 0, st, tf, n, box, ks, dst, ef); break;
      default: hipLaunchKernelGGL((k_pair_rows_stream<3, ROWS_NW, 8>), dim3(npairs, ks), dim3(NT), 0, st, tf, n, box, ks, dst, ef); break;
    }
    if (ks > 1) {
      const long long total = (long long)npairs * box * box;
      hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, part,
                         ks, box * box, total, boxmaps);
    }
    return hipGetLastError();
  }
  const float2* etab = (precision == 1 || variant != PAIR_ROWS_VALU) ? row_twiddles(n, box) : nullptr;
  if ((precision == 1 || variant != PAIR_ROWS_VALU) && !etab) return hipErrorOutOfMemory;
  if (precision == 1) {
    if (!tmax) return hipErrorInvalidValue;
    switch ((box + 31) / 32) {
#define MC_ROWS_MFMA16(B) \
  case B: hipLaunchKernelGGL(k_pair_rows_mfma16<B>, dim3(npairs), dim3(256), 0, st, tbuf, tmax, n, box, boxmaps, etab); return hipGetLastError();
      MC_ROWS_MFMA16(1) MC_ROWS_MFMA16(2) MC_ROWS_MFMA16(3)
#undef MC_ROWS_MFMA16
      default: return hipErrorInvalidValue;
    }
  }
  if (variant != PAIR_ROWS_VALU && box <= 96) {
    // one workgroup per (pair, 32-row block of the box)
    const int nrb = (box + 31) / 32;
    switch (nrb) {
#define MC_ROWS_MFMA(B) \
  case B: \
    if (variant == PAIR_ROWS_MFMA8) \
      hipLaunchKernelGGL((k_pair_rows_mfma<B, 512>), dim3(npairs, nrb), dim3(512), 0, st, tbuf, n, box, boxmaps, etab); \
    else \
      hipLaunchKernelGGL((k_pair_rows_mfma<B, 256>), dim3(npairs, nrb), dim3(256), 0, st, tbuf, n, box, boxmaps, etab); \
    break;
      MC_ROWS_MFMA(1) MC_ROWS_MFMA(2) MC_ROWS_MFMA(3)
#undef MC_ROWS_MFMA
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  const int bt = (box + 15) / 16;
  switch (bt) {
#define MC_ROWS_CASE(B) \
  case B: hipLaunchKernelGGL(k_pair_rows<B>, dim3(npairs), dim3(256), 0, st, tbuf, n, box, boxmaps, tw); break;
    MC_ROWS_CASE(1) MC_ROWS_CASE(2) MC_ROWS_CASE(3) MC_ROWS_CASE(4) MC_ROWS_CASE(5)
    MC_ROWS_CASE(6) MC_ROWS_CASE(7) MC_ROWS_CASE(8) MC_ROWS_CASE(9) MC_ROWS_CASE(10)
    MC_ROWS_CASE(11) MC_ROWS_CASE(12) MC_ROWS_CASE(13) MC_ROWS_CASE(14) MC_ROWS_CASE(15)
    MC_ROWS_CASE(16)
#undef MC_ROWS_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mc
