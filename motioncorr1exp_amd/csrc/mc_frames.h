// mc_frames.h -- element decoding and dark/gain/stretch correction of the
// frame loads, shared by the stage-1 row passes (mc_fft.hip, mc_czt.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_device.h"
#include "mc_kernels.h"

namespace mc {

template <int DT> struct Elem;
template <> struct Elem<DT_INT8> { typedef int8_t T; };
template <> struct Elem<DT_UINT8> { typedef uint8_t T; };
template <> struct Elem<DT_INT16> { typedef int16_t T; };
template <> struct Elem<DT_UINT16> { typedef uint16_t T; };
template <> struct Elem<DT_FLOAT32> { typedef float T; };

// two-level twiddle split for a table of size ntw
__host__ __device__ constexpr int tw_lbits(int ntw) { return ilog2c(ntw) / 2; }

// Corrected value of frame pixel (X, Y): dark/gain at the source pixel of
// the nearest-neighbour stretch (DFAlign.cu:718-766).  The float operations
// are the reference's: (raw - dark) * gain, raw * gain or raw - dark; the
// source index is int(X * xs + 0.5) with the product in float and the
// rounding addition in double.
MC_DEV int2 stretch_src(const R2CGeom& g, int X, int Y) {
  if ((g.xs < 1.f || g.ys < 1.f) && X > 0 && Y > 0) {
    X = (int)((double)((float)X * g.xs) + 0.5);
    Y = (int)((double)((float)Y * g.ys) + 0.5);
  }
  return make_int2(X, Y);
}
MC_DEV float correct(const R2CGeom& g, float v, float d, float gn) {
  if (g.dark && g.gain) return (v - d) * gn;
  if (g.gain) return v * gn;
  if (g.dark) return v - d;
  return v;
}

}  // namespace mc
