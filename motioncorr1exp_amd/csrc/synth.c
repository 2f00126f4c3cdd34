/* synth.c -- Poisson sampling for the synthetic movie generator
 * (motioncorr1exp_amd/synth.py).  Counter-based: the draw for element i
 * depends only on (seed, i), so results do not depend on thread count. */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

static inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static inline double u01(uint64_t seed, uint64_t i, uint64_t lane) {
  uint64_t z = mix64(seed ^ mix64(i * 4 + lane));
  return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void mc_synth_poisson(const float *lam, float *out, size_t n, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) {
    double l = lam[i] > 0 ? lam[i] : 0.0;
    double u = u01(seed, i, 0);
    int k = 0;
    if (l < 64.0) {
      double p = exp(-l), c = p;
      while (u > c && k < 1000) {
        k++;
        p *= l / k;
        c += p;
      }
    } else {
      double u2 = u01(seed, i, 1);
      double z = sqrt(-2.0 * log(u)) * cos(2.0 * 3.14159265358979323846 * u2);
      double v = floor(l + sqrt(l) * z + 0.5);
      k = v < 0 ? 0 : (int)v;
    }
    out[i] = (float)k;
  }
}
