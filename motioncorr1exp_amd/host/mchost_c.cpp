// mchost_c.cpp -- C entry points over the host I/O classes, for tests and
// scripting (format sniffing, section reads, MRC writes).
#include <string.h>

#include "mrcio.h"

extern "C" {

// 0 none, 1 MRC, 2 gzip MRC, 3 TIFF (DFAlign.cu:470-481 order)
int mch_sniff(const char* fn) {
  if (isTIFF(fn)) return 3;
  if (isMRCZ(fn)) return 2;
  if (isMRC(fn)) return 1;
  return 0;
}

// geometry: dims[0..3] = nx, ny, nz, mode.  Returns 1 on success.
int mch_info(const char* fn, int* dims) {
  MRC* s = openStack(fn, 0);
  if (!s) return 0;
  dims[0] = s->getNx();
  dims[1] = s->getNy();
  dims[2] = s->getNz();
  dims[3] = s->getMode();
  delete s;
  return 1;
}

// Section n as float (read2DIm_32bit).  Returns bytes read.
long long mch_read_float(const char* fn, int n, float* out) {
  MRC* s = openStack(fn, 0);
  if (!s) return -1;
  long long r = (long long)s->read2DIm_32bit(out, n);
  delete s;
  return r;
}

// Writes a float volume (mode 2) with header statistics; gz != 0 writes gzip.
int mch_write_float(const char* fn, const float* data, int nx, int ny, int nz, int gz) {
  if (gz) {
    MRCZ m;
    if (m.open(fn, "wb") <= 0) return 0;
    return m.createMRC(data, nx, ny, nz) > 0;
  }
  MRC m;
  if (m.open(fn, "wb") <= 0) return 0;
  return m.createMRC(data, nx, ny, nz) > 0;
}

// Header statistics as written: min, max, mean.
int mch_header_stats(const char* fn, float* mmm) {
  MRC* s = openStack(fn, 0);
  if (!s) return 0;
  mmm[0] = s->getMin();
  mmm[1] = s->getMax();
  mmm[2] = s->getMean();
  delete s;
  return 1;
}
}
