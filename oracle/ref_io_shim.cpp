// ref_io_shim.cpp -- C entry points over the reference's own stack readers
// (src-1125/mrc.cpp, mrcz.cpp, mrct.cpp, compiled from /root/reference by
// oracle/ref.mk).  Test infrastructure: used only to check the host I/O of
// the product against the reference code on the same files.
#include <string.h>

#include "mrc.h"
#include "mrct.h"
#include "mrcz.h"

extern "C" {

// kind: 1 MRC, 2 gzip MRC, 3 TIFF.  dims = nx, ny, nz, mode.
int ref_info(int kind, const char* fn, int* dims) {
  MRC* s = kind == 1 ? new MRC() : kind == 2 ? (MRC*)new MRCZ() : (MRC*)new MRCT();
  int ok = s->open(fn, "rb") > 0;
  if (ok) {
    dims[0] = s->getNx();
    dims[1] = s->getNy();
    dims[2] = s->getNz();
    dims[3] = s->getMode();
  }
  s->close();
  delete s;
  return ok;
}

// Reads sections 0..count-1 in order into out (float, nx*ny each).
int ref_read_float(int kind, const char* fn, int count, float* out) {
  MRC* s = kind == 1 ? new MRC() : kind == 2 ? (MRC*)new MRCZ() : (MRC*)new MRCT();
  if (s->open(fn, "rb") <= 0) {
    delete s;
    return 0;
  }
  const size_t np = (size_t)s->getNx() * s->getNy();
  int ok = 1;
  for (int n = 0; n < count; n++)
    if (s->read2DIm_32bit(out + (size_t)n * np, n) <= 0) ok = 0;
  s->close();
  delete s;
  return ok;
}

// MRC::createMRC of a float volume (header min/max/mean).
int ref_write_float(const char* fn, float* data, int nx, int ny, int nz) {
  MRC m;
  // open() in "wb" returns -1 (its header read fails on a write-only
  // stream); the reference callers ignore it, as does this shim.
  m.open(fn, "wb");
  if (!m.hasFile()) return 0;
  m.createMRC(data, nx, ny, nz);
  m.close();
  return 1;
}
}
