// mrcio.cpp -- MRC / gzip-MRC / TIFF stack I/O (see mrcio.h).
#include "mrcio.h"

#include <dlfcn.h>
#include <string.h>
#include <tiffio.h>
#include <zlib.h>

#include <mutex>
#include <vector>

// libtiff is loaded on first use (dlopen), so the driver does not depend on
// it unless a TIFF stack is read.
namespace {
struct TiffApi {
  TIFF* (*Open)(const char*, const char*) = nullptr;
  void (*Close)(TIFF*) = nullptr;
  int (*GetField)(TIFF*, uint32_t, ...) = nullptr;
  int (*ReadDirectory)(TIFF*) = nullptr;
  int (*SetDirectory)(TIFF*, uint16_t) = nullptr;
  tmsize_t (*StripSize)(TIFF*) = nullptr;
  uint32_t (*NumberOfStrips)(TIFF*) = nullptr;
  tmsize_t (*ReadEncodedStrip)(TIFF*, uint32_t, void*, tmsize_t) = nullptr;
  TIFFErrorHandler (*SetWarningHandler)(TIFFErrorHandler) = nullptr;
  bool ok = false;
};

const TiffApi& tiff_api() {
  static TiffApi api;
  static std::once_flag once;
  std::call_once(once, [] {
    const char* names[] = {"libtiff.so.6", "libtiff.so.5", "/opt/conda/lib/libtiff.so.5",
                           "libtiff.so"};
    void* h = nullptr;
    for (const char* n : names)
      if ((h = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) return;
    api.Open = (TIFF * (*)(const char*, const char*)) dlsym(h, "TIFFOpen");
    api.Close = (void (*)(TIFF*))dlsym(h, "TIFFClose");
    api.GetField = (int (*)(TIFF*, uint32_t, ...))dlsym(h, "TIFFGetField");
    api.ReadDirectory = (int (*)(TIFF*))dlsym(h, "TIFFReadDirectory");
    api.SetDirectory = (int (*)(TIFF*, uint16_t))dlsym(h, "TIFFSetDirectory");
    api.StripSize = (tmsize_t(*)(TIFF*))dlsym(h, "TIFFStripSize");
    api.NumberOfStrips = (uint32_t(*)(TIFF*))dlsym(h, "TIFFNumberOfStrips");
    api.ReadEncodedStrip = (tmsize_t(*)(TIFF*, uint32_t, void*, tmsize_t))dlsym(h, "TIFFReadEncodedStrip");
    api.SetWarningHandler = (TIFFErrorHandler(*)(TIFFErrorHandler))dlsym(h, "TIFFSetWarningHandler");
    api.ok = api.Open && api.Close && api.GetField && api.ReadDirectory && api.SetDirectory &&
             api.StripSize && api.NumberOfStrips && api.ReadEncodedStrip;
  });
  return api;
}
}  // namespace

// ------------------------------------------------------------------ MRC
MRC::MRC() { defaults(); }

MRC::MRC(const char* filename, const char* mode) {
  defaults();
  open(filename, mode);
}

MRC::~MRC() { close(); }

void MRC::defaults() {
  memset(&m_header, 0, sizeof m_header);
  m_header.mode = 2;
  m_header.cellb[0] = m_header.cellb[1] = m_header.cellb[2] = 90;
  m_header.mapc = 1;
  m_header.mapr = 2;
  m_header.maps = 3;
}

int MRC::open(const char* filename, const char* mode) {
  close();
  m_fp = fopen(filename, mode);
  if (!m_fp) return 0;
  if (mode[0] == 'w') return 1;  // header written by createMRC / updateHeader
  rewind(m_fp);
  if (fread(&m_header, 1, sizeof m_header, m_fp) < sizeof m_header) return -1;
  return 1;
}

void MRC::close() {
  if (m_fp) fclose(m_fp);
  m_fp = nullptr;
}

bool MRC::hasFile() const { return m_fp != nullptr; }

void MRC::getHeader(MRCHeader* h) const { memcpy(h, &m_header, sizeof m_header); }

int MRC::getWordLength() const {
  switch (m_header.mode) {
    case 0: return 1;
    case 1: return 2;
    case 2: return 4;
    case 3: return 4;
    case 4: return 8;
    case 6: return 2;
    default: return 0;
  }
}

size_t MRC::getImSize() const {
  return (size_t)m_header.nx * (size_t)m_header.ny * (size_t)getWordLength();
}

size_t MRC::read2DIm(void* buf, int n) {
  if (!m_fp || n < 0) return 0;
  const size_t im = getImSize();
  const long long off = 1024LL + m_header.nsymbt + (long long)n * (long long)im;
  if (fseeko(m_fp, off, SEEK_SET) != 0) return 0;
  return fread(buf, 1, im, m_fp);
}

size_t MRC::read2DIm_32bit(float* buf, int n) {
  const size_t np = (size_t)m_header.nx * m_header.ny;
  if (!np) return 0;
  const int mode = m_header.mode;
  if (mode == 0 || mode == 1 || mode == 6) {
    std::vector<unsigned char> raw(getImSize());
    const size_t r = read2DIm(raw.data(), n);
    if (r == 0) return 0;
    if (mode == 0) {
      const signed char* p = reinterpret_cast<const signed char*>(raw.data());
      for (size_t i = 0; i < np; i++) buf[i] = (float)p[i];
    } else if (mode == 1) {
      const int16_t* p = reinterpret_cast<const int16_t*>(raw.data());
      for (size_t i = 0; i < np; i++) buf[i] = (float)p[i];
    } else {
      const uint16_t* p = reinterpret_cast<const uint16_t*>(raw.data());
      for (size_t i = 0; i < np; i++) buf[i] = (float)p[i];
    }
    return r;
  }
  return read2DIm(buf, n);  // float32 (and TIFF, which decodes to float)
}

size_t MRC::write2DIm(const void* buf, int n) {
  if (!m_fp) return 0;
  const size_t im = getImSize();
  const long long off = 1024LL + m_header.nsymbt + (long long)n * (long long)im;
  if (fseeko(m_fp, off, SEEK_SET) != 0) return 0;
  return fwrite(buf, 1, im, m_fp);
}

void MRC::updateHeader() {
  if (!m_fp) return;
  rewind(m_fp);
  fwrite(&m_header, 1, sizeof m_header, m_fp);
}

template <class T> void MRC::stats(const T* data, size_t n) {
  T mn = data[0], mx = data[0];
  double mean = 0;
  for (size_t i = 0; i < n; i++) {
    if (mn > data[i]) mn = data[i];
    if (mx < data[i]) mx = data[i];
    mean += data[i];
  }
  m_header.dmin = (float)mn;
  m_header.dmax = (float)mx;
  m_header.dmean = (float)(mean / (double)n);
}

static void set_geometry(MRCHeader& h, int nx, int ny, int nz, int mode) {
  h.nx = nx;
  h.ny = ny;
  h.nz = nz;
  h.mode = mode;
  h.cellb[0] = h.cellb[1] = h.cellb[2] = 90;
  h.mapc = 1;
  h.mapr = 2;
  h.maps = 3;
}

size_t MRC::createMRC(const float* data, int nx, int ny, int nz) {
  set_geometry(m_header, nx, ny, nz, 2);
  const size_t n = (size_t)nx * ny * nz;
  stats(data, n);
  updateHeader();
  if (!m_fp || fseeko(m_fp, 1024LL + m_header.nsymbt, SEEK_SET) != 0) return 0;
  return fwrite(data, 1, n * sizeof(float), m_fp);
}

size_t MRC::createMRC(const short* data, int nx, int ny, int nz) {
  set_geometry(m_header, nx, ny, nz, 1);
  const size_t n = (size_t)nx * ny * nz;
  stats(data, n);
  updateHeader();
  if (!m_fp || fseeko(m_fp, 1024LL + m_header.nsymbt, SEEK_SET) != 0) return 0;
  return fwrite(data, 1, n * sizeof(short), m_fp);
}

// ----------------------------------------------------------------- MRCZ
MRCZ::MRCZ() {}
MRCZ::MRCZ(const char* filename, const char* mode) {
  if (open(filename, mode) <= 0) printf("Error opening file %s\n", filename);
}
MRCZ::~MRCZ() { close(); }

int MRCZ::open(const char* filename, const char* mode) {
  close();
  gzFile g = gzopen(filename, mode);
  if (!g) return 0;
  m_gz = g;
  if (mode[0] == 'w') return 1;
  gzrewind(g);
  if (gzread(g, &m_header, sizeof m_header) < (int)sizeof m_header) return -1;
  return 1;
}

void MRCZ::close() {
  if (m_gz) gzclose((gzFile)m_gz);
  m_gz = nullptr;
}

size_t MRCZ::read2DIm(void* buf, int n) {
  if (!m_gz || n < 0) return 0;
  const size_t im = getImSize();
  const z_off_t off = (z_off_t)(1024LL + m_header.nsymbt + (long long)n * (long long)im);
  if (gzseek((gzFile)m_gz, off, SEEK_SET) != off) return 0;
  const int r = gzread((gzFile)m_gz, buf, (unsigned)im);
  return r < 0 ? 0 : (size_t)r;
}

size_t MRCZ::write2DIm(const void* buf, int n) {
  if (!m_gz) return 0;
  const size_t im = getImSize();
  const z_off_t off = (z_off_t)(1024LL + m_header.nsymbt + (long long)n * (long long)im);
  if (gzseek((gzFile)m_gz, off, SEEK_SET) != off) return 0;
  const int r = gzwrite((gzFile)m_gz, buf, (unsigned)im);
  return r < 0 ? 0 : (size_t)r;
}

void MRCZ::updateHeader() {
  if (!m_gz) return;
  // gzip streams written sequentially: the header must be first
  gzwrite((gzFile)m_gz, &m_header, sizeof m_header);
}

// gzip output carries the "MAP " stamp: gzip stacks are recognised by it
// (func.cpp:597-599), not by their name.
size_t MRCZ::createMRC(const float* data, int nx, int ny, int nz) {
  set_geometry(m_header, nx, ny, nz, 2);
  memcpy(m_header.map, "MAP ", 4);
  const size_t n = (size_t)nx * ny * nz;
  stats(data, n);
  updateHeader();
  const int r = gzwrite((gzFile)m_gz, data, (unsigned)(n * sizeof(float)));
  return r < 0 ? 0 : (size_t)r;
}

size_t MRCZ::createMRC(const short* data, int nx, int ny, int nz) {
  set_geometry(m_header, nx, ny, nz, 1);
  memcpy(m_header.map, "MAP ", 4);
  const size_t n = (size_t)nx * ny * nz;
  stats(data, n);
  updateHeader();
  const int r = gzwrite((gzFile)m_gz, data, (unsigned)(n * sizeof(short)));
  return r < 0 ? 0 : (size_t)r;
}

// ----------------------------------------------------------------- MRCT
MRCT::MRCT() {}
MRCT::MRCT(const char* filename, const char* mode) { open(filename, mode); }
MRCT::~MRCT() { close(); }

int MRCT::open(const char* filename, const char* mode) {
  close();
  const TiffApi& T = tiff_api();
  if (!T.ok) {
    printf("TIFF support needs libtiff (not found)\n");
    return 0;
  }
  if (T.SetWarningHandler) T.SetWarningHandler(nullptr);
  TIFF* t = T.Open(filename, mode);
  if (!t) return 0;
  m_tif = t;
  defaults();
  m_header.mode = 7199;
  uint32_t w = 0, h = 0;
  uint16_t bits = 0;
  int count = 0;
  do {
    T.GetField(t, TIFFTAG_IMAGEWIDTH, &w);
    T.GetField(t, TIFFTAG_IMAGELENGTH, &h);
    count++;
  } while (T.ReadDirectory(t));
  T.SetDirectory(t, 0);
  T.GetField(t, TIFFTAG_BITSPERSAMPLE, &bits);
  m_header.nx = (int)w;
  m_header.ny = (int)h;
  m_header.nz = count;
  m_bits = bits;
  return 1;
}

void MRCT::close() {
  if (m_tif) tiff_api().Close((TIFF*)m_tif);
  m_tif = nullptr;
}

size_t MRCT::read2DIm(void* buf, int n) {
  TIFF* t = (TIFF*)m_tif;
  const TiffApi& T = tiff_api();
  if (!t || n < 0 || n >= m_header.nz) return 0;
  if (!T.SetDirectory(t, (uint16_t)n)) return 0;
  uint16_t bits = 0;
  T.GetField(t, TIFFTAG_BITSPERSAMPLE, &bits);
  if (bits != 8 && bits != 16) {
    printf("Only 8-bit or 16-bit TIFF is supported\n");
    return 0;
  }
  const tmsize_t ss = T.StripSize(t);
  const uint32_t ns = T.NumberOfStrips(t);
  std::vector<unsigned char> raw((size_t)ss * ns);
  size_t off = 0;
  for (uint32_t s = 0; s < ns; s++) {
    const tmsize_t r = T.ReadEncodedStrip(t, s, raw.data() + off, ss);
    if (r < 0) return 0;
    off += (size_t)r;
  }
  const size_t np = (size_t)m_header.nx * m_header.ny;
  float* f = (float*)buf;
  if (bits == 8)
    for (size_t i = 0; i < np; i++) f[i] = (float)raw[i];
  else
    for (size_t i = 0; i < np; i++) f[i] = (float)reinterpret_cast<const uint16_t*>(raw.data())[i];
  return np * sizeof(float);
}

size_t MRCT::write2DIm(const void*, int) {
  printf("Writing to TIFF not supported\n");
  return 0;
}
void MRCT::updateHeader() { printf("Writing to TIFF not supported\n"); }
size_t MRCT::createMRC(const float*, int, int, int) {
  printf("Writing to TIFF not supported\n");
  return 0;
}
size_t MRCT::createMRC(const short*, int, int, int) {
  printf("Writing to TIFF not supported\n");
  return 0;
}

// ------------------------------------------------------------- sniffing
static int first_two(const char* fn, int* a, int* b) {
  FILE* fp = fopen(fn, "rb");
  if (!fp) return 0;
  *a = fgetc(fp);
  *b = fgetc(fp);
  fclose(fp);
  return 1;
}

bool isMRCZ(const char* fn, int verb) {
  int a, b;
  if (!first_two(fn, &a, &b) || a != 0x1f || b != 0x8b) return false;
  MRCZ s;
  if (s.open(fn, "rb") > 0 && memcmp(s.m_header.map, "MAP ", 4) == 0) {
    if (verb) printf("Auto-detected file format: MRC-gz\n");
    return true;
  }
  return false;
}

bool isTIFF(const char* fn, int verb) {
  int a, b;
  if (!first_two(fn, &a, &b)) return false;
  if ((a == 'I' && b == 'I') || (a == 'M' && b == 'M')) {
    if (verb) printf("Auto-detected file format: TIFF\n");
    return true;
  }
  return false;
}

bool isMRC(const char* fn, int verb) {
  int a, b;
  if (!first_two(fn, &a, &b)) return false;
  if (a == 0x1f && b == 0x8b) return false;  // gzip: not a plain MRC
  MRC s;
  if (s.open(fn, "rb") <= 0) return false;
  if (memcmp(s.m_header.map, "MAP ", 4) == 0) {
    printf("Auto-detected file format: MRC2000\n");  // printed regardless of verb (func.cpp:566)
    return true;
  }
  if (strstr(fn, "mrc")) {
    if (verb) printf("Auto-detected file format: old MRC with mrc extension.\n");
    return true;
  }
  return false;
}

MRC* openStack(const char* fn, int verb) {
  MRC* s = nullptr;
  if (isTIFF(fn, verb)) s = new MRCT();
  else if (isMRCZ(fn, verb)) s = new MRCZ();
  else if (isMRC(fn, verb)) s = new MRC();
  else return nullptr;
  if (s->open(fn, "rb") <= 0) {
    delete s;
    return nullptr;
  }
  return s;
}

void buf2mrc(const char* filename, const float* buf, int nx, int ny, int nz) {
  MRC m;
  if (m.open(filename, "wb") > 0) m.createMRC(buf, nx, ny, nz);
  m.close();
}
