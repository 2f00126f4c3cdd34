// mrcio.h -- MRC, gzip-compressed MRC and TIFF stack I/O for the host driver.
//
// The class surface is the one the alignment driver of the reference uses
// (src-1125/mrc.h:96-147, mrcz.h:8-42, mrct.h:8-53): open/close, geometry
// getters, read2DIm / read2DIm_32bit, write2DIm, createMRC, updateHeader.
// The 1024-byte header layout is the MRC2000 format (src-1125/mrcheader.h).
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <string>

struct MRCHeader {
  int nx, ny, nz;   // columns, rows, sections
  int mode;         // 0 int8, 1 int16, 2 float32, 3/4 complex, 6 uint16
  int nxstart, nystart, nzstart;
  int mx, my, mz;
  float cella[3];
  float cellb[3];
  int mapc, mapr, maps;
  float dmin, dmax, dmean;
  int ispg;
  int nsymbt;       // bytes of extended header after the 1024-byte header
  char extra[100];
  float origin[3];
  char map[4];      // "MAP "
  int machst;
  float rms;
  int nlabels;
  char label[10][80];
};
static_assert(sizeof(MRCHeader) == 1024, "MRC header is 1024 bytes");

class MRC {
 public:
  MRC();
  MRC(const char* filename, const char* mode);
  virtual ~MRC();

  // 1 on success, 0 if the file cannot be opened, -1 if the header is short.
  virtual int open(const char* filename, const char* mode);
  virtual void close();

  void getHeader(MRCHeader* h) const;
  int getNx() const { return m_header.nx; }
  int getNy() const { return m_header.ny; }
  int getNz() const { return m_header.nz; }
  int getMode() const { return m_header.mode; }
  int getSymdatasize() const { return m_header.nsymbt; }
  float getMin() const { return m_header.dmin; }
  float getMax() const { return m_header.dmax; }
  float getMean() const { return m_header.dmean; }
  virtual int getWordLength() const;
  size_t getImSize() const;  // bytes of one section

  // Raw section n; returns bytes read.
  virtual size_t read2DIm(void* buf, int n);
  // Section n converted to float (mode 0 as signed bytes, mrc.cpp:338-348).
  size_t read2DIm_32bit(float* buf, int n);
  virtual size_t write2DIm(const void* buf, int n);
  virtual void updateHeader();
  // Writes a float32 (mode 2) or int16 (mode 1) volume with header statistics
  // (min, max, mean accumulated in double).
  virtual size_t createMRC(const float* data, int nx, int ny, int nz);
  virtual size_t createMRC(const short* data, int nx, int ny, int nz);
  bool hasFile() const;

  MRCHeader m_header;

 protected:
  void defaults();
  template <class T> void stats(const T* data, size_t n);
  FILE* m_fp = nullptr;
};

// gzip-compressed MRC through zlib (read and write).
class MRCZ : public MRC {
 public:
  MRCZ();
  MRCZ(const char* filename, const char* mode);
  ~MRCZ() override;
  int open(const char* filename, const char* mode) override;
  void close() override;
  size_t read2DIm(void* buf, int n) override;
  size_t write2DIm(const void* buf, int n) override;
  void updateHeader() override;
  size_t createMRC(const float* data, int nx, int ny, int nz) override;
  size_t createMRC(const short* data, int nx, int ny, int nz) override;

 private:
  void* m_gz = nullptr;
};

// Read-only multi-page TIFF (8- or 16-bit samples) presented as a stack; the
// header reports mode 7199 (mrct.cpp:80-95).  Sections are read in order.
class MRCT : public MRC {
 public:
  MRCT();
  MRCT(const char* filename, const char* mode);
  ~MRCT() override;
  int open(const char* filename, const char* mode) override;
  void close() override;
  int getWordLength() const override { return 4; }
  size_t read2DIm(void* buf, int n) override;  // fills floats
  size_t write2DIm(const void*, int) override;
  void updateHeader() override;
  size_t createMRC(const float*, int, int, int) override;
  size_t createMRC(const short*, int, int, int) override;
  int bitsPerSample() const { return m_bits; }

 private:
  void* m_tif = nullptr;
  int m_bits = 0;
};

// Format sniffing (func.cpp:557-625): gzip magic, "MAP " stamp or a name
// containing "mrc", TIFF byte-order mark.
bool isMRC(const char* fn, int verb = 0);
bool isMRCZ(const char* fn, int verb = 0);
bool isTIFF(const char* fn, int verb = 0);

// Opens the stack with the sniffing order of the alignment worker
// (DFAlign.cu:470-481): TIFF, then gzip MRC, then MRC.  nullptr if none.
MRC* openStack(const char* fn, int verb);

// Writes a float volume to a new MRC file (func.cpp:175-181).
void buf2mrc(const char* filename, const float* buf, int nx, int ny, int nz);
