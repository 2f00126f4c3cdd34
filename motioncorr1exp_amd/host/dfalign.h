// dfalign.h -- CDFAlign: the host interface of the alignment (reference
// src-1125/DFAlign.h:16-151).  A caller fills m_para, m_iter, m_iterthres and
// the file names, then calls RunAlign(), which blocks until the aligned sums
// and the log are written.  All device work goes through include/mc_abi.h.
#pragma once
#include <stdio.h>

#include <complex>
#include <string>
#include <vector>

#include "mrcio.h"

#define DISPDIM 400

// APARA (DFAlign.h:16-63), plus the geometry option of this build.
struct APARA {
  int crop_offsetx;
  int crop_offsety;
  int crop_nsam;
  float bin;
  float xscale, yscale;
  int nStart, nEnd;
  int nStartSum, nEndSum, nEndSum2;
  int GPUNum;
  float bfactor;
  int CCPeakSearchDim;
  int FrameDistOffset;
  int NoisePeakSize;
  float kiThresh;
  bool bSaveSubAreaCorrSum;
  bool bSaveRawSum;
  bool bSaveStackCorr;
  bool bSaveCCmap;
  bool bSaveLog;
  bool bAlignToMid;
  float fftscale;
  bool bDispSumCorr;
  bool bDispFFTRaw;
  bool bDispFFTCorr;
  bool bDispCCMap;
  bool bDispFSC;
  bool bLogFSC;
  float fscMax;
  // new in this build: 0 = reference square crop; N = pad to N x N
  int pad;
  // new in this build: pair-stage precision form (MC_PREC_*), adjoint
  // sharpness refinement steps (0 = off) and its first step in pixels
  int precision;
  int grad_steps;
  float grad_step_px;
};

class CDFAlign {
 public:
  CDFAlign();
  ~CDFAlign();

  bool m_bRun;
  int m_iter;
  float m_iterthres;

  char m_fnStack[512];
  char m_fnRawsum[512];
  char m_fnAlignsum[512];
  char m_fnAlignsum2[512];
  char m_fnStackCorr[512];
  char m_fnCCmap[512];
  char m_fnLog[512];
  char m_fnDark[512];
  char m_fnNorm[512];

  bool m_bSaveDisp;
  char m_dispRawFFT[512];
  char m_dispCorrSum[512];
  char m_dispCorrFFT[512];

  APARA m_para;
  std::string m_log;
  int m_nsam;
  int m_nsamRaw;

  // final per-frame shifts (binned pixels) after RunAlign
  std::vector<std::complex<double> > m_curve;
  // 0 on success, else an MC_ERR_* code or -100 for I/O errors
  int m_status;

  static void Message(const char* str);
  void TextOutput(const char* str);
  void RunAlign();
  int getNFrame();
  int getNFrame(const char* filename);
  // header of an MRC, gzip-MRC or TIFF stack (zeros if none opens),
  // DFAlign.h:150
  MRCHeader getMrcHeader(const char* filename);

 private:
  void AlignWorker();
};
