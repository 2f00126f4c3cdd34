// dosefgpu_driftcorr -- command-line drift correction of a dose-fractionated
// stack on one MI355X.  Options, defaults, clamps, default output names and
// the parameter block of the log follow the reference CLI
// (src-1125/dosefgpu_driftcorr.cpp:8-650); -pad, -prc, -grd and -grs are
// new.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <string>

#include "dfalign.h"

// dosefgpu_driftcorr.cpp:8-58
static APARA default_para() {
  APARA p;
  memset(&p, 0, sizeof p);
  p.bin = 1.0f;
  p.xscale = p.yscale = 1.0f;
  p.bfactor = 150;
  p.CCPeakSearchDim = 96;
  p.FrameDistOffset = 2;
  p.NoisePeakSize = 0;
  p.kiThresh = 1.0f;
  p.bSaveLog = true;
  p.bAlignToMid = true;
  p.fftscale = 0.0001f;
  p.bDispSumCorr = p.bDispFFTRaw = p.bDispFFTCorr = p.bDispCCMap = true;
  p.fscMax = 0.25f;
  p.pad = 0;
  p.precision = 0;
  p.grad_steps = 0;
  p.grad_step_px = 0.05f;
  return p;
}

// dosefgpu_driftcorr.cpp:60-118
static bool verify_para(CDFAlign& a) {
  const int nf = a.getNFrame();
  APARA& p = a.m_para;
  if (nf <= 1) {
    printf("Not a stack or wrong file. Abort\n");
    return false;
  }
  if (p.nStart < 0) p.nStart = 0;
  if (p.nEnd >= nf) p.nEnd = nf - 1;
  if (p.nEnd <= p.nStart) p.nEnd = nf - 1;
  if (p.nStartSum < p.nStart) p.nStartSum = p.nStart;
  if (p.nEndSum > p.nEnd) p.nEndSum = p.nEnd;
  if (p.nEndSum <= p.nStartSum) p.nEndSum = p.nEnd;
  if (p.nEndSum2 > p.nEnd) p.nEndSum2 = p.nEnd;
  if (p.nEndSum2 <= p.nStartSum) p.nEndSum2 = 0;
  if (p.FrameDistOffset >= nf / 2 || p.FrameDistOffset < 0) {
    printf("Frame distance offset is too big or less than 0. Abort\n");
    return false;
  }
  if (p.nEndSum2 >= p.nEndSum) {
    printf("Last frame # for the second sum image cannot be larger than or equal to that for the first sum image.\n");
    return false;
  }
  if (p.CCPeakSearchDim < 16) {
    printf("Peak search box (-pbx) must be at least 16. Abort\n");
    return false;
  }
  if (p.precision != 0 && p.precision != 1) {
    printf("Precision form (-prc) must be 0 or 1. Abort\n");
    return false;
  }
  if (p.grad_steps < 0 || p.grad_step_px < 0) {
    printf("Refinement steps (-grd) and step (-grs) must not be negative. Abort\n");
    return false;
  }
  return true;
}

static bool opt_is(const char* a, const char* b) { return strcmp(a, b) == 0; }

// dosefgpu_driftcorr.cpp:121-515
static bool get_para(int narg, char* argv[], CDFAlign& a) {
  APARA& p = a.m_para;
  p = default_para();
  a.m_fnNorm[0] = a.m_fnDark[0] = 0;
  a.m_bSaveDisp = true;
  snprintf(a.m_fnStack, sizeof a.m_fnStack, "%s", argv[1]);
  if (narg % 2 == 1) {
    printf("Wrong number of arguments. Abort.\n");
    return false;
  }
  bool dRaw = true, dSum = true, dSum2 = true, dStack = true, dCC = true, dLog = true;
  for (int i = 2; i < narg; i++) {
    const char* o = argv[i];
    const char* v = argv[i + 1];
    if (opt_is(o, "-crx")) p.crop_offsetx = atoi(v);
    else if (opt_is(o, "-cry")) p.crop_offsety = atoi(v);
    else if (opt_is(o, "-crd")) p.crop_nsam = atoi(v);
    else if (opt_is(o, "-bin")) {
      p.bin = (float)atof(v);
      if (p.bin > 2.0001) {
        printf("Larger than maximum binning 2. Set to 2.\n");
        p.bin = 2.0f;
      } else if (p.bin < 0.9999) {
        printf("Smaller than minimum binning 1. Set to 1.\n");
        p.bin = 1.0f;
      }
    } else if (opt_is(o, "-scx")) {
      p.xscale = (float)atof(v);
      if (p.xscale < 1.0) {
        printf("Stretching factor along X-axis needs to be >=1.0.\n");
        return false;
      }
    } else if (opt_is(o, "-scy")) {
      p.yscale = (float)atof(v);
      if (p.yscale < 1.0) {
        printf("Stretching factor along Y-axis needs to be >=1.0.\n");
        return false;
      }
    } else if (opt_is(o, "-nst")) p.nStart = atoi(v);
    else if (opt_is(o, "-ned")) p.nEnd = atoi(v);
    else if (opt_is(o, "-nss")) p.nStartSum = atoi(v);
    else if (opt_is(o, "-nes")) p.nEndSum = atoi(v);
    else if (opt_is(o, "-2es")) p.nEndSum2 = atoi(v);
    else if (opt_is(o, "-gpu")) p.GPUNum = atoi(v);
    else if (opt_is(o, "-bft")) p.bfactor = (float)atof(v);
    else if (opt_is(o, "-pbx")) p.CCPeakSearchDim = atoi(v);
    else if (opt_is(o, "-fod")) p.FrameDistOffset = atoi(v);
    else if (opt_is(o, "-nps")) p.NoisePeakSize = atoi(v);
    else if (opt_is(o, "-kit")) p.kiThresh = (float)atof(v);
    else if (opt_is(o, "-sub")) p.bSaveSubAreaCorrSum = atoi(v);
    else if (opt_is(o, "-srs")) p.bSaveRawSum = atoi(v);
    else if (opt_is(o, "-ssc")) p.bSaveStackCorr = atoi(v);
    else if (opt_is(o, "-scc")) p.bSaveCCmap = atoi(v);
    else if (opt_is(o, "-slg")) p.bSaveLog = atoi(v);
    else if (opt_is(o, "-atm")) p.bAlignToMid = atoi(v);
    else if (opt_is(o, "-dsp")) {
      if (atoi(v) == 0) {
        p.bDispSumCorr = p.bDispFFTRaw = p.bDispFFTCorr = p.bDispCCMap = false;
        a.m_bSaveDisp = false;
      }
    } else if (opt_is(o, "-fsc")) p.bLogFSC = atoi(v);
    else if (opt_is(o, "-frs")) { snprintf(a.m_fnRawsum, 512, "%s", v); dRaw = false; }
    else if (opt_is(o, "-fcs")) { snprintf(a.m_fnAlignsum, 512, "%s", v); dSum = false; }
    else if (opt_is(o, "-2cs")) { snprintf(a.m_fnAlignsum2, 512, "%s", v); dSum2 = false; }
    else if (opt_is(o, "-fct")) { snprintf(a.m_fnStackCorr, 512, "%s", v); dStack = false; }
    else if (opt_is(o, "-fcm")) { snprintf(a.m_fnCCmap, 512, "%s", v); dCC = false; }
    else if (opt_is(o, "-flg")) { snprintf(a.m_fnLog, 512, "%s", v); dLog = false; }
    else if (opt_is(o, "-fnm")) snprintf(a.m_fnNorm, 512, "%s", v);
    else if (opt_is(o, "-fdk")) snprintf(a.m_fnDark, 512, "%s", v);
    else if (opt_is(o, "-itr")) a.m_iter = atoi(v);
    else if (opt_is(o, "-ith")) a.m_iterthres = (float)atof(v);
    else if (opt_is(o, "-pad")) p.pad = atoi(v);
    else if (opt_is(o, "-prc")) p.precision = atoi(v);
    else if (opt_is(o, "-grd")) p.grad_steps = atoi(v);
    else if (opt_is(o, "-grs")) p.grad_step_px = (float)atof(v);
    else {
      printf("Undefined option: %s .Abort.\n", o);
      return false;
    }
    i++;
  }
  // default names in the working directory (dosefgpu_driftcorr.cpp:433-481)
  std::string fnp = argv[1];
  if (fnp.rfind('.') != std::string::npos) fnp = fnp.substr(0, fnp.rfind('.'));
  if (p.bin > 1.0001) fnp += "_binned";
  const std::string path = ".";
  const std::string prefix = fnp.rfind('/') == std::string::npos ? fnp : fnp.substr(fnp.rfind('/') + 1);
  if (dRaw) snprintf(a.m_fnRawsum, 512, "%s_SumRaw.mrc", prefix.c_str());
  if (dSum) snprintf(a.m_fnAlignsum, 512, "%s_SumCorr.mrc", prefix.c_str());
  if (dSum2) snprintf(a.m_fnAlignsum2, 512, "%s_SumCorr2.mrc", prefix.c_str());
  if (dStack) snprintf(a.m_fnStackCorr, 512, "%s_Corrected.mrc", prefix.c_str());
  if (dCC) snprintf(a.m_fnCCmap, 512, "%s_CCMap.mrc", prefix.c_str());
  if (dLog) snprintf(a.m_fnLog, 512, "%s_Log.txt", prefix.c_str());
  if (a.m_bSaveDisp) {
    struct stat st;
    if (stat("dosef_quick", &st) != 0) mkdir("dosef_quick", 0755);
    const std::string d = path + "/dosef_quick/" + prefix;
    snprintf(a.m_dispRawFFT, 512, "%s_RawFFT.mrc", d.c_str());
    snprintf(a.m_dispCorrSum, 512, "%s_CorrSum.mrc", d.c_str());
    snprintf(a.m_dispCorrFFT, 512, "%s_CorrFFT.mrc", d.c_str());
  }
  return verify_para(a);
}

// dosefgpu_driftcorr.cpp:517-579
static void show_para(CDFAlign& a) {
  APARA& p = a.m_para;
  if (!p.bSaveLog) return;
  FILE* fp = fopen(a.m_fnLog, "w");
  if (!fp) return;
  fprintf(fp, "******Parameter List******\n");
  fprintf(fp, "Input:  %s\n", a.m_fnStack);
  fprintf(fp, "  -crx  %d\n", p.crop_offsetx);
  fprintf(fp, "  -cry  %d\n", p.crop_offsety);
  fprintf(fp, "  -crd  %d\n", p.crop_nsam);
  fprintf(fp, "  -bin  %f\n", p.bin);
  fprintf(fp, "  -scx  %f\n", p.xscale);
  fprintf(fp, "  -scy  %f\n", p.yscale);
  fprintf(fp, "  -nst  %d\n", p.nStart);
  fprintf(fp, "  -ned  %d\n", p.nEnd);
  fprintf(fp, "  -nss  %d\n", p.nStartSum);
  fprintf(fp, "  -nes  %d\n", p.nEndSum);
  fprintf(fp, "  -2es  %d\n", p.nEndSum2);
  fprintf(fp, "  -gpu  %d\n", p.GPUNum);
  fprintf(fp, "  -bft  %f\n", p.bfactor);
  fprintf(fp, "  -pbx  %d\n", p.CCPeakSearchDim);
  fprintf(fp, "  -fod  %d\n", p.FrameDistOffset);
  fprintf(fp, "  -nps  %d\n", p.NoisePeakSize);
  fprintf(fp, "  -kit  %f\n", p.kiThresh);
  fprintf(fp, "  -sub  %d\n", (int)p.bSaveSubAreaCorrSum);
  fprintf(fp, "  -srs  %d\n", (int)p.bSaveRawSum);
  fprintf(fp, "  -ssc  %d\n", (int)p.bSaveStackCorr);
  fprintf(fp, "  -scc  %d\n", (int)p.bSaveCCmap);
  fprintf(fp, "  -slg  %d\n", (int)p.bSaveLog);
  fprintf(fp, "  -atm  %d\n", (int)p.bAlignToMid);
  fprintf(fp, "  -dsp  %d\n", (int)a.m_bSaveDisp);
  fprintf(fp, "  -fsc  %d\n", (int)p.bLogFSC);
  fprintf(fp, "  -fcs  %s\n", a.m_fnAlignsum);
  fprintf(fp, "  -2cs  %s\n", a.m_fnAlignsum2);
  if (p.bSaveRawSum) fprintf(fp, "  -frs  %s\n", a.m_fnRawsum);
  if (p.bSaveStackCorr) fprintf(fp, "  -fct  %s\n", a.m_fnStackCorr);
  if (p.bSaveCCmap) fprintf(fp, "  -fcm  %s\n", a.m_fnCCmap);
  if (p.bSaveLog) fprintf(fp, "  -flg  %s\n", a.m_fnLog);
  if (a.m_bSaveDisp) {
    fprintf(fp, "\nQuick results in dosef_quick folder:\n");
    fprintf(fp, "UnCorrected FFT: %s \n", a.m_dispRawFFT);
    fprintf(fp, "Corrected Sum:   %s \n", a.m_dispCorrSum);
    fprintf(fp, "Corrected FFT:   %s \n", a.m_dispCorrFFT);
  }
  if (strlen(a.m_fnDark)) fprintf(fp, "  -fdk  %s\n", a.m_fnDark);
  if (strlen(a.m_fnNorm)) fprintf(fp, "  -fnm  %s\n", a.m_fnNorm);
  fprintf(fp, "  -itr  %d\n", a.m_iter);
  fprintf(fp, "  -ith  %f\n", a.m_iterthres);
  if (p.pad) fprintf(fp, "  -pad  %d\n", p.pad);
  if (p.precision) fprintf(fp, "  -prc  %d\n", p.precision);
  if (p.grad_steps) fprintf(fp, "  -grd  %d\n  -grs  %f\n", p.grad_steps, p.grad_step_px);
  fprintf(fp, "******Parameter End******\n\n");
  fclose(fp);
}

static void usage() {
  static const char* rows[][3] = {
      {"-crx", "0", "Image crop offset X"},
      {"-cry", "0", "Image crop offset Y"},
      {"-crd", "0", "Image crop dimension. If 0, use maximum size."},
      {"-bin", "1.0", "Float number between 1x and 2x. Bin stack before processing."},
      {"-scx", "1.0", "Float number (>=1.0). Stretch image along X axis."},
      {"-scy", "1.0", "Float number (>=1.0). Stretch image along Y axis."},
      {"-nst", "0", "First frame (0-base) used in alignment."},
      {"-ned", "0", "Last frame (0-base) used in alignment. If 0, use maximum value."},
      {"-nss", "0", "First frame (0-base) used for final sum."},
      {"-nes", "0", "Last frame (0-base) used for final sum. If 0,use maximum value."},
      {"-2es", "0", "Last frame (0-base) used for the second sum."},
      {"-gpu", "0", "GPU device ID."},
      {"-bft", "150", "BFactor in pix^2."},
      {"-pbx", "96", "Box dimension for searching CC peak."},
      {"-fod", "2", "Number of frame offset for frame comparision."},
      {"-nps", "0", "Radius of noise peak."},
      {"-kit", "1.0", "Threshold of alignment error in pixel."},
      {"-sub", "0", "1: Save as sub-area corrected sum. 0: Not."},
      {"-srs", "0", "1: Save uncorrected sum. 0: Not."},
      {"-ssc", "0", "1: Save aligned stack. 0: Not."},
      {"-scc", "0", "1: Save CC Map. 0: Not."},
      {"-slg", "1", "1: Save Log. 0: Not."},
      {"-atm", "1", "1: Align to middle frame. 0: Not."},
      {"-dsp", "1", "1: Save quick results. 0: Not."},
      {"-fsc", "0", "1: Calculate and log FSC. 0: Not."},
      {"-frs", "FileName.mrc", "Uncorrected sum"},
      {"-fcs", "FileName.mrc", "Corrected sum"},
      {"-2cs", "FileName.mrc", "Corrected second sum"},
      {"-fct", "FileName.mrc", "Corrected stack"},
      {"-fcm", "FileName.mrc", "CC map"},
      {"-flg", "FileName.txt", "Log file"},
      {"-fdk", "FileName.mrc", "Dark reference"},
      {"-fnm", "FileName.mrc", "Gain reference"},
      {"-itr", "0", "Number of iteration for iterative alignment"},
      {"-ith", "0.0", "Convergence threshold for iterative alignment (RMSD shift)"},
      {"-pad", "0", "Pad the frame into an N x N canvas (N even) instead of square crop."},
      {"-prc", "0", "Pair-stage precision: 0 FP32, 1 FP16 hi/lo MFMA (n a multiple of 512)."},
      {"-grd", "0", "Sharpness-gradient refinement steps of the final shifts (0: off)."},
      {"-grs", "0.05", "First refinement step in pixels."},
  };
  printf("Dose Fractionation Tool:\nDrift correction (MI355X)\n\n");
  printf("    Input: InputStack.mrc [OPTION VALUE] ...\n");
  printf("          *Note: If not specify OPTION, the default value will be used.\n\n");
  printf("           OPTION     VALUE(Default)    Introduction  \n");
  for (auto& r : rows) printf("           %-10s %-17s %s\n", r[0], r[1], r[2]);
}

int main(int narg, char* argv[]) {
  if (narg == 1) {
    usage();
    return 0;
  }
  CDFAlign align;
  if (!get_para(narg, argv, align)) {
    printf("Error: Have wrong parameters. Abort!\n");
    return 1;
  }
  show_para(align);
  align.RunAlign();
  return align.m_status == 0 ? 0 : 2;
}
