// dfalign.cpp -- CDFAlign host driver: reads the stack, applies dark/gain
// and stretch, runs the resident GPU alignment through include/mc_abi.h and
// writes the sums, optional outputs and the log.  The log lines reproduce
// the formats of the reference worker (src-1125/DFAlign.cu:456-1355) so that
// tools that parse the reference log read this one.
#include "dfalign.h"

#include <math.h>
#include <stdarg.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/mc_abi.h"
#include "mrcio.h"

CDFAlign::CDFAlign() {
  m_bRun = false;
  m_iter = 0;
  m_iterthres = 0;
  m_nsam = 0;
  m_nsamRaw = 0;
  m_status = 0;
  m_bSaveDisp = false;
  m_fnStack[0] = m_fnRawsum[0] = m_fnAlignsum[0] = m_fnAlignsum2[0] = 0;
  m_fnStackCorr[0] = m_fnCCmap[0] = m_fnLog[0] = m_fnDark[0] = m_fnNorm[0] = 0;
  m_dispRawFFT[0] = m_dispCorrSum[0] = m_dispCorrFFT[0] = 0;
  memset(&m_para, 0, sizeof m_para);
}

CDFAlign::~CDFAlign() {}

void CDFAlign::Message(const char* str) { printf("%s\n", str); }

// DFAlign.cu:59-71
void CDFAlign::TextOutput(const char* str) {
  m_log = str;
  printf("%s", m_log.c_str());
  if (m_para.bSaveLog && m_fnLog[0]) {
    FILE* fp = fopen(m_fnLog, "a");
    if (fp) {
      fprintf(fp, "%s", m_log.c_str());
      fclose(fp);
    }
  }
}

static std::string fmt(const char* f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// DFAlign.cu:344-368 (gzip is sniffed before the name-based MRC test)
int CDFAlign::getNFrame() { return getNFrame(m_fnStack); }

int CDFAlign::getNFrame(const char* filename) {
  MRC* s = nullptr;
  if (isMRCZ(filename)) s = new MRCZ();
  else if (isMRC(filename)) s = new MRC();
  else if (isTIFF(filename)) s = new MRCT();
  if (!s) return 0;
  int n = 0;
  if (s->open(filename, "rb") > 0) n = s->getNz();
  delete s;
  return n;
}

// DFAlign.cu:395-421: the header of whichever format opens first.
MRCHeader CDFAlign::getMrcHeader(const char* filename) {
  MRCHeader h;
  memset(&h, 0, sizeof h);
  MRC* s = nullptr;
  if (isMRCZ(filename)) s = new MRCZ();
  else if (isMRC(filename)) s = new MRC();
  else if (isTIFF(filename)) s = new MRCT();
  if (!s) return h;
  if (s->open(filename, "rb") > 0) s->getHeader(&h);
  delete s;
  return h;
}

// DFAlign.cu:423-453: the worker runs on its own thread; RunAlign blocks.
void CDFAlign::RunAlign() {
  std::thread t(&CDFAlign::AlignWorker, this);
  t.join();
}

static int read_ref(const char* fn, int nx, int ny, std::vector<float>& a, std::vector<float>& b,
                    float fill2, const char* what, const char* two, const char* one) {
  MRC r;
  char str[1024];
  if (r.open(fn, "rb") <= 0) {
    snprintf(str, sizeof str, "Error: Failed to open %s reference %s .", what, fn);
    CDFAlign::Message(str);
    return -1;
  }
  if (nx != r.getNx() || ny != r.getNy()) {
    snprintf(str, sizeof str, "Error: Image dimension of %s reference %s differs from image stack .",
             what, fn);
    CDFAlign::Message(str);
    return -1;
  }
  a.resize((size_t)nx * ny);
  b.resize((size_t)nx * ny);
  if (r.read2DIm_32bit(a.data(), 0) != r.getImSize()) {
    snprintf(str, sizeof str, "Error: Failed to read %s reference %s .", what, fn);
    CDFAlign::Message(str);
    return -1;
  }
  if (r.getNz() > 1) {
    CDFAlign::Message(two);
    r.read2DIm_32bit(b.data(), 1);
    return 2;
  }
  CDFAlign::Message(one);
  std::fill(b.begin(), b.end(), fill2);
  return 1;
}

void CDFAlign::AlignWorker() {
  APARA& para = m_para;
  m_bRun = true;
  m_status = 0;
  // open stack (DFAlign.cu:468-490)
  MRC* stack = openStack(m_fnStack, 1);
  if (!stack) {
    Message("Only MRC, MRC-gz, TIFF file formats are supported.");
    m_bRun = false;
    m_status = -100;
    return;
  }
  const int nx = stack->getNx(), ny = stack->getNy(), nz = stack->getNz();
  TextOutput(fmt("Input Stack: Nx(%d) Ny(%d) Nz(%d) Mode(%d)\n\n", nx, ny, nz, stack->getMode()).c_str());
  const float bin = para.bin;
  if (bin <= 0) {
    delete stack;
    m_status = -1;
    return;
  }
  // DFAlign.cu:503-504: the quotient in double, stored as float
  const float xscale = (float)(1.0 / para.xscale), yscale = (float)(1.0 / para.yscale);
  const int offx = para.crop_offsetx, offy = para.crop_offsety;
  // geometry (DFAlign.cu:508-522; func.cpp:80-103)
  int nsamUnbin;
  if (para.pad > 0) {
    nsamUnbin = (offx >= 0 && offx < nx && offy >= 0 && offy < ny) ? para.pad : 0;
  } else {
    int s = para.crop_nsam;
    if (offx < 0 || offx >= nx || offy < 0 || offy >= ny) s = 0;
    else {
      if (s <= 0) s = std::min(nx - offx, ny - offy);
      s = std::min(s, std::min(nx - offx, ny - offy));
      if (fabs(bin - 1.0) < 0.0001) s = s / 2 * 2;
      else if (fabs(bin - 2.0) < 0.0001) s = s / 4 * 4;
      else s = s / 2 * 2;
    }
    nsamUnbin = s;
  }
  if (nsamUnbin <= 0) {
    Message("Error: Wrong image Size.");
    m_bRun = false;
    delete stack;
    m_status = MC_ERR_ARG;
    return;
  }
  const int nsam = (int)((float)nsamUnbin / bin) / 2 * 2;
  TextOutput(fmt("Crop Image: Offset(%d %d) RawDim(%d) BinnedDim(%d)\n", offx, offy, nsamUnbin, nsam).c_str());
  m_nsam = nsam;
  m_nsamRaw = (int)((float)nx / bin);
  if (para.nStart < 0) para.nStart = 0;
  if (para.nEnd >= nz) para.nEnd = nz - 1;
  const int nframe = para.nEnd - para.nStart + 1;
  const size_t sizeb = (size_t)(nsam + 2) * nsam;
  const size_t sizebUnbin = (size_t)(nsamUnbin + 2) * nsamUnbin;
  TextOutput(fmt("Allocate host memory: %f Gb\n",
                 (6.0 * nx * ny + sizeb * (2.0 * nframe + 6) + sizebUnbin) / 256.0 / 1024.0 / 1024.0)
                 .c_str());

  // device
  if (mc_init_gpu(para.GPUNum) != MC_OK) {
    Message(mc_last_error());
    Message(fmt("Failed to initialize GPU #%d.", para.GPUNum).c_str());
    m_bRun = false;
    delete stack;
    m_status = MC_ERR_DEVICE;
    return;
  }
  size_t fr = 0, tot = 0;
  mc_mem_check(&fr, &tot);
  TextOutput(fmt("GPU memory:  free:%.0fMb    total:%.0fMb\n", fr / 1024.0 / 1024.0, tot / 1024.0 / 1024.0).c_str());

  // dark / gain references (DFAlign.cu:614-695)
  std::vector<float> dark, dark2, norm, norm2;
  const bool hasDark = strlen(m_fnDark) > 0, hasNorm = strlen(m_fnNorm) > 0;
  if (hasDark && read_ref(m_fnDark, nx, ny, dark, dark2, 0.0f, "dark", "Processing two dark references.",
                          "Processing single dark reference.") < 0) {
    m_bRun = false;
    delete stack;
    m_status = -100;
    return;
  }
  if (hasNorm && read_ref(m_fnNorm, nx, ny, norm, norm2, 1.0f, "gain", "Processing two norm references.",
                          "Processing single norm reference.") < 0) {
    m_bRun = false;
    delete stack;
    m_status = -100;
    return;
  }
  const bool stretch = xscale < 1.0f || yscale < 1.0f;
  const bool correct = hasDark || hasNorm || stretch;
  // The conjugated pair operand differs from the other spectra only with
  // references whose second section differs from the first (a single-section
  // reference has dark2 = 0, gain2 = 1: the raw frame, DFAlign.cu:640-690).
  const bool dark_two = hasDark && dark2 != dark, norm_two = hasNorm && norm2 != norm;
  const bool two = dark_two || norm_two;

  // read stack (DFAlign.cu:696-820): raw sections are uploaded and decoded,
  // dark/gain corrected and stretched on the device (mc_corrections)
  TextOutput("\nRead stack:\n");
  int dtype = MC_DTYPE_FLOAT32;
  size_t esize = 4;
  const int mode = stack->getMode();
  const bool raw = mode == 0 || mode == 1 || mode == 2 || mode == 6;
  if (raw) {
    dtype = mode == 0 ? MC_DTYPE_INT8 : mode == 1 ? MC_DTYPE_INT16 : mode == 6 ? MC_DTYPE_UINT16 : MC_DTYPE_FLOAT32;
    esize = mode == 0 ? 1 : mode == 2 ? 4 : 2;
  }
  const size_t fbytes = (size_t)nx * ny * esize;
  std::vector<unsigned char> host((size_t)nframe * fbytes);
  std::vector<std::string> corr_lines(nframe);
  for (int j = para.nStart; j <= para.nEnd; j++) {
    const int pos = j - para.nStart;
    if (hasNorm && hasDark) corr_lines[pos] = "Correcting dark/gain...\n";
    else if (hasNorm) corr_lines[pos] = "Correcting gain...\n";
    else if (hasDark) corr_lines[pos] = "Correcting dark...\n";
    unsigned char* dst = host.data() + (size_t)pos * fbytes;
    const size_t got = raw ? stack->read2DIm(dst, j) : stack->read2DIm_32bit(reinterpret_cast<float*>(dst), j) * sizeof(float);
    if (got != (raw ? fbytes : (size_t)stack->getImSize() * sizeof(float)))
      TextOutput(fmt("Error when reading #%03d\n", j).c_str());
  }
  delete stack;

  // upload: the device frames hold [nStart, nEnd]; indices are made relative
  void* dframes = nullptr;
  void* dref[4] = {nullptr, nullptr, nullptr, nullptr};  // dark, gain, dark2, gain2
  const std::vector<float>* refs[4] = {hasDark ? &dark : nullptr, hasNorm ? &norm : nullptr,
                                       dark_two ? &dark2 : nullptr, norm_two ? &norm2 : nullptr};
  bool ok = mc_mem_alloc(&dframes, host.size()) == MC_OK &&
            mc_mem_h2d(dframes, host.data(), host.size()) == MC_OK;
  for (int i = 0; ok && i < 4; i++)
    if (refs[i])
      ok = mc_mem_alloc(&dref[i], refs[i]->size() * sizeof(float)) == MC_OK &&
           mc_mem_h2d(dref[i], refs[i]->data(), refs[i]->size() * sizeof(float)) == MC_OK;
  auto free_dev = [&]() {
    if (dframes) mc_mem_free(dframes);
    for (void* d : dref)
      if (d) mc_mem_free(d);
  };
  if (!ok) {
    Message(mc_last_error());
    m_status = MC_ERR_NOMEM;
    free_dev();
    m_bRun = false;
    return;
  }
  mc_corrections corr;
  memset(&corr, 0, sizeof corr);
  corr.dark = static_cast<const float*>(dref[0]);
  corr.gain = static_cast<const float*>(dref[1]);
  corr.dark2 = static_cast<const float*>(dref[2]);
  corr.gain2 = static_cast<const float*>(dref[3]);
  corr.xscale = stretch ? xscale : 1.0f;
  corr.yscale = stretch ? yscale : 1.0f;
  corr.second = two ? 1 : 0;
  mc_movie_desc md;
  memset(&md, 0, sizeof md);
  md.frames = dframes;
  md.frames2 = nullptr;
  md.dtype = dtype;
  md.nx = nx;
  md.ny = ny;
  md.nz = nframe;
  md.nz_file = nz;
  md.first_frame = para.nStart;
  md.corr = correct ? &corr : nullptr;
  mc_params mp;
  memset(&mp, 0, sizeof mp);
  mp.crop_offsetx = offx;
  mp.crop_offsety = offy;
  mp.crop_nsam = para.crop_nsam;
  mp.bin = bin;
  mp.nStart = para.nStart;
  mp.nEnd = para.nEnd;
  mp.nStartSum = para.nStartSum;
  mp.nEndSum = para.nEndSum;
  mp.nEndSum2 = para.nEndSum2;
  mp.bfactor = para.bfactor;
  mp.CCPeakSearchDim = para.CCPeakSearchDim;
  mp.FrameDistOffset = para.FrameDistOffset;
  mp.NoisePeakSize = para.NoisePeakSize;
  mp.kiThresh = para.kiThresh;
  mp.bAlignToMid = para.bAlignToMid;
  mp.iter = m_iter;
  mp.iterthres = m_iterthres;
  mp.pad = para.pad;
  mp.fftscale = para.fftscale;
  mp.precision = para.precision;
  mp.grad_steps = para.grad_steps;
  mp.grad_step_px = para.grad_step_px;

  const int box = para.CCPeakSearchDim;
  const int P = mc_pair_list(nframe, para.FrameDistOffset, nullptr, 0);
  const int K = nframe - 1;
  const int it_cap = std::max(m_iter, 1);
  std::vector<int> pairs(2 * (size_t)P), good(P);
  std::vector<double> pb(2 * (size_t)P), pcc(P), s1(2 * (size_t)K), ki1(P), s2(2 * (size_t)K), ki2(P),
      imeas((size_t)it_cap * 2 * nframe), icc((size_t)it_cap * nframe), rmsd(it_cap),
      fshift(2 * (size_t)K), frs(2 * (size_t)nframe), fmean(nframe);
  std::vector<float> sum((size_t)nsam * nsam), sum2, rawsum, boxes, iboxes;
  if (para.bSaveRawSum) rawsum.resize((size_t)nsam * nsam);
  if (para.nEndSum2 > 0 && para.nEndSum2 < para.nEndSum) sum2.resize((size_t)nsam * nsam);
  // quick-look images (dosef_quick/, DFAlign.cu:73-179, 842-848, 908-911, 1270-1285)
  std::vector<short> disp_raw, disp_sum, disp_fft;
  if (m_bSaveDisp && nsam >= MC_DISPDIM) {
    const size_t dd = (size_t)MC_DISPDIM * MC_DISPDIM;
    if (para.bDispFFTRaw) disp_raw.resize(dd);
    if (para.bDispSumCorr) disp_sum.resize(dd);
    if (para.bDispFFTCorr) disp_fft.resize(dd);
  } else if (m_bSaveDisp) {
    Message(fmt("Quick-look images need a transform size of at least %d (here %d): "
                "none written.", MC_DISPDIM, nsam).c_str());
  }
  std::vector<double> fsc;  // DFAlign.cu:1314-1318
  if (para.bLogFSC) {
    const int nb = mc_fsc_bins(nsam);
    if (nb > 0) fsc.resize((size_t)nb * 4);
  }
  if (para.bSaveCCmap) {
    boxes.resize((size_t)P * box * box);
    iboxes.resize((size_t)it_cap * nframe * box * box);
  }
  mc_outputs out;
  memset(&out, 0, sizeof out);
  out.pairs = pairs.data();
  out.pair_b = pb.data();
  out.pair_cc = pcc.data();
  out.boxmaps = boxes.empty() ? nullptr : boxes.data();
  out.shift1 = s1.data();
  out.ki1 = ki1.data();
  out.good = good.data();
  out.shift2 = s2.data();
  out.ki2 = ki2.data();
  out.iter_meas = imeas.data();
  out.iter_cc = icc.data();
  out.iter_boxmaps = iboxes.empty() ? nullptr : iboxes.data();
  out.rmsd = rmsd.data();
  out.final_shift = fshift.data();
  out.frame_shift = frs.data();
  out.frame_mean = fmean.data();
  out.sum_image = sum.data();
  out.sum2_image = sum2.empty() ? nullptr : sum2.data();
  out.rawsum_image = rawsum.empty() ? nullptr : rawsum.data();
  out.fsc = fsc.empty() ? nullptr : fsc.data();
  out.disp_raw_fft = disp_raw.empty() ? nullptr : disp_raw.data();
  out.disp_corr_sum = disp_sum.empty() ? nullptr : disp_sum.data();
  out.disp_corr_fft = disp_fft.empty() ? nullptr : disp_fft.data();
  auto write_disp = [&](const char* fn, const std::vector<short>& im) {
    if (im.empty() || !fn[0]) return;
    MRC m;
    if (m.open(fn, "wb") > 0) {
      m.createMRC(im.data(), MC_DISPDIM, MC_DISPDIM, 1);
      m.close();
    }
  };
  mc_report rep;
  mc_context ctx = nullptr;
  int rc = mc_context_create(para.GPUNum, &ctx);
  if (rc == MC_OK) rc = mc_align_movie(ctx, &md, &mp, &out, &rep);
  if (rc != MC_OK) {
    Message(mc_last_error());
    m_status = rc;
    mc_context_destroy(ctx);
    free_dev();
    m_bRun = false;
    return;
  }

  // ---- log: stage 1 (DFAlign.cu:818-872)
  for (int j = para.nStart; j <= para.nEnd; j++) {
    const int pos = j - para.nStart;
    if (!corr_lines[pos].empty()) TextOutput(corr_lines[pos].c_str());
    TextOutput(fmt("......Read and sum frame #%03d   mean:%f\n", j, (float)fmean[pos]).c_str());
  }
  if (para.bSaveRawSum && rep.rawsum_written) {
    MRC m;
    if (m.open(m_fnRawsum, "wb") > 0) {
      m.createMRC(rawsum.data(), nsam, nsam, 1);
      TextOutput(fmt("Mean=%f   Min=%f   Max=%f\n", m.m_header.dmean, m.m_header.dmin, m.m_header.dmax).c_str());
      m.close();
    }
    TextOutput(fmt("Save Unaligned Sum to: %s\n", m_fnRawsum).c_str());
  }

  // ---- pairs (DFAlign.cu:875-948)
  TextOutput("\nCalculate relative drift between frames\n");
  for (int q = 0; q < P; q++)
    TextOutput(fmt("......%03d Frame #%03d VS #%03d xy-shift: %8.4f %8.4f      CC:%f\n", q,
                   pairs[2 * q] + para.nStart, pairs[2 * q + 1] + para.nStart, (float)pb[2 * q],
                   (float)pb[2 * q + 1], (float)pcc[q])
                   .c_str());
  write_disp(m_dispRawFFT, disp_raw);  // DFAlign.cu:907-911
  TextOutput("\n......ki: First round \n");
  double kis = 0;
  for (int q = 0; q < P; q++) {
    TextOutput(fmt("......ki #%03d of Frame #%03d VS #%03d: %8.4lf \n", q + para.nStart,
                   pairs[2 * q] + para.nStart, pairs[2 * q + 1] + para.nStart, ki1[q])
                   .c_str());
    kis += ki1[q];
  }
  TextOutput(fmt("................................Average ki: %8.4lf \n\n", kis / P).c_str());
  TextOutput("......ki: Second round \n");
  kis = 0;
  for (int q = 0; q < rep.ngood; q++) {
    TextOutput(fmt("......ki #%03d of Frame #%03d VS #%03d: %8.4f \n", q + para.nStart,
                   pairs[2 * good[q]] + para.nStart, pairs[2 * good[q] + 1] + para.nStart, (float)ki2[q])
                   .c_str());
    kis += ki2[q];
  }
  TextOutput(fmt("................................Average ki: %8.4lf \n\n", kis / rep.ngood).c_str());

  // ---- iterations (DFAlign.cu:950-1071)
  TextOutput("Begin Iterative Alignment... \n");
  std::vector<double> old(2 * (size_t)nframe, 0.0);
  {
    double ax = 0, ay = 0;
    for (int k = 0; k < K; k++) {
      ax += s2[2 * k];
      ay += s2[2 * k + 1];
      old[2 * (k + 1)] = ax;
      old[2 * (k + 1) + 1] = ay;
    }
  }
  int jj;
  for (jj = 0; jj < rep.niter; jj++) {
    TextOutput(fmt("Iteration [%d/%d]\n", jj, m_iter).c_str());
    for (int k = 0; k < nframe; k++)
      TextOutput(fmt("......Add Frame #%03d with xy shift: %8.4lf %8.4lf\n", k + para.nStart, -old[2 * k],
                     -old[2 * k + 1])
                     .c_str());
    TextOutput("Sum complete... \n");
    const double* meas = imeas.data() + (size_t)jj * 2 * nframe;
    double avgcc = 0;
    for (int k = 0; k < nframe; k++) {
      avgcc += (float)icc[(size_t)jj * nframe + k];
      TextOutput(fmt("......%03d Frame #%03d VS <sum> xy-shift: %8.4f %8.4f (%8.4f %8.4f) CC:%f\n", k,
                     k + para.nStart, (float)meas[2 * k], (float)meas[2 * k + 1], old[2 * k] - meas[2 * k],
                     old[2 * k + 1] - meas[2 * k + 1], (float)icc[(size_t)jj * nframe + k])
                     .c_str());
    }
    if (para.bSaveCCmap) {
      const std::string fn = fmt("iterCC%03d.mrc", jj);
      buf2mrc(fn.c_str(), iboxes.data() + (size_t)jj * nframe * box * box, box, box, nframe);
      TextOutput(fmt("Save CC map to: %s\n", fn.c_str()).c_str());
    }
    for (int k = 0; k < nframe; k++) {
      old[2 * k] = meas[2 * k] - meas[0];
      old[2 * k + 1] = meas[2 * k + 1] - meas[1];
    }
    TextOutput(fmt("Average CC : %8.4f\n", avgcc / nframe).c_str());
    if (jj == rep.niter - 1 && rep.converged)
      TextOutput(fmt("Converged at iteration [%d/%d]\n", jj, m_iter).c_str());
    TextOutput(fmt("RMSD to <sum> : %8.4f  (%8.4f for convergence)\n", sqrt(rmsd[jj]), m_iterthres).c_str());
  }
  TextOutput(fmt("Final RMSD to <sum> : %8.4f at iteration [%d/%d]\n\nShifts (delta) after iterative aligment\n",
                 rep.niter ? sqrt(rmsd[rep.niter - 1]) : 0.0, jj - 1, m_iter)
                 .c_str());
  TextOutput(fmt("...Frame #%03d xy-shift %8.4f %8.4f (%8.4f %8.4f)\n", 0, 0.0, 0.0, 0.0, 0.0).c_str());
  {
    double ax = 0, ay = 0, px = 0, py = 0;
    for (int k = 0; k < K; k++) {
      ax += s2[2 * k];
      ay += s2[2 * k + 1];
      px += fshift[2 * k];
      py += fshift[2 * k + 1];
      TextOutput(fmt("...Frame #%03d xy-shift %8.4f %8.4f (%8.4f %8.4f)\n", k + 1, px, py, ax - px, ay - py).c_str());
    }
  }
  // final shifts (DFAlign.cu:1074-1092)
  TextOutput("Final shift:\n");
  m_curve.clear();
  double tx = 0, ty = 0;
  TextOutput(fmt("......Shift of Frame #%03d : %8.4f %8.4f\n", para.nStart, tx, ty).c_str());
  m_curve.push_back(std::complex<double>(tx, ty));
  for (int k = 0; k < K; k++) {
    tx += fshift[2 * k];
    ty += fshift[2 * k + 1];
    TextOutput(fmt("......Shift of Frame #%03d : %8.4f %8.4f\n", k + para.nStart + 1, tx, ty).c_str());
    m_curve.push_back(std::complex<double>(tx, ty));
  }
  if (para.bSaveCCmap) {
    buf2mrc(m_fnCCmap, boxes.data(), box, box, P);
    TextOutput(fmt("Save CC map to: %s\n", m_fnCCmap).c_str());
  }

  // ---- sums (DFAlign.cu:1114-1312)
  int nss = para.nStartSum - para.nStart, nes = para.nEndSum - para.nStart;
  if (nss <= 0) nss = 0;
  if (nes <= nss || nes >= nframe) nes = nframe - 1;
  TextOutput(fmt("\nSum Frame #%03d - #%03d\n", nss + para.nStart, nes + para.nStart).c_str());
  for (int k = nss; k <= nes; k++) {
    // the first summed frame is logged with its shift even when -atm 0 leaves it unshifted
    double sx = frs[2 * k], sy = frs[2 * k + 1];
    if (k == nss && !para.bAlignToMid) {
      double ax = 0, ay = 0;
      for (int q = 0; q < nss; q++) {
        ax += fshift[2 * q];
        ay += fshift[2 * q + 1];
      }
      sx = -ax;
      sy = -ay;
    }
    TextOutput(fmt("......Add Frame #%03d with xy shift: %8.4lf %8.4lf\n", k + para.nStart, sx, sy).c_str());
  }
  std::string tempname;
  if (para.bSaveStackCorr) {
    MRC sc;
    if (sc.open(m_fnStackCorr, "wb") > 0) {
      sc.m_header.nx = nsam;
      sc.m_header.ny = nsam;
      sc.m_header.nz = para.nEndSum - para.nStartSum + 1;
      sc.m_header.mode = 2;
      sc.updateHeader();
      std::vector<double> one(2 * (size_t)nframe, NAN);
      std::vector<float> img((size_t)nsam * nsam);
      for (int k = nss; k <= nes; k++) {
        one[2 * k] = frs[2 * k];
        one[2 * k + 1] = frs[2 * k + 1];
        if (mc_sum_image(ctx, one.data(), img.data()) == MC_OK) sc.write2DIm(img.data(), k - nss);
        one[2 * k] = one[2 * k + 1] = NAN;
      }
      sc.close();
    }
  }
  auto write_sum = [&](const char* fnbase, const std::vector<float>& img) {
    std::string fn = fnbase;
    if (para.bSaveSubAreaCorrSum) {
      const size_t L = strlen(fnbase);
      fn = fmt("%s_ox%d_oy%d_dim%d.mrc", std::string(fnbase, L > 4 ? L - 4 : 0).c_str(), offx, offy, nsamUnbin);
    }
    MRC m;
    if (m.open(fn.c_str(), "wb") > 0) {
      m.createMRC(img.data(), nsam, nsam, 1);
      TextOutput(fmt("Mean=%f   Min=%f   Max=%f\n", m.m_header.dmean, m.m_header.dmin, m.m_header.dmax).c_str());
      m.close();
    }
    TextOutput(fmt("Save Sum to: %s\n", fn.c_str()).c_str());
  };
  if (!sum2.empty() && rep.sum2_written) write_sum(m_fnAlignsum2, sum2);
  write_sum(m_fnAlignsum, sum);
  write_disp(m_dispCorrFFT, disp_fft);  // DFAlign.cu:1271-1285
  write_disp(m_dispCorrSum, disp_sum);
  if (!fsc.empty()) {
    // DFAlign.cu:311-320
    const int nb = (int)(fsc.size() / 4);
    TextOutput("\nFSC parallel(D) and perpendicular(U) to drift direction:\n");
    TextOutput("   Nq%       D_Raw       U_Raw       D_Corr       U_Corr\n");
    for (int i = 0; i < nb; i++)
      TextOutput(fmt("%7.2f %12.4f %12.4f %12.4f %12.4f\n", i * 100.0 / nb, fsc[4 * i],
                     fsc[4 * i + 1], fsc[4 * i + 2], fsc[4 * i + 3]).c_str());
    TextOutput("\n");
  }
  TextOutput("Done.\n");
  mc_context_destroy(ctx);
  free_dev();
  m_bRun = false;
}
