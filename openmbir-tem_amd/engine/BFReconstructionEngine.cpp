// BFReconstructionEngine.cpp -- the bright-field engine's problem, measurement and
// selector output (Code/MBIRLib/BrightField/BFReconstructionEngine.cpp:295-970,
// BFForwardModel.cpp:682-740, 939-972, 1001-1045).
#include "BFReconstructionEngine.h"

#include <cmath>

#include "MBIRIO.h"

namespace mbirtem {

int BFReconstructionEngine::configureProblem(mbir_problem& p)
{
  if (!m_BF) {
    fail(-1, "Error: The ForwardModel Structure was NULL.");
    return -1;
  }
  const BFForwardModel& F = *m_BF;
  p.bf_mode = 1;
  p.bragg_T = F.BraggThreshold;
  p.bragg_delta = F.BraggDelta;
  p.bf_offset = F.BfOffset;
  p.mrf_c = 0.01;  // BrightField/BF_QGGMRFPriorModel.cpp:49-58
  p.target_gain = F.TargetGain;
  p.default_offset = F.DefaultOffset;
  p.default_variance = F.DefaultVariance;
  p.bf_nhicd = 1;
  return 0;
}

// -log((counts + BfOffset) / BF_MAX), BF_MAX = 1 (BFForwardModel.cpp:939-972), from the
// float counts the device job is given
void BFReconstructionEngine::measurement(std::vector<double>& y) const
{
  const Sinogram& S = *getSinogram();
  const double off = m_BF ? m_BF->BfOffset : 0.0;
  y.resize(S.counts.size());
  for (size_t i = 0; i < y.size(); ++i) y[i] = -std::log(((double)(float)S.counts[i] + off) / 1.0);
}

// offsets: the file's, else the default, else fitted (BFForwardModel.cpp:682-740)
bool BFReconstructionEngine::useDefaultOffset() const
{
  return m_BF && m_BF->UseDefaultOffset && getTomoInputs()->offsetsInputFile.empty();
}

// the selector [N_theta][N_r][N_t] as a volume with N_z = N_theta, N_x = N_r,
// N_y = N_t, all columns (BFForwardModel.cpp:1001-1045)
void BFReconstructionEngine::writeExtraOutputs(mbir_job* job)
{
  const TomoInputs& T = *getTomoInputs();
  if (T.braggSelectorFile.empty()) return;
  const Sinogram& S = *getSinogram();
  const size_t M = (size_t)S.N_theta * S.N_r * S.N_t;
  std::vector<uint8_t> sel(M);
  if (mbir_job_get_selector(job, sel.data()) < 0) return fail(-100, "selector readback failed");
  std::vector<float> v(sel.begin(), sel.end());
  if (io::writeMRCVolume(T.braggSelectorFile, v.data(), S.N_theta, S.N_r, S.N_t, 0, S.N_r,
                         getGeometry()->LengthX) < 0)
    fail(-1, "Error writing Selector MRC file");
}

}  // namespace mbirtem
