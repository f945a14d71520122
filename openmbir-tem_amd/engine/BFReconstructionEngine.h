// BFReconstructionEngine.h -- the reference's bright-field engine and forward model
// interface (Code/MBIRLib/BrightField/BFReconstructionEngine.h:60-140,
// BFForwardModel.h:102-134) over the MI355X engine: the same setup, outer/inner loop
// and files as HAADF_ReconstructionEngine, with the bright-field data term of the
// C-ABI (bf_mode: -log(counts + BfOffset), Bragg selector, offset-only joint
// estimation, one shared variance) and the reference's NHICD schedule
// (BrightField/BFConstants.h:51 defines NHICD).  Extra output: the final Bragg
// selector as an MRC file (TomoInputs::braggSelectorFile).
#pragma once
#include <memory>
#include <vector>

#include "HAADF_ReconstructionEngine.h"

namespace mbirtem {

class BFForwardModel : public HAADF_ForwardModel {
 public:
  using Pointer = std::shared_ptr<BFForwardModel>;
  static Pointer New() { return Pointer(new BFForwardModel()); }
  // defaults of BrightField/BFReconstructionArgsParser.cpp:176-182
  Real_t BraggThreshold = 3;   // T of the Bragg rejector
  Real_t BraggDelta = 0.5;     // its delta
  Real_t BfOffset = 0;         // added to the counts before the log
  Real_t DefaultOffset = 0;    // per-view offset when UseDefaultOffset (log domain)
  bool UseDefaultOffset = false;  // BFForwardModel.cpp:93
  Real_t DefaultVariance = 1;
};

class BFReconstructionEngine : public HAADF_ReconstructionEngine {
 public:
  using Pointer = std::shared_ptr<BFReconstructionEngine>;
  static Pointer New() { return Pointer(new BFReconstructionEngine()); }

  void setForwardModel(BFForwardModel::Pointer v)
  {
    m_BF = v;
    HAADF_ReconstructionEngine::setForwardModel(v);
  }
  BFForwardModel::Pointer getForwardModel() const { return m_BF; }

 protected:
  BFReconstructionEngine() = default;
  int configureProblem(mbir_problem& p) override;
  void measurement(std::vector<double>& y) const override;
  bool useDefaultOffset() const override;
  void writeExtraOutputs(mbir_job* job) override;

 private:
  BFForwardModel::Pointer m_BF;
};

}  // namespace mbirtem
