// BFMultiResolutionReconstruction.cpp -- per-level bright-field engines
// (Code/MBIRLib/BrightField/BFMultiResolutionReconstruction.cpp:175-470).
#include "BFMultiResolutionReconstruction.h"

namespace mbirtem {

void BFMultiResolutionReconstruction::execute()
{
  if (DeviceHandover && NumberResolutions > 1) {
    m_ErrorCondition = -1;
    m_ErrorMessage = "the device hand-over is offered for HAADF levels only";
    return;
  }
  HAADF_MultiResolutionReconstruction::execute();
}

HAADF_ReconstructionEngine::Pointer BFMultiResolutionReconstruction::makeEngine(
    int level, TomoInputs& T, HAADF_ForwardModel::Pointer* fm)
{
  auto f = BFForwardModel::New();
  f->TargetGain = TargetGain;  // :355
  f->BraggDelta = BraggDelta;  // :357-359
  f->BraggThreshold = BraggThreshold;
  f->BfOffset = BfOffset;
  if (level == 0) {  // :342-353
    f->DefaultVariance = DefaultVariance;
    f->DefaultOffset = DefaultOffsetValue;
    f->UseDefaultOffset = UseDefaultOffset;
  }
  if (!T.tempDir.empty()) T.braggSelectorFile = T.tempDir + "/BraggSelector.mrc";  // :282-285
  auto eng = BFReconstructionEngine::New();
  eng->setForwardModel(f);
  *fm = f;
  return eng;
}

}  // namespace mbirtem
