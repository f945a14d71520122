// BFMultiResolutionReconstruction.h -- the reference's bright-field coarse-to-fine
// driver (Code/MBIRLib/BrightField/BFMultiResolutionReconstruction.{h,cpp}): the same
// levels, parameters and file hand-over as the HAADF driver, each level a
// BFReconstructionEngine; the default offset and variance go to the first level's
// forward model, the Bragg parameters and BfOffset to every level's, and each level
// writes its Bragg selector to <level dir>/BraggSelector.mrc.
#pragma once
#include <memory>

#include "BFReconstructionEngine.h"
#include "HAADF_MultiResolutionReconstruction.h"

namespace mbirtem {

class BFMultiResolutionReconstruction : public HAADF_MultiResolutionReconstruction {
 public:
  using Pointer = std::shared_ptr<BFMultiResolutionReconstruction>;
  static Pointer New() { return Pointer(new BFMultiResolutionReconstruction()); }

  // BrightField/BFMultiResolutionReconstruction.h:105-107
  float BraggThreshold = 3;
  float BraggDelta = 0.5f;
  float BfOffset = 0;

  // The device hand-over is offered for HAADF levels only: -1 when it is asked for.
  void execute() override;

 protected:
  BFMultiResolutionReconstruction() = default;
  HAADF_ReconstructionEngine::Pointer makeEngine(int level, TomoInputs& T,
                                                 HAADF_ForwardModel::Pointer* fm) override;
};

}  // namespace mbirtem
