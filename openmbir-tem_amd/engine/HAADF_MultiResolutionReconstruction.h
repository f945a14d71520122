// HAADF_MultiResolutionReconstruction.h -- the coarse-to-fine driver of the reference
// (Code/MBIRLib/HAADF/HAADF_MultiResolutionReconstruction.{h,cpp}) over the MI355X
// engine: level i of NumberResolutions reconstructs with voxels 2^(levels-1-i) times
// the final resolution, starts from the previous level's volume replicated 2 x 2 x 2
// and its gains, offsets and variances, and only the last level writes the MRC
// volume.  Levels hand over through the files the engine writes into one temporary
// directory per level ("4x", "2x", "1x" ...), as the reference's do, or on the device
// (DeviceHandover).
#pragma once
#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "HAADF_ReconstructionEngine.h"

namespace mbirtem {

class HAADF_MultiResolutionReconstruction {
 public:
  using Pointer = std::shared_ptr<HAADF_MultiResolutionReconstruction>;
  static Pointer New() { return Pointer(new HAADF_MultiResolutionReconstruction()); }
  virtual ~HAADF_MultiResolutionReconstruction() = default;

  // properties of the reference driver (HAADF_MultiResolutionReconstruction.h:66-110)
  std::string InputFile;                 // MRC tilt series
  std::string TempDir;
  std::string OutputFile;                // MRC volume of the last level
  std::string InitialReconstructionFile; // optional .bin volume for the first level
  std::string BrightFieldFile;           // optional bright-field tilt series (MRC)
  bool DeleteTempFiles = false;
  int NumberResolutions = 1;
  float SampleThickness = 0;
  float TargetGain = 1;
  float StopThreshold = 0.001f;
  int OuterIterations = 1;
  int InnerIterations = 1;
  float SigmaX = 0;
  float MRFShapeParameter = 0.2f;        // diffuseness: the prior's q = this + 1
  float DefaultOffsetValue = 0;
  bool UseDefaultOffset = true;
  int FinalResolution = 1;
  bool ExtendObject = false;
  bool InterpolateInitialReconstruction = false;
  float DefaultVariance = 1;
  float InitialReconstructionValue = 0;
  Real_t DefaultPixelSize = 1;           // used when the file has no FEI header
  std::vector<float> Tilts;              // degrees per used view (empty: from the file)
  std::vector<uint16_t> Subvolume;       // xMin, yMin, zMin, xMax, yMax, zMax (or empty)
  std::vector<uint8_t> ViewMasks;        // excluded file sections
  AdvancedParametersPtr AdvParams;
  uint32_t Seed = 12345;
  // Hand the levels over on the device (the finished coarse job's volume replicated
  // 2 x 2 x 2 into the next level's job, the nuisance parameters in memory) instead
  // of through the level files; intermediate levels then write no files.  Not in the
  // reference; off by default.
  bool DeviceHandover = false;

  // Runs every level; 0 on success, else the failing level's error code.
  virtual void execute();
  int getErrorCondition() const { return m_ErrorCondition; }
  const std::string& getErrorMessage() const { return m_ErrorMessage; }
  void setCancel(bool v);
  bool getCancel() const { return m_Cancel.load(); }

  // The inputs each level ran with, and its geometry and sweep count (after execute()).
  const std::vector<TomoInputs>& getLevelInputs() const { return m_LevelInputs; }
  const std::vector<Geometry>& getLevelGeometry() const { return m_LevelGeometry; }
  const std::vector<int>& getLevelSweeps() const { return m_LevelSweeps; }

 protected:
  HAADF_MultiResolutionReconstruction() = default;
  // The engine of level `level` with its forward model (*fm), given the level's
  // inputs T (the bright-field driver makes a bright-field engine here).
  virtual HAADF_ReconstructionEngine::Pointer makeEngine(int level, TomoInputs& T,
                                                         HAADF_ForwardModel::Pointer* fm);
  int m_ErrorCondition = 0;
  std::string m_ErrorMessage;

 private:
  std::atomic<bool> m_Cancel{false};
  HAADF_ReconstructionEngine::Pointer m_Current;
  std::vector<TomoInputs> m_LevelInputs;
  std::vector<Geometry> m_LevelGeometry;
  std::vector<int> m_LevelSweeps;
};

}  // namespace mbirtem
