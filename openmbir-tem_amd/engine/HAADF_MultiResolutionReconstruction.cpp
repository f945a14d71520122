// HAADF_MultiResolutionReconstruction.cpp -- per-level parameters and hand-over of the
// reference's coarse-to-fine driver
// (Code/MBIRLib/HAADF/HAADF_MultiResolutionReconstruction.cpp:175-455).
#include "HAADF_MultiResolutionReconstruction.h"

#include <cmath>
#include <filesystem>
#include <sstream>

namespace mbirtem {

namespace fs = std::filesystem;

void HAADF_MultiResolutionReconstruction::setCancel(bool v)
{
  m_Cancel.store(v);
  HAADF_ReconstructionEngine::Pointer e = m_Current;
  if (e) e->setCancel(v);
}

HAADF_ReconstructionEngine::Pointer HAADF_MultiResolutionReconstruction::makeEngine(
    int level, TomoInputs& T, HAADF_ForwardModel::Pointer* fm)
{
  (void)level;
  (void)T;
  *fm = HAADF_ForwardModel::New();
  (*fm)->TargetGain = TargetGain;
  auto eng = HAADF_ReconstructionEngine::New();
  eng->setForwardModel(*fm);
  return eng;
}

void HAADF_MultiResolutionReconstruction::execute()
{
  m_ErrorCondition = 0;
  m_ErrorMessage.clear();
  m_LevelInputs.clear();
  m_LevelGeometry.clear();
  m_LevelSweeps.clear();
  if (!AdvParams) AdvParams = std::make_shared<AdvancedParameters>();
  if (NumberResolutions < 1 || FinalResolution < 1) {
    m_ErrorCondition = -1;
    m_ErrorMessage = "NumberResolutions and FinalResolution must be at least 1";
    return;
  }
  std::vector<std::string> temps;
  TomoInputs prev;
  const int L = NumberResolutions;
  // device-resident hand-over: the previous level's finished job (volume on the
  // device) and its nuisance parameters instead of the level files
  const bool dev = DeviceHandover && L > 1;
  mbir_job* prevJob = nullptr;
  mbir_ctx* prevCtx = nullptr;
  std::vector<Real_t> prevI0, prevMu, prevAlpha;
  auto dropPrev = [&]() {
    if (prevJob) mbir_job_destroy(prevJob);
    if (prevCtx) mbir_destroy(prevCtx);
    prevJob = nullptr;
    prevCtx = nullptr;
  };
  for (int i = 0; i < L; ++i) {
    const bool last = i == L - 1;
    if (m_Cancel.load()) {
      dropPrev();
      m_ErrorCondition = -999;
      return;
    }
    auto in = std::make_shared<TomoInputs>();
    HAADF_ReconstructionEngine::InitializeTomoInputs(in);
    TomoInputs& T = *in;
    T.extendObject = ExtendObject;
    // the previous level's outputs are this level's inputs (:211-229)
    T.gainsInputFile = prev.gainsOutputFile;
    T.offsetsInputFile = prev.offsetsOutputFile;
    T.varianceInputFile = prev.varianceOutputFile;
    T.initialReconFile = i == 0 ? InitialReconstructionFile
                                : prev.tempDir + "/ReconstructedObject.bin";
    T.InterpFlag = i == 0 ? (InterpolateInitialReconstruction ? 1 : 0) : 1;
    T.interpolateFactor = std::pow(2.0f, (float)L - 1) * (float)FinalResolution;  // :231
    T.sinoFile = InputFile;
    {
      std::ostringstream d;
      const double f = T.interpolateFactor / (double)(1 << i);
      if (f == std::floor(f)) d << (long long)f;
      else d << f;
      T.tempDir = TempDir + "/" + d.str() + "x";
    }
    if (dev && i > 0) {  // no level files: the coarse job hands over on the device
      T.gainsInputFile.clear();
      T.offsetsInputFile.clear();
      T.varianceInputFile.clear();
      T.initialReconFile.clear();
    }
    if (dev && !last) {
      T.tempDir.clear();  // intermediate levels write nothing
    } else {
      std::error_code ec;
      fs::create_directories(T.tempDir, ec);
      if (ec) {
        dropPrev();
        m_ErrorCondition = -1;
        m_ErrorMessage = "Could not create path: " + T.tempDir;
        return;
      }
      if (last) T.mrcOutputFile = OutputFile;  // only the last level writes the volume
      T.gainsOutputFile = T.tempDir + "/FinalGainParameters.bin";
      T.offsetsOutputFile = T.tempDir + "/FinalOffsetParameters.bin";
      T.varianceOutputFile = T.tempDir + "/FinalVariances.bin";
      for (const char* f : {"/FinalGainParameters.bin", "/FinalOffsetParameters.bin",
                            "/FinalVariances.bin", "/ReconstructedSinogram.bin",
                            "/ReconstructedObject.bin", "/BraggSelector.mrc"})
        temps.push_back(T.tempDir + f);
    }
    // iterations, prior and stop rule per level (:325-349)
    T.NumOuterIter = (uint16_t)OuterIterations;
    T.NumIter = (uint16_t)(i == 0 ? InnerIterations : 1);
    T.p = MRFShapeParameter + 1;
    T.StopThreshold = StopThreshold;
    if (i >= 2) T.StopThreshold = StopThreshold * 2.0f;
    T.SigmaX = std::pow(2.0, (L - 1 - i) * (1 - 3 / T.p)) * SigmaX;
    if (i == 0) {
      T.defaultInitialRecon = InitialReconstructionValue;
      T.defaultVariance = DefaultVariance;
      T.defaultOffset = DefaultOffsetValue;
      T.useDefaultOffset = UseDefaultOffset;
    }
    // voxel size multiple of this level (:350-351)
    T.delta_xy = std::pow(2.0f, (float)(L - i - 1)) * (Real_t)FinalResolution;
    T.delta_xz = std::pow(2.0f, (float)(L - i - 1)) * (Real_t)FinalResolution;
    T.LengthZ = SampleThickness;
    T.targetGain = TargetGain;
    T.tilts = Tilts;
    if (Subvolume.size() == 6) {
      T.useSubvolume = true;
      T.xStart = Subvolume[0];
      T.xEnd = Subvolume[3];
      T.yStart = Subvolume[1];
      T.yEnd = Subvolume[4];
      T.zStart = Subvolume[2];
      T.zEnd = Subvolume[5];
    }
    T.excludedViews = ViewMasks;
    T.seed = Seed;

    auto sino = std::make_shared<Sinogram>();
    sino->delta_r = DefaultPixelSize;  // replaced by the FEI pixel size when present
    sino->delta_t = DefaultPixelSize;
    auto geom = std::make_shared<Geometry>();
    HAADF_ForwardModel::Pointer fm;
    auto eng = makeEngine(i, T, &fm);
    eng->setTomoInputs(in);
    eng->setSinogram(sino);
    eng->setGeometry(geom);
    eng->setAdvParams(AdvParams);
    if (!BrightFieldFile.empty()) {
      // the bright-field tilt series with the level's selection
      // (HAADF/HAADF_MultiResolutionReconstruction.cpp:189-203, 366-390, 421-422)
      auto bf = std::make_shared<TomoInputs>();
      HAADF_ReconstructionEngine::InitializeTomoInputs(bf);
      bf->sinoFile = BrightFieldFile;
      if (Subvolume.size() == 6) {
        bf->useSubvolume = true;
        bf->xStart = Subvolume[0];
        bf->xEnd = Subvolume[3];
        bf->yStart = Subvolume[1];
        bf->yEnd = Subvolume[4];
        bf->zStart = Subvolume[2];
        bf->zEnd = Subvolume[5];
      }
      bf->interpolateFactor = T.interpolateFactor;
      bf->excludedViews = ViewMasks;
      eng->setBFTomoInputs(bf);
      eng->setBFSinogram(std::make_shared<Sinogram>());
    }
    if (dev && i > 0) {
      // what the level files would carry: the nuisance files exist only for the
      // estimated parameters (gains and offsets with joint estimation, variances
      // with noise estimation), and an offsets file turns the default offset off
      if (AdvParams->JOINT_ESTIMATION) {
        fm->InitialGain = prevI0;
        fm->InitialOffset = prevMu;
        T.useDefaultOffset = false;
      }
      if (AdvParams->NOISE_ESTIMATION) fm->InitialVariance = prevAlpha;
      eng->setInitialJob(prevJob);
    }
    if (dev && !last) eng->setKeepJob(true);
    m_Current = eng;
    if (m_Cancel.load()) eng->setCancel(true);
    eng->execute();
    m_Current.reset();
    dropPrev();  // the level's job has taken the coarse volume (or failed)
    if (dev && !last && eng->getErrorCondition() >= 0) {
      prevJob = eng->takeJob(&prevCtx);
      prevI0 = fm->I_0;
      prevMu = fm->Mu;
      prevAlpha = fm->Alpha;
    }
    m_LevelInputs.push_back(T);
    Geometry gcopy = *geom;
    gcopy.Object.clear();
    m_LevelGeometry.push_back(gcopy);
    m_LevelSweeps.push_back(eng->getSweeps());
    if (eng->getErrorCondition() < 0) {
      m_ErrorCondition = eng->getErrorCondition();
      m_ErrorMessage = "level " + std::to_string(i) + ": " + eng->getErrorMessage();
      return;
    }
    prev = T;
    if (!T.tempDir.empty()) temps.push_back(T.tempDir);  // directories last: removed once empty
  }
  dropPrev();
  if (DeleteTempFiles) {
    for (const std::string& p : temps) {
      std::error_code ec;
      fs::remove(p, ec);  // a directory is removed only when empty
    }
  }
}

}  // namespace mbirtem
