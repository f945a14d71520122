// HAADF_ReconstructionEngine.h -- C++ host adapter with the reference engine's
// interface, running the reconstruction through the MI355X C-ABI (mbir_tem.h).
//
// Mirrors (Code/MBIRLib/ paths of the reference):
//   Sinogram / Geometry / TomoInputs / AdvancedParameters
//       Reconstruction/ReconstructionStructures.h:68-187
//   HAADF_ForwardModel nuisance arrays (TargetGain, InitialGain/Offset/Variance,
//       I_0, Mu, Alpha)                         HAADF/HAADF_ForwardModel.h:84-108
//   HAADF_ReconstructionEngine properties, execute(), getErrorCondition(), setCancel()
//       HAADF/HAADF_ReconstructionEngine.h:67-99, Common/AbstractFilter.h
// Error codes follow the reference: 0 ok, -1 bad input, -2 detector response,
// -999 cancelled (HAADF/HAADF_ReconstructionEngine.cpp:369-393,401,522-528);
// -100 and below are device failures of this engine.
// Inputs come either as arrays (the caller fills the sinogram's counts and tilt
// angles in degrees) or as the reference's files (TomoInputs::sinoFile, an MRC tilt
// series; initialReconFile; gains/offsets/varianceInputFile); after the run the
// engine writes the files the reference writes when their names are set
// (mrcOutputFile, gains/offsets/varianceOutputFile, and ReconstructedSinogram.bin /
// ReconstructedObject.bin in tempDir), see MBIRIO.h.
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mbir_tem.h"

namespace mbirtem {

using Real_t = double;

struct Sinogram {
  uint16_t N_r = 0, N_t = 0, N_theta = 0;
  Real_t delta_r = 0, delta_t = 0;
  std::vector<Real_t> counts;  // [N_theta][N_r][N_t]
  std::vector<Real_t> angles;  // tilt angles in degrees
  Real_t R0 = 0, RMax = 0, T0 = 0, TMax = 0;
};

struct Geometry {
  std::vector<Real_t> Object;  // [N_z][N_x][N_y], written by execute()
  Real_t LengthX = 0, LengthY = 0;
  uint16_t N_x = 0, N_z = 0, N_y = 0;
  Real_t x0 = 0, z0 = 0, y0 = 0;
};

struct TomoInputs {
  uint16_t NumIter = 0;
  uint16_t NumOuterIter = 0;
  Real_t SigmaX = 0;
  Real_t p = 0;  // MRF_Q = diffuseness + 1
  Real_t StopThreshold = 0;
  bool extendObject = false;
  Real_t LengthZ = 0;          // sample thickness
  Real_t delta_xz = 1;         // voxel size multiple set by the multi-resolution driver
  Real_t delta_xy = 1;
  Real_t interpolateFactor = 1;
  Real_t defaultInitialRecon = 0;
  std::vector<Real_t> initialRecon;  // optional initial volume [N_z][N_x][N_y]
  Real_t targetGain = 1;
  bool useDefaultOffset = true;
  Real_t defaultOffset = 0;
  Real_t defaultVariance = 1;
  uint32_t seed = 12345;  // update order (the reference seeds from the clock)
  // the reference's file fields (Reconstruction/ReconstructionStructures.h:103-160)
  uint8_t InterpFlag = 0;  // 1: initialReconFile is half size, replicated 2 x 2 x 2
  bool useSubvolume = false;
  uint16_t xStart = 0, xEnd = 0, yStart = 0, yEnd = 0, zStart = 0, zEnd = 0;
  uint16_t fileXSize = 0, fileYSize = 0, fileZSize = 0;
  std::vector<float> tilts;             // degrees, per used view (empty: from the file)
  std::string sinoFile;                 // MRC tilt series
  std::string initialReconFile;         // .bin volume, y x z order
  std::string gainsInputFile, offsetsInputFile, varianceInputFile;
  std::string tempDir;                  // ReconstructedSinogram.bin, ReconstructedObject.bin
  std::string gainsOutputFile, offsetsOutputFile, varianceOutputFile;
  std::string mrcOutputFile;
  std::string braggSelectorFile;        // bright field: the final Bragg selector (MRC)
  std::vector<uint8_t> excludedViews;   // file sections left out
  std::vector<int> goodViews;           // file sections used (set when read from a file)
};

struct AdvancedParameters {
  unsigned int JOINT_ESTIMATION = 1;
  unsigned int ZERO_SKIPPING = 1;
  unsigned int NOISE_MODEL = 1;
  unsigned int NOISE_ESTIMATION = 1;  // write the variances (HAADF/HAADF_ForwardModel.cpp:62)
  unsigned int ESTIMATE_PRIOR = 0;    // re-estimate SigmaX per outer iteration
                                      // (HAADF/HAADF_ReconstructionEngine.cpp:257, 807-813)
};

class HAADF_ForwardModel {
 public:
  using Pointer = std::shared_ptr<HAADF_ForwardModel>;
  static Pointer New() { return Pointer(new HAADF_ForwardModel()); }
  Real_t TargetGain = 1;
  std::vector<Real_t> InitialGain, InitialOffset, InitialVariance;  // optional inputs
  std::vector<Real_t> I_0, Mu, Alpha;                               // outputs of execute()
};

using SinogramPtr = std::shared_ptr<Sinogram>;
using GeometryPtr = std::shared_ptr<Geometry>;
using TomoInputsPtr = std::shared_ptr<TomoInputs>;
using AdvancedParametersPtr = std::shared_ptr<AdvancedParameters>;

class HAADF_ReconstructionEngine {
 public:
  using Pointer = std::shared_ptr<HAADF_ReconstructionEngine>;
  static Pointer New() { return Pointer(new HAADF_ReconstructionEngine()); }
  virtual ~HAADF_ReconstructionEngine();

  void setTomoInputs(TomoInputsPtr v) { m_TomoInputs = v; }
  TomoInputsPtr getTomoInputs() const { return m_TomoInputs; }
  void setSinogram(SinogramPtr v) { m_Sinogram = v; }
  SinogramPtr getSinogram() const { return m_Sinogram; }
  void setGeometry(GeometryPtr v) { m_Geometry = v; }
  GeometryPtr getGeometry() const { return m_Geometry; }
  void setAdvParams(AdvancedParametersPtr v) { m_AdvParams = v; }
  AdvancedParametersPtr getAdvParams() const { return m_AdvParams; }
  void setForwardModel(HAADF_ForwardModel::Pointer v) { m_ForwardModel = v; }
  // Bright-field tilt series (HAADF/HAADF_ReconstructionEngine.h:75-84): when
  // BFTomoInputs::sinoFile names one, every projection entry is scaled by its
  // element's bright-field counts + BF_OFFSET and gains and offsets are not estimated
  void setBFTomoInputs(TomoInputsPtr v) { m_BFTomoInputs = v; }
  TomoInputsPtr getBFTomoInputs() const { return m_BFTomoInputs; }
  void setBFSinogram(SinogramPtr v) { m_BFSinogram = v; }
  SinogramPtr getBFSinogram() const { return m_BFSinogram; }
  HAADF_ForwardModel::Pointer getForwardModel() const { return m_ForwardModel; }

  // HIP device the engine runs on (default 0)
  void setDevice(int d) { m_Device = d; }

  // Runs setup, the outer/inner ICD loop with joint gain/offset and noise
  // estimation, and fills Geometry::Object and the forward model's I_0, Mu, Alpha.
  void execute();
  int getErrorCondition() const { return m_ErrorCondition; }
  const std::string& getErrorMessage() const { return m_ErrorMessage; }
  // Thread-safe: a running execute() stops at its next sweep with -999.
  void setCancel(bool v);
  bool getCancel() const { return m_Cancel.load(); }

  int getSweeps() const { return m_Sweeps; }
  // The cost after the setup, every sweep and every nuisance step, recorded only when
  // asked for (off by default): the reference computes these only in its COST_CALCULATE
  // debugging build (HAADF/HAADF_ReconstructionEngine.cpp:448, 700, 780), and each costs
  // a pass over the sinogram and the volume.
  void setCostRecording(bool on) { m_RecordCosts = on; }
  const std::vector<Real_t>& getCosts() const { return m_Costs; }

  // Device-resident multi-resolution hand-over (no counterpart in the reference,
  // whose levels hand over through files): start from a finished coarse job's volume
  // replicated 2 x 2 x 2 on the device (mbir_job_upsample_volume) instead of
  // TomoInputs::initialReconFile, and keep this run's job for the next level.
  void setInitialJob(mbir_job* coarse) { m_InitialJob = coarse; }
  void setKeepJob(bool v) { m_KeepJob = v; }
  // The kept job and its context (after a successful execute() with setKeepJob(true));
  // the caller destroys both (mbir_job_destroy, then mbir_destroy).
  mbir_job* takeJob(mbir_ctx** ctx);

  static void InitializeTomoInputs(TomoInputsPtr v) { *v = TomoInputs(); }
  static void InitializeAdvancedParams(AdvancedParametersPtr v) { *v = AdvancedParameters(); }

 protected:
  HAADF_ReconstructionEngine() = default;
  void fail(int code, const std::string& msg);

  // Hooks of the bright-field engine (BFReconstructionEngine), which runs the same
  // setup, loop and outputs with another data term:
  // the filled problem before the job is created (< 0: failed, the error is set)
  virtual int configureProblem(mbir_problem& p) { (void)p; return 0; }
  // the measurement the model fits, [N_theta][N_r][N_t] (the counts here): the
  // initial-offset fit and ReconstructedSinogram.bin use it
  virtual void measurement(std::vector<double>& y) const { y = m_Sinogram->counts; }
  // offsets neither from a file nor estimated: the default offset (else fitted)
  virtual bool useDefaultOffset() const { return m_TomoInputs->useDefaultOffset; }
  // files beyond the engine's common outputs (after they are written)
  virtual void writeExtraOutputs(mbir_job* job) { (void)job; }

 private:
  int readInputData();
  int readNuisanceInputs();
  int readBFData();
  void writeOutputs(mbir_job* job, const std::vector<float>& vol);

  TomoInputsPtr m_TomoInputs;
  SinogramPtr m_Sinogram;
  GeometryPtr m_Geometry;
  AdvancedParametersPtr m_AdvParams;
  HAADF_ForwardModel::Pointer m_ForwardModel;
  TomoInputsPtr m_BFTomoInputs;
  SinogramPtr m_BFSinogram;
  bool m_BFFlag = false;
  int m_Device = 0;
  int m_ErrorCondition = 0;
  std::string m_ErrorMessage;
  std::atomic<bool> m_Cancel{false};
  std::mutex m_JobMutex;
  mbir_job* m_Job = nullptr;
  mbir_job* m_InitialJob = nullptr;
  bool m_KeepJob = false;
  mbir_job* m_KeptJob = nullptr;
  mbir_ctx* m_KeptCtx = nullptr;
  int m_Sweeps = 0;
  bool m_RecordCosts = false;
  std::vector<Real_t> m_Costs;
};

}  // namespace mbirtem
