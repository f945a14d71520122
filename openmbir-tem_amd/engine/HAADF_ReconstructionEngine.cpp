// HAADF_ReconstructionEngine.cpp -- execute() of the reference engine interface on
// the MI355X C-ABI.  The setup order follows HAADF_ReconstructionEngine::execute
// (Code/MBIRLib/HAADF/HAADF_ReconstructionEngine.cpp:327-913): input checks,
// geometry sizing (GenericFilters/InitialReconstructionInitializer.cpp:98-167),
// nuisance initialisation, the reconstruction loop, then the outputs.
#include "HAADF_ReconstructionEngine.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "MBIRIO.h"
#include "mbir_estimate.h"

namespace mbirtem {

HAADF_ReconstructionEngine::~HAADF_ReconstructionEngine()
{
  if (m_KeptJob) mbir_job_destroy(m_KeptJob);  // kept for a next level that never came
  if (m_KeptCtx) mbir_destroy(m_KeptCtx);
}

void HAADF_ReconstructionEngine::fail(int code, const std::string& msg)
{
  m_ErrorCondition = code;
  m_ErrorMessage = msg;
}

void HAADF_ReconstructionEngine::setCancel(bool v)
{
  m_Cancel.store(v);
  std::lock_guard<std::mutex> lk(m_JobMutex);
  if (v && m_Job) mbir_job_cancel(m_Job);
}

// The tilt series named by TomoInputs::sinoFile into the sinogram
// (HAADF/HAADF_ReconstructionEngine_Extra.cpp:43-79, MRC files only: the reference's
// raw .bin reader reads nothing).
int HAADF_ReconstructionEngine::readInputData()
{
  TomoInputs& T = *m_TomoInputs;
  Sinogram& S = *m_Sinogram;
  io::TiltSeriesSelect sel;
  sel.useSubvolume = T.useSubvolume;
  sel.xStart = T.xStart; sel.xEnd = T.xEnd;
  sel.yStart = T.yStart; sel.yEnd = T.yEnd;
  sel.zStart = T.zStart; sel.zEnd = T.zEnd;
  sel.interpolateFactor = T.interpolateFactor;
  sel.excludedViews.assign(T.excludedViews.begin(), T.excludedViews.end());
  sel.tilts = T.tilts;
  io::TiltSeries ts;
  const int rc = io::readMRCTiltSeries(T.sinoFile, sel, &ts);
  if (rc < 0) return rc;
  if (ts.N_theta > 65535 || ts.N_r > 65535 || ts.N_t > 65535) return -1;
  S.N_theta = (uint16_t)ts.N_theta;
  S.N_r = (uint16_t)ts.N_r;
  S.N_t = (uint16_t)ts.N_t;
  if (ts.delta_r > 0) S.delta_r = ts.delta_r;
  if (ts.delta_t > 0) S.delta_t = ts.delta_t;
  if (!(S.delta_r > 0)) S.delta_r = 1;
  if (!(S.delta_t > 0)) S.delta_t = 1;
  S.counts = std::move(ts.counts);
  S.angles = std::move(ts.angles);
  T.xStart = (uint16_t)ts.xStart; T.xEnd = (uint16_t)ts.xEnd;
  T.yStart = (uint16_t)ts.yStart; T.yEnd = (uint16_t)ts.yEnd;
  T.fileXSize = (uint16_t)ts.fileX; T.fileYSize = (uint16_t)ts.fileY; T.fileZSize = (uint16_t)ts.fileZ;
  T.goodViews = ts.goodViews;
  return 0;
}

// The bright-field tilt series (HAADF/HAADF_ReconstructionEngine_Extra.cpp:85-175):
// read with its own selection (the HAADF series' tilts), once more over the full
// ranges if that fails; it must have the HAADF sinogram's sizes.  The counts stay raw:
// the C-ABI adds BF_OFFSET.
int HAADF_ReconstructionEngine::readBFData()
{
  m_BFFlag = false;
  if (!m_BFTomoInputs || !m_BFSinogram || m_BFTomoInputs->sinoFile.empty()) return 0;
  TomoInputs& B = *m_BFTomoInputs;
  Sinogram& BS = *m_BFSinogram;
  const Sinogram& S = *m_Sinogram;
  const std::string& f = B.sinoFile;
  const size_t dot = f.find_last_of('.');
  const std::string ext = dot == std::string::npos ? std::string() : f.substr(dot + 1);
  if (ext != "mrc" && ext != "ali") {
    fail(-1, "A supported file reader for the Bright Field file was not found.");
    return -1;
  }
  B.tilts = m_TomoInputs->tilts;
  auto read = [&]() -> int {
    io::TiltSeriesSelect sel;
    sel.useSubvolume = B.useSubvolume;
    sel.xStart = B.xStart; sel.xEnd = B.xEnd;
    sel.yStart = B.yStart; sel.yEnd = B.yEnd;
    sel.zStart = B.zStart; sel.zEnd = B.zEnd;
    sel.interpolateFactor = B.interpolateFactor;
    sel.excludedViews.assign(B.excludedViews.begin(), B.excludedViews.end());
    sel.tilts = B.tilts;
    io::TiltSeries ts;
    const int rc = io::readMRCTiltSeries(f, sel, &ts);
    if (rc < 0) return rc;
    BS.N_theta = (uint16_t)ts.N_theta;
    BS.N_r = (uint16_t)ts.N_r;
    BS.N_t = (uint16_t)ts.N_t;
    BS.counts = std::move(ts.counts);
    BS.angles = std::move(ts.angles);
    return 0;
  };
  if (read() < 0) {
    // the sizes may not match: the full bright-field sinogram (:113-141)
    B.xStart = 0; B.xEnd = (uint16_t)(S.N_r - 1);
    B.yStart = 0; B.yEnd = (uint16_t)(S.N_t - 1);
    B.zStart = 0; B.zEnd = (uint16_t)(S.N_theta - 1);
    if (read() < 0) {
      fail(-1, "Error reading Input Sinogram Data file");
      return -1;
    }
  }
  if (BS.N_r != S.N_r || BS.N_t != S.N_t || BS.N_theta != S.N_theta) {
    fail(-1, "The two file sizes are not matched");
    return -1;
  }
  m_BFFlag = true;
  return 0;
}

// Initial gains, offsets and variances from the files named in TomoInputs, one double
// per used view (IOFilters/GainsOffsetsReader.cpp:64-230): a file that cannot be
// opened leaves the defaults, a short one is an error.
int HAADF_ReconstructionEngine::readNuisanceInputs()
{
  TomoInputs& T = *m_TomoInputs;
  HAADF_ForwardModel& F = *m_ForwardModel;
  const size_t n = m_Sinogram->N_theta;
  struct In {
    const std::string* path;
    std::vector<Real_t>* dst;
  } ins[3] = {{&T.gainsInputFile, &F.InitialGain}, {&T.offsetsInputFile, &F.InitialOffset},
              {&T.varianceInputFile, &F.InitialVariance}};
  for (const In& in : ins) {
    if (in.path->empty()) continue;
    std::vector<Real_t> v(n);
    const int rc = io::readDoubles(*in.path, v.data(), n);
    if (rc == -1) continue;
    if (rc < 0) return -1;
    *in.dst = std::move(v);
    if (in.dst == &F.InitialOffset) T.useDefaultOffset = false;
  }
  return 0;
}

// The files HAADF_ReconstructionEngine::execute() writes after the reconstruction
// (HAADF/HAADF_ReconstructionEngine.cpp:836-902): nuisance parameters, the
// reconstructed sinogram and raw volume in tempDir, the MRC volume over the x range
// that lies over the detector.  Only files whose names are set are written.
void HAADF_ReconstructionEngine::writeOutputs(mbir_job* job, const std::vector<float>& vol)
{
  const TomoInputs& T = *m_TomoInputs;
  const Sinogram& S = *m_Sinogram;
  const Geometry& G = *m_Geometry;
  const HAADF_ForwardModel& F = *m_ForwardModel;
  const AdvancedParameters& A = *m_AdvParams;
  auto note = [&](int rc, const char* what) {
    if (rc < 0) fail(-1, std::string("Error writing ") + what);
  };
  if (A.JOINT_ESTIMATION) {
    if (!T.gainsOutputFile.empty()) note(io::writeDoubles(T.gainsOutputFile, F.I_0.data(), S.N_theta), "gains");
    if (!T.offsetsOutputFile.empty()) note(io::writeDoubles(T.offsetsOutputFile, F.Mu.data(), S.N_theta), "offsets");
  }
  if (A.NOISE_ESTIMATION && !T.varianceOutputFile.empty())
    note(io::writeDoubles(T.varianceOutputFile, F.Alpha.data(), S.N_theta), "variances");
  if (!T.tempDir.empty()) {
    const size_t M = (size_t)S.N_theta * S.N_r * S.N_t;
    std::vector<float> e(M);
    if (mbir_job_get_error_sinogram(job, e.data()) < 0) return fail(-100, "error sinogram readback failed");
    std::vector<double> y;
    measurement(y);
    note(io::writeReconstructedSinogram(T.tempDir + "/ReconstructedSinogram.bin", S.N_theta, S.N_r,
                                        S.N_t, y.data(), e.data(), F.I_0.data(),
                                        F.Mu.data()),
         "ReconstructedSinogram.bin");
    note(io::writeRawVolume(T.tempDir + "/ReconstructedObject.bin", vol.data(), G.N_z, G.N_x, G.N_y),
         "ReconstructedObject.bin");
  }
  if (!T.mrcOutputFile.empty()) {
    int c0 = 0, c1 = G.N_x;
    // voxel size in detector units: the multiple times delta_r
    // (GenericFilters/InitialReconstructionInitializer.cpp:109)
    io::originalXDims(G.N_x, G.x0, T.delta_xz * S.delta_r, S.N_r, S.delta_r, &c0, &c1);
    // the reference widens the range by one column here (:894)
    c1 = std::min<int>(c1 + 1, G.N_x);
    note(io::writeMRCVolume(T.mrcOutputFile, vol.data(), G.N_z, G.N_x, G.N_y, c0, c1, G.LengthX), "MRC");
  }
}

mbir_job* HAADF_ReconstructionEngine::takeJob(mbir_ctx** ctx)
{
  mbir_job* j = m_KeptJob;
  if (ctx) *ctx = m_KeptCtx;
  m_KeptJob = nullptr;
  m_KeptCtx = nullptr;
  return j;
}

void HAADF_ReconstructionEngine::execute()
{
  m_ErrorCondition = 0;
  m_ErrorMessage.clear();
  m_Sweeps = 0;
  m_Costs.clear();
  // :369-393
  if (!m_TomoInputs) return fail(-1, "Error: The TomoInput Structure was NULL.");
  if (!m_Sinogram) return fail(-1, "Error: The Sinogram Structure was NULL.");
  if (!m_Geometry) return fail(-1, "Error: The Geometry Structure was NULL.");
  if (!m_ForwardModel) return fail(-1, "Error: The ForwardModel Structure was NULL.");
  if (!m_AdvParams) m_AdvParams = std::make_shared<AdvancedParameters>();
  if (m_Cancel.load()) return fail(-999, "cancelled");

  Sinogram& S = *m_Sinogram;
  TomoInputs& T = *m_TomoInputs;
  if (S.counts.empty() && !T.sinoFile.empty()) {
    const int rd = readInputData();
    if (rd < 0) return fail(rd, "Error reading Input Sinogram Data file");
  }
  const size_t M = (size_t)S.N_theta * S.N_r * S.N_t;
  if (S.N_theta == 0 || S.counts.size() != M || S.angles.size() != S.N_theta)
    return fail(-1, "Error reading Input Sinogram Data");
  if (readBFData() < 0) return;  // the error is set

  mbir_geometry g{};
  int rc = mbir_geometry_init(S.N_r, S.N_t, S.N_theta, S.delta_r, S.delta_t, T.LengthZ,
                              T.delta_xz, T.interpolateFactor, T.extendObject ? 1 : 0,
                              S.angles.data(), &g);
  if (rc < 0) return fail(-1, "Error computing the reconstruction geometry");
  S.R0 = g.R0;
  S.RMax = g.RMax;
  S.T0 = g.T0;
  S.TMax = g.TMax;
  Geometry& G = *m_Geometry;
  G.N_x = (uint16_t)g.N_x;
  G.N_y = (uint16_t)g.N_y;
  G.N_z = (uint16_t)g.N_z;
  G.LengthX = g.LengthX;
  G.LengthY = g.LengthY;
  G.x0 = g.x0;
  G.y0 = g.y0;
  G.z0 = g.z0;
  T.delta_xy = g.delta_xy;  // :109-110 (delta_xz kept as the caller's multiple)
  const size_t V = (size_t)g.N_z * g.N_x * g.N_y;

  std::vector<float> counts(M);
  for (size_t i = 0; i < M; ++i) counts[i] = (float)S.counts[i];
  if (!T.initialReconFile.empty()) {
    // HAADF/HAADF_ReconstructionEngine_Extra.cpp:182-222
    const std::string& f = T.initialReconFile;
    const std::string ext = f.size() > 4 ? f.substr(f.size() - 4) : std::string();
    if (ext == ".mrc") return fail(-1, "We are not dealing with mrc volume files.");
    if (ext != ".bin") return fail(-1, "Could not find a compatible reader for the initial reconstruction data file.");
    T.initialRecon.assign(V, 0.0);
    if (io::readRawVolume(f, T.initialRecon.data(), g.N_z, g.N_x, g.N_y, T.InterpFlag) < 0)
      return fail(-1, "Error reading Initial Reconstruction Data from File");
  }
  if (readNuisanceInputs() < 0) return fail(-1, "Error Reading Gains from File");
  if (!useDefaultOffset() && m_ForwardModel->InitialOffset.size() != S.N_theta) {
    // offsets neither given nor defaulted: least-squares fit of the view means
    // (HAADF/HAADF_ForwardModel.cpp:327-345, GenericFilters/ComputeInitialOffsets.cpp)
    std::vector<double> y;
    measurement(y);
    m_ForwardModel->InitialOffset.assign(S.N_theta, 0.0);
    if (mbir_initial_offsets(y.data(), S.N_theta, S.N_r, S.N_t, S.angles.data(),
                             m_ForwardModel->InitialOffset.data(), nullptr) < 0)
      return fail(-1, "Error initializing Input Offsets Data file");
  }
  std::vector<float> init;
  if (!T.initialRecon.empty()) {
    if (T.initialRecon.size() != V) return fail(-1, "Error reading Initial Reconstruction Data");
    init.resize(V);
    for (size_t i = 0; i < V; ++i) init[i] = (float)T.initialRecon[i];
  }
  const HAADF_ForwardModel& F = *m_ForwardModel;
  auto opt = [&](const std::vector<Real_t>& v) -> const double* {
    return v.size() == S.N_theta ? v.data() : nullptr;
  };

  mbir_problem p{};
  p.geom = g;
  p.angles_deg = S.angles.data();
  p.counts = counts.data();
  p.init_volume = init.empty() ? nullptr : init.data();
  std::vector<float> bfc;
  if (m_BFFlag) {
    bfc.assign(m_BFSinogram->counts.begin(), m_BFSinogram->counts.end());
    p.bf_sinogram = bfc.data();
  }
  p.default_recon = T.defaultInitialRecon;
  p.init_gain = opt(F.InitialGain);
  p.init_offset = useDefaultOffset() ? nullptr : opt(F.InitialOffset);
  p.init_variance = opt(F.InitialVariance);
  p.target_gain = T.targetGain;
  p.default_offset = T.defaultOffset;
  p.default_variance = T.defaultVariance;
  p.sigma_x = T.SigmaX;
  p.mrf_q = T.p;
  p.mrf_c = 1e-3;  // HAADF/HAADF_QGGMRFPriorModel.cpp:15
  p.stop_threshold = T.StopThreshold;
  p.num_outer = T.NumOuterIter;
  p.num_inner = T.NumIter;
  p.joint_estimation = (int)m_AdvParams->JOINT_ESTIMATION;
  p.noise_model = (int)m_AdvParams->NOISE_MODEL;
  p.zero_skipping = (int)m_AdvParams->ZERO_SKIPPING;
  p.estimate_prior = (int)m_AdvParams->ESTIMATE_PRIOR;
  p.seed = T.seed;
  p.schedule = 0;
  p.rows_per_step = 1;
  p.kernel_mode = 0;
  if (configureProblem(p) < 0) return;  // the error is set

  mbir_opts o{};
  o.device = m_Device;
  mbir_ctx* ctx = nullptr;
  if (mbir_create(&o, &ctx) < 0) return fail(-100, "no usable GPU");
  mbir_job* job = nullptr;
  rc = mbir_job_create(ctx, &p, &job);
  if (rc < 0) {
    mbir_destroy(ctx);
    return fail(rc, rc == -2 ? "Error Calling function detectorResponse" : "job setup failed");
  }
  if (m_InitialJob && (rc = mbir_job_upsample_volume(job, m_InitialJob)) < 0) {
    mbir_job_destroy(job);
    mbir_destroy(ctx);
    return fail(-1, "Error reading Initial Reconstruction Data from the coarse level");
  }
  {
    std::lock_guard<std::mutex> lk(m_JobMutex);
    m_Job = job;
    if (m_Cancel.load()) mbir_job_cancel(job);
  }
  mbir_run_stats rs{};
  std::vector<double> costs(m_RecordCosts ? 4096 : 0);
  int32_t nc = 0;
  rc = mbir_job_run(job, &rs, m_RecordCosts ? costs.data() : nullptr, (int32_t)costs.size(), &nc);
  if (rc == 0) {
    m_Sweeps = rs.sweeps;
    m_Costs.assign(costs.begin(), costs.begin() + nc);
    std::vector<float> vol(V);
    HAADF_ForwardModel& Fw = *m_ForwardModel;
    Fw.I_0.assign(S.N_theta, 0);
    Fw.Mu.assign(S.N_theta, 0);
    Fw.Alpha.assign(S.N_theta, 0);
    Fw.TargetGain = T.targetGain;
    if (mbir_job_get_volume(job, vol.data()) < 0 ||
        mbir_job_get_nuisance(job, Fw.I_0.data(), Fw.Mu.data(), Fw.Alpha.data()) < 0) {
      rc = -100;
    } else {
      G.Object.assign(vol.begin(), vol.end());
      // the re-estimated prior scale, as the reference writes it back (:811)
      if (m_AdvParams->ESTIMATE_PRIOR) (void)mbir_job_get_sigma_x(job, &T.SigmaX);
      writeOutputs(job, vol);
      if (m_ErrorCondition == 0) writeExtraOutputs(job);
    }
  }
  {
    std::lock_guard<std::mutex> lk(m_JobMutex);
    m_Job = nullptr;
  }
  if (m_KeepJob && rc == 0 && m_ErrorCondition == 0) {
    m_KeptJob = job;
    m_KeptCtx = ctx;
  } else {
    mbir_job_destroy(job);
    mbir_destroy(ctx);
  }
  if (rc == -999) return fail(-999, "cancelled");
  if (rc < 0) return fail(rc, "reconstruction failed");
}

}  // namespace mbirtem
