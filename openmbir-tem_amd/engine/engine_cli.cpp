// engine_cli -- run HAADF_ReconstructionEngine::execute() on raw binary inputs or on an
// MRC tilt series.
//
//   engine_cli <sino.f64> <angles.f64> N_r N_t N_theta thickness sigma_x p
//              outer inner joint noise seed <out_prefix>
// sino.f64 is [N_theta][N_r][N_t] little-endian FP64, angles.f64 the N_theta tilts in
// degrees.  Writes <out_prefix>.vol.f64 ([N_z][N_x][N_y]), <out_prefix>.nuis.f64 (I_0,
// Mu, Alpha) and <out_prefix>.txt (dimensions, sweeps, error code, costs).
//
//   engine_cli --mrc <tilts.mrc> thickness sigma_x p outer inner joint noise seed <out_dir>
//              [bright_field.mrc]
// reads the tilt series (tilt angles and pixel size from its FEI extended header) and
// lets the engine write the reference's outputs into out_dir: recon.mrc, gains.bin,
// offsets.bin, variances.bin, ReconstructedSinogram.bin, ReconstructedObject.bin, and
// run.txt as above.  With a bright-field tilt series of the same size every projection
// entry is scaled by its element's bright-field counts + 32768 (no gain or offset
// estimation then).
//
//   engine_cli --multires <tilts.mrc> levels thickness sigma_x diffuseness outer inner
//              stop seed <temp_dir> <out.mrc> [device_handover 0|1]
// runs HAADF_MultiResolutionReconstruction (coarse to fine, levels >= 1; the levels
// hand over through files, or on the device with device_handover 1) and writes
// <temp_dir>/levels.txt: one line per level with its voxel size multiple, sigma_x, q,
// stop threshold, inner iterations, interpolation flag, geometry and sweeps.
//
//   engine_cli --bf-multires <tilts.mrc> levels thickness sigma_x diffuseness outer inner
//              stop seed <temp_dir> <out.mrc> bragg_T bragg_delta bf_offset default_offset
//              use_default_offset
// the same with BFMultiResolutionReconstruction (transmitted counts, bright-field
// engines; each level also writes BraggSelector.mrc).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "BFMultiResolutionReconstruction.h"
#include "HAADF_MultiResolutionReconstruction.h"
#include "HAADF_ReconstructionEngine.h"

using namespace mbirtem;

static std::vector<double> read_f64(const char* path, size_t n)
{
  std::vector<double> v(n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
  if (!f) v.clear();
  return v;
}

static int run_mrc(int argc, char** argv)
{
  if (argc != 12 && argc != 13) {
    std::fprintf(stderr, "usage: %s --mrc tilts.mrc thickness sigma_x p outer inner joint noise "
                         "seed out_dir [bright_field.mrc]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[11];
  auto sino = std::make_shared<Sinogram>();
  auto ti = std::make_shared<TomoInputs>();
  HAADF_ReconstructionEngine::InitializeTomoInputs(ti);
  ti->sinoFile = argv[2];
  ti->LengthZ = std::atof(argv[3]);
  ti->SigmaX = std::atof(argv[4]);
  ti->p = std::atof(argv[5]);
  ti->NumOuterIter = (uint16_t)std::atoi(argv[6]);
  ti->NumIter = (uint16_t)std::atoi(argv[7]);
  ti->StopThreshold = 1e-4;
  ti->seed = (uint32_t)std::atoi(argv[10]);
  ti->tempDir = dir;
  ti->mrcOutputFile = dir + "/recon.mrc";
  ti->gainsOutputFile = dir + "/gains.bin";
  ti->offsetsOutputFile = dir + "/offsets.bin";
  ti->varianceOutputFile = dir + "/variances.bin";
  auto adv = std::make_shared<AdvancedParameters>();
  adv->JOINT_ESTIMATION = (unsigned)std::atoi(argv[8]);
  adv->NOISE_MODEL = (unsigned)std::atoi(argv[9]);
  auto geom = std::make_shared<Geometry>();
  auto fm = HAADF_ForwardModel::New();
  auto eng = HAADF_ReconstructionEngine::New();
  eng->setTomoInputs(ti);
  eng->setSinogram(sino);
  eng->setGeometry(geom);
  eng->setAdvParams(adv);
  eng->setForwardModel(fm);
  if (argc == 13) {  // bright-field tilt series with the same selection
    auto bf = std::make_shared<TomoInputs>();
    HAADF_ReconstructionEngine::InitializeTomoInputs(bf);
    bf->sinoFile = argv[12];
    bf->interpolateFactor = ti->interpolateFactor;
    bf->excludedViews = ti->excludedViews;
    eng->setBFTomoInputs(bf);
    eng->setBFSinogram(std::make_shared<Sinogram>());
  }
  eng->setCostRecording(true);  // the costs go to the run's text file
  eng->execute();
  std::ofstream txt(dir + "/run.txt");
  txt << std::setprecision(17);
  txt << "error " << eng->getErrorCondition() << "\n";
  txt << "dims " << geom->N_z << " " << geom->N_x << " " << geom->N_y << "\n";
  txt << "sinogram " << sino->N_theta << " " << sino->N_r << " " << sino->N_t << "\n";
  txt << "sweeps " << eng->getSweeps() << "\n";
  for (double c : eng->getCosts()) txt << "cost " << c << "\n";
  std::printf("engine_cli: error %d, %d sweeps\n", eng->getErrorCondition(), eng->getSweeps());
  return eng->getErrorCondition() == 0 ? 0 : 1;
}

static int run_multires(int argc, char** argv, bool bf)
{
  if (bf ? argc != 18 : (argc != 13 && argc != 14)) {
    std::fprintf(stderr, "usage: %s --multires tilts.mrc levels thickness sigma_x diffuseness "
                         "outer inner stop seed temp_dir out.mrc [device_handover 0|1]\n"
                         "       %s --bf-multires tilts.mrc levels thickness sigma_x diffuseness "
                         "outer inner stop seed temp_dir out.mrc bragg_T bragg_delta bf_offset "
                         "default_offset use_default_offset\n", argv[0], argv[0]);
    return 2;
  }
  HAADF_MultiResolutionReconstruction::Pointer mr;
  if (bf) {
    auto b = BFMultiResolutionReconstruction::New();
    b->BraggThreshold = (float)std::atof(argv[13]);
    b->BraggDelta = (float)std::atof(argv[14]);
    b->BfOffset = (float)std::atof(argv[15]);
    b->DefaultOffsetValue = (float)std::atof(argv[16]);
    b->UseDefaultOffset = std::atoi(argv[17]) != 0;
    mr = b;
  } else {
    mr = HAADF_MultiResolutionReconstruction::New();
  }
  mr->InputFile = argv[2];
  mr->NumberResolutions = std::atoi(argv[3]);
  mr->SampleThickness = (float)std::atof(argv[4]);
  mr->SigmaX = (float)std::atof(argv[5]);
  mr->MRFShapeParameter = (float)std::atof(argv[6]);
  mr->OuterIterations = std::atoi(argv[7]);
  mr->InnerIterations = std::atoi(argv[8]);
  mr->StopThreshold = (float)std::atof(argv[9]);
  mr->Seed = (uint32_t)std::atoi(argv[10]);
  mr->TempDir = argv[11];
  mr->OutputFile = argv[12];
  mr->DeviceHandover = !bf && argc == 14 && std::atoi(argv[13]) != 0;
  mr->execute();
  std::ofstream txt(std::string(argv[11]) + "/levels.txt");
  txt << std::setprecision(17);
  txt << "error " << mr->getErrorCondition() << "\n";
  const auto& ins = mr->getLevelInputs();
  for (size_t i = 0; i < ins.size(); ++i) {
    const Geometry& g = mr->getLevelGeometry()[i];
    txt << "level " << i << " delta " << ins[i].delta_xz << " sigma " << ins[i].SigmaX << " q "
        << ins[i].p << " stop " << ins[i].StopThreshold << " inner " << ins[i].NumIter
        << " interp " << (int)ins[i].InterpFlag << " dims " << g.N_z << " " << g.N_x << " "
        << g.N_y << " sweeps " << mr->getLevelSweeps()[i] << " dir " << ins[i].tempDir << "\n";
  }
  std::printf("engine_cli: multires error %d (%s)\n", mr->getErrorCondition(),
              mr->getErrorMessage().c_str());
  return mr->getErrorCondition() == 0 ? 0 : 1;
}

int main(int argc, char** argv)
{
  if (argc > 1 && std::string(argv[1]) == "--mrc") return run_mrc(argc, argv);
  if (argc > 1 && std::string(argv[1]) == "--multires") return run_multires(argc, argv, false);
  if (argc > 1 && std::string(argv[1]) == "--bf-multires") return run_multires(argc, argv, true);
  if (argc != 15) {
    std::fprintf(stderr, "usage: %s sino.f64 angles.f64 N_r N_t N_theta thickness sigma_x p "
                         "outer inner joint noise seed out_prefix\n", argv[0]);
    return 2;
  }
  auto sino = std::make_shared<Sinogram>();
  sino->N_r = (uint16_t)std::atoi(argv[3]);
  sino->N_t = (uint16_t)std::atoi(argv[4]);
  sino->N_theta = (uint16_t)std::atoi(argv[5]);
  sino->delta_r = sino->delta_t = 1.0;
  sino->counts = read_f64(argv[1], (size_t)sino->N_theta * sino->N_r * sino->N_t);
  sino->angles = read_f64(argv[2], sino->N_theta);
  auto ti = std::make_shared<TomoInputs>();
  HAADF_ReconstructionEngine::InitializeTomoInputs(ti);
  ti->LengthZ = std::atof(argv[6]);
  ti->SigmaX = std::atof(argv[7]);
  ti->p = std::atof(argv[8]);
  ti->NumOuterIter = (uint16_t)std::atoi(argv[9]);
  ti->NumIter = (uint16_t)std::atoi(argv[10]);
  ti->StopThreshold = 1e-4;
  ti->seed = (uint32_t)std::atoi(argv[13]);
  auto adv = std::make_shared<AdvancedParameters>();
  adv->JOINT_ESTIMATION = (unsigned)std::atoi(argv[11]);
  adv->NOISE_MODEL = (unsigned)std::atoi(argv[12]);
  auto geom = std::make_shared<Geometry>();
  auto fm = HAADF_ForwardModel::New();
  auto eng = HAADF_ReconstructionEngine::New();
  eng->setTomoInputs(ti);
  eng->setSinogram(sino);
  eng->setGeometry(geom);
  eng->setAdvParams(adv);
  eng->setForwardModel(fm);
  eng->setCostRecording(true);  // the costs go to the run's text file
  eng->execute();
  const std::string pre = argv[14];
  std::ofstream txt(pre + ".txt");
  txt << std::setprecision(17);
  txt << "error " << eng->getErrorCondition() << "\n";
  txt << "dims " << geom->N_z << " " << geom->N_x << " " << geom->N_y << "\n";
  txt << "sweeps " << eng->getSweeps() << "\n";
  for (double c : eng->getCosts()) txt << "cost " << c << "\n";
  if (eng->getErrorCondition() == 0) {
    std::ofstream v(pre + ".vol.f64", std::ios::binary);
    v.write(reinterpret_cast<const char*>(geom->Object.data()),
            (std::streamsize)(geom->Object.size() * sizeof(double)));
    std::ofstream n(pre + ".nuis.f64", std::ios::binary);
    for (auto* a : {&fm->I_0, &fm->Mu, &fm->Alpha})
      n.write(reinterpret_cast<const char*>(a->data()), (std::streamsize)(a->size() * sizeof(double)));
  }
  std::printf("engine_cli: error %d, %d sweeps\n", eng->getErrorCondition(), eng->getSweeps());
  return eng->getErrorCondition() == 0 ? 0 : 1;
}
