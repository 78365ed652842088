// channelflat: drop-in for the reference's main/channelflat.cc:16-35 on one
// MI355X.  Usage: channelflat [configuration.ini] [logcomputation]
//   --parse-only <ini>   print the configuration the path reads, then exit
//                        (no GPU needed)
//   --postprocess <ini>  swirling strength of the saved indices (the
//                        reference's FlatChannelPostprocessor)
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>

#include "solver_channel_flat.h"

using namespace tbdhost;

static int ParseOnly(const char* path) {
  IniFile ini;
  ini.Load(path);
  const int nphase = ini.GetValueInt("System", "total_phase_number");
  const tbd_config c = LoadConfig(ini, nphase, 0);
  RK3TimeStepper::Configuration tc;
  tc.Load(ini);
  std::cout << std::setprecision(17);
  std::cout << "{\"nx\": " << c.nx << ", \"ny\": " << c.ny << ", \"nz\": " << c.nz << ", \"nphase\": " << c.nphase
            << ", \"lx\": " << c.lx << ", \"ly\": " << c.ly << ", \"lz\": " << c.lz << ", \"slope\": " << c.slope
            << ", \"slope_direction\": " << c.slope_direction << ", \"mapping\": " << c.mapping
            << ", \"dt\": " << c.dt << ", \"output_on_step\": " << c.output_on_step
            << ", \"force_type\": " << c.force_type << ", \"force_direction\": [" << c.force_direction[0] << ", "
            << c.force_direction[1] << "], \"force_amplitude\": [" << c.force_amplitude[0] << ", "
            << c.force_amplitude[1] << ", " << c.force_amplitude[2] << ", " << c.force_amplitude[3]
            << "], \"reynolds\": " << c.reynolds << ", \"froude\": " << c.froude
            << ", \"averaged_concentration\": " << c.averaged_concentration
            << ", \"specific_gravity\": " << c.specific_gravity
            << ", \"affected_by_scalar_phase\": " << c.affected_by_scalar_phase;
  auto vec = [&](const char* k, const double* v) {
    std::cout << ", \"" << k << "\": [" << v[0] << ", " << v[1] << ", " << v[2] << ", " << v[3] << ", " << v[4] << "]";
  };
  vec("vt_u1", c.vt_u1);
  vec("vt_u2", c.vt_u2);
  vec("vt_u3", c.vt_u3);
  vec("vt_p", c.vt_p);
  std::cout << ", \"schmidt\": " << c.schmidt << ", \"settling_velocity\": " << c.settling_velocity
            << ", \"erosion_rate\": " << c.erosion_rate << ", \"critical_shear_stress\": " << c.critical_shear_stress
            << ", \"scalar_bc\": " << c.scalar_bc << ", \"maximum_cfl_number\": " << tc.maxcfl
            << ", \"maximum_task_step\": " << tc.maxstep << ", \"result_write_interval\": " << tc.writeinterval
            << ", \"task_start_step\": " << tc.startstep << "}" << std::endl;
  return 0;
}

// --copy-state <ini> <from> <to>: Load the phase variables of <from>/ and Save
// them to <to>/ with the host's cube I/O (no GPU needed)
static int CopyState(const char* inipath, const char* from, const char* to) {
  IniFile ini;
  ini.Load(inipath);
  const int nphase = ini.GetValueInt("System", "total_phase_number");
  const tbd_config c = LoadConfig(ini, nphase, 0);
  const int nyc = c.ny / 3 * 2, mx = (c.nx / 3 * 2) / 2 + 1;
  std::vector<cx> uvw, p, cc;
  const std::string a = std::string("./") + from + "/", b = std::string("./") + to + "/";
  LoadCarrierVariable(a + "phase_00_variable.dat", uvw, p, c.nz, nyc, mx);
  if (nphase > 1) LoadScalarVariable(a + "phase_01_variable.dat", cc, c.nz, nyc, mx);
  SaveCarrierVariable(b + "phase_00_variable.dat", uvw.data(), p.data(), c.nz, nyc, mx);
  if (nphase > 1) SaveScalarVariable(b + "phase_01_variable.dat", cc.data(), c.nz, nyc, mx);
  return 0;
}

// --postprocess <ini>: FlatChannelPostprocessor::Execute
// (postprocessor_channel_flat.cc:183-296): for every index in
// [PostProcessor] work_index_range, load <index>/phase_00_variable.dat and
// write the swirling strength of its velocity to <index>/phase_00_vortex.dat
// (an Armadillo real cube Nx x Ny x Nz).
static int Postprocess(const char* inipath) {
  IniFile ini;
  ini.Load(inipath);
  const int nphase = ini.GetValueInt("System", "total_phase_number");
  tbd_config c = LoadConfig(ini, nphase, 0);
  c.nphase = 1;   // the postprocessor reads the carrier only
  const std::vector<int> range = ini.GetVectorInt("PostProcessor", "work_index_range");
  if (range.size() != 2) throw std::runtime_error("FlatChannelPostprocessor > work_index_range needs two values");
  const int nyc = c.ny / 3 * 2, mx = (c.nx / 3 * 2) / 2 + 1;
  tbd_ctx* ctx = nullptr;
  Check(tbd_create(&c, &ctx));
  try {
    std::vector<cx> uvw, p;
    std::vector<double> out((size_t)c.nx * c.ny * c.nz);
    for (int index = range[0]; index <= range[1]; ++index) {
      const std::string path = "./" + std::to_string((long long)index) + "/";
      LoadCarrierVariable(path + "phase_00_variable.dat", uvw, p, c.nz, nyc, mx);
      Check(tbd_upload_state(ctx, reinterpret_cast<const double*>(uvw.data()),
                             reinterpret_cast<const double*>(p.data()), nullptr));
      Check(tbd_swirling_strength(ctx, out.data()));
      std::ofstream f(path + "phase_00_vortex.dat", std::ios::out | std::ios::binary);
      if (!f) throw std::runtime_error("FlatChannelPostprocessor::Execute > Failed to Open " + path);
      WriteCubeReal(f, out.data(), c.nx, c.ny, c.nz);
      std::cout << "postprocess: " << path << "phase_00_vortex.dat" << std::endl;
    }
  } catch (...) {
    tbd_destroy(ctx);
    throw;
  }
  tbd_destroy(ctx);
  return 0;
}

int main(int argc, char** argv) {
  try {
    if (argc > 2 && std::strcmp(argv[1], "--parse-only") == 0) return ParseOnly(argv[2]);
    if (argc > 2 && std::strcmp(argv[1], "--postprocess") == 0) return Postprocess(argv[2]);
    if (argc > 4 && std::strcmp(argv[1], "--copy-state") == 0) return CopyState(argv[2], argv[3], argv[4]);
    FlatChannelSolver solver(argc, argv);
    solver.Configure();
    solver.Execute();
  } catch (std::runtime_error& err) {
    std::cout << err.what() << std::endl;
    return 1;
  }
  return 0;
}
