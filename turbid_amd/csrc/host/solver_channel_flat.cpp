#include "solver_channel_flat.h"

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>

namespace tbdhost {

void Check(int rc) {
  if (rc != 0) throw std::runtime_error(tbd_last_error());
}

// ---------------------------------------------------------------------------
// RK3TimeStepper
// ---------------------------------------------------------------------------
void RK3TimeStepper::Configuration::Load(const IniFile& ini) {   // stepper_time.cc:20-41
  const char* S = "Time Stepper";
  if (UpperCase(ini.GetValueString(S, "applied_stepper_type")) != "RUNGEKUTTA3")
    throw std::runtime_error("FlatChannelSolver::FlatChannelSolver() > Unsupported Time Stepper!");
  maxcfl = ini.GetValueDouble(S, "maximum_cfl_number");
  output_based_on_step = ini.GetValueInt(S, "output_based_on_step");
  maxstep = ini.GetValueInt(S, "maximum_task_step");
  writeinterval = ini.GetValueInt(S, "result_write_interval");
  timestep = ini.GetValueDouble(S, "marching_time_step");
  startstep = ini.GetValueInt(S, "task_start_step");
  max_marching_time = ini.GetValueDouble(S, "max_marching_time");
  output_interval = ini.GetValueDouble(S, "output_interval");
  initial_time_step = ini.GetValueDouble(S, "initial_time_step");
  start_time_index = ini.GetValueInt(S, "task_start_time_index");
  cfl_adjust_scale = ini.GetVectorDouble(S, "cfl_adjust_scale");
  writeinitial = ini.GetValueBool(S, "result_write_initial");
}

RK3TimeStepper::RK3TimeStepper(const Configuration& c) : cfg_(c) {
  const double c1[3] = {0.0 / 1.0, -5.0 / 9.0, -153.0 / 128.0};   // stepper_time_rk3.cc:29-31
  const double c2[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
  const double c3[3] = {1.0 / 6.0, 5.0 / 24.0, 1.0 / 8.0};
  for (int s = 0; s < 3; ++s) { coef_[0][s] = c1[s]; coef_[1][s] = c2[s]; coef_[2][s] = c3[s]; }
}

bool RK3TimeStepper::Configure() {   // stepper_time.cc:215-246
  if (cfg_.output_based_on_step) {
    step_ = std::max(0, cfg_.startstep);
    dt_ = cfg_.timestep;
    time_ = step_ * dt_;
    if (step_ <= cfg_.maxstep) forward_ = true;
  } else {
    step_ = 0;
    dt_ = cfg_.initial_time_step;
    time_ = cfg_.start_time_index > 0 ? cfg_.start_time_index * cfg_.output_interval : 0;
    output_time_index_ = cfg_.start_time_index;
    if (time_ <= cfg_.max_marching_time) forward_ = true;
  }
  output_ = cfg_.writeinitial;
  return true;
}

bool RK3TimeStepper::Update(double maxulen) {   // stepper_time.cc:72-113
  if (!forward_) return true;
  step_ += 1;
  time_ = time_ + dt_;
  forward_ = step_ < cfg_.maxstep;
  output_ = (step_ % cfg_.writeinterval == 0);
  cfl_ = maxulen * dt_;
  if (cfl_ > cfg_.maxcfl) {
    forward_ = false;
    output_ = true;
  }
  return true;
}

bool RK3TimeStepper::Update2(double maxulen) {   // stepper_time.cc:120-208
  if (!forward_) return true;
  step_ += 1;
  time_ = time_ + dt_;
  forward_ = time_ < cfg_.max_marching_time;
  cfl_ = maxulen * dt_;
  if (cfl_ > 0) {
    const double top = cfg_.cfl_adjust_scale.size() > 0 ? cfg_.cfl_adjust_scale[0] : 0.95;
    const double bot = cfg_.cfl_adjust_scale.size() > 1 ? cfg_.cfl_adjust_scale[1] : 0.9;
    if (cfl_ > top * cfg_.maxcfl) {
      dt_ = dt_ * top;
      while (maxulen * dt_ > top * cfg_.maxcfl) dt_ = dt_ * top;
    } else if (cfl_ < bot * cfg_.maxcfl) {
      dt_ = dt_ * 1.05;
    }
    if ((time_ + dt_) / cfg_.output_interval - output_time_index_ - 1 >= 0) {
      const double dttmp = (output_time_index_ + 1) * cfg_.output_interval - time_;
      if (dttmp < cfg_.initial_time_step / 100) dt_ = dt_ / 20;
      else dt_ = dttmp;
      output_time_index_ = output_time_index_ + 1;
      output_ = true;
    } else {
      output_ = false;
    }
  }
  if (cfl_ > cfg_.maxcfl * 1.1 || dt_ < cfg_.initial_time_step / 1000) {
    forward_ = false;
    output_ = true;
  }
  return true;
}

// ---------------------------------------------------------------------------
// configuration
// ---------------------------------------------------------------------------
static void Copy5(double* dst, const std::vector<double>& v, const char* key) {
  if (v.size() < 5) throw std::runtime_error(std::string("FlatCarrierPhase::Configuration::Load(IniFile) > ") + key +
                                             " needs five entries");
  for (int i = 0; i < 5; ++i) dst[i] = v[i];
}

tbd_config LoadConfig(const IniFile& ini, int nphase, int precision) {
  tbd_config c{};
  c.nphase = nphase;
  c.precision = precision;
  // [Domain] (domain.cc:20-35, system.cc:58)
  if (UpperCase(ini.GetValueString("Domain", "applied_domain_type")) != "FIXED")
    throw std::runtime_error("FlatChannelSolver::FlatChannelSolver() > Unsupported Domain Type!");
  const std::vector<double> len = ini.GetVectorDouble("Domain", "dimensionless_length");
  const std::vector<int> grid = ini.GetVectorInt("Domain", "grid_point_number");
  if (len.size() < 3 || grid.size() < 3)
    throw std::runtime_error("Domain::Configuration::Load(IniFile) > length and grid need three entries");
  c.lx = len[0]; c.ly = len[1]; c.lz = len[2];
  c.nx = grid[0]; c.ny = grid[1]; c.nz = grid[2];
  c.slope = ini.GetValueDouble("Domain", "top_bottom_slope");
  c.slope_direction = ini.GetValueInt("Domain", "slope_direction");
  if (ini.GetValueBool("Domain", "read_profile_from_file"))
    throw std::runtime_error("Domain::Configure(void) > bottom/top profiles are not supported (flat bed only)");
  // [Space Stepper] (stepper_space.cc:20-37)
  if (UpperCase(ini.GetValueString("Space Stepper", "applied_stepper_type")) != "CHEBYSHEV")
    throw std::runtime_error("FlatChannelSolver::FlatChannelSolver() > Unsupported Space Stepper!");
  const std::string map = UpperCase(ini.GetValueString("Space Stepper", "mapping_function"));
  if (map == "NONE") c.mapping = 0;
  else if (map == "TWOENDTANGENT") c.mapping = 1;
  else if (map == "TWOENDSIN") c.mapping = 2;
  else if (map == "ONEENDTANGENT") c.mapping = 3;
  else throw std::runtime_error("SpaceStepper::Configuration::Load(IniFile) > Unsupported Mapping Function!");
  const std::vector<double> mc = ini.GetVectorDouble("Space Stepper", "mapping_coefficient");
  c.mapping_coefficient = mc.empty() ? 0.0 : mc[0];
  // [Time Stepper] dt
  c.output_on_step = ini.GetValueInt("Time Stepper", "output_based_on_step");
  c.dt = c.output_on_step ? ini.GetValueDouble("Time Stepper", "marching_time_step")
                          : ini.GetValueDouble("Time Stepper", "initial_time_step");
  // [Phase Carrier] (phase_carrier_flat.cc:27-61)
  const char* PC = "Phase Carrier";
  c.reynolds = ini.GetValueDouble(PC, "reynolds_number");
  c.froude = ini.GetValueDouble(PC, "froude_number");
  c.averaged_concentration = ini.GetValueDouble(PC, "averaged_concentration");
  c.specific_gravity = ini.GetValueDouble(PC, "specific_gravity");
  c.affected_by_scalar_phase = ini.GetValueBool(PC, "affected_by_scalar_phase") ? 1 : 0;
  Copy5(c.vt_u1, ini.GetVectorDouble(PC, "vertical_type_u1"), "vertical_type_u1");
  Copy5(c.vt_u2, ini.GetVectorDouble(PC, "vertical_type_u2"), "vertical_type_u2");
  Copy5(c.vt_u3, ini.GetVectorDouble(PC, "vertical_type_u3"), "vertical_type_u3");
  Copy5(c.vt_p, ini.GetVectorDouble(PC, "vertical_type_p"), "vertical_type_p");
  const std::string force = UpperCase(ini.GetValueString(PC, "applied_force_type"));
  if (force == "NONE") c.force_type = 0;
  else if (force == "CONSTANT") c.force_type = 1;
  else if (force == "OSCILLATORY") c.force_type = 2;
  else if (force == "PULSATILE") c.force_type = 3;
  else if (force == "DAMPING") c.force_type = 4;
  else throw std::runtime_error("FlatCarrierPhase::Configuration::Load(IniFile) > Unsupported Force Type!");
  const std::vector<int> fd = ini.GetVectorInt(PC, "force_direction");
  const std::vector<double> fa = ini.GetVectorDouble(PC, "force_amplitude");
  for (int i = 0; i < 2 && i < (int)fd.size(); ++i) c.force_direction[i] = fd[i];
  for (int i = 0; i < 4 && i < (int)fa.size(); ++i) c.force_amplitude[i] = fa[i];
  // [Phase Scalar] (phase_scalar_flat.cc:27-54)
  if (nphase > 1) {
    const char* PS = "Phase Scalar";
    c.schmidt = ini.GetValueDouble(PS, "schmidt_number");
    c.settling_velocity = ini.GetValueDouble(PS, "settling_velocity");
    c.erosion_rate = ini.GetValueDouble(PS, "erosion_rate");
    const std::string bc = UpperCase(ini.GetValueString(PS, "vertical_boundary_type"));
    if (bc == "ZEROFLUX") {
      c.scalar_bc = 0;
    } else if (bc == "EROSIONDEPOSITION") {
      c.scalar_bc = 1;
      const std::vector<double> tc = ini.GetVectorDouble(PS, "critical_shear_stress");
      c.critical_shear_stress = tc.empty() ? 0.0 : tc[0];
    } else {
      throw std::runtime_error("FlatScalarPhase::Configuration::Load(IniFile) > Unsupported Boundary Type!");
    }
  }
  return c;
}

// ---------------------------------------------------------------------------
// phases: Save / Load (phase_carrier_flat.cc:107-171, variable_carrier_flat.cc:143-377)
// ---------------------------------------------------------------------------
static std::string PhaseFile(int index, const char* what) {
  std::ostringstream s;
  s << "phase_" << std::setw(2) << std::setfill('0') << index << "_" << what;
  return s.str();
}

static void MakeFolder(const std::string& path, const char* who) {
  struct stat st = {};
  const mode_t mode = S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH;
  if (stat(path.c_str(), &st) == -1 && mkdir(path.c_str(), mode) == -1)
    throw std::runtime_error(std::string(who) + "Save(int,ioctrl_t,string) > Failed to Create Target Folder!");
}

bool FlatCarrierPhase::Save(const std::string& sub) {
  const std::string path = solver_->directory() + sub;
  MakeFolder(path, "FlatCarrierPhase::");
  std::ofstream(path + PhaseFile(index_, "phase.dat"), std::ios::out | std::ios::binary).close();
  const tbd_config& c = solver_->config();
  const long long S = (long long)c.nz * solver_->nyc() * solver_->mx();
  solver_->uvw_.resize(3 * S);
  solver_->p_.resize(S);
  Check(tbd_download_state(solver_->ctx(), reinterpret_cast<double*>(solver_->uvw_.data()),
                           reinterpret_cast<double*>(solver_->p_.data()), nullptr));
  SaveCarrierVariable(path + PhaseFile(index_, "variable.dat"), solver_->uvw_.data(), solver_->p_.data(), c.nz,
                      solver_->nyc(), solver_->mx());
  std::ofstream(path + PhaseFile(index_, "boundary.dat"), std::ios::out | std::ios::binary).close();
  return true;
}

bool FlatCarrierPhase::Load(const std::string& sub) {
  const tbd_config& c = solver_->config();
  LoadCarrierVariable(solver_->directory() + sub + PhaseFile(index_, "variable.dat"), solver_->uvw_, solver_->p_,
                      c.nz, solver_->nyc(), solver_->mx());
  return true;
}

bool FlatScalarPhase::Save(const std::string& sub) {
  const std::string path = solver_->directory() + sub;
  MakeFolder(path, "FlatScalarPhase::");
  std::ofstream(path + PhaseFile(index_, "phase.dat"), std::ios::out | std::ios::binary).close();
  const tbd_config& c = solver_->config();
  const long long S = (long long)c.nz * solver_->nyc() * solver_->mx();
  solver_->c_.resize(S);
  Check(tbd_download_state(solver_->ctx(), nullptr, nullptr, reinterpret_cast<double*>(solver_->c_.data())));
  SaveScalarVariable(path + PhaseFile(index_, "variable.dat"), solver_->c_.data(), c.nz, solver_->nyc(),
                     solver_->mx());
  std::ofstream(path + PhaseFile(index_, "boundary.dat"), std::ios::out | std::ios::binary).close();
  return true;
}

bool FlatScalarPhase::Load(const std::string& sub) {
  const tbd_config& c = solver_->config();
  LoadScalarVariable(solver_->directory() + sub + PhaseFile(index_, "variable.dat"), solver_->c_, c.nz,
                     solver_->nyc(), solver_->mx());
  return true;
}

// ---------------------------------------------------------------------------
// FlatChannelSolver (solver_channel_flat.cc:31-414, system.cc:19-64)
// ---------------------------------------------------------------------------
FlatChannelSolver::FlatChannelSolver(int& argc, char**& argv) {
  std::string inipath = "configuration.ini";
  if (argc > 1) inipath = argv[1];
  if (argc > 2) logname_ = argv[2];
  ini_.Load(inipath);
  printlog_ = ini_.GetValueBool("System", "print_log_when_running");
  writelog_ = ini_.GetValueBool("System", "write_log_when_running");
  phasenum_ = ini_.GetValueInt("System", "total_phase_number");
  (void)ini_.GetVectorInt("System", "mpi_grid_point_number");   // mandatory keys (one GPU here)
  (void)ini_.GetValueInt("System", "thread_number_openmp");
  if (phasenum_ < 1 || phasenum_ > 2)
    throw std::runtime_error("FlatChannelSolver::FlatChannelSolver() > total_phase_number must be 1 or 2");
  // precision of the GPU path: environment TURBID_PRECISION = fp64 | mixed
  int precision = 0;
  if (const char* p = std::getenv("TURBID_PRECISION")) precision = (std::string(p) == "mixed") ? 1 : 0;
  cfg_ = LoadConfig(ini_, phasenum_, precision);
  RK3TimeStepper::Configuration tc;
  tc.Load(ini_);
  tstepper_ = new RK3TimeStepper(tc);
  carrier_ = new FlatCarrierPhase(this, 0);
  if (phasenum_ > 1) scalar_ = new FlatScalarPhase(this, 1);
  if (writelog_) logfile_.open(logname_, std::ios::out);
  Check(tbd_create(&cfg_, &ctx_));
}

FlatChannelSolver::~FlatChannelSolver() {
  if (ctx_) tbd_destroy(ctx_);
  delete tstepper_;
  delete carrier_;
  delete scalar_;
}

bool FlatChannelSolver::Configure() {   // solver_channel_flat.cc:148-172
  tstepper_->Configure();
  const int input = tstepper_->isOutputOnStep() ? tstepper_->GetStartStep() : tstepper_->GetStartTimeIndex();
  const long long S = (long long)cfg_.nz * nyc() * mx();
  uvw_.assign(3 * S, cx(0, 0));
  p_.assign(S, cx(0, 0));
  c_.assign(phasenum_ > 1 ? S : 0, cx(0, 0));
  if (input >= 0) {      // variable_carrier_flat.cc:395-401
    const std::string sub = std::to_string((long long)input) + "/";
    carrier_->Load(sub);
    if (scalar_) scalar_->Load(sub);
  }
  Check(tbd_upload_state(ctx_, reinterpret_cast<const double*>(uvw_.data()), reinterpret_cast<const double*>(p_.data()),
                         phasenum_ > 1 ? reinterpret_cast<const double*>(c_.data()) : nullptr));
  configured_ = true;
  return true;
}

bool FlatChannelSolver::Execute() {   // solver_channel_flat.cc:179-313
  std::ostringstream message;
  PrintLogHead(message);
  double feature[50] = {0};
  while (tstepper_->isForward()) {
    if (tstepper_->isOutputOnStep()) {
      tstepper_->Update(feature[19]);
    } else {
      tstepper_->Update2(feature[19]);
      Check(tbd_set_dt(ctx_, tstepper_->GetDt()));   // Phase::Update reassembly
    }
    const int step = tstepper_->GetStep();
    // the three stages of carrier Predict/Correct and scalar Predict/Correct
    // (solver_channel_flat.cc:223-234) as one call: the library replays them
    // from a captured HIP graph
    Check(tbd_step(ctx_, tstepper_->GetTime(), tstepper_->GetAlternationFlag() ? 1 : 0));
    tstepper_->SetAlternationFlag();
    Check(tbd_features(ctx_, feature));
    if (tstepper_->isOutput()) {
      const int index = tstepper_->isOutputOnStep() ? step : tstepper_->GetOutputTimeIndex();
      const std::string sub = std::to_string((long long)index) + "/";
      carrier_->Save(sub);
      if (scalar_) scalar_->Save(sub);
    }
    PrintLogBody(message, feature);
    LogSave(message, false);
  }
  LogSave(message, true);
  return true;
}

void FlatChannelSolver::PrintLogHead(std::ostringstream& s) const {
  s << std::string(120, '*') << std::endl;
  s << std::setw(30) << std::right << "FlatChannelSolver > " << "GPU (MI355X) substage, "
    << (cfg_.precision ? "mixed FP32/FP64" : "FP64") << ", grid " << cfg_.nx << " " << cfg_.ny << " " << cfg_.nz
    << ", phases " << phasenum_ << ", Re " << cfg_.reynolds << ", dt " << cfg_.dt << std::endl;
  s << std::string(120, '*') << std::endl;
}

void FlatChannelSolver::PrintLogBody(std::ostringstream& s, const double* f) const {   // :372-414
  s << std::setw(12) << tstepper_->GetStep() << std::setw(16) << std::setprecision(10) << tstepper_->GetTime()
    << std::setw(12) << std::setprecision(6) << tstepper_->GetCfl();
  for (int i : {0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 14, 15, 21, 22, 23, 24, 25, 28, 29}) s << std::setw(12) << f[i];
  s << std::endl;
}

void FlatChannelSolver::LogSave(std::ostringstream& s, bool force) {   // file_log.cc:54-73
  ++logcount_;
  if (logcount_ > 10 || force) {
    if (printlog_) std::cout << s.str() << std::flush;
    if (writelog_ && logfile_.is_open()) logfile_ << s.str() << std::flush;
    s.str("");
    s.clear();
    logcount_ = 1;
  }
}

}  // namespace tbdhost
