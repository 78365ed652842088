// GPU-backed drop-in for the reference's solver driver: FlatChannelSolver,
// RK3TimeStepper and the two phases' Save/Load (reference
// src/solver_channel_flat.cc:31-414, src/stepper_time.cc:20-246,
// src/phase_carrier_flat.cc:107-171, src/phase_scalar_flat.cc:102-169).
// The stage loop stays on the host; each substage of both phases runs on
// the device through the C ABI (include/turbid_amd.h).
#pragma once

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/turbid_amd.h"
#include "arma_io.h"
#include "ini_file.h"

namespace tbdhost {

// TimeStepper + RK3TimeStepper (stepper_time.h:54-112, stepper_time_rk3.cc:28-31)
class RK3TimeStepper {
 public:
  struct Configuration {
    double maxcfl = 0;
    int output_based_on_step = 1;
    int maxstep = 0, writeinterval = 1, startstep = 0;
    double timestep = 0;
    double max_marching_time = 0, output_interval = 0, initial_time_step = 0;
    int start_time_index = -1;
    std::vector<double> cfl_adjust_scale;
    bool writeinitial = false;
    void Load(const IniFile& ini);
  };
  explicit RK3TimeStepper(const Configuration& c);
  bool Configure();
  bool Update(double maxulen);    // step mode (stepper_time.cc:72-113)
  bool Update2(double maxulen);   // time mode, adaptive dt (stepper_time.cc:120-208)
  void SetAlternationFlag() { alternation_ = !alternation_; }
  bool GetAlternationFlag() const { return alternation_; }
  double GetCoefficients(int id, int stage) const { return coef_[id][stage]; }
  int GetStep() const { return step_; }
  double GetDt() const { return dt_; }
  double GetTime() const { return time_; }
  double GetCfl() const { return cfl_; }
  bool isForward() const { return forward_; }
  bool isOutput() const { return output_; }
  bool isOutputOnStep() const { return cfg_.output_based_on_step != 0; }
  int GetStartStep() const { return cfg_.startstep; }
  int GetStartTimeIndex() const { return cfg_.start_time_index; }
  int GetOutputTimeIndex() const { return output_time_index_; }
  const Configuration& cfg() const { return cfg_; }

 private:
  Configuration cfg_;
  double coef_[3][3];
  bool forward_ = false, output_ = false, alternation_ = false;
  int step_ = 0, output_time_index_ = 0;
  double dt_ = 0, time_ = 0, cfl_ = 0;
};

// Configuration of [Domain], [Space Stepper], both phases -> tbd_config
tbd_config LoadConfig(const IniFile& ini, int nphase, int precision);

class FlatChannelSolver;

// Phase Save/Load of the device-resident variables (phase_NN_{phase,variable,boundary}.dat)
class FlatCarrierPhase {
 public:
  FlatCarrierPhase(FlatChannelSolver* s, int index) : solver_(s), index_(index) {}
  bool Save(const std::string& subfolder);
  bool Load(const std::string& subfolder);
 private:
  FlatChannelSolver* solver_;
  int index_;
};

class FlatScalarPhase {
 public:
  FlatScalarPhase(FlatChannelSolver* s, int index) : solver_(s), index_(index) {}
  bool Save(const std::string& subfolder);
  bool Load(const std::string& subfolder);
 private:
  FlatChannelSolver* solver_;
  int index_;
};

class FlatChannelSolver {
 public:
  // argv[1]: ini path (default configuration.ini), argv[2]: log name
  // (default logcomputation) -- reference src/system.cc:27-34.
  FlatChannelSolver(int& argc, char**& argv);
  ~FlatChannelSolver();
  bool Configure();
  bool Execute();
  bool isConfigured() const { return configured_; }

  // shared with the phases
  tbd_ctx* ctx() { return ctx_; }
  const tbd_config& config() const { return cfg_; }
  std::string directory() const { return "./"; }
  std::vector<cx> uvw_, p_, c_;   // host staging of the state
  int nyc() const { return cfg_.ny / 3 * 2; }
  int mx() const { return (cfg_.nx / 3 * 2) / 2 + 1; }

 private:
  void PrintLogHead(std::ostringstream& s) const;
  void PrintLogBody(std::ostringstream& s, const double* feature) const;
  void LogSave(std::ostringstream& s, bool force);

  IniFile ini_;
  tbd_config cfg_{};
  RK3TimeStepper* tstepper_ = nullptr;
  FlatCarrierPhase* carrier_ = nullptr;
  FlatScalarPhase* scalar_ = nullptr;
  tbd_ctx* ctx_ = nullptr;
  int phasenum_ = 1;
  bool printlog_ = true, writelog_ = true, configured_ = false;
  std::string logname_ = "logcomputation";
  std::ofstream logfile_;
  int logcount_ = 0;
};

// throws std::runtime_error with the library's message when rc != 0
void Check(int rc);

}  // namespace tbdhost
