// Plan: the device-resident engine behind the C ABI.  Pure C++ interface
// (no HIP types) so the C-ABI layer compiles with the host compiler.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "config.hpp"

namespace fwiflow {

struct DeviceState;  // defined in engine.hip

class Plan {
 public:
  Plan();
  ~Plan();
  Status create(const std::string& para_fname, int gpu_id, int group_size, const int* shot_ids);
  Status set_model(const double* Lambda, const double* Mu, const double* Den, bool zfast);
  Status set_stf(const double* stf);
  Status set_obs(const float* obs);
  Status load_obs();
  Status run(int calc_id);
  Status misfit(double* out);
  Status gradients(double* gL, double* gM, double* gD, double* gS);
  Status data(float* syn);
  Status residual(float* res);  // adjoint source of the last calc 0/1 run, [R][nSteps]
  Status source(float* rows);   // source rows [group][nSteps] after the last run
  Status write_data();
  Status write_scratch();
  Status set_checkpoint(int k);
  Status set_timing(bool on);
  Status timing(const std::string& name, double* ms, long long* count);
  long long device_bytes() const;
  int shots_per_launch() const;      // forward sweep
  int shots_per_launch_bwd() const;  // backward sweep
  Status set_illumination(bool on);
  Status illumination(double* out);

  Para para;
  Survey survey;
  int gpu_id = 0;
  int last_calc = -1;
  float cp_max = 0.f;

 private:
  std::unique_ptr<DeviceState> d_;
};

}  // namespace fwiflow
