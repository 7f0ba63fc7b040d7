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
struct JobCache;     // defined in engine.hip
class Job;

class Plan {
 public:
  Plan();
  ~Plan();
  // in_scratch: every device buffer carved from the device's fixed scratch
  // array (FWIFLOW_SCRATCH_BYTES) instead of hipMalloc
  Status create(const std::string& para_fname, int gpu_id, int group_size, const int* shot_ids,
                bool in_scratch = false);
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
  // backward boundary store: 0 automatic (the reference's ring of every step
  // when the sweep's ring fits the memory budget, else checkpoints at the
  // interval that minimises memory); k > 0 checkpoints every k steps
  Status set_checkpoint(int k);
  Status set_adjoint(int mode);  // 0: the reference's adjoint, 1: the exact adjoint
  int adjoint() const;
  Status set_memory_budget(double bytes);  // ring + checkpoints of a gradient run
  // launch shape: tile width in columns (1-4), shots per forward / backward
  // launch (0: all, -1: automatic); replaces the environment defaults
  Status set_launch(int tile_x, int shots_fwd, int shots_bwd);
  Status launch(int* tile_x, int* shots_fwd, int* shots_bwd) const;
  // the last gradient run's checkpoint interval (0: ring of every step), shots
  // per sweep, and device bytes of the ring and of the checkpoints
  Status memory_plan(int* k, int* shots_per_sweep, long long* ring_bytes,
                     long long* checkpoint_bytes) const;
  Status set_timing(bool on);
  Status timing(const std::string& name, double* ms, long long* count);
  long long device_bytes() const;
  int shots_per_launch() const;      // forward sweep (largest group of the last run)
  int shots_per_launch_bwd() const;  // backward sweep
  int launch_groups(int backward) const;  // shot groups of the last run
  Status set_illumination(bool on);
  Status illumination(double* out);

  Para para;
  Survey survey;
  int gpu_id = 0;
  int last_calc = -1;
  float cp_max = 0.f;

 private:
  friend class Job;
  std::unique_ptr<DeviceState> d_;
};

// Several plans (different grids or surveys on one GPU, same nSteps) run
// together: every time-step launch covers a group of shots of every plan.
class Job {
 public:
  Job();
  ~Job();
  Status add(Plan* p);
  Status run(int calc_id);
  int size() const;
  bool fused() const;  // the last run's launches covered every plan

 private:
  bool last_fused_ = true;
  std::vector<Plan*> plans_;
  std::unique_ptr<JobCache> cache_;
};

}  // namespace fwiflow
