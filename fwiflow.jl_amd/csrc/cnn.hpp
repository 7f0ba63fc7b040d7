// Learned preconditioner network behind the C ABI (fwiflow_cnn_*).  Pure C++
// interface (no HIP types) so the C-ABI layer compiles with the host compiler.
#pragma once
#include <memory>

#include "config.hpp"

namespace fwiflow {

struct CnnState;  // defined in cnn.hip

// 4 -> 32 -> 32 -> 32 -> 3 channels, 3x3 convolutions, GELU, on one H x W grid
class Cnn {
 public:
  Cnn();
  ~Cnn();
  Status create(int gpu_id, int height, int width, unsigned seed);
  int num_params() const;
  Status get_params(float* out) const;
  Status set_params(const float* in);  // also resets the optimiser state
  Status forward(const float* input, float* output);
  Status gradients(const float* input, const float* target, const float* mask, double* loss,
                   float* grad);
  Status train_step(const float* input, const float* target, const float* mask, float lr,
                    double* loss);

 private:
  Status stage(const float* input, const float* target, const float* mask);
  std::unique_ptr<CnnState> s_;
};

}  // namespace fwiflow
