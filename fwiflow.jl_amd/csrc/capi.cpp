// extern "C" entry points declared in include/fwiflow.h.
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fwiflow.h"
#include "cnn.hpp"
#include "config.hpp"
#include "engine.hpp"

using fwiflow::Para;
using fwiflow::Plan;
using fwiflow::Status;
using fwiflow::Survey;

struct fwiflow_plan {
  Plan impl;
};

struct fwiflow_cnn {
  fwiflow::Cnn impl;
};

struct fwiflow_job {
  fwiflow::Job impl;
};

namespace {

thread_local std::string g_last_error;

int report(const Status& s) {
  g_last_error = s.ok() ? std::string() : s.msg;
  return s.code;
}

#define TRY(expr)            \
  do {                       \
    Status s_ = (expr);      \
    if (!s_.ok()) return s_; \
  } while (0)

// Everything fwiflow_cufd checks before it touches a GPU, in the reference's
// order: calc_id (libCUFD.cu:45), parameter file (libCUFD.cu:51), survey file
// and shot ids (libCUFD.cu:85), Courant number (libCUFD.cu:96), observed data
// files (libCUFD.cu:189-192).
Status host_checks(int calc_id, int group_size, const int* shot_ids, const char* para_fname,
                   const double* Lambda, const double* Mu, const double* Den, Para& para,
                   Survey& sv) {
  TRY(fwiflow::check_calc_id(calc_id));
  if (!para_fname) return Status::Err(FWIFLOW_ERR_ARG, "null para_fname");
  TRY(fwiflow::parse_para(para_fname, para));
  if (!Lambda || !Mu || !Den) return Status::Err(FWIFLOW_ERR_ARG, "null model pointer");
  TRY(fwiflow::parse_survey(para, group_size, shot_ids, sv));
  std::vector<float> lam, mu, den;
  float cp_max = 0.f;
  fwiflow::model_to_zfast(para, Lambda, Mu, Den, false, lam, mu, den, cp_max);
  TRY(fwiflow::check_courant(para, cp_max));
  if (calc_id != 2) {
    for (const auto& sh : sv.shots) {
      const std::string fn = para.data_dir + "/Shot" + std::to_string(sh.id) + ".bin";
      FILE* fp = std::fopen(fn.c_str(), "rb");
      if (!fp) return Status::Err(FWIFLOW_ERR_IO, "cannot open \"" + fn + "\"");
      std::fseek(fp, 0, SEEK_END);
      const long sz = std::ftell(fp);
      std::fclose(fp);
      const long want = (long)(sh.z_rec.size() * (size_t)para.nSteps * sizeof(float));
      if (sz < want)
        return Status::Err(FWIFLOW_ERR_IO, "\"" + fn + "\" is " + std::to_string(sz) +
                                               " bytes, expected " + std::to_string(want));
    }
  }
  return Status::Ok();
}

Status cufd_impl(double* misfit, double* gL, double* gM, double* gD, double* gS,
                 const double* Lambda, const double* Mu, const double* Den, const double* stf,
                 int calc_id, int gpu_id, int group_size, const int* shot_ids,
                 const char* para_fname, bool in_scratch) {
  Para para;
  Survey sv;
  TRY(host_checks(calc_id, group_size, shot_ids, para_fname, Lambda, Mu, Den, para, sv));
  if (!stf) return Status::Err(FWIFLOW_ERR_ARG, "null stf pointer");
  if (calc_id == 0 && !misfit) return Status::Err(FWIFLOW_ERR_ARG, "null misfit pointer");
  Plan plan;
  TRY(plan.create(para_fname, gpu_id, group_size, shot_ids, in_scratch));
  TRY(plan.set_model(Lambda, Mu, Den, false));
  TRY(plan.set_stf(stf));
  if (calc_id != 2) TRY(plan.load_obs());
  TRY(plan.run(calc_id));
  if (calc_id == 0) TRY(plan.misfit(misfit));
  if (calc_id == 1) TRY(plan.gradients(gL, gM, gD, gS));
  if (calc_id == 1 && plan.para.save_scratch) TRY(plan.write_scratch());
  if (calc_id == 2) TRY(plan.write_data());
  return Status::Ok();
}

}  // namespace

extern "C" {

int fwiflow_cufd(double* misfit, double* grad_Lambda, double* grad_Mu, double* grad_Den,
                 double* grad_stf, const double* Lambda, const double* Mu, const double* Den,
                 const double* stf, int calc_id, int gpu_id, int group_size,
                 const int* shot_ids, const char* para_fname) {
  return report(cufd_impl(misfit, grad_Lambda, grad_Mu, grad_Den, grad_stf, Lambda, Mu, Den,
                          stf, calc_id, gpu_id, group_size, shot_ids, para_fname, false));
}

int fwiflow_cufd_noalloc(double* misfit, double* grad_Lambda, double* grad_Mu,
                         double* grad_Den, double* grad_stf, const double* Lambda,
                         const double* Mu, const double* Den, const double* stf, int calc_id,
                         int gpu_id, int group_size, const int* shot_ids,
                         const char* para_fname) {
  return report(cufd_impl(misfit, grad_Lambda, grad_Mu, grad_Den, grad_stf, Lambda, Mu, Den,
                          stf, calc_id, gpu_id, group_size, shot_ids, para_fname, true));
}

int fwiflow_check(int calc_id, int group_size, const int* shot_ids, const char* para_fname,
                  const double* Lambda, const double* Mu, const double* Den) {
  Para para;
  Survey sv;
  return report(host_checks(calc_id, group_size, shot_ids, para_fname, Lambda, Mu, Den, para, sv));
}

int fwiflow_host_tables(const char* para_fname, float* cpml_z, float* cpml_x, float* win_data,
                        float* win_src) {
  if (!para_fname) return report(Status::Err(FWIFLOW_ERR_ARG, "null para_fname"));
  Para p;
  Status s = fwiflow::parse_para(para_fname, p);
  if (!s.ok()) return report(s);
  if (cpml_z) fwiflow::cpml_profiles(p.nzA(), p.nPml, p.dz, p.f0, p.dt, cpml_z);
  if (cpml_x) fwiflow::cpml_profiles(p.nx, p.nPml, p.dx, p.f0, p.dt, cpml_x);
  if (win_data) fwiflow::window_squared(p.nSteps, p.dt, 0.005f, win_data);
  if (win_src) fwiflow::window_squared(p.nSteps, p.dt, 0.001f, win_src);
  return report(Status::Ok());
}

const char* fwiflow_last_error(void) { return g_last_error.c_str(); }

const char* fwiflow_version(void) { return "fwiflow-mi355x 0.1.0 (gfx950)"; }

int fwiflow_plan_create(fwiflow_plan** plan, const char* para_fname, int gpu_id, int group_size,
                        const int* shot_ids) {
  if (!plan || !para_fname) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  *plan = nullptr;
  fwiflow_plan* p = new fwiflow_plan();
  Status s = p->impl.create(para_fname, gpu_id, group_size, shot_ids);
  if (!s.ok()) {
    delete p;
    return report(s);
  }
  *plan = p;
  return report(s);
}

int fwiflow_plan_destroy(fwiflow_plan* plan) {
  delete plan;
  return report(Status::Ok());
}

#define PLAN_CALL(expr)                                                   \
  do {                                                                    \
    if (!plan) return report(Status::Err(FWIFLOW_ERR_ARG, "null plan")); \
    return report(expr);                                                  \
  } while (0)

int fwiflow_plan_set_model(fwiflow_plan* plan, const double* Lambda, const double* Mu,
                           const double* Den) {
  PLAN_CALL(plan->impl.set_model(Lambda, Mu, Den, false));
}
int fwiflow_plan_set_model_zfast(fwiflow_plan* plan, const double* Lambda, const double* Mu,
                                 const double* Den) {
  PLAN_CALL(plan->impl.set_model(Lambda, Mu, Den, true));
}
int fwiflow_plan_set_stf(fwiflow_plan* plan, const double* stf) {
  PLAN_CALL(plan->impl.set_stf(stf));
}
int fwiflow_plan_set_obs(fwiflow_plan* plan, const float* obs) {
  PLAN_CALL(plan->impl.set_obs(obs));
}
int fwiflow_plan_load_obs(fwiflow_plan* plan) { PLAN_CALL(plan->impl.load_obs()); }
int fwiflow_plan_run(fwiflow_plan* plan, int calc_id) { PLAN_CALL(plan->impl.run(calc_id)); }
int fwiflow_plan_misfit(fwiflow_plan* plan, double* misfit) {
  PLAN_CALL(plan->impl.misfit(misfit));
}
int fwiflow_plan_gradients(fwiflow_plan* plan, double* grad_Lambda, double* grad_Mu,
                           double* grad_Den, double* grad_stf) {
  PLAN_CALL(plan->impl.gradients(grad_Lambda, grad_Mu, grad_Den, grad_stf));
}
int fwiflow_plan_data(fwiflow_plan* plan, float* syn) { PLAN_CALL(plan->impl.data(syn)); }
int fwiflow_plan_write_data(fwiflow_plan* plan) { PLAN_CALL(plan->impl.write_data()); }
int fwiflow_plan_residual(fwiflow_plan* plan, float* res) { PLAN_CALL(plan->impl.residual(res)); }
int fwiflow_plan_source(fwiflow_plan* plan, float* rows) { PLAN_CALL(plan->impl.source(rows)); }
int fwiflow_plan_dims(fwiflow_plan* plan, int* dims) {
  if (!plan || !dims) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  const Plan& p = plan->impl;
  dims[0] = p.para.nz;
  dims[1] = p.para.nx;
  dims[2] = p.para.nSteps;
  dims[3] = (int)p.survey.shots.size();
  dims[4] = p.survey.total_rec();
  dims[5] = p.para.nPml;
  return report(Status::Ok());
}
int fwiflow_plan_set_checkpoint(fwiflow_plan* plan, int ckpt_every) {
  PLAN_CALL(plan->impl.set_checkpoint(ckpt_every));
}
int fwiflow_plan_set_adjoint(fwiflow_plan* plan, int mode) {
  PLAN_CALL(plan->impl.set_adjoint(mode));
}
int fwiflow_plan_set_timing(fwiflow_plan* plan, int enable) {
  PLAN_CALL(plan->impl.set_timing(enable != 0));
}
int fwiflow_plan_timing(fwiflow_plan* plan, const char* name, double* total_ms,
                        long long* count) {
  if (!name) return report(Status::Err(FWIFLOW_ERR_ARG, "null name"));
  PLAN_CALL(plan->impl.timing(name, total_ms, count));
}
int fwiflow_plan_device_bytes(fwiflow_plan* plan, long long* bytes) {
  if (!plan || !bytes) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  *bytes = plan->impl.device_bytes();
  return report(Status::Ok());
}
int fwiflow_plan_set_illumination(fwiflow_plan* plan, int enable) {
  PLAN_CALL(plan->impl.set_illumination(enable != 0));
}
int fwiflow_plan_illumination(fwiflow_plan* plan, double* illum) {
  PLAN_CALL(plan->impl.illumination(illum));
}
int fwiflow_plan_shots_per_launch(fwiflow_plan* plan, int* forward, int* backward) {
  if (!plan) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  if (forward) *forward = plan->impl.shots_per_launch();
  if (backward) *backward = plan->impl.shots_per_launch_bwd();
  return report(Status::Ok());
}
int fwiflow_plan_launch_groups(fwiflow_plan* plan, int* forward, int* backward) {
  if (!plan) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  if (forward) *forward = plan->impl.launch_groups(0);
  if (backward) *backward = plan->impl.launch_groups(1);
  return report(Status::Ok());
}
int fwiflow_plan_set_memory_budget(fwiflow_plan* plan, double bytes) {
  PLAN_CALL(plan->impl.set_memory_budget(bytes));
}
int fwiflow_plan_set_launch(fwiflow_plan* plan, int tile_x, int shots_fwd, int shots_bwd) {
  PLAN_CALL(plan->impl.set_launch(tile_x, shots_fwd, shots_bwd));
}
int fwiflow_plan_launch(fwiflow_plan* plan, int* tile_x, int* shots_fwd, int* shots_bwd) {
  PLAN_CALL(plan->impl.launch(tile_x, shots_fwd, shots_bwd));
}
int fwiflow_plan_memory_plan(fwiflow_plan* plan, int* ckpt_every, int* shots_per_sweep,
                             long long* ring_bytes, long long* checkpoint_bytes) {
  PLAN_CALL(plan->impl.memory_plan(ckpt_every, shots_per_sweep, ring_bytes, checkpoint_bytes));
}

int fwiflow_job_create(fwiflow_job** job) {
  if (!job) return report(Status::Err(FWIFLOW_ERR_ARG, "null output handle"));
  *job = new fwiflow_job();
  return report(Status::Ok());
}
int fwiflow_job_destroy(fwiflow_job* job) {
  delete job;
  return report(Status::Ok());
}
int fwiflow_job_add(fwiflow_job* job, fwiflow_plan* plan) {
  if (!job || !plan) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  return report(job->impl.add(&plan->impl));
}
int fwiflow_job_run(fwiflow_job* job, int calc_id) {
  if (!job) return report(Status::Err(FWIFLOW_ERR_ARG, "null job"));
  return report(job->impl.run(calc_id));
}

int fwiflow_job_fused(fwiflow_job* job, int* fused) {
  if (!job || !fused) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  *fused = job->impl.fused() ? 1 : 0;
  return report(Status::Ok());
}

#define CNN_CALL(expr)                                                      \
  do {                                                                      \
    if (!net) return report(Status::Err(FWIFLOW_ERR_ARG, "null network")); \
    return report(expr);                                                    \
  } while (0)

int fwiflow_cnn_create(fwiflow_cnn** net, int gpu_id, int height, int width, unsigned seed) {
  if (!net) return report(Status::Err(FWIFLOW_ERR_ARG, "null output handle"));
  *net = nullptr;
  fwiflow_cnn* n = new fwiflow_cnn();
  const Status s = n->impl.create(gpu_id, height, width, seed);
  if (!s.ok()) {
    delete n;
    return report(s);
  }
  *net = n;
  return report(Status::Ok());
}
int fwiflow_cnn_destroy(fwiflow_cnn* net) {
  delete net;
  return report(Status::Ok());
}
int fwiflow_cnn_num_params(fwiflow_cnn* net, int* n) {
  if (!net || !n) return report(Status::Err(FWIFLOW_ERR_ARG, "null argument"));
  *n = net->impl.num_params();
  return report(Status::Ok());
}
int fwiflow_cnn_get_params(fwiflow_cnn* net, float* params) {
  CNN_CALL(net->impl.get_params(params));
}
int fwiflow_cnn_set_params(fwiflow_cnn* net, const float* params) {
  CNN_CALL(net->impl.set_params(params));
}
int fwiflow_cnn_forward(fwiflow_cnn* net, const float* input, float* output) {
  CNN_CALL(net->impl.forward(input, output));
}
int fwiflow_cnn_gradients(fwiflow_cnn* net, const float* input, const float* target,
                          const float* mask, double* loss, float* grad_params) {
  CNN_CALL(net->impl.gradients(input, target, mask, loss, grad_params));
}
int fwiflow_cnn_train_step(fwiflow_cnn* net, const float* input, const float* target,
                           const float* mask, float lr, double* loss) {
  CNN_CALL(net->impl.train_step(input, target, mask, lr, loss));
}

}  // extern "C"
