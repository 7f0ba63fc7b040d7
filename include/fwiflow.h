/*
 * fwiflow.h -- C ABI of the MI355X (gfx950) elastic FWI operator.
 *
 * Drop-in for FwiFlow.jl's `cufd` driver and the three TensorFlow ops built
 * on it (`FwiOp`, `FwiOpGrad`, `FwiObsOp`).  Plain pointers and sizes only;
 * every entry point returns an `int` status (FWIFLOW_OK = 0) and never exits
 * the process.  The text of the last error of the calling thread is returned
 * by fwiflow_last_error().
 *
 * Conventions shared with the reference (deps/CustomOps/FWI/FwiOp.cpp and
 * deps/CustomOps/FWI/Src/libCUFD.cu):
 *   - Lambda and Mu in MPa, Den in kg/m^3, each a row-major [NZ][NX] array of
 *     double (NZ, NX = "nz", "nx" of the parameter file: the padded grid);
 *   - stf is row-major [nShotsTotal][nSteps]; row = shot id (Src_Rec.cu:135);
 *   - shot_ids are 0-based int32, group_size of them;
 *   - grad_Lambda/grad_Mu/grad_Den are [NZ][NX] (dJ/dLambda and dJ/dMu per MPa);
 *   - grad_stf rows 0..group_size-1 are written by group position
 *     (libCUFD.cu:454-456); other rows are not touched;
 *   - the upstream gradient dy of the misfit is not applied (FwiOp.cpp:253-255
 *     never passes it): callers that need dy scale the outputs themselves.
 */
#ifndef FWIFLOW_H_
#define FWIFLOW_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Each replaces an exit(1)/printf path of the reference. */
enum fwiflow_status {
  FWIFLOW_OK = 0,
  FWIFLOW_ERR_CALC_ID = 1,     /* calc_id not 0/1/2 (libCUFD.cu:45-48, Parameter.cpp:141-143) */
  FWIFLOW_ERR_IO = 2,          /* file missing/unreadable (Parameter.cpp:22-25, Src_Rec.cu:26-29,
                                  utilities.cu:12-16) */
  FWIFLOW_ERR_JSON = 3,        /* malformed JSON or missing/ill-typed key (Parameter.cpp:41-113) */
  FWIFLOW_ERR_COURANT = 4,     /* Courant number > 1 (utilities.cu:225-240) */
  FWIFLOW_ERR_SHAPE = 5,       /* shot id / geometry / source or receiver outside the grid */
  FWIFLOW_ERR_UNSUPPORTED = 6, /* band-pass or source-update transform longer than 20480
                                  points (2*nSteps) or with a prime factor above 61; a grid
                                  spacing whose reciprocal division is not exact */
  FWIFLOW_ERR_HIP = 7,         /* HIP runtime error, including "no GPU" */
  FWIFLOW_ERR_ARG = 8,         /* null pointer or invalid argument */
  FWIFLOW_ERR_MEMORY = 9       /* device or pinned-host allocation failed */
};

/*
 * Replaces `void cufd(double* misfit, double* grad_Lambda, double* grad_Mu,
 * double* grad_Den, double* grad_stf, const double* Lambda, const double* Mu,
 * const double* Den, const double* stf, int calc_id, const int gpu_id,
 * int group_size, const int* shot_ids, const std::string para_fname)`
 * (deps/CustomOps/FWI/Src/libCUFD.cu:34-38, declared in FwiOp.h:9-12; bound
 * by ccall in deps/CustomOps/FWI/Obcalc.jl:16-20).
 *   calc_id 0: misfit 0.5*sum(res^2) over the group (libCUFD.cu:528-536);
 *              reads {data_dir}/Shot{id}.bin as observed data.
 *   calc_id 1: gradients (misfit not written); reads the Shot files.
 *   calc_id 2: observed data: writes {data_dir}/Shot{id}.bin, f32 [nrec][nSteps].
 * Output pointers that the calc_id does not use may be NULL.
 */
int fwiflow_cufd(double* misfit, double* grad_Lambda, double* grad_Mu, double* grad_Den,
                 double* grad_stf, const double* Lambda, const double* Mu, const double* Den,
                 const double* stf, int calc_id, int gpu_id, int group_size,
                 const int* shot_ids, const char* para_fname);

/*
 * Device scratch of fwiflow_cufd_noalloc on each GPU, in bytes: a fixed
 * array in the library's device code, held by every GPU the library runs on
 * from the code object's load to the end of the process.
 */
#define FWIFLOW_SCRATCH_BYTES (128ull << 20)

/*
 * fwiflow_cufd without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, files and results as fwiflow_cufd; every device buffer of the
 * call (model, wavefields, seismograms, boundary ring or checkpoints) is
 * carved from the FWIFLOW_SCRATCH_BYTES array of the GPU, and the call makes
 * no hipMalloc, hipFree or graph.  A gradient run whose boundary ring does
 * not fit in what is left keeps checkpoints instead (the interval chosen to
 * fit), so its gradients can differ from fwiflow_cufd's within the FP32
 * tolerance of that scheme.  A call whose buffers exceed the array returns
 * FWIFLOW_ERR_MEMORY naming the shortfall.  Calls on one GPU share its
 * array and its one stream: calls that overlap on one GPU (from several
 * threads) race; the array keeps the last call's contents.
 */
int fwiflow_cufd_noalloc(double* misfit, double* grad_Lambda, double* grad_Mu,
                         double* grad_Den, double* grad_stf, const double* Lambda,
                         const double* Mu, const double* Den, const double* stf, int calc_id,
                         int gpu_id, int group_size, const int* shot_ids,
                         const char* para_fname);

/* Host-only validation of a cufd call: parses the JSON files, checks calc_id,
 * shot ids, source/receiver positions and the Courant number exactly as
 * fwiflow_cufd does before touching a GPU.  Needs no GPU.  No counterpart in
 * the reference (its checks exit the process, see the status codes above). */
int fwiflow_check(int calc_id, int group_size, const int* shot_ids, const char* para_fname,
                  const double* Lambda, const double* Mu, const double* Den);

/* Host tables the GPU path uses, for inspection (no GPU needed):
 * cpml_z: 6 x (NZ-nPad) floats K,a,b,K_half,a_half,b_half (Cpml.cu:7-117),
 * cpml_x: 6 x NX floats, win_data: nSteps floats of w(t)^2 for data
 * (ratio 0.005, libCUFD.cu:63), win_src: nSteps floats of w(t)^2 for the source
 * (ratio 0.001, Src_Rec.cu:140).  Any pointer may be NULL. */
int fwiflow_host_tables(const char* para_fname, float* cpml_z, float* cpml_x, float* win_data,
                        float* win_src);

/* Last error text of the calling thread ("" if none). */
const char* fwiflow_last_error(void);

/* Library version string. */
const char* fwiflow_version(void);

/* ------------------------------------------------------------------------
 * Persistent plan: parse the JSON files once, keep the model, observed data,
 * wavefields and boundary store resident in HBM, run the shots of the group
 * batched (one launch per kernel per time step covers a group of shots sized
 * for the last-level cache).  No counterpart in the reference, which
 * rebuilds everything on each call (libCUFD.cu:51-145) and runs one shot at
 * a time; fwiflow_cufd is implemented on top of it.
 * ---------------------------------------------------------------------- */
typedef struct fwiflow_plan fwiflow_plan;

int fwiflow_plan_create(fwiflow_plan** plan, const char* para_fname, int gpu_id,
                        int group_size, const int* shot_ids);
int fwiflow_plan_destroy(fwiflow_plan* plan);
/* Model in the reference layout (row-major [NZ][NX] double; MPa, MPa, kg/m^3). */
int fwiflow_plan_set_model(fwiflow_plan* plan, const double* Lambda, const double* Mu,
                           const double* Den);
/* Model in the Julia-native layout: column-major (NZ,NX) == z fastest, no
 * transpose; same units. */
int fwiflow_plan_set_model_zfast(fwiflow_plan* plan, const double* Lambda, const double* Mu,
                                 const double* Den);
/* Source time functions [nShotsTotal][nSteps]; row shot_ids[i] feeds shot i. */
int fwiflow_plan_set_stf(fwiflow_plan* plan, const double* stf);
/* Observed data from host memory, concatenated over the group in group order,
 * each shot f32 [nrec][nSteps] (the Shot{id}.bin layout). */
int fwiflow_plan_set_obs(fwiflow_plan* plan, const float* obs);
/* Observed data from {data_dir}/Shot{id}.bin. */
int fwiflow_plan_load_obs(fwiflow_plan* plan);
/* Run calc_id 0, 1 or 2 on the device; results stay resident. */
int fwiflow_plan_run(fwiflow_plan* plan, int calc_id);
/* Misfit 0.5 * sum(res^2) of the last calc_id 0 or 1 run (fwiflow_cufd
 * reports it for calc_id 0 only, as the reference does; a plan also keeps the
 * one its gradient run computed, so an optimiser needs one run per step). */
int fwiflow_plan_misfit(fwiflow_plan* plan, double* misfit);
/* Gradients of the last calc_id 1 run; layouts as in fwiflow_cufd
 * (grad_stf: group_size rows).  Any pointer may be NULL. */
int fwiflow_plan_gradients(fwiflow_plan* plan, double* grad_Lambda, double* grad_Mu,
                           double* grad_Den, double* grad_stf);
/* Synthetic data of the last run, concatenated [nrec][nSteps] per shot. */
int fwiflow_plan_data(fwiflow_plan* plan, float* syn);
/* Write the synthetic data to {data_dir}/Shot{id}.bin (calc_id 2's side effect). */
int fwiflow_plan_write_data(fwiflow_plan* plan);
/* Adjoint source of the last calc_id 0/1 run, every shot [nrec][nSteps] f32,
 * concatenated: the residual after the conditioning stages (the reference's
 * host copy vec_res, libCUFD.cu:315-318, written as Residual_Shot{id}.bin). */
int fwiflow_plan_residual(fwiflow_plan* plan, float* res);
/* Source rows [group_size][nSteps] f32 after the last run: the stf row times
 * its window (Src_Rec.cu:132-143), replaced by the updated signature after a
 * run with "if_src_update" (vec_source, libCUFD.cu:326-329; src_updated{id}.bin). */
int fwiflow_plan_source(fwiflow_plan* plan, float* rows);
/* Geometry of the plan: dims[0..5] = NZ, NX, nSteps, group_size, total receivers,
 * nPml. */
int fwiflow_plan_dims(fwiflow_plan* plan, int* dims);
/* Boundary store of the backward sweep.  A gradient run takes the shots in
 * sweeps: each sweep runs the forward of its shots, their residual and their
 * backward before the next starts, so the store holds one sweep's shots.
 * ckpt_every = 0 (default): automatic.  The reference's store (a 5-layer
 * boundary ring of every step, Boundary.cu:17-41, and time-reversed
 * re-materialisation of the forward wavefield) when one backward shot group's
 * ring fits the memory budget; otherwise checkpoints at the interval k that
 * minimises ring + checkpoint memory.  k > 0: checkpoints every k steps: the
 * full forward state (5 wavefield components and 8 C-PML memory variables)
 * before steps k, 2k, ..., plus the ring of one k-step segment, which the
 * backward sweep recomputes from its checkpoint (one more forward sweep).
 * The reverse sweep then starts every segment from the recomputed state
 * instead of the reversed one.  No reference counterpart (the reference's
 * scheme is the ring of every step for one shot at a time). */
int fwiflow_plan_set_checkpoint(fwiflow_plan* plan, int ckpt_every);
/* Adjoint of the plan's gradient runs.  mode 0 (default): the reference's
 * (el_velocity_adj.cu, el_stress_adj.cu and the sprays of el_stress.cu /
 * el_velocity.cu, as cufd computes them).  mode 1: the exact adjoint, the
 * transpose of the forward time step, whose gradients (lambda, mu, rho and
 * stf) are the derivatives of the discrete misfit also where waves enter the
 * C-PML and in zero-shear layers; the stf gradient then includes the source
 * scale 1500^2 dt and window of add_source.  No counterpart in the reference.
 * Mode 1 runs one plan at a time (a job runs its plans one after another),
 * refuses if_src_update, and holds 8 more arrays per backward-launch shot.
 * FWIFLOW_EXACT_ADJOINT=1 makes it the default of new plans (and of
 * fwiflow_cufd). */
int fwiflow_plan_set_adjoint(fwiflow_plan* plan, int mode);
/* Device bytes a gradient run may hold for the boundary ring and the
 * checkpoints (default 32e9).  It sets the shots per sweep. */
int fwiflow_plan_set_memory_budget(fwiflow_plan* plan, double bytes);
/* Launch shape of the plan's runs: tile width in columns (1-4; a workgroup
 * is 64 cells x tile_x), shots per forward and per backward launch (0: all
 * of the plan's shots, -1: automatic, the Infinity-Cache rule).  Replaces
 * the FWIFLOW_TILE_X / FWIFLOW_SHOT_GROUP(_BWD) defaults; what the autotuner
 * (fwiflow/tune.py) chooses and caches per grid and shot count.  No
 * counterpart in the reference (its launch shape is fixed). */
int fwiflow_plan_set_launch(fwiflow_plan* plan, int tile_x, int shots_fwd, int shots_bwd);
int fwiflow_plan_launch(fwiflow_plan* plan, int* tile_x, int* shots_fwd, int* shots_bwd);
/* The memory plan of the last gradient run: checkpoint interval (0: ring of
 * every step), shots per sweep, device bytes of the ring and of the
 * checkpoints.  Any pointer may be NULL. */
int fwiflow_plan_memory_plan(fwiflow_plan* plan, int* ckpt_every, int* shots_per_sweep,
                             long long* ring_bytes, long long* checkpoint_bytes);
/* Kernel timing with HIP events on the plan's stream (off by default).
 * fwiflow_plan_timing(name) returns the summed milliseconds and launch count
 * of kernel `name` since timing was enabled. */
int fwiflow_plan_set_timing(fwiflow_plan* plan, int enable);
int fwiflow_plan_timing(fwiflow_plan* plan, const char* name, double* total_ms, long long* count);
/* Device bytes held by the plan. */
int fwiflow_plan_device_bytes(fwiflow_plan* plan, long long* bytes);
/* Shots covered by one launch of the time-step kernels of the forward and of
 * the backward sweep (the engine runs the plan's shots in groups of these
 * sizes; each group runs every time step before the next).  Either pointer
 * may be NULL.  No reference counterpart: it sizes per-launch byte counts for
 * kernel timing. */
int fwiflow_plan_shots_per_launch(fwiflow_plan* plan, int* forward, int* backward);
/* Number of shot groups (launches per time step) of the plan's last run in
 * the forward and the backward sweep.  Either pointer may be NULL. */
int fwiflow_plan_launch_groups(fwiflow_plan* plan, int* forward, int* backward);
/* Forward illumination: with enable != 0, later runs also sum, over the
 * plan's shots and the steps it = 0..nSteps-2, vz^2 + vx^2 after each
 * velocity update on the update region [2, NZ-nPad-3] x [2, NX-3] (zero
 * elsewhere).  fwiflow_plan_illumination writes the sum of the last run,
 * row-major [NZ][NX].  No reference counterpart: it is the preconditioner
 * input of the inversion driver (gradient / illumination). */
int fwiflow_plan_set_illumination(fwiflow_plan* plan, int enable);
int fwiflow_plan_illumination(fwiflow_plan* plan, double* illum);

/* ------------------------------------------------------------------------
 * Job: several plans run together on one GPU, e.g. the baseline and
 * monitor surveys of a time-lapse study on grids of different sizes.  Fused,
 * every time-step launch covers a group of shots of every plan (a workgroup
 * map assigns each workgroup a plan and a tile), so the job takes as many
 * launches per step as one plan.  A fused launch holds fewer shots of each
 * plan than the plan's own launches (its groups are sized for the Infinity
 * Cache on the plans' summed bytes), so the job fuses only when that is no
 * fewer than the smallest plan's own group, and otherwise runs the plans one
 * after another (FWIFLOW_JOB_FUSE = 1 / 0 forces either);
 * fwiflow_job_fused reports what the last run did.  The plans must share
 * gpu_id and nSteps;
 * each keeps its own model, data, misfit and gradients, read with the plan
 * calls above after fwiflow_job_run.  No reference counterpart: its callers
 * run one fwi_op per survey (docs/codes/src_fwi_coupled/
 * main_two_phase_flow_inversion.jl:82-92).  At most 31 plans per job; a job
 * does not own its plans, which must outlive it.
 * ---------------------------------------------------------------------- */
typedef struct fwiflow_job fwiflow_job;
int fwiflow_job_create(fwiflow_job** job);
int fwiflow_job_destroy(fwiflow_job* job);
int fwiflow_job_add(fwiflow_job* job, fwiflow_plan* plan);
int fwiflow_job_run(fwiflow_job* job, int calc_id);
int fwiflow_job_fused(fwiflow_job* job, int* fused);

/* ------------------------------------------------------------------------
 * Learned preconditioner network.  No reference counterpart: FwiFlow's plain
 * method is L-BFGS on the raw gradient.  A fixed 4-layer 3x3 convolutional
 * network (4 -> 32 -> 32 -> 32 -> 3 channels, GELU after the first three,
 * zero padding) on one height x width grid, FP32 weights, BF16 activations.
 * Feature maps are planar FP32 [channel][height][width], width fastest (the
 * row-major [NZ][NX] of the operator's arrays, one plane per channel):
 * input 4 planes (cp, cs, rho gradients and the illumination), output 3
 * planes (a search direction for cp, cs, rho).  Parameters are one flat FP32
 * vector, per layer W [Cout][Cin][3][3] then b [Cout].  Training minimises
 * the mask-weighted mean squared error
 *   0.5 / (3 * #{mask != 0}) * sum mask^2 (output - target)^2
 * with Adam (beta 0.9 / 0.999, eps 1e-8). */
typedef struct fwiflow_cnn fwiflow_cnn;
int fwiflow_cnn_create(fwiflow_cnn** net, int gpu_id, int height, int width, unsigned seed);
int fwiflow_cnn_destroy(fwiflow_cnn* net);
int fwiflow_cnn_num_params(fwiflow_cnn* net, int* n);
int fwiflow_cnn_get_params(fwiflow_cnn* net, float* params);
/* also resets the optimiser state */
int fwiflow_cnn_set_params(fwiflow_cnn* net, const float* params);
int fwiflow_cnn_forward(fwiflow_cnn* net, const float* input, float* output);
/* loss and its gradient with respect to the parameters (no update) */
int fwiflow_cnn_gradients(fwiflow_cnn* net, const float* input, const float* target,
                          const float* mask, double* loss, float* grad_params);
/* one Adam step at learning rate lr; loss is the value before the step */
int fwiflow_cnn_train_step(fwiflow_cnn* net, const float* input, const float* target,
                           const float* mask, float lr, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* FWIFLOW_H_ */
