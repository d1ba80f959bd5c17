/*
 * microcosmos_hip.h -- C ABI of the MI355X (gfx950) coupled fluid-filament step.
 *
 * Drop-in boundary for the reference's step API (ProjectProgramAMark/microcosmos):
 * plain pointers and sizes, no framework types.  Device pointers are HIP device
 * memory; `stream` is a hipStream_t passed as void*.  Every function returns
 * an mc_status; mc_last_error() gives a message for the calling thread.
 *
 * Array layouts are the reference's pytrees (structs/nodes.py:5-15,
 * structs/edges.py:6-21, structs/fields.py:5-16), all float32 unless noted:
 *   position, velocity, debug_vector   (N, 2)  x, y interleaved
 *   pairs                              (E, 2)  int32 (src, tgt)
 *   theta, rest_lengths                (E,)
 *   bending_pairs                      (B, 2)  int32 (edge_in, edge_out)
 *   bending_rest_angles, stiffness     (B,)
 *   steric                             (H, W)          Fields.grid_shape = (H, W)
 *   fluid_velocity                     (2, Hf, Wf)     lattice shape = f_grid.shape[1:]
 *   f_grid                             (9, Hf, Wf)     D2Q9 order of fluid.py:15-18
 */
#ifndef MICROCOSMOS_HIP_H
#define MICROCOSMOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mc_status {
  MC_OK = 0,
  MC_ERR_BAD_SHAPE = 1,       /* inconsistent or unsupported sizes */
  MC_ERR_OOM = 2,             /* device allocation failed */
  MC_ERR_NON_FINITE = 3,      /* reserved: non-finite input detected */
  MC_ERR_UNSUPPORTED = 4,     /* option not implemented on the device path */
  MC_ERR_INVALID = 5,         /* null or aliasing argument that is not allowed */
  MC_ERR_HIP = 6              /* HIP runtime error (message has the detail) */
} mc_status;

typedef enum mc_solver { MC_SOLVER_PBD = 0, MC_SOLVER_COSSERAT = 1 } mc_solver;

/* Which rod constraint the step runs (ConstraintSolverConfig.bending_constraint,
 * config.py:12, called at simulate.py:74-77).  The reference runs whatever
 * bending_constraint is and uses solver_type only to decide whether the
 * inertial target y = debug_vector is set to the prediction (simulate.py:
 * 71-72); a Cosserat constraint under solver_type "pbd" therefore sees y = 0.
 * MC_CONSTRAINT_AUTO (0, what a zero-filled struct holds) follows solver_type. */
typedef enum mc_constraint {
  MC_CONSTRAINT_AUTO = 0,
  MC_CONSTRAINT_PBD = 1,       /* pbd_rod_constraint (pbd.py:14-91) */
  MC_CONSTRAINT_COSSERAT = 2   /* stable_cosserat_constraint (stable_cosserat.py:13-119) */
} mc_constraint;

/* Every numeric field of ConstraintSolverConfig (solver/config.py:9-71), in
 * double precision: constants derived from them (omega+, 1/dt^2, 2 sigma^2, ...)
 * are computed in double and rounded to float32 the way the reference's
 * Python-float arithmetic is. */
typedef struct mc_params {
  int32_t solver_type;            /* mc_solver (config.py:28): sets the Cosserat target */
  int32_t cycles_per_step;        /* config.py:15 */
  double stiffness_stretch;       /* config.py:16 */
  double stiffness_shear;         /* config.py:17 */
  double damping;                 /* config.py:20 */
  int32_t enable_fluid;           /* config.py:23 */
  double viscosity;               /* config.py:24 */
  double max_fluid_velocity;      /* config.py:25 */
  double config_dt;               /* config.py:32 (Cosserat inertia) */
  int32_t ibm_iterations;         /* config.py:35 */
  double synthetic_node_width;    /* config.py:36; > 0: synthetic IB points (fluid.py:246-263) */
  int32_t ibm_kernel_size;        /* config.py:37, 1..6 */
  double ibm_relaxation;          /* config.py:38 */
  double node_mass;               /* config.py:45 */
  double lambda_trt;              /* config.py:49 */
  int32_t enable_force_clamp;     /* config.py:59 */
  int32_t enable_steric;          /* config.py:63 */
  double steric_strength;         /* config.py:64 */
  double steric_sigma;            /* config.py:65 */
  int32_t steric_neighbor_skip;   /* config.py:68 */
  double steric_scatter_value;    /* config.py:71 */
  int32_t bending_constraint;     /* mc_constraint (config.py:12) */
} mc_params;

/* Device pointers of one state.  `in` arrays are read, `out` arrays written.
 * Node and edge arrays may alias between in and out; f_grid may not. */
typedef struct mc_state {
  float* position;
  float* velocity;
  float* debug_vector;
  float* theta;
  float* steric;
  float* fluid_velocity;
  float* f_grid;
} mc_state;

/* Per-step edge parameters (device pointers, read only). */
typedef struct mc_edge_params {
  const float* rest_lengths;          /* (E,) */
  const float* bending_rest_angles;   /* (B,) */
  const float* bending_stiffness;     /* (B,) */
} mc_edge_params;

/* Optional in-kernel actuation of the bending rest angles for rollouts:
 * rest_b(t) = amp_b * sin(omega_b * t + phase_b), evaluated in double, with
 * t = t0 + step index.  This is the caller's per-step rewrite of
 * bending_rest_angles (gym/base.py:169-173) moved onto the device. */
typedef struct mc_actuation {
  const float* amp;
  const float* omega;
  const float* phase;
  int64_t t0;
} mc_actuation;

typedef struct mc_plan mc_plan;

/* Library version string. */
const char* mc_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* mc_last_error(void);

/* Builds the device plan for one topology and grid (like an FFT plan): the
 * per-component rod layout (graph.py:156-401 topology, host arrays), FFT
 * twiddles, and all workspace.  grid_h/grid_w = Fields.grid_shape; fluid_h/
 * fluid_w = f_grid.shape[1:]; max_kernel_size bounds ibm_kernel_size.
 * Replaces the jit trace of simulate.step (simulate.py:37-38) for a shape. */
mc_status mc_plan_create(mc_plan** plan, int32_t grid_h, int32_t grid_w, int32_t fluid_h,
                         int32_t fluid_w, int32_t num_nodes, const int32_t* pairs_host,
                         int32_t num_edges, const int32_t* bending_pairs_host, int32_t num_hinges,
                         int32_t max_kernel_size);
/* Batched plan: num_instances independent simulations (each its own periodic
 * grid, lattice and filaments -- one reference `simulate` each) stepped
 * together, every kernel launch covering all instances.  Instance i has
 * instance_shapes[4i..4i+3] = (grid_h, grid_w, fluid_h, fluid_w) and nodes
 * [node_offsets[i], node_offsets[i+1]); pairs / bending pairs use global node
 * / edge ids and never join two instances.  State buffers are the
 * concatenation of the per-instance arrays in instance order: positions and
 * velocities (N, 2); steric (sum of grid_h*grid_w); fluid_velocity and f_grid
 * instance by instance as (2 | 9, fluid_h, fluid_w) blocks.  The steric
 * window (forces.py:66-70) runs over the instance's own nodes.  A plan made by
 * mc_plan_create is the one-instance case.  Replaces the reference's jax.vmap
 * over rollouts (examples/01_es_locomotion.ipynb), for ragged shapes. */
mc_status mc_batch_plan_create(mc_plan** plan, int32_t num_instances,
                               const int32_t* instance_shapes, const int32_t* node_offsets,
                               const int32_t* pairs_host, int32_t num_edges,
                               const int32_t* bending_pairs_host, int32_t num_hinges,
                               int32_t max_kernel_size);
/* Number of instances; offsets (may be NULL) receives per instance
 * (grid cell offset, lattice cell offset, first node) as int64 triples. */
mc_status mc_plan_instances(const mc_plan* plan, int32_t* num_instances, int64_t* offsets);
/* Waits for all work on the device, then frees the plan. */
mc_status mc_plan_destroy(mc_plan* plan);
/* Frees the plan once the work enqueued so far on `stream` has finished,
 * without synchronising the rest of the device: every call that used the plan
 * must have been enqueued on `stream` (or have completed).  Plan memory comes
 * from a per-device stream-ordered pool and returns to it.  No counterpart in
 * the reference (jit traces are freed by the JAX runtime); it lets a request
 * server retire plans while other requests keep running. */
mc_status mc_plan_destroy_after(mc_plan* plan, void* stream);
/* Waits for `stream`, then reports whether the plan's work so far completed
 * as specified.  The wide immersed-boundary solve (clusters of more than 768
 * nodes) synchronises its workgroups with grid barriers whose waits are
 * bounded; if a wait gives up (its workgroups were not all resident, e.g.
 * under heavy concurrent use of the device), the steps' node velocities and
 * IB forces are invalid.  That is returned as MC_ERR_HIP, here and by the
 * next mc_step / mc_rollout / mc_batch_rollout / mc_batch_advance /
 * mc_ib_lbm_step / mc_track_read / mc_plan_destroy_after on the plan, once,
 * and the plan's IB state is restored so later calls compute normally.  No
 * counterpart in the reference (XLA has no device-side failure mode here).
 * MC_IB_WIDE_POLL_LIMIT (environment, read at plan creation) sets the poll
 * bound of those waits (default 2^22). */
mc_status mc_plan_status(mc_plan* plan, void* stream);
/* Device bytes held by the plan (topology + workspace). */
mc_status mc_plan_workspace_bytes(const mc_plan* plan, size_t* bytes);

/* Autotuning of a plan's launch choices (no counterpart in the reference:
 * jax's compiler picks its own): times `steps` coupled steps, on a scratch
 * copy of `state` (allocated and freed here; state itself is not changed),
 * for each candidate -- the cluster size above which the multi-workgroup
 * immersed-boundary solve takes over (768 or 1024 nodes, when the plan has
 * that solve and a cluster of the timed steps is that large) and where each step's LBM pass runs (see
 * mc_plan_tuning) -- and keeps the fastest for the plan's later
 * rollouts.  Every candidate computes the same result bit for bit.  Choices
 * are cached per (instances, physics) for the process: a later plan of the
 * same shape takes them without timing.  Waits for `stream`. */
mc_status mc_plan_autotune(mc_plan* plan, const mc_params* params, const mc_state* state,
                           const mc_edge_params* edges, double dt, int32_t steps, void* stream);
/* The plan's current choices: the wide-solve cluster size, and where a
 * rollout runs each step's LBM pass: 0 a launch of its own after the step's
 * IB, 1 inside the next step's rod launch, 2 on a second stream beside the
 * next step's steric and rod launches. */
mc_status mc_plan_tuning(const mc_plan* plan, int32_t* wide_min, int32_t* lbm_mode);
/* Sets those choices by hand (wide_min 64 .. 1024, taken only when the plan
 * has the wide solve; lbm_mode 0, 1 or 2).  No counterpart in the reference. */
mc_status mc_plan_set_tuning(mc_plan* plan, int32_t wide_min, int32_t lbm_mode);

/* One coupled step: simulate.step (simulate.py:37-100).
 * in -> out with dt; f_grid of out must differ from in. */
mc_status mc_step(mc_plan* plan, const mc_params* params, const mc_state* in, const mc_state* out,
                  const mc_edge_params* edges, double dt, void* stream);

/* num_steps coupled steps in place on `state` (f_grid ping-pongs through the
 * plan's second lattice buffer and ends in state->f_grid).  simulate.py:15-35.
 * act may be NULL (then edges->bending_rest_angles is used every step).
 * record_position (num_steps, N, 2) and record_fluid_velocity
 * (num_steps, 2, Hf, Wf) may be NULL; when given they receive the per-step
 * trajectory that simulate() returns.  use_graph != 0 (and no recording)
 * captures 8 steps in one HIP graph and replays it: within the graph each
 * step's LBM collide-stream runs inside the next step's rod launch, so every
 * kernel is enqueued on one stream; remaining steps run unfused.  The graph
 * stays instantiated in the plan and later calls with the same arguments
 * replay it; the call returns without waiting for the steps. */
mc_status mc_rollout(mc_plan* plan, const mc_params* params, const mc_state* state,
                     const mc_edge_params* edges, const mc_actuation* act, double dt,
                     int64_t num_steps, float* record_position, float* record_fluid_velocity,
                     int32_t use_graph, void* stream);

/* ES evaluation of every instance of a (batched) plan: num_steps coupled
 * steps in place, graph-replayed, with the in-kernel actuation `act` (may be
 * NULL).  displacement_out (num_instances, 2), device, may be NULL: the
 * cumulative minimum-image displacement of each instance's periodic centre
 * of mass over the rollout (swim_base.py:167-209), tracked per step on the
 * device.  No host synchronisation per step. */
mc_status mc_batch_rollout(mc_plan* plan, const mc_params* params, const mc_state* state,
                           const mc_edge_params* edges, const mc_actuation* act, double dt,
                           int64_t num_steps, float* displacement_out, void* stream);

/* Trajectory of simulate() (simulate.py:15-35), recorded on the device:
 * after every stride-th step one frame of each non-NULL array, frame
 * f = (step + 1) / stride - 1, F = num_steps / stride frames:
 *   position, velocity, debug_vector  (F, N, 2)
 *   steric                            (F, H, W)
 *   fluid_velocity                    (F, 2, H, W): the lattice velocity,
 *       linearly resized to the grid as jax.image.resize(..., "linear")
 *       does when the lattice and the grid differ (simulate.py:24-28).
 * The frame index is a device counter, so the recording is captured inside
 * the replayed rollout graphs (use_graph != 0).  Single-instance plans. */
typedef struct mc_record {
  float* position;
  float* velocity;
  float* debug_vector;
  float* steric;
  float* fluid_velocity;
  int64_t stride;
} mc_record;
mc_status mc_rollout_record(mc_plan* plan, const mc_params* params, const mc_state* state,
                            const mc_edge_params* edges, const mc_actuation* act, double dt,
                            int64_t num_steps, const mc_record* record, int32_t use_graph,
                            void* stream);

/* Per-step actions of the gym step (gym/base.py:163-176): before step k of
 * the rollout, rest_lengths = edges->rest_lengths + d_rest_length[k] and
 * bending_rest_angles = edges->bending_rest_angles + d_bending_angle[k]
 * (float32 adds), the rewrite the reference's env step makes from the base
 * values and an arbitrary controller's action.  Rows are device arrays
 * (num_rows, E) and (num_rows, B) over the plan's whole edge and hinge
 * numbering (every instance of a batched plan); either may be NULL (a zero
 * action); num_rows >= num_steps. */
typedef struct mc_actions {
  const float* d_rest_length;
  const float* d_bending_angle;
  int64_t num_rows;
} mc_actions;
/* num_steps coupled steps in place on every instance of the plan, graph-
 * replayed, step k taking action row k; displacement_out as in
 * mc_batch_rollout (may be NULL).  Stream-ordered. */
mc_status mc_rollout_actions(mc_plan* plan, const mc_params* params, const mc_state* state,
                             const mc_edge_params* edges, const mc_actions* actions, double dt,
                             int64_t num_steps, float* displacement_out, void* stream);

/* Continuous batching: mc_batch_rollout for instances that each run on their
 * own clock.  Instance i's actuation uses t = act->t0 + instance_t0[i] + step
 * index (instance_t0: host, num_instances entries, may be NULL), so
 * instances that joined the batch at different points of their own rollouts
 * keep their own actuation phase.  displacement_io (num_instances, 2),
 * device, may be NULL: on entry each instance's cumulative centre-of-mass
 * displacement so far, on return that value advanced over these steps
 * (swim_base.py:179-209) -- the same sum, in the same order, as one
 * uninterrupted rollout, so a rollout split over any sequence of calls and
 * batches gives bit-identical results.  Stream-ordered: returns without
 * waiting for the steps (the S-step graph is instantiated once and replayed
 * by later calls with the same arguments).  Replaces the per-request jit
 * calls of the reference's ES callers (gym/base.py:163-196). */
mc_status mc_batch_advance(mc_plan* plan, const mc_params* params, const mc_state* state,
                           const mc_edge_params* edges, const mc_actuation* act,
                           const int64_t* instance_t0, double dt, int64_t num_steps,
                           float* displacement_io, void* stream);

/* Differentiation of the step: the configurations the reference
 * differentiates with jax.value_and_grad through lax.scan of simulate.step
 * (experiments/filament_folding/experiment.py:985-1018: the rod-only step;
 * tests/test_differentiability.py:51-71: PBD_SCHEME, the fluid-coupled step).
 *
 * mc_rod_adjoint_forward runs num_steps steps in place on `state` (as
 * mc_rollout without actuation: the rest angles edges->bending_rest_angles
 * are the parameters, constant over the rollout) and keeps a checkpoint of
 * (position, velocity, theta) every checkpoint_every steps in the plan.
 *
 * mc_rod_adjoint_backward then takes position_bar = dL/d position after the
 * last step (N, 2) and velocity_bar = dL/d velocity (may be NULL: zero), and
 * writes rest_angle_bar (B) = dL/d bending_rest_angles, plus (each may be
 * NULL) the gradient with respect to the initial position, velocity and
 * theta.  The PBD and stable Cosserat constraints are covered, with or
 * without the steric term, and with or without the fluid.  With the fluid
 * (single-instance plans; MC_ERR_UNSUPPORTED for batched ones), as with the
 * steric term, every step's state is kept, and its input lattice too (9 Hf Wf
 * floats per step); the forward runs the coupled steps as mc_rollout does
 * (bit-identical), and each reverse step recomputes the rod half from its
 * checkpoint, then runs the adjoint of the TRT/Guo collide-and-stream, the
 * drag, the force and post clamps and the multi-direct forcing iterations
 * (the IB positions held constant: the reference's stop_gradient,
 * fluid.py:239; the node velocities carry gradient), giving the cotangents
 * of the step's input lattice and of the rod's output velocities, then the
 * rod half's adjoint.  Without it the
 * filaments are independent: one workgroup per filament runs a whole
 * checkpoint segment per launch, forward (bit-identical to mc_rollout) or
 * reverse (the segment recomputed from its checkpoint, then the hand-written
 * adjoint of each step backwards).  With it the filaments couple through the
 * field every step: every step's state is kept (checkpoint_every is taken as
 * 1; memory (4 N + E) floats per step), the forward runs the coupled step's
 * steric pipeline each step (bit-identical to mc_rollout, field included) and
 * keeps each step's field, force and prediction (grid cells + 4 N floats per
 * step, up to 2 GiB -- the environment variable MC_ADJ_STERIC_TAPE_MAX sets
 * the limit in bytes; beyond it the reverse pass recomputes them,
 * bit-identically), and each reverse step runs the adjoint of the
 * constraint, the steric force (deposit, Gaussian filter, samples, window)
 * and the prediction.  Every filament must fit the LDS budget
 * (MC_ERR_UNSUPPORTED otherwise).
 * params, dt and the edge arrays must be those of the recorded rollout.
 * Minimum-image, angle and periodic wraps are differentiated as identity. */
mc_status mc_rod_adjoint_forward(mc_plan* plan, const mc_params* params, const mc_state* state,
                                 const mc_edge_params* edges, double dt, int64_t num_steps,
                                 int32_t checkpoint_every, void* stream);
mc_status mc_rod_adjoint_backward(mc_plan* plan, const mc_params* params,
                                  const mc_edge_params* edges, double dt,
                                  const float* position_bar, const float* velocity_bar,
                                  float* rest_angle_bar, float* position0_bar,
                                  float* velocity0_bar, float* theta0_bar, void* stream);
/* The recorded state after `step` steps (a multiple of the checkpoint
 * interval below num_steps): position and velocity (N, 2), theta (E); each
 * output may be NULL.  The frames of a multi-frame loss are read here. */
mc_status mc_rod_adjoint_state(mc_plan* plan, int64_t step, float* position_out,
                               float* velocity_out, float* theta_out, void* stream);
/* The same with position cotangents on several frames -- a loss over the
 * positions after frame_steps[f] steps, as the folding experiment's shape and
 * stability terms over evenly spaced frames (experiment.py:166-330):
 * frame_position_bars holds num_frames (N, 2) arrays (device); every frame
 * step lies in 1 .. num_steps and is a multiple of the recorded checkpoint
 * interval or num_steps itself (MC_ERR_BAD_SHAPE otherwise).  velocity_bar is
 * the cotangent of the final velocity (may be NULL). */
mc_status mc_rod_adjoint_backward_frames(mc_plan* plan, const mc_params* params,
                                         const mc_edge_params* edges, double dt,
                                         int32_t num_frames, const int64_t* frame_steps,
                                         const float* frame_position_bars,
                                         const float* velocity_bar, float* rest_angle_bar,
                                         float* position0_bar, float* velocity0_bar,
                                         float* theta0_bar, void* stream);
/* As mc_rod_adjoint_backward_frames, and also d loss / d rest lengths (E,
 * device, required): the folding experiment optimises the rest lengths with
 * the rest angles (experiments/filament_folding/experiment.py:739-745, its
 * rest_lengths_raw).  The rest lengths enter the position pass (PBD: the
 * stretch and the shear target, pbd.py:55-89; stable Cosserat: the edge force
 * and the diagonal h, stable_cosserat.py:94-117); with the stable Cosserat
 * solver the reverse pass records each cycle's node update (checkpoint
 * interval x cycles x N float2 of device memory, allocated on first use). */
mc_status mc_rod_adjoint_backward_params(mc_plan* plan, const mc_params* params,
                                         const mc_edge_params* edges, double dt,
                                         int32_t num_frames, const int64_t* frame_steps,
                                         const float* frame_position_bars,
                                         const float* velocity_bar, float* rest_angle_bar,
                                         float* rest_length_bar, float* position0_bar,
                                         float* velocity0_bar, float* theta0_bar, void* stream);

/* The direct filter's taps along an axis of n cells for sigma (host only, no
 * device work; no counterpart in the reference -- it exposes the plan's
 * tabulation of utils.py:41-47's filter for checking): returns the tap count
 * and writes the first offset to *lo and up to `capacity` float32 taps, 0
 * when the axis needs the FFT (more than 129 taps), -1 on a bad argument. */
int32_t mc_filter_taps(int32_t n, double sigma, int32_t* lo, float* taps, int32_t capacity);

/* The kernels' float32 math routines evaluated on the host (no device work;
 * no counterpart in the reference -- it exposes the routines that replace
 * jnp.sin / jnp.cos / jnp.exp in the filament cycles (pbd.py:63-64,
 * stable_cosserat.py:111-112), the IB weights (fluid.py:31-38) and the steric
 * window (forces.py:72-75), for checking): fn 0 writes sin(x[i]) to out0 and
 * cos(x[i]) to out1, fn 1 writes exp(x[i]) to out0.  Returns MC_OK or
 * MC_ERR_INVALID. */
int32_t mc_math_f32(int32_t fn, const float* x, int64_t n, float* out0, float* out1);

/* The linear-resize taps of mc_rollout_record along one axis (host only, no
 * device work; no counterpart in the reference -- it exposes the tabulation
 * of jax.image.resize's linear weights, simulate.py:24-28, for checking):
 * writes n_out + 1 CSR offsets and up to `capacity` (input index, weight)
 * taps; returns the tap count, -1 on a bad argument. */
int32_t mc_resize_taps(int32_t n_in, int32_t n_out, int32_t* offsets, int32_t* index,
                       float* weight, int32_t capacity);

/* Folding losses of the filament-folding experiment between particles (n, 2)
 * and targets (m, 2), float32 device arrays, and their gradient with respect
 * to the particles (grad (n, 2), may be NULL); loss is one device float:
 *   MC_LOSS_CHAMFER     experiments/filament_folding/experiment.py:57-63
 *   MC_LOSS_DCD         density-aware Chamfer, experiment.py:105-128
 *   MC_LOSS_SMOOTH_MAX  logsumexp(alpha sqrt(min d^2 + 1e-8)) / alpha, :139-149
 * The derivative of a minimum over tied entries is split equally among them
 * (jax.grad's rule); the DCD counts carry no derivative (stop_gradient).
 * workspace: mc_point_loss_workspace(n, m) bytes of device memory.
 * Stream-ordered; no host synchronisation. */
typedef enum mc_loss_kind {
  MC_LOSS_CHAMFER = 0,
  MC_LOSS_DCD = 1,
  MC_LOSS_SMOOTH_MAX = 2
} mc_loss_kind;
size_t mc_point_loss_workspace(int64_t n, int64_t m);
mc_status mc_point_loss(int32_t kind, const float* particles, int64_t n, const float* targets,
                        int64_t m, double alpha, float* loss, float* grad, void* workspace,
                        void* stream);

/* mc_point_loss without a workspace argument and without allocating: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a static device array that each device
 * holds until the process ends (2.25 MiB); n and m are each at most
 * MC_POINT_LOSS_STATIC_MAX, larger counts return MC_ERR_BAD_SHAPE.  Every call
 * on one device shares that device's array: calls that overlap on one device
 * (other streams, other threads) race, so order them on one stream.  Results
 * are bit-identical to mc_point_loss's. */
#define MC_POINT_LOSS_STATIC_MAX 65536
mc_status mc_point_loss_static(int32_t kind, const float* particles, int64_t n,
                               const float* targets, int64_t m, double alpha, float* loss,
                               float* grad, void* stream);

/* Stability term of the folding loss (experiments/filament_folding/
 * experiment.py:878-893, `_frame_stab` averaged over the stability frames):
 *   loss = mean_s [ mean(d(frames[s], main)^2) + mean(d(frames[s], anchor)^2) ]
 * with d the minimum-image displacement on the height x width torus
 * (utils.py:18-23) and each mean over the (n, 2) array.  frames (num_frames,
 * n, 2), main (n, 2) and anchor (n, 2) are float32 device arrays; anchor may
 * be NULL (the reference's use_anchor false: that term is left out).  With
 * frames_grad non-NULL, writes d loss / d frames (num_frames, n, 2), d loss /
 * d main (n, 2) and, when anchor is given, d loss / d anchor (n, 2); a caller
 * whose main or anchor frame is also a stability frame adds the gradients.
 * loss is one device float.  Stream-ordered; no host synchronisation. */
mc_status mc_frame_stability(const float* frames, int32_t num_frames, int64_t n,
                             const float* main_pos, const float* anchor_pos, int32_t height,
                             int32_t width, float* loss, float* frames_grad, float* main_grad,
                             float* anchor_grad, void* stream);

/* Measurement: num_steps coupled steps in place (no graph, the LBM
 * collide-stream as its own launch) with HIP events carried by every LBM
 * dispatch and recorded after every step on `stream`; returns the average
 * LBM kernel duration and step duration in ms.  No counterpart in the
 * reference (its only timer is utils.py:53-59). */
mc_status mc_profile_step(mc_plan* plan, const mc_params* params, const mc_state* state,
                          const mc_edge_params* edges, const mc_actuation* act, double dt,
                          int64_t num_steps, double* lbm_ms_avg, double* step_ms_avg, void* stream);

/* Measurement of the launch that runs in the rollout graphs: num_steps (even,
 * >= 2) coupled steps in place with each step's LBM pass carried by the next
 * step's rod launch, as in those graphs, and HIP events carried by each such
 * rod launch; returns the average duration of those launches (filament solve
 * plus the previous step's LBM pass) and of a step in ms.  Needs a fluid
 * scenario whose plan fuses the two (MC_ERR_UNSUPPORTED otherwise).  No
 * counterpart in the reference. */
mc_status mc_profile_fused_step(mc_plan* plan, const mc_params* params, const mc_state* state,
                                const mc_edge_params* edges, const mc_actuation* act, double dt,
                                int64_t num_steps, double* kernel_ms_avg, double* step_ms_avg,
                                void* stream);

/* Sub-operators of the step, each a drop-in for one reference function. */

/* utils.py:35-49 gaussian_filter on an (H, W) field (H = grid_h, W = grid_w).
 * The filter is the circular convolution with the exact periodic kernel of
 * each axis when both kernels fit 129 taps (their dropped tails below 2^-27
 * of their mass; sigma 2: 25 taps), else an FFT.  MC_ERR_UNSUPPORTED when the
 * kernel is long and a side has a prime factor above 13 (no FFT plan); the
 * same holds for every entry point with the steric term on. */
mc_status mc_gaussian_filter(mc_plan* plan, const float* data, float* out, double sigma,
                             void* stream);

/* forces.py:41-81 compute_field_steric_force_corrected: steric (H, W) field and
 * per-node force (N, 2) at `position` (N, 2).  density_out (H, W) may be NULL
 * (the CIC deposit before filtering, forces.py:112-115). */
mc_status mc_steric_field(mc_plan* plan, const mc_params* params, const float* position,
                          float* steric_out, float* force_out, float* density_out, void* stream);

/* fluid.py:202-386 immersed_boundary_interaction: returns node velocity,
 * fluid_velocity (u*) and streamed f.  pos_out receives the wrapped positions
 * (simulate.py:97); it may be NULL.  With params->synthetic_node_width > 0 the
 * IB runs on each node's two synthetic points (fluid.py:246-263) and the drag
 * is the mean of their forces (fluid.py:364-368). */
mc_status mc_ib_lbm_step(mc_plan* plan, const mc_params* params, const float* position,
                         const float* velocity, const float* f_in, float* velocity_out,
                         float* fluid_velocity_out, float* f_out, float* pos_out, double dt,
                         void* stream);

/* cycles x (pbd_rod_constraint | stable_cosserat_constraint) (pbd.py:14-91,
 * stable_cosserat.py:13-119) on positions / thetas, in place.  `y` is the
 * Cosserat inertial target (debug_vector), may be NULL for PBD. */
mc_status mc_rod_solve(mc_plan* plan, const mc_params* params, float* position, float* theta,
                       const float* y, const mc_edge_params* edges, int32_t cycles, void* stream);

/* Swimmer fitness (the caller of the step).  Centre of mass of node groups
 * [group_offsets[g], group_offsets[g+1]) (device int32, n_groups + 1):
 * periodic != 0 -> reference-node minimum-image mean, wrapped
 * (experiments/swim_sweep/swimmers/swim_base.py:167-177 _periodic_com);
 * periodic == 0 -> plain mean (examples/01_es_locomotion.ipynb
 * rollout_fitness, position[s].mean(axis=0)).  com_out is (n_groups, 2). */
mc_status mc_center_of_mass(const float* position, int32_t n_groups, const int32_t* group_offsets,
                            int32_t grid_h, int32_t grid_w, int32_t periodic, float* com_out,
                            void* stream);

/* Per-step displacement tracking on a plan (swim_base.py:179-209
 * _init_displacement / _update_displacement): once enabled, every step that
 * mc_step / mc_rollout enqueues ends with, per group, com = periodic centre
 * of the new positions; count += 1; if count > start_step then
 * cumulative += displacement(com, prev); prev = com.  group_offsets_host
 * (n_groups + 1) are host int32 node offsets; n_groups = 0 turns tracking
 * off.  No host synchronisation per step. */
mc_status mc_track_enable(mc_plan* plan, int32_t n_groups, const int32_t* group_offsets_host,
                          int32_t start_step);
/* _init_displacement: prev = centre of `position`, cumulative = 0, count = 0. */
mc_status mc_track_reset(mc_plan* plan, const float* position, void* stream);
/* Copies cumulative displacement and the last centre, (n_groups, 2) each,
 * into device buffers (either may be NULL), stream-ordered. */
mc_status mc_track_read(mc_plan* plan, float* cumulative_out, float* com_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MICROCOSMOS_HIP_H */
