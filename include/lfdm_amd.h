/* C ABI of the MI355X batched half-implicit vehicle solver (11/18/24-DOF, TMeasy / TMeasyNr).
 *
 * Drop-in boundary for the reference's GPU solver classes d18SolverHalfImplicitGPU
 * (wheeled_vehicle_models/18dof-gpu/dof18_halfImplicit_gpu.cuh:15-343) and their 11/24-DOF twins.
 * One opaque handle replaces one solver object; every call returns a status code (LFDM_OK or a
 * negative LFDM_E_*) and lfdm_last_error() describes the last failure on the calling thread.
 * Plain pointers and sizes only: host buffers unless a name says "device".
 *
 * The 11- and 24-DOF classes are d11SolverHalfImplicitGPU (11dof-gpu/dof11_halfImplicit_gpu.cuh:16)
 * and d24SolverHalfImplicitGPU (24dof-gpu/dof24_halfImplicit_gpu.cuh:15); their entry points are
 * the ones cited below for the 18-DOF class at the corresponding lines of those headers.
 *
 * Records: a vehicle's full state is exchanged as a flat array of doubles in a fixed order
 * (lfdm_record_len gives its length):
 *   18-DOF, per tire LF RF LR RR, TMeasy  : xe ye xedot yedot omega dOmega xt rStat fx fy fz vsx vsy My engTor
 *                                 TMeasyNr: omega dOmega xt rStat fx fy fz vsx vsy My engTor
 *           then the vehicle              : x y u v psi wz phi wx udot vdot wxdot wzdot crankOmega dOmega_crank gear
 *   11-DOF, per tire F R (fields as 18-DOF), then x y u v psi wz udot vdot wzdot crankOmega dOmega_crank gear
 *   24-DOF, per tire LF RF LR RR, TMeasy  : xe ye xedot yedot omega dOmega xt dxt rStat fx fy fz fxg fyg fzg
 *                                           vgx vgy vgz vsx vsy vsz My engTor (TMeasyNr without xe..yedot)
 *           per strut LF RF LR RR         : xs xsi dxs ls lsi us vs ws dus dvs dws uu vu wu duu dvu dwu fxs fys
 *                                           fzs fzd mx my mz
 *           then the vehicle              : x y z u v w psi wz phi wx theta wy udot vdot wdot wxdot wydot wzdot
 *                                           dtheta dphi dpsi crankOmega dOmega_crank gear
 * The gear index is stored as a double.
 */
#ifndef LFDM_AMD_H
#define LFDM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define LFDM_OK 0
#define LFDM_E_ARG (-1)         /* invalid argument or handle */
#define LFDM_E_IO (-2)          /* file cannot be opened or holds no data */
#define LFDM_E_PARSE (-3)       /* malformed JSON or missing / mistyped key */
#define LFDM_E_STATE (-4)       /* call out of order (e.g. Solve before SetEndTime) */
#define LFDM_E_CAPACITY (-5)    /* more vehicles than the handle was created for, or a map too large */
#define LFDM_E_HIP (-6)         /* HIP runtime error (no GPU, launch failure, out of memory) */
#define LFDM_E_UNSUPPORTED (-7) /* model / option not available */

#define LFDM_TIRE_TMEASY 0   /* TireType::TMeasy   (dof18_gpu.cuh, enum TireType) */
#define LFDM_TIRE_TMEASYNR 1 /* TireType::TMeasyNr */

typedef struct lfdm_solver lfdm_solver;

/* Text of the last error on this thread ("" if none). */
const char* lfdm_last_error(void);

/* Length of the full state record of one vehicle. */
int lfdm_record_len(int dof, int tire_type);

/* ---- setup helpers (host only, no GPU needed) ------------------------------------------ */

/* Parses vehicle + tire (+ suspension for 24-DOF) JSON into the flat parameter vector the
 * handle and the fixtures exchange; replaces setVehParamsJSON / setTireParamsJSON / tireInit
 * (18dof-gpu/dof18_gpu.cu:1075-1460).  *len receives the length; fails with LFDM_E_CAPACITY if
 * cap is too small. */
int lfdm_params_from_json(int dof, const char* veh_json, const char* tire_json, const char* sus_json, int tire_type,
                          double* out, int cap, int* len);

/* Loads a driver-input table "time steering throttle braking" (LoadDriverData, utils_gpu.cu:7-31)
 * into rows[nrows][4]. */
int lfdm_driver_file_load(const char* path, double* rows, int cap_rows, int* nrows);

/* ---- solver handle: the reference class API ------------------------------------------- */

/* d18SolverHalfImplicitGPU(unsigned total_num_vehicles), dof18_halfImplicit_gpu.cuh:19 */
int lfdm_create(int dof, unsigned total_vehicles, lfdm_solver** out);
/* ~d18SolverHalfImplicitGPU(), .cuh:20 */
int lfdm_destroy(lfdm_solver* h);

/* Construct(veh, tire, n[, driver][, TireType]), .cuh:34-87; sus_json only for 24-DOF; driver_file
 * may be NULL (controller-driven fleets).  Appends a group of n vehicles. */
int lfdm_construct(lfdm_solver* h, const char* veh_json, const char* tire_json, const char* sus_json,
                   unsigned n, const char* driver_file, int tire_type);
/* Same from a flat parameter vector and an in-memory driver table rows[nrows][4] (may be NULL). */
int lfdm_construct_flat(lfdm_solver* h, const double* flat, int len, int tire_type, unsigned n,
                        const double* rows, int nrows);

/* Initialize(states..., n), .cuh:172-195: appends n vehicles with the given full record.  For the
 * 24-DOF model each appended record then goes through initializeTireSus as in
 * dof24_halfImplicit_gpu.cu:320-333 (static tire loads and compressions, strut preload and lengths,
 * unsprung heave velocity), using the vehicle's own mass from lfdm_set_overrides when that was
 * called before Initialize.  Records given to lfdm_set_states are taken as they are. */
int lfdm_initialize(lfdm_solver* h, const double* record, unsigned n);
/* Bulk form of Initialize (no reference counterpart; the reference appends one shared state per
 * call): appends n vehicles with their own records records[n][record_len], same 24-DOF rule. */
int lfdm_initialize_batch(lfdm_solver* h, const double* records, unsigned n);

/* Setters / getters, .cuh:93-129 */
int lfdm_set_end_time(lfdm_solver* h, double tend);
int lfdm_set_time_step(lfdm_solver* h, double step);
int lfdm_set_kernel_sim_time(lfdm_solver* h, double time);
int lfdm_set_host_dump_time(lfdm_solver* h, double time);
int lfdm_set_threads_per_block(lfdm_solver* h, unsigned threads);
int lfdm_get_step(lfdm_solver* h, double* step);
int lfdm_get_end_time(lfdm_solver* h, double* tend);
int lfdm_get_kernel_sim_time(lfdm_solver* h, double* time);
int lfdm_get_tire_type(lfdm_solver* h, int* tire_type);

/* SetOutput(file, freq, store_all, no_outs), .cuh:138-141: CSV files <file>_<vehicle>.csv */
int lfdm_set_output(lfdm_solver* h, const char* output_file, double output_freq, int store_all, unsigned no_outs);

/* Solve(), .cuh:148: integrates every vehicle to the end time; returns after the GPU finished. */
int lfdm_solve(lfdm_solver* h);

/* SolveStep(t, steering, throttle, braking), .cuh:158: one kernel-sim-time chunk with uniform
 * controls; asynchronous; *t_out = t + kernel_sim_time. */
int lfdm_solve_step(lfdm_solver* h, double t, double steering, double throttle, double braking, double* t_out);

/* GetSimState(idx), .cuh:204: full record of one vehicle (synchronizes). */
int lfdm_get_sim_state(lfdm_solver* h, unsigned vehicle, double* record);

/* ---- batched extensions (no counterpart in the reference) ------------------------------ */

/* All records at once, out[total][record] (synchronizes). */
int lfdm_get_states(lfdm_solver* h, double* records);
/* Replaces the records of all vehicles, in[total][record]. */
int lfdm_set_states(lfdm_solver* h, const double* records);

/* Per-vehicle parameter values ovr[7][total]: m, jz, jx, dfy0Pn, dfy0P2n, torque_scale (multiplies
 * the engine torque-map ordinates), maxBrakeTorque.  Defaults come from each vehicle's group.  A pending
 * adjoint (lfdm_adjoint_forward without its backward yet) reads them again, so changing them cancels it. */
int lfdm_set_overrides(lfdm_solver* h, const double* ovr);
/* A vehicle's parameters in full, as the flat vector of lfdm_params_from_json: the reference keeps one
 * VehicleParam / TMeasyParam copy per vehicle in SimData (dof18_halfImplicit_gpu.cuh:209-215,
 * dof18_gpu.cuh:821-838) and lets users edit any field of any vehicle.  Here vehicles with equal parameters
 * share one parameter set (so a wavefront never mixes sets).  lfdm_get_vehicle_params returns the
 * vehicle's set with its per-vehicle parameters applied (m, jz, jx, dfy0Pn, dfy0P2n, maxBrakeTorque, and
 * the torque map's ordinates times the torque scale); lfdm_set_vehicle_params gives the vehicle this
 * vector as its parameters: it joins the set with exactly these values and its own driver table, or
 * gets a set of its own, its per-vehicle parameters become the vector's, and the slot layout is rebuilt
 * at the next launch (records kept; a pending adjoint is cancelled).  Many vehicles with distinct sets
 * cost one wavefront each. */
int lfdm_get_vehicle_params(lfdm_solver* h, unsigned vehicle, double* flat, int cap, int* len);
int lfdm_set_vehicle_params(lfdm_solver* h, unsigned vehicle, const double* flat, int len);
/* Device-pointer forms of lfdm_set_overrides and lfdm_set_states (device_ovr[7][total],
 * device_records[total][record_len]): copied on the handle's stream without synchronizing and without
 * passing through the host (e.g. from torch tensors every optimization step); values are not checked. */
int lfdm_set_overrides_device(lfdm_solver* h, const double* device_ovr);
int lfdm_set_states_device(lfdm_solver* h, const double* device_records);

/* Control sources for Solve (the default is each group's driver table):
 *  knots: steering linear on a grid of ds seconds (ns knots), throttle and braking piecewise
 *         constant on a grid of dp seconds (np values each); knots[ns + 2 np][total].
 *  steps: one (steering, throttle, braking) per step, ctrl[nsteps][3][total]; host or device. */
int lfdm_set_controls_knots(lfdm_solver* h, int ns, double ds, int np, double dp, const double* knots);
int lfdm_set_controls_steps(lfdm_solver* h, int nsteps, const double* ctrl);
int lfdm_set_controls_steps_device(lfdm_solver* h, int nsteps, const double* device_ctrl);
int lfdm_set_controls_table(lfdm_solver* h);
/* Per-step controls shared by every vehicle, device_ctrl[nsteps][3] (steering, throttle, braking), e.g.
 * a controller's command for a whole fleet. */
int lfdm_set_controls_shared_device(lfdm_solver* h, int nsteps, const double* device_ctrl);
/* SolveStep(t, functor) of the reference (dof18_halfImplicit_gpu.cuh:233-290, kernels :454-640): the
 * controls functor runs on every vehicle's lane at every step, so the integration kernel is
 * instantiated on the functor's type in the caller's translation unit (include/lfdm_amd_functor.hpp).
 * lfdm_functor_chunk_begin prepares the handle's device image, sets its time base to t and writes the
 * launch record of nsteps steps (record_size: the caller's sizeof of the library's record type, which
 * must match), the kernel's block and grid sizes and the handle's stream; the caller launches its kernel
 * on that stream, then calls lfdm_functor_chunk_done(h, nsteps), which advances the time base.  No
 * synchronization; refused while CSV output is set (the reference exits in that case). */
int lfdm_functor_chunk_begin(lfdm_solver* h, double t, int nsteps, void* record, int record_size, int* block,
                             int* grid, void** stream);
int lfdm_functor_chunk_done(lfdm_solver* h, int nsteps);
/* Number of steps one kernel-sim-time chunk starting at t takes (the reference kernel's loop
 * t < t0 + kernel_sim_time - step/10, dof18_halfImplicit_gpu.cu:892-948). */
int lfdm_chunk_steps(lfdm_solver* h, double t, int* nsteps);
/* Number of vehicles initialized so far (Initialize / initialize_batch / set_states). */
int lfdm_initialized_count(lfdm_solver* h, unsigned* count);

/* Arithmetic of the integration step (no counterpart in the reference, which has one form).
 * LFDM_ARITH_FAST (the default) evaluates the reference's functions in a few other forms of the same
 * value: the lateral slip -tan(atan2(vsy, vta) - delta) as a tangent difference, hypot(a, b) as
 * sqrt(a a + b b), and divisions by a shared denominator as multiplications by its reciprocal
 * (dof18.cpp:412-539, 1019-1078 and their 11/24-DOF twins).  LFDM_ARITH_STRICT evaluates every one of
 * them as the reference writes it (atan2 and tan, hypot, one correctly rounded division per quotient),
 * so results differ from the reference CPU solver only where the device libm and glibc round
 * differently; it runs at the model's default occupancy and costs more instructions per step. */
#define LFDM_ARITH_FAST 0
#define LFDM_ARITH_STRICT 1
int lfdm_set_arithmetic(lfdm_solver* h, int arith);
int lfdm_get_arithmetic(lfdm_solver* h, int* arith);

/* Advances every vehicle nsteps steps from its current time in one launch (no output), on the
 * handle's stream, without synchronizing.  The time base and the position in per-step controls
 * continue across calls. */
int lfdm_advance(lfdm_solver* h, int nsteps);
/* Resets the handle's time base to t and the per-step control position to row 0 (Solve always
 * starts at t = 0 from row 0). */
int lfdm_set_time(lfdm_solver* h, double t);

/* ---- differentiation form (no counterpart in the reference) ------------------------------
 * Reverse-mode gradients of the 18-DOF half-implicit map (TMeasy tires, linear steering, no
 * throttleMod; LFDM_E_UNSUPPORTED otherwise), with discrete events frozen at the forward pass.
 * lfdm_adjoint_forward integrates nsteps steps from the handle's current records (which it does not
 * modify) with per-step controls device_ctrl[nsteps][3][total], writes the position after every step
 * to device_xy[nsteps][2][total] (x, y) and keeps a checkpoint of every vehicle's carried state every
 * checkpoint_every steps.  lfdm_adjoint_backward takes dLoss/d(xy) in device_xy_bar (layout of
 * device_xy) and writes dLoss/d(controls) (layout of device_ctrl), dLoss/d(initial record)
 * [record_len][total] and dLoss/d(per-vehicle parameters) [7][total] in the order of
 * lfdm_set_overrides.  device_ctrl stays borrowed until the backward call.  Both are asynchronous on
 * the handle's stream; lfdm_last_launch_ms times the last of them. */
int lfdm_adjoint_forward(lfdm_solver* h, int nsteps, int checkpoint_every, const double* device_ctrl,
                         double* device_xy);
int lfdm_adjoint_backward(lfdm_solver* h, const double* device_xy_bar, double* device_ctrl_bar,
                          double* device_state_bar, double* device_param_bar);
/* Device memory the adjoint may use for its tape, from the next lfdm_adjoint_forward on; 0 (the
 * default) is automatic: half of the device memory free when the forward starts (counting the tape the
 * handle already holds).  The backward walks the checkpoint segments in groups, recomputing a whole
 * group's steps at once (38 doubles per vehicle and step; the forward tapes the last group itself);
 * a group holds as many segments as the budget does, at least one.  A budget that holds the whole
 * horizon (steps x 38 x total x 8 bytes) makes the forward tape every step and the backward recompute
 * nothing.  Results do not depend on it. */
int lfdm_adjoint_set_tape_budget(lfdm_solver* h, unsigned long long bytes);
/* Bytes of tape the last lfdm_adjoint_forward set up and its segments per backward group (the whole
 * horizon's segments when the backward recomputes nothing); LFDM_E_STATE before any forward.  No
 * counterpart in the reference (which has no adjoint). */
int lfdm_adjoint_tape_bytes(lfdm_solver* h, unsigned long long* bytes, int* segments_per_group);
/* Counter of the handle's adjoint state: it changes with every lfdm_adjoint_forward and whenever a
 * pending adjoint is cancelled (new per-vehicle parameters, a new slot layout).  A caller holding a
 * forward's value can tell whether its backward is still the one that forward set up. */
int lfdm_adjoint_generation(lfdm_solver* h, unsigned long long* generation);
/* Exact Jacobian of one step at every vehicle's current record, for linearizing the model (e.g. MPC).
 * Replaces d18SolverHalfImplicit::IntegrateStepWithJacobian + GetJacobianState / GetJacobianControls
 * (18dof/dof18_halfImplicit.h:79-86, 159; dof18_halfImplicit.cpp:356-463), which differences the
 * right-hand side for one vehicle on the CPU (relative step 1e-3, at gear 0 with zero lagged
 * accelerations).  Here the step adjoint gives the discrete step's exact derivatives at the record as
 * it is, for every vehicle at once: device_jac[21][24][total], rows the stepped packed states, columns
 * the packed states then steering, throttle, braking, both in the reference's packY order of a vehicle
 * with a torque converter (xe ye per tire, omega x 4, crankOmega, x y u v psi wz phi wx).  Without a
 * converter crankOmega is no state: its column is zero and its row the derivative of the slaved
 * crank speed.  device_ctrl[3][total]: the step's steering, throttle, braking.  The records do not
 * change (advance the step separately).  *rows / *cols receive 21 / 24.  18-DOF TMeasy only, like the
 * adjoint; asynchronous on the handle's stream. */
int lfdm_step_jacobian(lfdm_solver* h, const double* device_ctrl, double* device_jac, int* rows, int* cols);

/* ---- batching form: per-vehicle horizons, layouts, mixed fleets (no counterpart in the
 * reference, whose solver has one end time for every vehicle and overshoots it to a whole number of
 * kernel-sim-time chunks, dof18_halfImplicit_gpu.cu:438-464) -------------------------------------- */

#define LFDM_LAYOUT_GROUPED 0    /* slots in Construct order (the default) */
#define LFDM_LAYOUT_BY_HORIZON 1 /* within each Construct group, longest horizon first */

/* Per-vehicle horizons steps[total]: vehicle i stops after steps[i] steps counted from t = 0 (a launch
 * starting at time t0 starts at step round(t0 / step)); later steps leave its record unchanged and
 * output rows after it repeat its final state at the row's time.  NULL removes the horizons. */
int lfdm_set_horizons(lfdm_solver* h, const int* steps);
/* Slot layout (results do not depend on it).  Grouping vehicles of similar horizons into the same
 * wavefronts lets a wavefront stop when its vehicles do; changing it re-lays the device image out
 * (the current records are kept). */
int lfdm_set_layout(lfdm_solver* h, int layout);
int lfdm_get_layout(lfdm_solver* h, int* layout);

/* A mixed fleet: solver handles of any models (buckets, normally one per model, so a wavefront never
 * mixes models) integrated together.  lfdm_fleet_run(f, nsteps) does for every bucket what Solve
 * does without output: nsteps steps from t = 0 from the current records with each bucket's control
 * source (driver tables, knots or per-step arrays from row 0) and horizons, as one HIP graph of one
 * chain of launches of steps_per_launch steps (0 = one launch) per bucket: the chains run one after
 * another when every bucket fills the GPU by itself (at least as many wavefronts as its kernels'
 * occupancy lets the GPU hold at once), else side by side on branches of the graph.  The graph
 * is captured on the first run and replayed while the launch configuration is unchanged.  Runs are
 * asynchronous on the fleet's stream; the buckets' streams are ordered after them.  The handles
 * stay owned by the caller and must outlive the fleet. */
typedef struct lfdm_fleet lfdm_fleet;
int lfdm_fleet_create(lfdm_fleet** out);
int lfdm_fleet_destroy(lfdm_fleet* f);
/* Adds a bucket; returns its index (>= 0) or a negative LFDM_E_*. */
int lfdm_fleet_add(lfdm_fleet* f, lfdm_solver* h);
/* Stream for the fleet's runs from now on (hipStream_t; NULL is the default stream).  A fleet whose
 * caller never chose one uses a non-blocking stream of its own. */
int lfdm_fleet_set_stream(lfdm_fleet* f, void* stream);
/* Launch configuration of a bucket: threads per block (64..256, a multiple of 64), steps per launch
 * and the kernel's waves per SIMD (0: the number of wavefronts per SIMD the bucket fills, up to the
 * model's default, the largest: 1 to 4 for the 11-DOF kernels, 1 to 3 for the 18-DOF ones, 1 or 2 for
 * the 24-DOF ones). */
int lfdm_fleet_set_config(lfdm_fleet* f, int bucket, int block, int steps_per_launch, int waves_per_simd);
int lfdm_fleet_get_config(lfdm_fleet* f, int bucket, int* block, int* steps_per_launch, int* waves_per_simd,
                          int* layout);
/* Picks every bucket's layout (from the lane-steps each layout would run over nsteps steps with the
 * bucket's horizons) and its block size, steps per launch and waves per SIMD (timed trials of
 * min(trial_steps, nsteps) steps from the current records, which are restored afterwards). */
int lfdm_fleet_autotune(lfdm_fleet* f, int nsteps, int trial_steps);
int lfdm_fleet_run(lfdm_fleet* f, int nsteps);
int lfdm_fleet_synchronize(lfdm_fleet* f);
/* Device time of the last lfdm_fleet_run, from HIP events around the graph launch. */
int lfdm_fleet_last_ms(lfdm_fleet* f, double* ms);

/* ---- no-allocation entry point (no counterpart in the reference) ---------------------------
 * For callers that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc): integrates n vehicles of one parameter set (the flat vector of
 * lfdm_params_from_json) nsteps steps of dt with per-step controls device_ctrl[nsteps][3][n] from the
 * records device_records_in[n][record_len] (taken as they are, like lfdm_set_states: no 24-DOF
 * initializeTireSus) with per-vehicle parameters device_ovr[7][n] (order of lfdm_set_overrides; NULL: the
 * set's own values), and writes the final records to device_records_out[n][record_len] (may alias
 * device_records_in).  The same kernels and results as a handle doing Initialize / set_states,
 * set_overrides, set_controls_steps_device, advance and get_states.  Asynchronous on `stream` (hipStream_t,
 * NULL = the default stream); calls no hipMalloc / hipFree: the parameter set, slot maps, SoA state image
 * and per-vehicle parameters live in a 64 MiB device array of the library, one per GPU, held until the
 * process ends.  Sizes beyond it (about 100,000 18-DOF, 190,000 11-DOF or 37,000 24-DOF vehicles) fail
 * with LFDM_E_CAPACITY.  Calls that overlap on one GPU share that array and race. */
int lfdm_integrate_noalloc(int dof, const double* flat, int len, int tire_type, unsigned n,
                           const double* device_records_in, const double* device_ovr, const double* device_ctrl,
                           int nsteps, double dt, int arith, double* device_records_out, void* stream);

/* Stream used for every launch and copy from now on (hipStream_t).  NULL is the default stream,
 * which is ordered with every blocking stream (torch's default stream is NULL).  A handle whose
 * caller never chose a stream uses a non-blocking stream of its own. */
int lfdm_set_stream(lfdm_solver* h, void* stream);
int lfdm_synchronize(lfdm_solver* h);
/* Device time of the last integration launch, from HIP events on the handle's stream. */
int lfdm_last_launch_ms(lfdm_solver* h, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* LFDM_AMD_H */
