/*
 * C ABI of the MI355X FastMonodomainSolver fibre library (libfms_<model>.so).
 *
 * Two layers:
 *
 * 1. Reference plugin boundary. OpenDiHu's FastMonodomainSolver with
 *    CellML.optimizationType = "gpu" and CellML.libraryFilename set dlopens a
 *    prebuilt library and binds exactly two symbols
 *    (core/src/specialized_solver/fast_monodomain_solver/fast_monodomain_solver_gpu.tpp:184-205):
 *      initializeArrays   (fast_monodomain_solver_gpu.tpp:356-365, called once at :1003-1007)
 *      computeMonodomain  (fast_monodomain_solver_gpu.tpp:478-481, called per advanceTimeSpan at :1078-1081)
 *    Their sizes are compile-time constants of the reference's generated source
 *    (fast_monodomain_solver_gpu.tpp:293-312); here they come from
 *    fms_reference_set_dims() or the OPENDIHU_FMS_DIMS environment variable.
 *
 * 2. Native API (fms_*): runtime sizes, fibres of uneven length, FP64 element
 *    lengths and per-fibre parameters, firing-table or current-pulse
 *    stimulation, FP64 or mixed precision, split or fused kernels. The handle
 *    owns all device state; calls on one handle are not re-entrant.
 *
 * All functions return 0 on success and a negative FMS_ERR_* code on failure;
 * fms_last_error() gives the message. Nothing throws across this boundary.
 * The reference-compatible void functions report failure through
 * fms_last_error() only (the reference binding has no return value).
 */
#ifndef OPENDIHU_FMS_H
#define OPENDIHU_FMS_H

#ifdef __cplusplus
extern "C" {
#endif

#define FMS_OK 0
#define FMS_ERR_ARGUMENT (-1)
#define FMS_ERR_HIP (-2)
#define FMS_ERR_STATE (-3)
#define FMS_ERR_UNSUPPORTED (-4)
#define FMS_ERR_INTERNAL (-5)   /* host memory exhausted or another host-side C++ failure */

/* diffusion time stepping (DiffusionTimeSteppingScheme of FastMonodomainSolverBase) */
#define FMS_IMPLICIT_EULER 0
#define FMS_CRANK_NICOLSON 1
/* state precision: FP64 everywhere, or FP32 gates with FP64 Vm and FP64 Thomas */
#define FMS_FP64 0
#define FMS_MIXED 1
/* exp/log: CellML.approximateExponentialFunction = True / False */
#define FMS_EXP_APPROXIMATE 0
#define FMS_EXP_LIBM 1
/* tridiagonal solver: reference recurrence (bit-exact) / partitioned + PCR (fast) */
#define FMS_THOMAS_SERIAL 0
#define FMS_THOMAS_PARTITIONED 1
/* kernel schedule: separate reaction/diffusion kernels, or one fused multi-step kernel */
#define FMS_KERNELS_SPLIT 0
#define FMS_KERNELS_FUSED 1
/* stimulation */
#define FMS_STIM_NONE 0
#define FMS_STIM_FIRING_TABLE 1
#define FMS_STIM_CURRENT 2

typedef struct FmsSolver FmsSolver;

typedef struct {
  int n_fibers;
  const int *n_nodes;              /* [n_fibers], each >= 2 */
  const double *element_lengths;   /* [sum(n_nodes-1)], fibre after fibre */
  const double *prefactor;         /* [n_fibers], sigma/(Am*Cm) (FiniteElementMethod.prefactor) */
  const double *parameters;        /* [n_fibers*n_params] or NULL for the model's defaults */
  const double *initial_states;    /* [n_states] or NULL for the model's defaults */
  int scheme;                      /* FMS_IMPLICIT_EULER / FMS_CRANK_NICOLSON */
  int precision;                   /* FMS_FP64 / FMS_MIXED */
  int exp_mode;                    /* FMS_EXP_APPROXIMATE / FMS_EXP_LIBM */
  int thomas;                      /* FMS_THOMAS_SERIAL / FMS_THOMAS_PARTITIONED */
  int kernels;                     /* FMS_KERNELS_SPLIT / FMS_KERNELS_FUSED */
  double value_for_stimulated_point; /* valueForStimulatedPoint (default 20) */
} FmsConfig;

typedef struct {
  int mode;                        /* FMS_STIM_* */
  const int *stim_begin;           /* [n_fibers] first stimulated node (fibre-local) */
  const int *stim_end;             /* [n_fibers] one past the last stimulated node */
  /* FMS_STIM_FIRING_TABLE: firingTimesFile table and setSpecificStates* settings */
  int firing_rows, firing_cols;
  const unsigned char *firing_events;   /* [rows*cols], row = sampling step, col = motor unit */
  const int *motor_unit;                /* [n_fibers] */
  const double *call_frequency;         /* [n_fibers] setSpecificStatesCallFrequency */
  const double *repeat_after_first_call;/* [n_fibers] setSpecificStatesRepeatAfterFirstCall */
  const double *call_enable_begin;      /* [n_fibers] setSpecificStatesCallEnableBegin */
  const double *jitter;                 /* concatenated setSpecificStatesFrequencyJitter lists */
  const int *jitter_offset;             /* [n_fibers] */
  const int *jitter_count;              /* [n_fibers] */
  const double *last_check_time;        /* [n_fibers] lastStimulationCheckTime, or NULL for
                                           enableBegin - 1e-13 - 1/frequency */
  int log_capacity;                     /* stimulation-onset log entries kept (0 = no log) */
  /* FMS_STIM_CURRENT: parameter stim_param := amplitude on [onset, onset+duration) */
  const double *onset, *duration, *amplitude;  /* [n_fibers] */
  int stim_param;
} FmsStimulation;

/* model compiled into this library */
int fms_model_info(int *n_states, int *n_algebraics, int *n_params, char *name, int name_len);
int fms_model_defaults(double *states /* [n_states] */, double *params /* [n_params] */);
/* algorithmic operations of one right-hand-side evaluation (each + - * / and
   each exp/log counts one; integer powers count their multiplications) */
int fms_model_ops(double *ops_rhs, double *divs_rhs, double *exps_rhs);

/* The model's right-hand side evaluated on the device for n independent
   points (the generated computeRates of the reference's CellML model source,
   GEN/03_generator_vc.cpp; CellmlAdapter evaluateTimesteppingRightHandSide):
   states [n_states][n] and params [n_params][n] in, rates [n_states][n] and
   algebraics [n_algebraics][n] (may be NULL) out, host buffers; exp_mode as in
   FmsConfig. Synchronous. */
int fms_eval_rhs(const double *states, const double *params, double t, int n, int exp_mode, double *rates,
                 double *algebraics);
/* n_steps Heun steps (the generated reaction step, GEN/03_generator_vc.cpp:1017-1130,
   without stimulation) of n independent points on the device, FP64; states
   [n_states][n] in and out (host buffers). Synchronous. */
int fms_eval_heun(double *states, const double *params, double t0, double dt, int n_steps, int n, int exp_mode);

/* CellML.initializeStatesToEquilibrium (core/src/cellml/03_cellml_adapter.tpp:446-520):
   classical RK4 of one cell (parameters [n_params], libm exp/log) from `states`
   with step dt (initializeStatesToEquilibriumTimestepWidth, default 1e-4) until
   the largest |dy/dt| of an RK4 step is below 1e-5; a step whose increment
   exceeds 1e4 restarts from the initial values with half the step. states
   [n_states] in and out (the last iterate if the search fails); iterations,
   max_increment and dt_out (may be NULL) receive the iterations taken, the last
   largest increment and the final step. Returns FMS_ERR_STATE if no equilibrium
   was found within max_iterations (the reference uses 1e7). Runs on the calling
   host thread: one serial recurrence, once, before the first step. */
int fms_equilibrium_states(double *states, const double *params, double dt, long long max_iterations,
                           long long *iterations, double *max_increment, double *dt_out);

int fms_create(const FmsConfig *config, FmsSolver **solver);
int fms_destroy(FmsSolver *solver);
int fms_set_stimulation(FmsSolver *solver, const FmsStimulation *stim);

/* computeMonodomain semantics: n_split Strang steps from start_time; one
   diffusion solve with dt1D per step (n1D is accepted and, as in the
   reference, not looped over). Enqueued on the solver's stream. */
int fms_advance(FmsSolver *solver, double start_time, double dt_split, int n_split,
                double dt0D, int n0D, double dt1D, int n1D);
/* same, with per-step start times and widths (RepeatedCall schedules):
   steps[k*4 + {0,1,2,3}] = {start_time, dt_split, dt0D, dt1D}; n0D Heun steps per half */
int fms_advance_schedule(FmsSolver *solver, const double *steps, int n_steps, int n0D);
int fms_synchronize(FmsSolver *solver);

/* Allocation-free variant of fms_advance_schedule (no counterpart in the
   reference; for callers that cannot allocate device memory: serving loops,
   harnesses that forbid hipMalloc). Calls no hipMalloc, hipFree or other
   allocating routine: the step table and the per-fibre stimulation flags of the
   call live in a fixed 4 MiB device scratch array that each GPU holds until the
   process ends (32 bytes per split step plus 2 bytes per fibre and Heun step;
   larger calls fail with FMS_ERR_UNSUPPORTED). Graph replay is off for the call
   (capturing would allocate); the serial solve with the split kernels needs the
   scratch fms_create or an earlier fms_advance made. Returns when the call's
   kernels are done, so calls made one after another on a GPU never overlap;
   calls from several host threads at once on one GPU share its scratch and
   race. Same results as fms_advance_schedule. */
int fms_advance_schedule_noalloc(FmsSolver *solver, const double *steps, int n_steps, int n0D);

int fms_get_states(FmsSolver *solver, int state_no, double *out /* [n_total] */);
int fms_set_states(FmsSolver *solver, int state_no, const double *in /* [n_total] */);
int fms_get_algebraics(FmsSolver *solver, double *out /* [n_alg_transfer*n_total] */);
int fms_set_algebraics_for_transfer(FmsSolver *solver, const int *indices, int count);
int fms_total_nodes(FmsSolver *solver, long long *n_total);

/* spike detector: first split step (1-based, counted from creation) at which
   Vm reaches threshold from below, and number of upward crossings, per node */
int fms_enable_spike_detection(FmsSolver *solver, double threshold);
int fms_get_spikes(FmsSolver *solver, int *first_step, int *count);
int fms_get_stimulation_log(FmsSolver *solver, int *count, double *times, int *fibers, int capacity);

/* RepeatedCall<FastMonodomainSolver> time-step bookkeeping
   (core/src/time_stepping_scheme/repeated_call.tpp:45-80,
   core/src/time_stepping_scheme/00_time_stepping_scheme.cpp:15-70):
   fills out[k*7 + {0..6}] = {startTime, dtSplit, nSplit, dt0D, n0D, dt1D, n1D}
   for each advanceTimeSpan call; returns the number of calls (or < 0). */
int fms_schedule(double rc_end, double rc_width, double strang_end, double strang_width,
                 double heun_end, double heun_width, double ie_end, double ie_width,
                 int max_calls, double *out);

int fms_last_error(char *buf, int len);

/* schedule knobs of the fused kernel (the autotuner's search space) */
#define FMS_TUNE_STEPS_PER_LAUNCH 1   /* split steps per fused launch (default 512) */
#define FMS_TUNE_RESIDENT_GATES 2     /* -1 auto, 0 stream gates from HBM, 1 keep them in registers */
/* onlyComputeIfHasBeenStimulated (0/1, default 0): a fibre's reaction runs from
   the Heun step of its first stimulation on; diffusion always runs */
#define FMS_TUNE_ONLY_IF_STIMULATED 3
#define FMS_TUNE_THREADS 4            /* threads per fibre workgroup of the fused kernel: 256 (default) or 128 */
#define FMS_TUNE_GRAPH_STEPS 5        /* split kernels: steps per captured hipGraph, 0 = launch eagerly (default) */
#define FMS_TUNE_REACTION_PER_FIBRE 6 /* split kernels, models of > 8 states: reaction with one workgroup per
                                         fibre (1, default) or one lane per node over the whole batch (0) */
int fms_set_tuning(FmsSolver *solver, int key, int value);

/* onlyComputeIfHasBeenStimulated in the reference's point-buffer order
   (fast_monodomain_solver_compute.tpp:131-216): the nodes of the concatenated
   fibres are swept in buffers of `width` nodes (Vc::double_v::size(), 4 on
   AVX2), buffer b belonging to fibre floor(b*width/n_nodes[0]); a fibre's
   stimulation is checked in its centre buffer only (the one holding
   nmj[f], its stimulation point; a fibre without one is never stimulated),
   and a buffer swept before the centre buffer in the half step of the first
   stimulation skips that whole half. Takes effect with
   FMS_TUNE_ONLY_IF_STIMULATED = 1; runs the split kernels with the reaction per
   node. width 0 returns to the per-fibre rule (a fibre computes from the Heun
   step of its first stimulation on). */
int fms_set_point_buffer_order(FmsSolver *solver, int width, const int *nmj /* [n_fibers] */);

#define FMS_TUNE_PROFILE_KERNELS 7    /* 1: time every launch with a start/stop event pair bound to the kernel
                                         (hipExtLaunchKernel); the host does not wait for the launches */
#define FMS_TUNE_PAIR_HALVES 8        /* split kernels: 1 (default) runs the second reaction half of step q and the
                                         first of step q+1 (no solve between them) as one launch that reads and
                                         writes the states once; 0: one launch per half. Same arithmetic. */
/* adjoint backward (fms_adjoint_backward): each checkpoint segment is recomputed
   (recording its tape) and then reversed. FMS_ADJOINT_OVERLAP: the recompute
   of the next segment runs on a second stream beside the reverse of this one;
   _OVERLAP_PRIORITY: the same, the reverse sweep's waves with raised issue
   priority; _SERIAL (default): recompute and reverse one after the other on the
   solver's stream. A recompute beside the reverse sweep streams its gates from
   HBM (fewer registers, so that it shares the SIMDs with the sweep); a serial
   one uses the solver's fused schedule. Same arithmetic, same gradients. */
#define FMS_TUNE_ADJOINT 10
#define FMS_ADJOINT_OVERLAP 0
#define FMS_ADJOINT_OVERLAP_PRIORITY 1
#define FMS_ADJOINT_SERIAL 2
#define FMS_ADJOINT_OVERLAP_RECOMPUTE_PRIORITY 3   /* beside it, the recompute's waves with raised priority */
/* threads per fibre workgroup of the adjoint's reverse sweep: 256, 512
   (default) or 1024 (fewer nodes per thread, more waves per SIMD). The prefactor
   gradient is summed per wave and then over the waves, so its last bits depend
   on this number. */
#define FMS_TUNE_ADJOINT_THREADS 11
/* where the adjoint backward finds the forward's tape (9 doubles per node and
   split step). FMS_ADJOINT_TAPE_SEGMENTS: fms_adjoint_forward keeps a
   checkpoint every checkpoint_interval steps and the backward recomputes each
   segment (two segment tapes in device memory, FMS_TUNE_ADJOINT schedules the
   recompute). FMS_ADJOINT_TAPE_WHOLE: fms_adjoint_forward records the whole
   tape (72 B per node and step, 92 GB for 256 x 1000 nodes over 5000 steps) and
   the backward only reverses it. FMS_ADJOINT_TAPE_AUTO (default): the whole
   tape when it fits in half of the device memory free at the forward, else
   segments. The mode of the last forward decides the backward's. A whole tape
   stays allocated until fms_destroy or the next forward with checkpoints. Same
   arithmetic, same gradients. */
#define FMS_TUNE_ADJOINT_TAPE 12
#define FMS_ADJOINT_TAPE_AUTO (-1)
#define FMS_ADJOINT_TAPE_SEGMENTS 0
#define FMS_ADJOINT_TAPE_WHOLE 1
#define FMS_TUNE_REACTION_THREADS 9   /* split kernels, models of > 8 states, reaction per fibre: threads per fibre
                                         workgroup, 256 (default) or 64 (one wave: four fibres share a CU and their
                                         loads do not coincide). Same arithmetic. */

/* accumulated launch durations in ms (each the kernel's own begin to end) and
   launch counts per kernel class while FMS_TUNE_PROFILE_KERNELS is on (graph
   replays are not timed); waits for the launches recorded so far; reset != 0
   clears them */
#define FMS_KERNEL_REACTION 0
#define FMS_KERNEL_DIFFUSION 1
#define FMS_KERNEL_FUSED 2
#define FMS_KERNEL_OTHER 3
#define FMS_KERNEL_REACTION_PAIR 4    /* paired reaction halves (FMS_TUNE_PAIR_HALVES) */
#define FMS_KERNEL_CLASSES 5
int fms_kernel_times(FmsSolver *solver, double ms[FMS_KERNEL_CLASSES], long long launches[FMS_KERNEL_CLASSES],
                     int reset);

/* graphs captured and graph replays issued so far (split kernels) */
int fms_graph_counts(FmsSolver *solver, long long *captures, long long *replays);

/* Autotuner (no counterpart in the reference, which has one fixed CPU schedule
   and an OpenMP-offload loop): times every schedule that gives bit-identical
   states on this batch - fused kernel with gates resident or streamed, 128 or
   256 threads per fibre, 64 or 512 steps per launch; with allow_split also the
   split kernels launched eagerly or replayed as hipGraphs of 8 or 32 steps
   (models of > 8 states: reaction per fibre with 256 or 64 threads, or per node,
   eager or graphs of 8) -
   over the n_steps records of `steps` (4 doubles each, as
   fms_advance_schedule), restores the states, stimulation state and spike
   counters after every trial and keeps the fastest schedule. Every trial's
   end state (states, stimulation state machine and log, spike detector) is
   compared, by a digest of its bits, with that of the solver's own schedule
   over the same records: a candidate that differs is rejected, never chosen,
   and counted in n_rejected (if all differ, the solver keeps its schedule and
   ms_per_step is -1). candidate_ms (optional, `capacity` entries) receives each
   candidate's ms per step in the order tried, -1 for a rejected one. */
typedef struct {
  int kernels;            /* FMS_KERNELS_SPLIT / FMS_KERNELS_FUSED */
  int steps_per_launch;
  int resident_gates;
  int threads;
  int graph_steps;
  double ms_per_step;
  int n_candidates;
  int reaction_per_fibre;
  int reaction_threads;
  int n_rejected;         /* candidates whose result differed from the solver's own schedule */
} FmsTuneResult;
int fms_autotune(FmsSolver *solver, const double *steps, int n_steps, int n0D, int allow_split,
                 FmsTuneResult *result, double *candidate_ms, int capacity);

/* timing on the solver's stream: events in slots 0..7, elapsed time between
   two recorded slots, and the number of kernel launches issued so far */
int fms_record_event(FmsSolver *solver, int slot);
int fms_event_elapsed_ms(FmsSolver *solver, int from_slot, int to_slot, double *ms);
int fms_launch_counts(FmsSolver *solver, long long *fused, long long *reaction, long long *thomas);

/* per-fibre parameters [n_fibers*n_params] and prefactors [n_fibers] (NULL keeps them) */
int fms_set_parameters(FmsSolver *solver, const double *params, const double *prefactor);

/* sensor nodes: n_sensors fibre-local node numbers per fibre, [n_fibers*n_sensors];
   Vm at the sensors is recorded after every split step of fms_adjoint_forward */
int fms_set_sensors(FmsSolver *solver, const int *sensor_nodes, int n_sensors);

/* Discrete adjoint (Hodgkin-Huxley model, FP64, current-pulse or no stimulation,
   fibres of <= 2048 nodes; no counterpart in the reference).
   fms_adjoint_forward runs n_steps split steps from the current state
   (steps[k*4 + {0..3}] = {t, dt_split, dt0D, dt1D}), records the whole tape or
   stores the state every checkpoint_interval steps (FMS_TUNE_ADJOINT_TAPE) and
   returns the sensor traces traces[(k*n_fibers + f)*n_sensors + j] (host
   buffer, may be NULL).
   fms_adjoint_backward takes dL/dtraces in the same layout and returns
   grad_params[f*3 + {0,1,2}] = dL/d(g_Na, g_K, g_L) and grad_prefactor[f] = dL/dprefactor
   (dL/dsigma = grad_prefactor / (Am*Cm)). With checkpoints it recomputes each
   segment with a tape first; it leaves the solver's states unspecified. */
int fms_adjoint_forward(FmsSolver *solver, const double *steps, int n_steps, int n0D, int checkpoint_interval,
                        double *traces);
int fms_adjoint_backward(FmsSolver *solver, const double *dtraces, double *grad_params, double *grad_prefactor);
/* *whole = 1 if the last fms_adjoint_forward recorded the whole tape, 0 if it
   kept checkpoints */
int fms_adjoint_whole_tape(FmsSolver *solver, int *whole);
/* the same with the traces, their adjoints and the gradients in device memory
   (e.g. torch tensors on the GPU): no host copies */
int fms_adjoint_forward_device(FmsSolver *solver, const double *steps, int n_steps, int n0D,
                               int checkpoint_interval, double *d_traces);
int fms_adjoint_backward_device(FmsSolver *solver, const double *d_dtraces, double *d_grad_params,
                                double *d_grad_prefactor);

/* ---- reference plugin boundary ---- */
typedef struct {
  int n_fibers_to_compute;
  int n_instances_per_fiber;
  int n_parameters_per_instance;
  int firing_events_rows;
  int firing_events_cols;
  int frequency_jitter_cols;
  int n_algebraics_for_transfer;
  int n_states_for_transfer;
  int scheme;                      /* FMS_IMPLICIT_EULER / FMS_CRANK_NICOLSON */
  int precision;                   /* useSinglePrecision -> FMS_MIXED */
  int exp_mode;
} FmsReferenceDims;

int fms_reference_set_dims(const FmsReferenceDims *dims);

void initializeArrays(const double *statesOneInstance, const int *algebraicsForTransferIndices,
                      const int *statesForTransferIndices, const char *firingEvents,
                      const double *setSpecificStatesFrequencyJitter, const int *motorUnitNo,
                      const int *fiberStimulationPointIndex, const double *lastStimulationCheckTime,
                      const double *setSpecificStatesCallFrequency,
                      const double *setSpecificStatesRepeatAfterFirstCall,
                      const double *setSpecificStatesCallEnableBegin);

void computeMonodomain(const float *parameters, double *algebraicsForTransfer, double *statesForTransfer,
                       const float *elementLengths, double startTime, double timeStepWidthSplitting,
                       int nTimeStepsSplitting, double dt0D, int nTimeSteps0D, double dt1D, int nTimeSteps1D,
                       double prefactor, double valueForStimulatedPoint);

#ifdef __cplusplus
}
#endif

#endif /* OPENDIHU_FMS_H */
