/* gridsim_hip.h - C ABI of the MI355X sparse Newton-Raphson load-flow engine (libgridsim_hip.so).
 *
 * Flat C: plain pointers and sizes, caller-owned host arrays, opaque handles for device-resident state. Status codes
 * follow the reference's P/Invoke convention (int, 0 = success; RocmInterop.cs:37-96) and its solver seams:
 *
 *   ISparseLinearSolver.Solve(CsrMatrix, double[]) -> double[]?          src/GridSim.Core/Abstractions/ISparseLinearSolver.cs:12-16
 *       -> gs_sparse_solve
 *   IBatchedSparseLinearSolver.SolveBatch(matrices, rhs, errors)          src/GridSim.Core/Abstractions/IBatchedSparseLinearSolver.cs:20-37
 *       -> gs_sparse_solve_batch
 *   IPowerFlowSolver.Solve(GridModel, PowerFlowOptions?)                  src/GridSim.Core/Abstractions/IPowerFlowSolver.cs:10-12
 *       -> gs_network_upload + gs_nr_solve_batch (one or many instances)
 *   ContingencyAnalysis.ScreenBranchesBatched / ScreenBranchesParallel    src/GridSim.Core/Analysis/ContingencyAnalysis.Batched.cs:55-136,
 *       -> gs_nr_solve_batch with one outage per instance                 ContingencyAnalysis.cs:323-347
 *   RpcCommand.Solve (gridsim-rpc/1 "solve", sequential)                  src/GridSim.Cli/RpcCommand.cs:668-700
 *       -> gs_server_* (continuous batching in front of the same contract)
 *
 * Threading: a gs_ctx serialises its own calls with an internal lock (the reference serialises device access with a
 * process-global lock, RocmContext.cs:21). Handles belong to the context that made them.
 */
#ifndef GRIDSIM_HIP_H
#define GRIDSIM_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_API __attribute__((visibility("default")))

typedef int32_t gs_status;
enum {
  GS_OK = 0,
  GS_SINGULAR = 1,          /* the sparse LU refused the matrix: the seam's null (caller's CPU fallback) */
  GS_NONFINITE_INPUT = 2,   /* NaN / Inf in the matrix or right-hand side: the seam's ArgumentException */
  GS_OUT_OF_MEMORY = 3,     /* device allocation failed: the only status the reference falls back on */
  GS_INVALID_ARGUMENT = 4,  /* malformed call (bad sizes, pattern, handle) */
  GS_DEVICE_ERROR = 5,      /* any other HIP failure */
  GS_NO_DEVICE = 6,         /* no usable GPU */
  GS_BUSY = 7,              /* the server's admission queue is full */
  GS_PENDING = 8            /* gs_server_wait: the request has not finished within the timeout */
};

/* Bus types in the reference enum order (src/GridSim.Core/Model/Model.cs:4-12). */
enum { GS_BUS_SLACK = 0, GS_BUS_PV = 1, GS_BUS_PQ = 2 };

/* PowerFlowStatus (src/GridSim.Core/PowerFlow/PowerFlowResult.cs; precedence NewtonShell.cs:671-674). */
enum { GS_PF_CONVERGED = 0, GS_PF_MAX_ITERATIONS = 1, GS_PF_DIVERGED = 2, GS_PF_NONFINITE = 3 };

/* Arithmetic of the Jacobian factorisation and triangular solves. The mismatch, convergence test, voltage update and
 * refinement residual are always FP64. */
enum { GS_LU_FP64 = 0, GS_LU_FP32 = 1 };

/* A network as the reference's GridModel holds it (Model.cs:20-35 Bus, :91-115 Generator, :253-276 Branch), with
 * buses identified by position. Loads and shunts in MW / MVAr, angles in degrees, impedances in pu. */
typedef struct gs_network {
  int32_t n_bus, n_branch, n_gen;
  double base_mva;
  const int32_t* bus_type;
  const double *pd, *qd, *gs, *bs, *vm, *va_deg;
  const int32_t *br_from, *br_to;
  const double *br_r, *br_x, *br_b, *br_tap, *br_shift_deg;
  const uint8_t* br_in_service;
  const int32_t* gen_bus;
  const double *gen_pg, *gen_qg, *gen_vg;
  /* For screening (may be null): branch rate A in MVA (0 = unrated; Branch.RateAMva), and each bus's own voltage
   * limits in pu (NaN = none: the operational band 0.94 / 1.06 applies on that side; Bus.VminPu / VmaxPu). */
  const double* br_rate_a;
  const double *bus_vmin, *bus_vmax;
  /* Generator reactive limits in MVAr (may be null: none), for GS_FLAG_Q_LIMITS (Generator.Qmin / Qmax). */
  const double *gen_qmin, *gen_qmax;
} gs_network;

/* PowerFlowOptions (src/GridSim.Core/PowerFlow/PowerFlowOptions.cs:6-170) for the engine. Reactive limits are
 * GS_FLAG_Q_LIMITS (EnforceQLimits; off by default), the DC-start retry GS_FLAG_DC_RETRY (RetryFromDcStart; off unless
 * the flag is set). */
typedef struct gs_nr_options {
  double tolerance;         /* PowerFlowOptions.Tolerance, pu on ||mismatch||_inf (1e-8) */
  int32_t max_iterations;   /* PowerFlowOptions.MaxIterations (30) */
  int32_t flat_start;       /* PowerFlowOptions.FlatStart */
  double diverge_limit;     /* PowerFlowOptions.DivergeLimitPu (10) */
  int32_t lu_precision;     /* GS_LU_FP64 or GS_LU_FP32 */
  int32_t refine_steps;     /* FP64 iterative-refinement steps after each FP32 solve (0..4) */
  double growth_tol;        /* a static pivot is kept while |pivot| >= growth_tol * max |candidate| */
  int32_t schedule;         /* 0 = a wavefront per instance, all instances of a call in one launch; k > 0 = the
                               same in launches of at most k instances one after another; GS_SCHEDULE_LEVELS = every
                               instance spread over the whole GPU, level by level; GS_SCHEDULE_GROUP = a work-group of
                               several wavefronts per instance (gs_nr_autotune chooses) */
  int32_t flags;            /* GS_FLAG_*: placement choices and GS_GROUP_WAVES (gs_nr_autotune chooses), reactive
                               limits, the DC-start retry */
} gs_nr_options;

/* gs_nr_options.flags: keep the triangular-solve vector in global memory even where it fits in LDS (smaller LDS per
 * instance, so more instances resident per compute unit). */
enum { GS_FLAG_X_GLOBAL = 1 };

/* gs_nr_options.flags: one wavefront per work-group. By default, where the solve vector sits in LDS and two instances
 * per work-group keep more of them resident per compute unit, the two wavefronts of a work-group share one LDS solve
 * vector and take it in turn for their triangular solves (same arithmetic, same results). */
enum { GS_FLAG_SINGLE_WAVE = 4 };

/* gs_nr_options.flags: enforce generator reactive limits (PowerFlowOptions.EnforceQLimits, the outer loop of
 * SparseNewtonRaphsonPowerFlow.SolveAcAttempt :118-160 with NewtonShell.ApplyQLimits / RevertBrokenPins :519-595,
 * RevertQLimits off, at most 100 passes as MaxQLimitPasses). After a converged pass the host switches each PV bus
 * whose generators' reactive output crossed a limit to PQ at that limit (or back, once, when its voltage shows the
 * limit no longer binds), groups the instances by their bus types and solves them again on the device from their
 * current state; iterations add up over the passes. gs_nr_solve_batch and gs_nr_screen_batch only. */
enum { GS_FLAG_Q_LIMITS = 2 };

/* gs_nr_options.flags: the automatic retry from a DC power-flow angle start (PowerFlowOptions.RetryFromDcStart,
 * NewtonShell.SolveWithRetry / Stalled / TrySeedDcAngles, src/GridSim.Core/PowerFlow/NewtonShell.cs:142-299). An
 * instance whose Newton run did not converge and stalled or diverged (not stopped by the iteration cap while still
 * converging; never when max_iterations <= 1) is solved again on the device from its ordinary start magnitudes and
 * the DC angles of its own network (reactance-only view, one island referenced at the single slack), and the retry's
 * state is reported when it converges with no in-service branch angle above 90 degrees; gs_nr_result.dc_retry says
 * what happened. gs_nr_solve_batch and gs_nr_screen_batch only (with or without GS_FLAG_Q_LIMITS). */
enum { GS_FLAG_DC_RETRY = 8 };

/* gs_nr_options.schedule: level-parallel. Each Newton step is a captured hipGraph of short kernels, one per phase and
 * one per elimination level of the refactor and of each triangular solve, a work-group per chunk of columns (per pass
 * of rows) of a level and per instance, replayed until every instance has finished. For large networks (tens of
 * thousands of buses), where one wavefront per instance leaves the GPU idle. gs_nr_solve_batch only. */
enum { GS_SCHEDULE_LEVELS = -1 };

/* gs_nr_options.schedule: a work-group per instance. The W wavefronts of one work-group (W = GS_GROUP_WAVES_OF(flags),
 * set with GS_GROUP_WAVES(W), 2..16, default 8) share the instance's whole Newton loop in one launch: the mismatch and
 * Jacobian over their slices of the cells and buses, then the refactor and both triangular solves level by level, the
 * chunks (passes) of a wide level split between the waves and a work-group barrier between levels, runs of narrow
 * levels walked by one wave. The same device functions over the same streams as the other schedules, so the same
 * bits. Each wave takes its own refactor workspace in LDS (W is lowered until the W workspaces fit); the solve vector
 * shares LDS with them when it fits too (GS_FLAG_X_GLOBAL: never). For networks of thousands to tens of thousands of
 * buses, where one wavefront per instance leaves the levels' parallelism unused and a kernel per level costs a
 * dispatch per level. Everywhere a wavefront schedule runs: gs_nr_solve_batch, gs_nr_screen_batch, gs_batch_* and the
 * serving loop (instances resumed after a fresh host plan run one wavefront each). */
enum { GS_SCHEDULE_GROUP = -2 };
#define GS_GROUP_WAVES(w) ((int32_t)(w) << 8)
#define GS_GROUP_WAVES_OF(flags) (((flags) >> 8) & 15)

/* One solve of a batch: the uploaded network with some branches out of service and every load and generator P scaled
 * by load_scale, as ContingencyAnalysis.ScreenOutageCore solves WithoutMany(model, indices)
 * (src/GridSim.Core/Analysis/ContingencyAnalysis.cs:1087-1097,1839-1845). The outaged set is outage_branch (when
 * >= 0) together with outages[0 .. n_outages): N-1 sets outage_branch alone, N-2 / N-k add the list (at most
 * GS_MAX_OUTAGES branches in all, each once). vm0 / va0 (radians), when non-null, replace the start point
 * BuildBusSpec would give. */
#define GS_MAX_OUTAGES 8
typedef struct gs_instance {
  int32_t outage_branch;    /* index into the network's branches, or -1 */
  int32_t n_outages;        /* further outaged branches in `outages` (0: none) */
  double load_scale;        /* 1.0 = as uploaded */
  const double* vm0;
  const double* va0;
  const int32_t* outages;   /* n_outages branch indices (may be null when n_outages is 0) */
} gs_instance;

/* Why the engine, rather than the Newton iteration, ended a run (gs_nr_result.engine_stop). The reference has no such
 * limits; a run that ends on one is reported with status GS_PF_MAX_ITERATIONS and this reason, so callers and parity
 * checks can tell it from a real verdict. */
enum {
  GS_STOP_NONE = 0,
  GS_STOP_LOCK_TIMEOUT = 1,     /* a wavefront never obtained its work-group's shared solve vector (bounded spin) */
  GS_STOP_REFRESH_CAP = 2,      /* the fresh symbolic passes of one instance kept failing the pivot-growth check */
  GS_STOP_FALLBACK_CAP = 3      /* more ILU(0) + BiCGSTAB steps than max_iterations + 1 */
};

/* NewtonRetryOutcome (src/GridSim.Core/PowerFlow/PowerFlowResult.cs:24-41): what the retry from the DC power-flow
 * angle start did (GS_FLAG_DC_RETRY). */
enum { GS_RETRY_NOT_TRIED = 0, GS_RETRY_CONVERGED = 1, GS_RETRY_DID_NOT_CONVERGE = 2, GS_RETRY_REJECTED = 3,
       GS_RETRY_UNAVAILABLE = 4 };

typedef struct gs_nr_result {
  int32_t status;               /* GS_PF_* */
  int32_t iterations;           /* Newton steps taken */
  double max_mismatch;          /* ||f||_inf at the final state, pu */
  int32_t symbolic_refreshes;   /* fresh CPU symbolic passes the pivot-growth check asked for */
  int32_t step_failures;        /* steps whose factorisation was refused (singular) */
  double min_pivot_ratio;       /* smallest |pivot| / max |candidate| seen (1 = the pivot was the largest) */
  double phase_us[4];           /* device time of this instance: mismatch + Jacobian, refactor, solves, update */
  int32_t engine_stop;          /* GS_STOP_*: an engine limit ended the run (status then GS_PF_MAX_ITERATIONS) */
  int32_t dc_retry;             /* GS_RETRY_*: PowerFlowResult.NewtonRetry */
} gs_nr_result;

typedef struct gs_device_info {
  char name[64];
  char arch[32];
  int32_t compute_units;
  int32_t fp64_ok;              /* (1 + 2^-40) - 1 computed on the device equals 2^-40 (GpuContext.cs:227-256) */
  size_t hbm_bytes;
  int32_t fdlibm_ok;            /* a device sincos of pi/7 matches the host's FDLIBM bits */
  int32_t reserved;
} gs_device_info;

typedef struct gs_ctx gs_ctx;
typedef struct gs_net gs_net;

/* ---- context ------------------------------------------------------------------------------------------------ */
GS_API gs_status gs_init(int32_t device, size_t hbm_budget_bytes, gs_ctx** out);
GS_API void gs_shutdown(gs_ctx* ctx);
GS_API const char* gs_last_error(const gs_ctx* ctx);
GS_API gs_status gs_device_info_get(gs_ctx* ctx, gs_device_info* out);
GS_API const char* gs_version(void);

/* ---- ISparseLinearSolver / IBatchedSparseLinearSolver --------------------------------------------------------
 * Solve A x = b for a square CSR matrix (columns sorted per row). Same contract as CpuSparseLuSolver.Solve
 * (src/GridSim.Core/PowerFlow/CpuSparseLuSolver.cs:30-48): minimum-degree ordering cached per pattern, Gilbert-
 * Peierls LU with partial pivoting (explicit zeros dropped), GS_SINGULAR for the seam's null, GS_NONFINITE_INPUT
 * for a NaN/Inf matrix or right-hand side. Arguments are read only; x (length n) is written on GS_OK only. */
GS_API gs_status gs_sparse_solve(gs_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                                 const double* val, const double* b, double* x);

/* gs_sparse_solve that allocates no device memory (no counterpart in the reference: for callers that cannot allocate,
 * such as code under HIP graph capture or a serving loop that must not call hipMalloc). The same contract, arithmetic
 * and results as gs_sparse_solve; the plan and the call's device buffers go into a fixed scratch array of 64 MiB that
 * each GPU holds until the process ends. The array keeps the last plan built in it: a call on the same structure
 * replays it (as long as its pivots hold) instead of building and uploading it again. A system needing more returns
 * GS_INVALID_ARGUMENT (gs_last_error says how much it needs). Calls on contexts of the same GPU share the array: they
 * must not overlap (calls on one context never do, the context serialises them). */
GS_API gs_status gs_sparse_solve_noalloc(gs_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                                         const double* val, const double* b, double* x);

typedef struct gs_csr {
  int32_t n;
  int32_t reserved;
  const int32_t* rowptr;
  const int32_t* colidx;
  const double* val;
} gs_csr;

/* One status per system in per_system (a failing system never aborts the others); returns GS_OK unless the call
 * itself is malformed (GS_INVALID_ARGUMENT) or the device failed. The systems whose structure (pattern and nonzero
 * positions) has a cached device plan are solved together, one launch per structure and a wavefront per system; the
 * others, and any whose pivots that plan cannot keep, one at a time as gs_sparse_solve solves them. */
GS_API gs_status gs_sparse_solve_batch(gs_ctx* ctx, int32_t count, const gs_csr* systems, const double* const* b,
                                       double* const* x, gs_status* per_system);

/* ---- whole Newton-Raphson on the device --------------------------------------------------------------------- */
/* Upload a network: Y-bus, branch stamps, bus specification, Jacobian pattern and the symbolic plan (minimum-degree
 * order and Gilbert-Peierls pivot sequence of the Jacobian at the network's start point under `opts`), resident in
 * HBM until gs_network_free. */
GS_API gs_status gs_network_upload(gs_ctx* ctx, const gs_network* net, const gs_nr_options* opts, gs_net** out);
GS_API void gs_network_free(gs_net* net);

/* Solve `count` instances of the network in one persistent launch per Newton iteration group. results[count];
 * vm_out / va_out (radians), when non-null, receive count * n_bus values. */
GS_API gs_status gs_nr_solve_batch(gs_net* net, int32_t count, const gs_instance* inst, const gs_nr_options* opts,
                                   gs_nr_result* results, double* vm_out, double* va_out);

/* ---- screening: line flows and the contingency classification of solved states ---------------------------------
 * gs_nr_solve_batch followed, on the device, by NewtonShell.ComputeLineFlows (NewtonShell.cs:699-734) over every
 * instance's solved voltages and the numbers ContingencyAnalysis.ScreenStateCore classifies a converged outage by
 * (ContingencyAnalysis.cs:1150-1236), with AssembleResult's losses (NewtonShell.cs:649-690). One summary per instance
 * (meaningful for converged instances); loading (count * n_branch, percent of rate A) and flows (count * n_branch * 4:
 * P_from, Q_from, P_to, Q_to in MW / MVAr) when non-null, index-aligned with the network's branches (the outaged and
 * out-of-service branches carry 0). vm_out / va_out as gs_nr_solve_batch; p_inj / q_inj (count * n_bus, MW / MVAr,
 * PowerFlowResult.PInjMw / QInjMvar) when non-null. */
typedef struct gs_screen_summary {
  double max_loading_pct;       /* largest branch loading, percent of rate A (0: no rated branch) */
  int32_t max_loading_branch;   /* its branch index, or -1 */
  int32_t overloads;            /* branches above 100 % */
  double vmin, vmax;            /* pu */
  int32_t undervoltage, overvoltage;  /* buses below / above their limits */
  double loss_mw, loss_mvar;    /* sums of the bus injections (PowerFlowResult.TotalLossMw / TotalLossMvar) */
} gs_screen_summary;
GS_API gs_status gs_nr_screen_batch(gs_net* net, int32_t count, const gs_instance* inst, const gs_nr_options* opts,
                                   gs_nr_result* results, gs_screen_summary* summaries, double* loading,
                                   double* flows, double* vm_out, double* va_out, double* p_inj, double* q_inj);

/* The plain path: one instance per launch, FP64 LU, the host driving each Newton iteration (the baseline the batched
 * engine is measured against). Same arguments as gs_nr_solve_batch; opts->lu_precision is ignored. */
GS_API gs_status gs_nr_solve_plain(gs_net* net, int32_t count, const gs_instance* inst, const gs_nr_options* opts,
                                   gs_nr_result* results, double* vm_out, double* va_out);

/* The symbolic plan of an uploaded network, for parity checks: sizes first (any pointer may be null). */
GS_API gs_status gs_network_plan_sizes(const gs_net* net, int32_t* m, int32_t* nnz_j, int32_t* nnz_l, int32_t* nnz_u);
GS_API gs_status gs_network_plan_get(const gs_net* net, int32_t* q, int32_t* pivot_row, int32_t* lp, int32_t* li,
                                     int32_t* up, int32_t* ui, int32_t* j_rowptr, int32_t* j_colidx);

/* Shape of the compiled device plan (no counterpart in the reference, which has no device plan): what every instance
 * of a batch walks. stream_bytes counts the refactor and solve record streams and the per-network index arrays the
 * kernel reads; they are shared by the instances of a launch, so their HBM cost per instance depends on cache reuse. */
typedef struct gs_plan_stats {
  int32_t m, nnz_j, nnz_l, nnz_u, levels;
  int32_t refactor_rows;        /* rows of the refactor stream (64 lanes each, padding included) */
  int32_t forward_steps, back_steps;
  int32_t workspace_stride;     /* refactor workspace slots per track */
  int32_t factor_slots;         /* per-instance factor values: 64 per WRITE row of the refactor stream */
  int64_t stream_bytes;
  /* refactor rows by kind (no-op, assign, update, pivot, write; padding not counted) and the lanes of the update
   * rows that carry an update (of 64 per row) */
  int32_t rows_by_kind[5];
  int32_t update_lanes;
  /* solve steps' lanes that carry an entry (forward, back; of 64 per step) */
  int32_t forward_lanes, back_lanes;
  /* 1: the Y-bus is symmetric bit for bit, so each Y cell's lane forms its transposed cell's Jacobian values and a
   * Jacobian column's entries lie together where the refactor gathers them; 0: every lane forms its own cell's (a
   * phase shifter, or a stamp whose zero is signed differently from its transpose's, as a reactance-only transformer's
   * conductance can be) */
  int32_t jacobian_transposed;
} gs_plan_stats;
GS_API gs_status gs_network_plan_stats(const gs_net* net, gs_plan_stats* out);

/* The same symbolic pass on the host alone (no device): the plan gs_network_upload would upload. For callers that
 * prepare plans on CPU threads, and for checking the ordering and pivot sequence without a GPU. */
typedef struct gs_plan gs_plan;
GS_API gs_status gs_plan_build(const gs_network* net, const gs_nr_options* opts, gs_plan** out);
GS_API gs_status gs_plan_sizes(const gs_plan* plan, int32_t* m, int32_t* nnz_j, int32_t* nnz_l, int32_t* nnz_u,
                               int32_t* n_levels);
GS_API gs_status gs_plan_get(const gs_plan* plan, int32_t* q, int32_t* pivot_row, int32_t* lp, int32_t* li,
                             int32_t* up, int32_t* ui, int32_t* j_rowptr, int32_t* j_colidx);
GS_API void gs_plan_free(gs_plan* plan);
/* The fresh symbolic passes the engine runs for paused instances (a static pivot failed its check), on the host alone:
 * for each instance its own network (outages out, loads scaled), Jacobian pattern, minimum-degree order and
 * Gilbert-Peierls plan at its state (vm / va: count * n_bus values, or null for each instance's start point), on up to
 * `threads` host threads, with the orders shared through a cache keyed by the Jacobian pattern (the order depends on
 * the pattern alone, as CpuSparseLuSolver.cs:30-39 caches it). plans[count] receive handles (null where the LU refuses
 * that state; the call then returns GS_SINGULAR); *order_cache_hits (may be null) the orders the cache supplied. No
 * counterpart in the reference, which orders every outage's pattern in its own solver. */
GS_API gs_status gs_plan_fresh_batch(const gs_network* net, const gs_nr_options* opts, int32_t count,
                                     const gs_instance* inst, const double* vm, const double* va, int32_t threads,
                                     gs_plan** plans, int64_t* order_cache_hits);
/* gs_network_plan_stats of the plan gs_network_upload would upload. */
GS_API gs_status gs_plan_shape(const gs_plan* plan, gs_plan_stats* out);
/* Runs the device factorisation and solves of the plan in a host emulation with every index checked, on the
 * Jacobian at the start point, and returns ||J x - f||_inf / ||f||_inf; also checks the level-parallel schedule's
 * split of the same streams (GS_SCHEDULE_LEVELS). GS_INVALID_ARGUMENT (with the reason in *why, 128 bytes) when an
 * index falls outside its buffer or the split is unsound. */
GS_API gs_status gs_plan_selfcheck(const gs_plan* plan, double* rel_residual, char* why);

/* ---- multi-network batches ------------------------------------------------------------------------------------
 * Instances of several uploaded networks (groups) solved together, as ContingencyAnalysis.ScreenBranchesBatched
 * drives its lanes in lock-step (src/GridSim.Core/Analysis/ContingencyAnalysis.Batched.cs:55-266, BatchedStepHub):
 * gs_batch_create lays every group's instances out in HBM once (outages, load-scaled specifications, start points);
 * each gs_batch_run launches every group's kernel on its own stream, forked from and joined into the context's
 * stream, or with use_graph replays one captured hipGraph of those launches. opts: one per group (null: defaults).
 * A group with schedule GS_SCHEDULE_LEVELS runs the level-parallel schedule with its Newton steps unrolled (the
 * start kernel and max_iterations + 1 steps; steps after its last instance ends only dispatch their kernels).
 * Results of the last run per group: results[count], vm_out / va_out (radians, count * n_bus, may be null; the
 * voltages stay in HBM after a run and are copied out by this call). */
typedef struct gs_batch gs_batch;
GS_API gs_status gs_batch_create(gs_ctx* ctx, int32_t n_groups, gs_net* const* nets, const int32_t* counts,
                                 const gs_instance* const* inst, const gs_nr_options* opts, int32_t use_graph,
                                 gs_batch** out);
GS_API gs_status gs_batch_run(gs_batch* batch, double* device_ms);
GS_API gs_status gs_batch_results(gs_batch* batch, int32_t group, gs_nr_result* results, double* vm_out,
                                  double* va_out);
GS_API void gs_batch_free(gs_batch* batch);

/* Autotuner (no counterpart in the reference, whose CPU path has nothing to tune): solves the `count` sample
 * instances under each candidate - FP32 LU with 1 or 2 FP64 refinement steps, or FP64 LU; all instances in one
 * launch or slices of 512 or 256; the solve vector in LDS or in global memory - and writes into *opts the fastest
 * whose verdicts equal the FP64 LU's, whose
 * converged instances meet the tolerance and whose converged voltages lie within 1e-9 (pu, rad) of the FP64 LU's.
 * With opts->schedule = GS_SCHEDULE_GROUP on entry a work-group of 2, 4 or 8 waves per instance (FP64 LU, or FP32
 * LU with one refinement step) is a candidate too, and with GS_SCHEDULE_LEVELS that and the level-parallel schedule
 * (the same arithmetic; the serving loop does not run it). The candidates are timed alone:
 * a caller that runs several groups at once (gs_batch) offers the wider schedules where they do not compete.
 * *best_ms: that candidate's kernel time on the sample. */
GS_API gs_status gs_nr_autotune(gs_net* net, int32_t count, const gs_instance* sample, gs_nr_options* opts,
                                double* best_ms);

/* ---- serving: continuous batching in front of the rpc "solve" contract ----------------------------------------
 * The reference serves gridsim-rpc/1 "solve" requests one at a time (src/GridSim.Cli/RpcCommand.cs:460-497,668-700:
 * each request runs SolverBackend.Solver(...).Solve(model, options) to completion before the next line is read). A
 * gs_server keeps resident networks (Y-bus, symbolic plan, base voltages) in HBM and runs a dispatcher thread over up
 * to `streams` slots, each a running batch of one network on its own stream: a slot opens for the oldest waiting
 * request's network with room for that network's waiting requests (at least 16, at most max_batch positions, as the
 * HBM and pinned-memory budgets admit), and each of its launches runs a few Newton steps of every instance in it (as
 * many as keep a launch near a millisecond). Between launches, instances that finished leave and waiting requests of
 * the same network take the free positions, starting in the next launch; a slot closes when it empties. A request thus
 * joins the device's work at the next Newton-step boundary of a running batch of its network (or in a free slot), and
 * a slot stops taking its own network's requests while an older request of another network waits with no slot free.
 * Every instance is solved by its own wavefront (or work-group), so a result does not depend on what else shares its
 * launch or when it joined. Networks added to a server must outlive it. */
typedef struct gs_server gs_server;

typedef struct gs_server_config {
  int32_t max_batch;        /* instance positions per slot at most (0: 256) */
  int32_t streams;          /* slots, each with its stream and launches in flight, 1..4 (0: 4) */
  int32_t queue_capacity;   /* waiting requests before gs_server_submit returns GS_BUSY (0: 65536) */
  int32_t reserved;
  size_t state_bytes;       /* HBM for the state of in-flight instances (0: 4 GiB) */
  size_t pinned_bytes;      /* pinned host memory for results in flight (0: 256 MiB) */
} gs_server_config;

typedef struct gs_server_stats {
  int64_t submitted, completed, rejected;
  int64_t launches;                                      /* kernel launches (each a few Newton steps of a slot) */
  int64_t queued, in_flight;
  double mean_batch;                                     /* instances per launch */
  double latency_p50_ms, latency_p99_ms, latency_max_ms; /* submit -> result ready, since the last reset */
  double busy_ms;                                        /* wall time with at least one launch in flight */
  int64_t state_bytes_peak;
  int64_t joined;   /* requests that took a position of a batch whose other instances were between Newton steps */
} gs_server_stats;

GS_API gs_status gs_server_create(gs_ctx* ctx, const gs_server_config* cfg, gs_server** out);
/* Registers an uploaded network with the options its requests run under (null: the engine defaults). */
GS_API gs_status gs_server_add_network(gs_server* srv, gs_net* net, const gs_nr_options* opts, int32_t* network_id);
/* Queues one request (its vm0 / va0, if any, are copied). GS_BUSY when queue_capacity requests are waiting. */
GS_API gs_status gs_server_submit(gs_server* srv, int32_t network_id, const gs_instance* inst, uint64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for a request's result: GS_OK and the result (vm / va in pu and radians,
 * p_inj / q_inj in MW / MVAr, n_bus each, any may be null; latency_ms may be null), after which the ticket is
 * forgotten; GS_PENDING when the time ran out; the request's own error status when its launch failed;
 * GS_INVALID_ARGUMENT for an unknown ticket. */
GS_API gs_status gs_server_wait(gs_server* srv, uint64_t ticket, int32_t timeout_ms, gs_nr_result* result,
                                double* vm, double* va, double* p_inj, double* q_inj, double* latency_ms);
GS_API gs_status gs_server_stats_get(gs_server* srv, gs_server_stats* out);
GS_API gs_status gs_server_stats_reset(gs_server* srv);
/* Finishes every queued and in-flight request, then stops the dispatcher and frees the server. */
GS_API void gs_server_free(gs_server* srv);

/* Kernel timing of the last gs_nr_solve_batch on this network (milliseconds, HIP events on its stream). */
GS_API gs_status gs_last_timing(const gs_net* net, double* total_ms, double* kernel_ms, int32_t* launches);

/* How gs_nr_solve_batch would launch the batched kernel for these options (no counterpart in the reference): LU
 * arithmetic, whether the triangular-solve vector sits in LDS, wavefronts (instances) per work-group and the LDS
 * bytes per work-group. Any pointer may be null. */
GS_API gs_status gs_network_launch_shape(const gs_net* net, const gs_nr_options* opts, int32_t* lu_fp32,
                                         int32_t* x_in_lds, int32_t* waves_per_group, int64_t* lds_bytes);

/* Device FDLIBM sincos over n host values (for the bit-parity test against the StrictMath vectors). */
GS_API gs_status gs_device_sincos(gs_ctx* ctx, int64_t n, const double* x, double* s, double* c);

/* Host FDLIBM sincos (the same source compiled for the host). */
GS_API void gs_host_sincos(int64_t n, const double* x, double* s, double* c);

#ifdef __cplusplus
}
#endif
#endif
