/* C ABI of the CPU oracle: a C++ restatement of GridSim's CPU sparse Newton-Raphson load flow and N-1 driver.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is the checker the tests, the smoke run and the benchmark's CPU-baseline leg
 * compare the HIP engine against; nothing in the product links, loads or calls it.
 *
 * Every function restates a reference file (cited in oracle.cpp); numbering, summation order and branch structure
 * follow the reference line by line so the FP64 results are the reference's bits wherever the inputs are. */
#ifndef GRIDSIM_ORACLE_H
#define GRIDSIM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bus types follow the reference enum order (Model.cs:4-12). */
enum { ORC_SLACK = 0, ORC_PV = 1, ORC_PQ = 2 };

/* PowerFlowStatus (PowerFlowResult.cs) as the engine reports it. */
enum { ORC_CONVERGED = 0, ORC_MAX_ITERATIONS = 1, ORC_DIVERGED = 2, ORC_NONFINITE = 3 };

typedef struct orc_network {
  int32_t n_bus, n_branch, n_gen;
  double base_mva;
  const int32_t* bus_type;
  const double *pd, *qd, *gs, *bs, *vm, *va_deg;
  const int32_t *br_from, *br_to;
  const double *br_r, *br_x, *br_b, *br_tap, *br_shift_deg;
  const uint8_t* br_in_service;
  const int32_t* gen_bus;
  const double *gen_pg, *gen_qg, *gen_vg;
  const double *gen_qmin, *gen_qmax;   /* MVAr; may be null (no limits) */
} orc_network;

typedef struct orc_options {
  double tolerance;      /* PowerFlowOptions.Tolerance (1e-8) */
  int32_t max_iterations;
  double diverge_limit;  /* PowerFlowOptions.DivergeLimitPu (10) */
  int32_t flat_start;
  int32_t enforce_q_limits;  /* PowerFlowOptions.EnforceQLimits: the reactive-limit outer loop */
  int32_t max_q_passes;      /* PowerFlowOptions.MaxQLimitPasses (100) */
  int32_t retry_dc_start;    /* PowerFlowOptions.RetryFromDcStart: the retry from a DC-angle start */
} orc_options;

/* NewtonRetryOutcome (PowerFlowResult.cs:24-41). */
enum { ORC_RETRY_NOT_TRIED = 0, ORC_RETRY_CONVERGED = 1, ORC_RETRY_DID_NOT_CONVERGE = 2, ORC_RETRY_REJECTED = 3,
       ORC_RETRY_UNAVAILABLE = 4 };

typedef struct orc_result {
  int32_t status;          /* ORC_CONVERGED ... */
  int32_t iterations;
  double max_mismatch;
  double solve_seconds;    /* wall time of this solve on its worker thread */
  int32_t step_failures;   /* steps the LU could not solve (ILU0/BiCGSTAB fallback taken) */
  int32_t dc_retry;        /* ORC_RETRY_*: PowerFlowResult.NewtonRetry */
} orc_result;

/* FDLIBM sincos over an array. */
void orc_sincos(const double* x, int64_t n, double* s, double* c);

/* .NET System.Numerics.Complex division (Smith's formula), for the stamp test. */
void orc_complex_div(double a, double b, double c, double d, double* re, double* im);

/* Branch stamp (AdmittanceStamp.Branch): out[8] = Yff, Ytt, Yft, Ytf as (re, im) pairs. */
void orc_branch_stamp(double r, double x, double b, double tap, double shift_deg, double* out8);

/* Y-bus of the network with branch `skip` (or -1) removed: call with null arrays to get nnz. */
int32_t orc_ybus(const orc_network* net, int32_t skip, int32_t* rowptr, int32_t* colidx, double* g, double* b);

/* Bus spec (BuildBusSpec): effective types, start Vm, Va (rad), Pspec, Qspec. */
void orc_bus_spec(const orc_network* net, int32_t flat_start, int32_t* type, double* vm, double* va, double* pspec,
                  double* qspec);

/* Mismatch f and P, Q at a state (ComputeInjections + residual); returns m. */
int32_t orc_mismatch(const orc_network* net, int32_t skip, const double* vm, const double* va, double* p, double* q,
                     double* f);

/* Jacobian structure + refill at a state: call with null arrays for (m, nnz). */
int32_t orc_jacobian(const orc_network* net, int32_t skip, const double* vm, const double* va, int32_t* m_out,
                     int32_t* rowptr, int32_t* colidx, double* val);

/* Minimum-degree ordering of a CSR pattern (MinimumDegreeOrdering.Order). */
void orc_md_order(int32_t n, const int32_t* rowptr, const int32_t* colidx, int32_t* q);

/* Gilbert-Peierls factorisation of CSR matrix values with ordering q.
 * keep_zeros: 0 = SparseLu.ToCsc (explicit zeros dropped), 1 = ToCscKeepingZeros.
 * Returns an opaque handle, null when singular. */
typedef struct orc_lu orc_lu;
orc_lu* orc_lu_factor(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val, const int32_t* q,
                      int32_t keep_zeros, double diagonal_preference);
void orc_lu_sizes(const orc_lu* h, int32_t* lnz, int32_t* unz, int32_t* reach_nnz);
/* lp[n+1], li[lnz] (pivot steps), lx, up[n+1], ui[unz], ux, pinv[n], q[n], pivot_row[n], reach_ptr[n+1], reach[] */
void orc_lu_get(const orc_lu* h, int32_t* lp, int32_t* li, double* lx, int32_t* up, int32_t* ui, double* ux,
                int32_t* pinv, int32_t* q, int32_t* pivot_row, int32_t* reach_ptr, int32_t* reach);
void orc_lu_solve(const orc_lu* h, const double* b, double* x);
/* Plan replay (SparseLuFactorization.Refactor) on new CSR values of the same structure, zeros kept (the
 * fixed-pattern solver's CSC); null if the pattern differs, a pivot would move, or a pivot is singular. */
orc_lu* orc_lu_refactor(const orc_lu* plan, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                        const double* val);
void orc_lu_free(orc_lu* h);

/* CpuSparseLuSolver.Solve: returns 0 on success, 1 singular (null), 2 non-finite right-hand side. */
int32_t orc_sparse_solve(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val,
                         const double* b, double* x);
/* ILU(0) of a CSR matrix with sorted columns (Sparse.cs:230-300): lu_out gets the factors in the pattern's positions
   (L below the diagonal, unit diagonal implied; U from the diagonal on). 0 when a row has no diagonal entry. */
int32_t orc_ilu0(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val, double* lu_out);
/* BiCGSTAB preconditioned by that ILU(0) (Sparse.cs:302-405): 1 when x solves A x = b to tol (true residual). */
int32_t orc_bicgstab(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* val, const double* b,
                     double tol, int32_t max_iter, double* x);

/* One power-flow solve (SparseNewtonRaphsonPowerFlow.Solve: reactive limits and the DC-start retry as the options say).
 * vm_out / va_out (rad) receive the final state, may be null. */
void orc_solve(const orc_network* net, int32_t skip, const orc_options* opt, double* vm_out, double* va_out,
               orc_result* res);

/* N-1 screen over `count` branch outages on `threads` worker threads (ScreenBranchesParallel's shape: one solver
 * per worker, one outage per task). vm_out / va_out may be null, else count * n_bus each. */
void orc_screen(const orc_network* net, const int32_t* outages, int32_t count, const orc_options* opt,
                int32_t threads, orc_result* res, double* vm_out, double* va_out);

/* N-k screen: outage set k is branches[offsets[k] .. offsets[k + 1]) (WithoutMany's index list); as orc_screen. */
void orc_screen_sets(const orc_network* net, const int32_t* offsets, const int32_t* branches, int32_t count,
                     const orc_options* opt, int32_t threads, orc_result* res, double* vm_out, double* va_out);

/* NewtonShell.TrySeedDcAngles on the ordinary start: 1 and va (rad, n_bus) the seeded angles, or 0 (unavailable). */
int32_t orc_dc_seed(const orc_network* net, int32_t flat_start, double* va);

#ifdef __cplusplus
}
#endif
#endif
