// Device-side data layout of the batched Newton-Raphson engine (shared by nr_kernels.hip and engine.cpp).
#pragma once
#include <cstdint>

namespace gs {

// One network pattern + symbolic plan, resident in HBM. Every array is a device pointer.
struct DevPlan {
  int32_t n, m, nnz_y, nnz_j, nnz_l, nnz_u, max_col, n_levels;
  // network pattern
  const int32_t *y_rp, *y_ci, *y_diag;
  const double *y_g, *y_b;            // base Y values (instances patch at most kMaxPatch cells)
  const int32_t* slots;               // [4 nnz_y] J CSR position of H, N, M, L of each Y cell, or -1
  const int32_t *ppos, *qpos;         // equation numbering
  const int32_t *j_rp, *j_ci;         // Jacobian CSR pattern (for the FP64 refinement residual)
  // base bus specification
  const double *vm0, *va0;            // start point (BuildBusSpec under the upload options)
  double base_mva;
  // plan (see model.h DevicePlanHost)
  const int32_t *col_order, *level_end;     // columns sorted by dependency level; per level one-past-end index
  const int32_t *col_a_ptr, *a_src, *a_slot;
  const int32_t *upd_ptr, *upd_slot;
  const int32_t *lp, *li, *up, *ui;
  const int32_t *pinv, *q, *prow;     // row permutation, column permutation, pivot row of each step
  // level chunks of the refactor (model.h DevicePlanHost)
  int32_t wstride, n_chunks;
  const int32_t *ch_track_ptr, *ch_row_ptr, *ch_a_ptr, *ch_fence, *tr_col, *ch_a_ws, *ch_a_src;
  const int32_t *row_rec, *row_lsrc;
  // level-scheduled triangular solves
  int32_t fw_levels, bw_levels;
  const int32_t *fw_order, *fw_level_end, *lr_ptr, *lr_col, *lr_idx;
  const int32_t *bw_order, *bw_level_end, *ur_ptr, *ur_col, *ur_idx;
  const int32_t* piv_pos;             // per step: candidates before the pivot in reach order
};

constexpr int kTracks = 8, kTrackLanes = 8;

constexpr int kMaxPatch = 8;

struct Patch {
  int32_t idx;      // Y cell
  int32_t pad;
  double g, b;
};

// Per-instance mutable state: lives in device memory across launches, so the host can resume an instance.
struct InstState {
  int32_t iter;         // Newton steps taken so far
  int32_t done;         // 1: finished (status valid); 2: paused for a fresh symbolic pass
  int32_t converged, diverged;
  int32_t nonfinite, step_failures;
  int32_t fail_column, pad;
  double start_mis, max_mis, min_ratio, pad2;
  // time in each phase (device wall-clock ticks): mismatch + Jacobian, refactor, solves + refinement, update
  uint64_t t_mismatch, t_refactor, t_solve, t_update;
};

struct InstDesc {
  const DevPlan* plan;
  InstState* state;
  const double *pspec, *qspec;        // specified injections: the network's, or this instance's load-scaled copy
  const double *vm_start, *va_start;  // start point: the network's, or this instance's own
  int32_t n_patch, pad;
  Patch patch[kMaxPatch];
  // per-instance buffers (device)
  double *vm, *va, *p, *q;
  double *f, *dx, *r, *jv;            // f, dx, residual, FP64 Jacobian values (CSR order)
  void *lx, *ux;                      // L / U values in the LU precision
  void* xg;                           // global triangular-solve workspace (used when LDS cannot hold x)
};

// One CSR system of the sparse-solver seam, solved on a plan made from its own factorisation.
struct SeamDesc {
  const DevPlan* plan;
  const double* a;         // CSR values
  const double* b;
  double* x;
  double *lx, *ux, *xs, *rowfloor;
  int32_t* status;         // 0 solved, 1 refused (a pivot would move, or is at or below its row floor)
  const int32_t* piv_pos;  // per step: candidates (in reach order) before the pivot
};

struct NrParams {
  double tolerance, diverge_limit, growth_tol;
  int32_t max_iterations, refine_steps;
  int32_t steps_this_launch;          // Newton steps one launch may take (the plain path takes 1)
  int32_t x_in_lds;                   // 1: triangular-solve workspace in LDS
  int32_t init;                       // 1: reset the instance's counters before stepping
  int32_t load_start;                 // 1: copy the start point into the iterate before stepping
};

}  // namespace gs
