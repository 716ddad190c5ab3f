// Device-side data layout of the batched Newton-Raphson engine (shared by nr_kernels.hip and engine.cpp).
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace gs {

// One network pattern + symbolic plan, resident in HBM. Every array is a device pointer. The struct itself is passed
// to the kernels by value (kernel argument), so its fields sit in scalar registers and its pointers are known to
// address global memory.
struct DevPlan {
  int32_t n, m, nnz_y, nnz_j, nnz_l, nnz_u, max_col, n_levels;
  // network pattern
  const int32_t *y_rp, *y_diag;
  const int2* y_ij;                   // per Y cell: (row, column)
  // per Y cell, then per bus again for its diagonal cell: base value (g, b) (instances patch at most kMaxPatch
  // cells), and per Y cell's (bus's) lane the J CSR positions of the H, N, M, L entries it forms, and their A
  // positions (-1: not in the Jacobian)
  const double2* y_gb;
  const int4* slots;
  const int4* aslots;
  const int32_t *ppos, *qpos;         // equation numbering
  const int32_t *j_rp, *j_ci;         // Jacobian CSR pattern (for the FP64 refinement residual)
  // The wavefront schedule keeps the Newton step dx in pivot-step order (the triangular solves' own order, so the
  // solve stores it contiguously): each bus's P and Q unknowns' and each Jacobian column's positions there (qinv of
  // ppos / qpos / j_ci; -1 where ppos / qpos are)
  const int32_t *ppos_k, *qpos_k, *j_ci_k;
  // base bus specification
  const double *vm0, *va0;            // start point (BuildBusSpec under the upload options)
  double base_mva;
  // plan (model.h DevicePlanHost)
  const int32_t *pinv, *q, *prow;     // row permutation, column permutation, pivot row of each step
  const int32_t* piv_pos;             // per step: candidates before the pivot in reach order
  // The refactor: one stream of rows, each one operation of the wavefront with one int4 record per lane (ASSIGN,
  // UPDATE, PIVOT, WRITE; model.h DevicePlanHost), padded with no-op rows to a multiple of kStreamPad rows plus two
  // more kStreamPad, so the kernel loads ahead without bounds checks. n_a: A entries (the first part of vals).
  // y_tr: the lane of Y cell (i, j) forms the Jacobian values of cell (j, i) (a symmetric Y-bus; slots and aslots
  // then name that cell's entries), so that a Jacobian column's values sit together.
  int32_t wstride, n_rs, n_a, y_tr;
  // [(n_rs + kRsPadRings kStreamPad) / 2][64] pairs of rows: lane l of pair q holds (x, y) of rows 2q and 2q + 1;
  // y = 24-bit field | (kind | flags) << 24
  const int4* rs_rec;
  const int32_t* rs_wr0;              // per row: the rows before it that write factors (WRITE, PIVOT + WRITE)
  const int32_t* ch_a_src;            // per A entry: J CSR position (the seam fills vals from its matrix)
  // Triangular-solve step streams: per lane and step (x, y): PRODUCT (col, value index | kind << 24) and CHAIN
  // (start | len << 16, row | (kind | flags) << 24) steps (model.h DevicePlanHost). The factors live where the
  // refactor's WRITE rows stored them, 64 per WRITE row (the factor part of vals, n_fac slots after the n_a A values;
  // a PRODUCT step's value index is a position there), the pivots also in dval by pivot step.
  // ws_min: the refactor workspace's slots at least (the blocked mismatch pass's scratch: four slots per cell of the
  // longest Y row, at least kFusedScratch)
  int32_t fs_steps, bs_steps, n_fac, ws_min;
  const int2 *fs_rec, *bs_rec;
  // Level-parallel schedules: first row of each refactor chunk (+ the end), first step of each solve pass (+ the end).
  // The streams carry two kStreamPad rings of padding, so one chunk's or pass's range rounded up to whole rings can
  // load a ring ahead.
  const int32_t *rs_chunk_row, *fs_pass_step, *bs_pass_step;
  // Screening: per branch its end buses (from = -1: out of service), pi-model stamp (Yff, Yft, Ytf, Ytt as re, im)
  // and rate A; per bus its own voltage limits (NaN = the operational band).
  int32_t nb, n_mb;
  // The blocked mismatch pass's blocks of buses: block b is buses [mb_ptr[b], mb_ptr[b + 1]), at most 64, whose rows
  // hold at most ws_slots x 8 / 16 cells (mb_ptr4, n_mb4: x 4 / 16, for FP32 workspaces); n_mb = 0: a network without
  // the Newton pattern (the wavefront and work-group schedules do not run then)
  const int32_t *mb_ptr, *mb_ptr4;
  // The group schedule's segments (x, y, z) = (first, end, split) of the refactor's chunks, then the forward and back
  // solves' passes: a level of two or more chunks (passes) split between the waves (z = 1), or a run of one-chunk levels
  // walked by one wave (z = 0).
  int32_t n_gseg_rs, n_gseg_fs, n_gseg_bs, n_mb4;
  const int4* gseg;
  const int32_t* j_diag;              // per Jacobian row: CSR position of its diagonal (the iterative fallback)
  const int2* br_ends;
  const double* br_stamp;
  const double *br_rate, *bus_lo, *bus_hi;
};

// refactor row kinds and flags, write kinds (model.h DevicePlanHost)
constexpr int kRsNop = 0, kRsAssign = 1, kRsUpdate = 2, kRsPivot = 3, kRsWrite = 4;
constexpr int kRsKindMask = 7, kRsClear = 8, kRsLevelEnd = 16, kRsPivotWrite = 128;
constexpr int kRsWidthShift = 5;   // a chunk's lanes per track: kTrackLanes << ((meta >> kRsWidthShift) & 3)
constexpr int kRsMetaShift = 24, kRsIndexMask = (1 << kRsMetaShift) - 1;
constexpr int kWriteU = 1, kWritePivot = 2, kWriteL = 3;
// solve step kinds and CHAIN flags (model.h DevicePlanHost)
constexpr int kStepProduct = 1, kStepChain = 2, kStepKindMask = 15;
constexpr int kChainBegin = 16, kChainEnd = 32, kChainIdle = 64;
constexpr int kProductLate = 16, kStepXAhead = 8;   // a PRODUCT step that loads x at the step, not ahead
// load-ahead distance of the row and step streams, and the padding each stream carries; rows of the refactor load
// their value kRsValueAhead rows ahead (the plan places its rows for that distance)
constexpr int kStreamPad = 16;
constexpr int kRsPadRings = 3;   // rings of no-op rows after the refactor stream (rounded to whole rings)
constexpr int kRsValueAhead = 8;

constexpr int kTracks = 8, kTrackLanes = 8;
// the refactor workspace: kTracks columns of wstride slots, then one spare slot per lane (model.h DevicePlanHost)
constexpr int kWsSpare = 64;
// slots of the refactor workspace (the triangular-solve vector follows it in LDS)
// (at least ws_min slots: the blocked mismatch pass keeps a block's injection terms there, two FP64 per cell, and
// every Y row fits a block)
constexpr int kFusedScratch = 512;
__host__ __device__ inline int ws_slots(const DevPlan& p) {
  const int w = kTracks * p.wstride + kWsSpare;   // even: wstride is odd; ws_min even
  return w > p.ws_min ? w : p.ws_min;
}
// Entries of each factor array per instance. Every stream record's 24-bit field indexes the array its value load
// reads (vals for the refactor rows, its factor part for the solve steps), also in the rows and steps that take no
// value (pivot steps, solve rows: below m, and the factor part holds at least one slot per column), so the loads need
// no select; the pivots' array is read at a CHAIN step's row only.
__host__ __device__ inline int imax3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }
__host__ __device__ inline int vals_slots(const DevPlan& p) { return imax3(p.n_a + p.n_fac, p.m, 1); }
__host__ __device__ inline int dval_slots(const DevPlan& p) { return imax3(p.m, 1, 1); }

// Level-parallel schedule: a level of at most this many chunks (solve passes) is cheaper walked in sequence by one
// work-group, together with its narrow neighbours, than given a kernel of its own (measured on the 70,000-bus grid).
constexpr int kNarrow = 2, kNarrowSolve = 4;
// The work-group-per-instance schedule splits levels of more chunks (passes) than these between its waves; the
// narrower ones are walked by one wave without barriers.
constexpr int kGroupNarrow = 1, kGroupNarrowSolve = 1;

// Where a level-parallel step may fork: the refactor's levels from trunk_level on (the final run of narrow levels)
// go to the side stream while the forward solve's levels before fs_level (which read no factor of those columns)
// proceed; the two branches join before forward level fs_level. side == nullptr: no fork.
struct WideSplit {
  int trunk_level = 0, fs_level = 0;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// Y cells an instance patches in place (every cell an N-1 or N-2 outage touches fits); an outage touching more cells
// gives its instance a Y value array of its own (InstDesc::y_gb_own).
constexpr int kMaxPatch = 8;
// outaged branches of one instance (the ABI's GS_MAX_OUTAGES)
constexpr int kMaxOutages = 8;

struct Patch {
  int32_t idx;      // Y cell
  int32_t pad;
  double g, b;
};

// Per-instance mutable state: lives in device memory across launches, so the host can resume an instance.
struct InstState {
  int32_t iter;         // Newton steps taken so far
  int32_t done;         // 0: running (paused between launches); 1: finished (status valid); 2: paused for a fresh
                        // symbolic pass; kDoneFresh: not started (the next launch starts it, whatever its init)
  int32_t converged, diverged;
  int32_t nonfinite, step_failures;
  int32_t fail_column;
  int32_t engine_stop;  // GS_STOP_* set on the device (the shared solve vector never came free)
  double start_mis, max_mis, min_ratio;
  double last_from;     // the mismatch before the last step taken (NaN: none), for the DC retry's stall test
  // time in each phase (device wall-clock ticks): mismatch + Jacobian, refactor, solves + refinement, update
  uint64_t t_mismatch, t_refactor, t_solve, t_update;
  // level-parallel schedules: reductions across the work-groups of one instance (bit patterns of non-negative
  // doubles, which order as unsigned integers; a NaN orders above +inf)
  uint64_t mis_bits, dx_bits, ratio_bits;
  int32_t jbad, wpad;
  uint64_t jn_bits;     // level-parallel: max |J entry| of the step being formed (bits), and of the current step
  double jnorm;
};

// InstState::done of an instance the next launch starts afresh (its start point, its counters reset) even when the
// launch itself resumes the others: how the serving loop adds requests to a running batch between Newton steps.
constexpr int kDoneFresh = 3;

// Control block of a level-parallel run: how many of its instances have finished (or paused).
struct WideCtl {
  int32_t finished;
  int32_t pad[15];
};

struct InstDesc {
  InstState* state;
  const double *pspec, *qspec;        // specified injections: the network's, or this instance's load-scaled copy
  const double *vm_start, *va_start;  // start point: the network's, or this instance's own
  int32_t n_patch, n_out;             // patched Y cells; outaged branches
  Patch patch[kMaxPatch];
  int32_t out[kMaxOutages];           // the outaged branches (their flows are not reported)
  const double* y_gb_own;             // this instance's own Y values (laid out as DevPlan::y_gb), or null: the
                                      // network's with the patches
  // per-instance buffers (device)
  double *vm, *va, *p, *q;
  double *f, *dx, *r, *jv;            // f, dx, residual, FP64 Jacobian values (CSR order)
  double* tpq;                        // per Y cell: the cell's two injection terms (P, Q)
  void* vals;                         // LU precision: [Jacobian in the refactor's record order | the factors]
  void* dval;                         // the pivots in the LU precision, by pivot step
  void* xg;                           // global triangular-solve workspace (used when LDS cannot hold x)
};

// One CSR system of the sparse-solver seam, solved on a plan made from its own factorisation.
struct SeamDesc {
  const double* a;         // CSR values
  const double* b;
  double* x;
  double *vals, *dval, *xs, *rowfloor;
  int32_t* status;         // 0 solved, 1 refused (a pivot would move, or is at or below its row floor)
};

struct NrParams {
  double tolerance, diverge_limit, growth_tol;
  int32_t max_iterations, refine_steps;
  int32_t steps_this_launch;          // Newton steps one launch may take (the plain path takes 1)
  int32_t init;                       // 1: reset the instance's counters before stepping
  int32_t load_start;                 // 1: copy the start point into the iterate before stepping
  int32_t n_inst;                     // instances of the launch (set by the launcher)
  int32_t trust_first;                // 1: the launch's first refactor runs on a plan the host made at this very
                                      // state (its pivots are the factorisation's own): no growth or floor check
  int32_t pad;
};

// The pivot floor of the Newton refactor: a static pivot is kept only above kPivotFloorRel x the largest |entry| of
// the Jacobian (at or above every row's own floor, NumericTolerances.RelativePivotFloor x max |A(r, :)|, so the device
// never accepts a pivot the reference's factorisation would refuse; a pivot below it goes to the host).
constexpr double kPivotFloorRel = 1e-12;
// In FP32 LU a pivot that rounding alone made non-zero (a column the FP64 factorisation would find singular) stands
// about eps32 x its column's scale, far above that floor: the FP32 refactor keeps a pivot only above 2^-20 x max |J|
// and hands anything smaller to the host's FP64 symbolic pass (real Jacobians' smallest pivots lie near 1e-4 x max |J|).
constexpr double kPivotFloorRelF32 = 9.5367431640625e-07;
template <typename T>
__host__ __device__ constexpr double pivot_floor_rel() { return sizeof(T) == 4 ? kPivotFloorRelF32 : kPivotFloorRel; }

// Instances per work-group of the batched kernel: 1, or 2 wavefronts sharing one LDS solve vector (each wave takes
// it in turn for its triangular solves), which halves the solve vector's LDS per instance.
constexpr int kPairWaves = 2;
// Spins of a wave waiting for the shared solve vector before it gives up (the instance then ends as failed); the
// holder releases it within one pair of triangular solves, far below this bound.
constexpr int kPairSpinLimit = 1 << 24;

}  // namespace gs
