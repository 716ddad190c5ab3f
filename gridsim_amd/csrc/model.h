// Host-side model and symbolic analysis of the sparse Newton-Raphson engine.
//
// Everything here runs once per network pattern (or once per fallback instance): the Y-bus with its per-cell
// contribution lists, the bus specification, the equation numbering, the Jacobian's structural pattern, the
// minimum-degree column ordering and the Gilbert-Peierls plan (pivot order, reach order, L/U patterns). The
// per-iteration numeric work runs on the GPU (nr_kernels.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "gridsim_hip.h"

namespace gs {

using std::vector;

struct Cx { double re, im; };

// The pi-model stamp of one branch (AdmittanceStamp.cs:41-75).
struct Stamp { Cx yff, ytt, yft, ytf; };
Stamp branch_stamp(double r, double x, double b, double tap_ratio, double shift_deg);

// A Y-bus in CSR with sorted columns and an explicit diagonal in every row (ComplexCsr, Sparse.cs:134-222),
// plus, per stored cell, the list of stamps summed into it in insertion order (branches in model order, then the bus
// shunt). The list lets one cell be re-summed without one branch, which is bit-identical to rebuilding the Y-bus
// without that branch (ContingencyAnalysis.WithoutMany rebuilds Y from scratch).
struct YBus {
  int n = 0;
  vector<int> rp, ci, diag;
  vector<double> g, b;
  vector<int> contrib_ptr;     // [nnz + 1]
  vector<int> contrib_branch;  // branch index, or -1 for the bus shunt
  vector<double> contrib_g, contrib_b;
  int find(int row, int col) const;
};

// Network copied from the C ABI description (only in-service branches matter to the solve).
struct Network {
  int n = 0, nb = 0, ng = 0;
  double base = 100.0;
  vector<int> bus_type;
  vector<double> pd, qd, gs, bs, vm, va_deg;
  vector<int> br_from, br_to;
  vector<double> br_r, br_x, br_b, br_tap, br_shift;
  vector<uint8_t> br_in;
  vector<int> gen_bus;
  vector<double> gen_pg, gen_qg, gen_vg;
  vector<double> br_rate, bus_lo, bus_hi;   // rate A (MVA, 0 = unrated); own voltage limits (NaN = the band)
  vector<double> gen_qmin, gen_qmax;        // MVAr (-inf / +inf: none)
};

Network copy_network(const gs_network& src);
// The Y-bus of `net`; `skip` (or -1) removes one branch as an outage does.
YBus build_ybus(const Network& net, int skip);

// NewtonShell.BuildBusSpec (NewtonShell.cs:433-486) for constant-power buses.
struct BusSpec {
  vector<int> type;                       // effective types (a PV bus without a generator reads PQ)
  vector<double> vm, va, pspec, qspec;    // start magnitudes, start angles (rad), specified injections (pu)
  vector<double> qd, vg, qmin, qmax;      // load MVAr, voltage set point, summed generator limits (MVAr)
};
BusSpec build_bus_spec(const Network& net, bool flat_start, double load_scale);

// Equation numbering (NewtonShell.cs:327-336) and the Jacobian's structural pattern with the map from each Y cell's
// four blocks (H, N, M, L) to their CSR positions, or -1 (SparseNewtonRaphsonPowerFlow.cs:296-332).
struct JacobianPattern {
  int m = 0;
  vector<int> ppos, qpos;
  vector<int> rp, ci;
  vector<int> slots;   // [4 * nnzY]
};
JacobianPattern build_jacobian_pattern(const YBus& y, const vector<int>& type);

// Injections and Jacobian values at a state (ComputeInjections :269-289, RefillJacobian :342-380), on the host:
// the numbers the symbolic pass factors.
void host_injections(const YBus& y, const double* vm, const double* va, double* p, double* q);
void host_jacobian(const YBus& y, const JacobianPattern& jp, const double* vm, const double* va, const double* p,
                   const double* q, double* val);

// Minimum-degree ordering of a square CSR pattern (MinimumDegreeOrdering.cs:23-136).
vector<int> minimum_degree(int n, const int* rowptr, const int* colidx);

// The Gilbert-Peierls plan of one factorisation (SparseLuFactorization.Factor with keepPlan, SparseLuPlan):
// CSC pattern of A with every structural cell kept (SparseLu.ToCscKeepingZeros), the ordering, each column's reach in
// DFS topological order, each step's pivot row, and the L and U patterns (rows as pivot steps).
struct Plan {
  int m = 0;
  double diagonal_preference = 0.0;
  vector<int> ap, ai, a_src;     // CSC of A; a_src[e] = CSR position of CSC entry e
  vector<int> q, pinv, pivot_row;
  vector<int> reach_ptr, reach;
  vector<int> lp, li, up, ui;    // li, ui: pivot steps
  vector<int> li_orig;           // L rows as original rows
  vector<double> lx, ux;         // the factors of the values the plan was made from (L unit diagonal first)
};

// Factor A (CSR values `val`) with ordering q and partial pivoting under the given diagonal preference, on the CSC
// that keeps every structural cell (keep_zeros, the fixed-pattern solver's) or drops exact zeros (the default
// solver's). Returns false when a pivot is at or below its row's floor (singular) or an entry is not finite.
bool gp_plan(int m, const vector<int>& rowptr, const vector<int>& colidx, const double* val, const vector<int>& q,
             double diagonal_preference, bool keep_zeros, Plan& plan, std::string& why);

// SparseLuFactorization.Solve on the factors a plan was made from: out = A^{-1} b.
void plan_solve(const Plan& f, const double* b, double* out);

// The DC power-flow angle start of the retry (NewtonShell.TrySeedDcAngles): va (start angles, radians, in and out) gets
// the reactance-only DC angles of the network offset by the slack's start angle; false (va untouched) when the
// network has no single slack, is not one DC island, or its susceptance does not factor.
bool dc_seed_angles(const Network& net, const vector<int>& type, vector<double>& va);
// NewtonShell.MaxBranchAngleDeg: the largest in-service branch angle (degrees) of the state va (radians).
double max_branch_angle_deg(const Network& net, const double* va);

// The plan compiled into the flat arrays the refactor and solve kernels walk (see nr_kernels.hip).
struct DevicePlanHost {
  int m = 0, nnz_l = 0, nnz_u = 0, max_col = 0, n_upd = 0;
  vector<int> col_a_ptr;     // [m+1] into a_src/a_slot, per pivot step
  vector<int> a_src;         // J CSR position of each A entry of the column
  vector<int> a_slot;        // workspace slot of that entry
  vector<int> upd_ptr;       // [nnz_u + 1]: per U entry, range of update slots (L(:,j) off-diagonals)
  vector<int> upd_slot;      // slot in column k of each L(i,j) row
  vector<int> lp, up;        // L, U column pointers (L unit diagonal first, U pivot last)
  vector<int> li, ui;        // pivot-step rows
  vector<int> pivot_slot;    // [m] slot of the pivot (= nU_k - 1)
  vector<int> pinv, q, prow;
  vector<int> col_order, level_end;  // columns by dependency level (a column needs the L columns its U row set names)
  vector<int> piv_pos;       // per step: candidates before the pivot in reach order (the exact pivot-search check)
  int n_levels = 0;

  // Level chunks of the refactor: up to kTracks columns of one dependency level factored side by side by one
  // wavefront, kTrackLanes lanes per column. Each chunk's updates are laid out in rows of 64 records (lane l of a row
  // serves track l / kTrackLanes), so every lane knows its record's address without a search.
  static constexpr int kTracks = 8, kTrackLanes = 8;
  int wstride = 0, n_chunks = 0, n_rows = 0;
  // Lanes per track of each chunk: kTrackLanes, or more for a level of fewer than kTracks columns (16, 32, 64; the
  // chunk's tracks then number 64 / TL). Every row of a chunk carries its width code in its flags.
  vector<int> ch_tl;
  static constexpr int kRsWidthShift = 5;
  static int width_code(int tl) { return (tl == 8 ? 0 : tl == 16 ? 1 : tl == 32 ? 2 : 3) << kRsWidthShift; }
  static int track_lanes(int meta) { return 8 << ((meta >> kRsWidthShift) & 3); }
  static constexpr int kSpareSlots = 64;   // one spare workspace slot per lane, after the kTracks workspaces
  int spare(int lane) const { return kTracks * wstride + lane; }
  vector<int> ch_track_ptr, ch_row_ptr, ch_a_ptr, ch_fence;  // per chunk; ch_fence: last chunk of its level
  vector<int> tr_col;                                        // column of each track, chunk order
  vector<int> ch_a_ws, ch_a_src;                             // A entries: workspace index, J CSR position
  vector<int> row_rec;                                       // (dst ws) | (xj ws << 16), 0xFFFF dst = no update
  vector<int> row_lsrc;                                      // L value index (CSC) of the record
  // The refactor as one stream of rows (every chunk has at least one row) with per-row flags, so the kernel can
  // keep loads for the next rows in flight: kRowBegin / kRowEnd mark a chunk's first / last row, kRowLevelEnd the
  // last row of a dependency level. Per-track metadata and the A values' order are laid out for one-hop loads.
  static constexpr int kRowBegin = 1, kRowEnd = 2, kRowLevelEnd = 4;
  vector<int> row_flags, row_level;                 // [n_rows]
  vector<int> tr_u0, tr_nu, tr_l0, tr_nl;           // per track (chunk order)
  vector<int> j2a;                                  // [nnz_j]: A position of each Jacobian CSR entry, or -1
  vector<int> a_pos;                                // per A record: its position in the A part of vals
  vector<int> y_tr;                                 // Newton plans: the Y cells' transposes (symmetric Y), or empty
  // The refactor as the device runs it: one stream of rows, each row one operation of the wavefront with one record
  // (x, y: 2 ints) per lane, so the kernel loads every row's record and value a fixed number of rows ahead. The
  // record's y holds a 24-bit field (y & kRsIndexMask) and the row's kind and flags (the same in every lane) above it:
  //   ASSIGN  ws[x] = vals[y]                     (a chunk's A entries, y = a_pos; kRsClear on its first: zero the
  //                                               workspace)
  //   UPDATE  ws[x & 0xFFFF] -= vals[y] * ws[x >> 16]
  //   (a lane with nothing to assign or update is pointed at its own spare slot spare(lane) past the tracks'
  //   workspaces, so the device runs these rows with every lane active and no branch)
  //   PIVOT   x = pivot slot ps | (candidates nc << 16) (the candidates are slots ps + 1 .. ps + nc), y = pivot step;
  //           x = -1: idle track. With kRsPivotWrite (every chunk's) it also does the chunk's write row 0, as a WRITE
  //           row would (pivot_write_lane), so the chunk's WRITE rows start at w = 1
  //   WRITE   slot: track base + w * TL + lane % TL for the chunk's w-th WRITE row (TL lanes per track); every lane
  //           stores its slot to vals[y] (y = n_a + 64 * the stream's WRITE-row count + lane, so one WRITE row is one
  //           contiguous 64-value store): an L slot divided by its column's pivot, a U slot or the pivot as it is, an
  //           idle lane whatever its slot holds (a position nothing reads); kind x >> 24: U, PIVOT (also dval[x &
  //           kRsIndexMask] = ws, by pivot step), L, 0: idle
  // vals = [A values (n_a slots) | the factors, 64 per WRITE row (n_fac)], so y indexes one array for every row kind.
  // kRsLevelEnd marks the last row of a dependency level (the device fences there). No-op rows are inserted so that an
  // UPDATE row stands at least kRsValueAhead rows after the end of every level it reads: its value, loaded that far
  // ahead, is then final.
  static constexpr int kRsNop = 0, kRsAssign = 1, kRsUpdate = 2, kRsPivot = 3, kRsWrite = 4;
  static constexpr int kRsKindMask = 7, kRsClear = 8, kRsLevelEnd = 16;
  // kRsPivotWrite (on a PIVOT row): the row also does the chunk's first WRITE row (write row 0), its lanes' kinds and
  // positions derived (pivot_write_lane), so a chunk costs one row less
  static constexpr int kRsPivotWrite = 128;
  static constexpr int kWriteU = 1, kWritePivot = 2, kWriteL = 3;
  static constexpr int kRsValueAhead = 8;
  static constexpr int kRsMetaShift = 24, kRsIndexMask = (1 << kRsMetaShift) - 1;
  int n_rs = 0, n_a = 0;
  vector<int> rs_rec, rs_meta, rs_level;            // [2 * 64 * n_rs] (meta folded into y), [n_rs], [n_rs]
  vector<int> rs_wr0;                               // [n_rs]: rows that write factors (WRITE, PIVOT+WRITE) before r
  // The WRITE a PIVOT+WRITE row's lane does: kind (kWrite*, 0: idle) and factor position, from its PIVOT record (x,
  // y), the row's lanes per track and the stream's write rows before it.
  struct WriteLane { int wk, k, pos; };
  WriteLane pivot_write_lane(int x, int y, int lane, int tl, int wr0) const {
    const int t = lane % tl;
    if (x == -1) return {0, 0, n_a + 64 * wr0 + lane};
    const int pt = (x & 0xFFFF) - (lane / tl) * wstride, nc = (int)((unsigned)x >> 16);
    const int wk = t < pt ? kWriteU : t == pt ? kWritePivot : t <= pt + nc ? kWriteL : 0;
    return {wk, y & kRsIndexMask, n_a + 64 * wr0 + lane};
  }
  // For level-parallel schedules: chunk c's rows are [rs_chunk_row[c], rs_chunk_row[c + 1]) of the row stream, and
  // level l's chunks are [rs_level_chunk[l], rs_level_chunk[l + 1]).
  vector<int> rs_chunk_row, rs_level_chunk;
  // Triangular solves as step streams. A pass is up to 64 rows of one level, one lane per row (rows by decreasing
  // length). Its entries' products are formed first, 64 at a time, into a scratch area of the wavefront's LDS (the
  // refactor workspace, free during the solves); then each lane subtracts its row's products from the row's value
  // one by one, in the order the reference's column sweeps apply them (forward: columns ascending; back: columns
  // descending, then the division by the pivot). A product rounds the same whoever forms it, so every entry is the
  // reference's value, and a long row no longer costs one step per entry. A pass whose products exceed the scratch
  // area runs in segments, each lane's running value kept across them. Records are 2 ints per lane and step, the
  // second holding a 24-bit field and the step's kind and flags above it:
  //   PRODUCT  (col, value index | kStepProduct << 24): scratch[s + lane] = vals[value index] * x[col] (col = -1:
  //            idle), s = 64 * (PRODUCT steps since the last CHAIN)
  //   CHAIN    (start | len << 16, row | (kStepChain | kChainBegin? | kChainEnd? | kChainIdle?) << 24): kChainBegin:
  //            acc = x[row]; then acc -= scratch[start + t] for t < len; kChainEnd: x[row] = acc (back: acc / pivot).
  //            kChainIdle: a lane with no row (its row field 0).
  // A device whose solve vector sits in global memory loads x kStepXAhead steps ahead with the values: x[row] for a
  // CHAIN step (only the row's own last CHAIN step writes it), x[col] for a PRODUCT step unless kProductLate marks it
  // (a column of the step ends fewer than kStepXAhead steps before it in the stream): that step loads x at the step.
  // The factors stay where the refactor's WRITE rows put them (the factor part of vals, 64 per WRITE row): a PRODUCT
  // step's value index is its entry's position there (fpos / bpos map an L / U entry, CSC index, to it; -1 for a
  // diagonal), so the solves gather the values they need while the refactor stores each WRITE row in one contiguous
  // piece; dval holds U's diagonal by pivot step.
  static constexpr int kStepProduct = 1, kStepChain = 2, kStepKindMask = 15;
  static constexpr int kChainBegin = 16, kChainEnd = 32, kChainIdle = 64;
  static constexpr int kProductLate = 16, kStepXAhead = 8;
  // scratch slots the solves may use: the refactor workspace rounded down to whole steps (or fewer, when the plan was
  // compiled with a smaller cap: scratch_cap, 0 = none; tests use it to make passes run in several segments)
  int scratch_cap = 0;
  int solve_scratch() const {
    const int full = (kTracks * wstride + kSpareSlots) / 64 * 64;
    return scratch_cap > 0 && scratch_cap < full ? scratch_cap / 64 * 64 : full;
  }
  int fs_steps = 0, bs_steps = 0;
  int n_wrows = 0, n_fac = 0;                       // WRITE rows of the refactor stream; factor slots (64 per row)
  vector<int> fs_rec, bs_rec;                       // 2 ints per lane per step
  // pass p of a stream covers steps [pass_step[p], pass_step[p + 1]); level l's passes are [level_pass[l], [l + 1])
  vector<int> fs_pass_step, fs_level_pass, bs_pass_step, bs_level_pass;
  vector<int> fpos, bpos;                           // [nnz_l], [nnz_u]
  // Level-scheduled triangular solves, rows pulled in the reference's update order.
  int fw_levels = 0, bw_levels = 0;
  vector<int> fw_order, fw_level_end, lr_ptr, lr_col, lr_idx;  // L rows: columns ascending
  vector<int> bw_order, bw_level_end, ur_ptr, ur_col, ur_idx;  // U rows: columns descending, diagonal excluded
};
// The Y cells' transposes (cell (i, j) -> cell (j, i)) when the Y-bus is symmetric bit for bit: every off-diagonal
// cell's stamp list (branches, values, order) equal to its transpose's, so that every outage's re-summed value is
// too. Empty otherwise (a phase shifter, or a stamp whose zero differs in sign).
vector<int> symmetric_cell_transpose(const YBus& y);
// The Newton path's positions of the Jacobian values in the A part of vals (the order the mismatch pass writes them),
// per Jacobian CSR entry; returns the A part's size, 4 nnz_y (entries absent from the Jacobian leave holes nothing
// reads). The pass's lane for Y cell k writes four values, H, N, M, L, to 4k .. 4k + 3 in one contiguous store: those
// of cell k itself, or with a transpose map (tr non-empty) those of cell tr[k], which a symmetric Y-bus lets it form
// from its own Y value and angle difference. Transposed, Jacobian column j's entries (Y column j's cells) sit in Y row
// j's stretch, next to each other, where the refactor's ASSIGN rows gather a column's entries. Bus i's diagonal block
// goes to its own cell's place, 4 y.diag[i] .. + 3, in the same stretch.
int jacobian_value_positions(const YBus& y, const JacobianPattern& jp, const vector<int>& tr, vector<int>& pos);

// scratch_cap > 0 caps the solves' scratch area (slots), so that passes run in several segments (for tests).
// a_pos: the A part's position of each Jacobian CSR entry and a_slots its size (jacobian_value_positions); null: the
// A entries in the refactor's record order (what the sparse-solve seam fills from its matrix).
void compile_plan(const Plan& plan, DevicePlanHost& out, int scratch_cap = 0, const vector<int>* a_pos = nullptr,
                  int a_slots = 0);

// Host emulation of the device refactor (level chunks, rows of 64 records) and level-scheduled solves, with every
// index checked against its buffer: out = A^{-1} rhs for the CSR values jv. Returns false (and why) on an
// out-of-range index or a zero pivot. The check the kernels' indexing gets on the CPU.
bool emulate_device_lu(const DevicePlanHost& d, int nnz_j, const double* jv, const double* rhs, double* out,
                       std::string& why);

// The level-parallel schedule's decomposition of the same streams: every chunk (solve pass) of a level starts its own
// workspace (row), reads only factors (rows) that earlier levels wrote, writes positions no other chunk of its level
// writes, and its look-ahead of `ring` rows past its range rounded up to whole rings stays inside the padded stream
// (padded_rows / padded_fs / padded_bs rows or steps). False (and why) otherwise.
bool check_level_schedule(const DevicePlanHost& d, int ring, int padded_rows, int padded_fs, int padded_bs,
                          std::string& why);

}  // namespace gs
