// libgridsim_hip.so host engine: device-resident networks and plans, batched and plain Newton-Raphson, the sparse
// linear-solver seams, and the C ABI of include/gridsim_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine_internal.h"
#include "fdlibm.h"

using namespace gs;

static_assert(DevicePlanHost::kRsValueAhead == kRsValueAhead, "plan and kernel disagree on the refactor's look-ahead");
static_assert(DevicePlanHost::kTracks == kTracks && DevicePlanHost::kTrackLanes == kTrackLanes, "track shape");

// Device memory of gs_sparse_solve_noalloc: the plan image and the per-call buffers of one system, placed in this
// file-scope array instead of allocated. Each GPU the library runs on holds its own copy (kNoAllocScratchBytes) until
// the process ends. In host code `scratch` only names the symbol: its device address on each device comes from
// hipGetSymbolAddress (scratch_address) and is the only form host code passes to a kernel or a copy.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_addr;   // device -> device address of `scratch` there
// The plan image each device's scratch array holds (the last one a call built there), so that the next call on the
// same structure replays it instead of building and uploading it again. Its key is empty while none is valid.
struct ScratchPlan {
  std::vector<int> key;
  gs::DevPlan h{};
  size_t plan_bytes = 0;
};
std::map<int, ScratchPlan> g_scratch_plan;
}  // namespace

// The device address of `scratch` on `device`, looked up once per device with that device current.
static gs_status scratch_address(int device, char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) { *out = it->second; return GS_OK; }
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) return GS_DEVICE_ERROR;
  if (prev != device && hipSetDevice(device) != hipSuccess) return GS_DEVICE_ERROR;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess || !ptr) return GS_DEVICE_ERROR;
  g_scratch_addr[device] = static_cast<char*>(ptr);
  *out = static_cast<char*>(ptr);
  return GS_OK;
}

// ================================================================================================ plans
namespace gs {

// GRIDSIM_SOLVE_SCRATCH_CAP (slots, for tests): compile plans whose triangular-solve passes run in segments of at most
// that many products, to exercise the segment carry on small networks. Unset or 0: the whole refactor workspace.
int solve_scratch_cap() {
  const char* v = std::getenv("GRIDSIM_SOLVE_SCRATCH_CAP");
  return v ? std::max(0, std::atoi(v)) : 0;
}

// The Y cells as the Jacobian fill reads them, one 8-byte (row, column) and one 16-byte (g, b) record per cell.
// Per Y cell its (row, column) and base value (g, b); gb then repeats each bus's diagonal cell, by bus, for the
// mismatch pass's bus loop (one coalesced load where the cell's own place is one scattered line per bus).
static void pack_cells(const YBus& y, std::vector<int>& ij, std::vector<double>& gb) {
  const int nnz = y.rp[y.n];
  ij.assign(2 * (size_t)nnz, 0);
  gb.assign(2 * ((size_t)nnz + y.n), 0.0);
  for (int i = 0; i < y.n; i++)
    for (int k = y.rp[i]; k < y.rp[i + 1]; k++) {
      ij[2 * (size_t)k] = i; ij[2 * (size_t)k + 1] = y.ci[k];
      gb[2 * (size_t)k] = y.g[k]; gb[2 * (size_t)k + 1] = y.b[k];
    }
  for (int i = 0; i < y.n; i++) {
    gb[2 * ((size_t)nnz + i)] = y.g[y.diag[i]];
    gb[2 * ((size_t)nnz + i) + 1] = y.b[y.diag[i]];
  }
}

// Blocks of consecutive buses for the blocked mismatch pass: at most 64 buses (one per lane), whose rows hold at most
// `cap` cells (the LDS scratch's two FP64 terms per cell); block b is buses [out[b], out[b + 1]). Empty (the two
// passes run apart) when one row alone holds more.
static std::vector<int> mismatch_blocks(const YBus& y, int cap) {
  std::vector<int> out(1, 0);
  for (int i = 0; i < y.n; i++) {
    if (y.rp[i + 1] - y.rp[i] > cap) return {};
    if (i - out.back() == 64 || y.rp[i + 1] - y.rp[out.back()] > cap) out.push_back(i);
  }
  out.push_back(y.n);
  return out;
}

// Four slots per Y cell, then each bus's diagonal cell's four again, by bus (as pack_cells).
static std::vector<int> with_diag_slots(const YBus& y, const std::vector<int>& cell_slots) {
  const size_t nnz = y.rp[y.n];
  std::vector<int> out(cell_slots);
  out.resize(4 * (nnz + y.n));
  for (int i = 0; i < y.n; i++)
    for (int b = 0; b < 4; b++) out[4 * (nnz + i) + b] = cell_slots[4 * (size_t)y.diag[i] + b];
  return out;
}

WideSplit wide_split_of(const DevicePlanHost& d) {
  WideSplit sp;
  const int nl = d.n_levels;
  // the trunk: the final run of levels the schedule walks on one work-group (at most kNarrow chunks each)
  int t = nl;
  while (t > 0 && d.rs_level_chunk[t] - d.rs_level_chunk[t - 1] <= kNarrow) t--;
  sp.trunk_level = t;
  if (t == 0 || t == nl) { sp.trunk_level = nl; return sp; }
  std::vector<int> col_level(d.m, 0);
  for (int l = 0, c = 0; l < nl; l++)
    for (; c < d.level_end[l]; c++) col_level[d.col_order[c]] = l;
  // the first forward level with a row that reads a factor column of the trunk
  int f = d.fw_levels;
  for (int l = 0, b0 = 0; l < d.fw_levels && f == d.fw_levels; b0 = d.fw_level_end[l], l++)
    for (int k = b0; k < d.fw_level_end[l] && f == d.fw_levels; k++) {
      const int i = d.fw_order[k];
      for (int e = d.lr_ptr[i]; e < d.lr_ptr[i + 1]; e++)
        if (col_level[d.lr_col[e]] >= t) { f = l; break; }
    }
  sp.fs_level = f;
  return sp;
}

// A Newton plan: the Jacobian values at the positions the mismatch pass writes them (Y-cell order).
void compile_newton_plan(const Plan& plan, DevicePlanHost& d, const YBus& y, const JacobianPattern& jp) {
  std::vector<int> pos;
  std::vector<int> tr = symmetric_cell_transpose(y);
  const int slots = jacobian_value_positions(y, jp, tr, pos);
  compile_plan(plan, d, solve_scratch_cap(), &pos, slots);
  d.y_tr = std::move(tr);
}

// The host image of a plan: the network pattern (when y is given), the start point (when spec is given), the Jacobian
// pattern and the compiled plan's streams, padded as the kernels expect; h's scalar fields are set, its pointer fields
// are recorded in the image for bind().
void build_plan_image(PlanImage& img, DevPlan& h, DevicePlanHost& d, const YBus* y, const JacobianPattern& jp,
                      const BusSpec* spec, const Plan& plan, double base_mva) {
  if (y) compile_newton_plan(plan, d, *y, jp);
  else compile_plan(plan, d, solve_scratch_cap());
  h = DevPlan{};
  img = PlanImage{};
  h.m = jp.m;
  h.nnz_j = jp.rp[jp.m];
  h.nnz_l = d.nnz_l;
  h.nnz_u = d.nnz_u;
  h.max_col = d.max_col;
  h.n_levels = d.n_levels;
  h.base_mva = base_mva;
  // the A position of each Jacobian CSR entry (every structural cell is kept, so every entry has one)
  std::vector<int> j2a(d.j2a);
  j2a.resize(std::max<size_t>((size_t)h.nnz_j, j2a.size()), -1);
  if (y) {
    h.n = y->n;
    h.nnz_y = y->rp[y->n];
    std::vector<int> y_ij, aslots(jp.slots.size()), slots(jp.slots);
    std::vector<double> y_gb;
    pack_cells(*y, y_ij, y_gb);
    // per Y cell's lane: the CSR positions of the values it forms (its own cell's, or its transpose's)
    h.y_tr = d.y_tr.empty() ? 0 : 1;
    if (h.y_tr)
      for (size_t k = 0; k < d.y_tr.size(); k++)
        for (int b = 0; b < 4; b++) slots[4 * k + b] = jp.slots[4 * (size_t)d.y_tr[k] + b];
    for (size_t e = 0; e < slots.size(); e++) aslots[e] = slots[e] >= 0 ? j2a[slots[e]] : -1;
    img.put(&DevPlan::y_rp, y->rp); img.put(&DevPlan::y_ij, y_ij); img.put(&DevPlan::y_diag, y->diag);
    img.put(&DevPlan::y_gb, y_gb); img.put(&DevPlan::slots, with_diag_slots(*y, slots));
    img.put(&DevPlan::aslots, with_diag_slots(*y, aslots));
    // the blocked mismatch pass's bus blocks, for FP64 and FP32 workspaces (ws_slots x 8 or x 4 bytes, 16 per cell),
    // with the workspace grown where one row alone would not fit it
    h.wstride = d.wstride;
    int longest = 0;
    for (int i = 0; i < y->n; i++) longest = std::max(longest, y->rp[i + 1] - y->rp[i]);
    h.ws_min = std::max(kFusedScratch, 4 * longest);
    const std::vector<int> mb8 = mismatch_blocks(*y, ws_slots(h) * 8 / 16);
    const std::vector<int> mb4 = mismatch_blocks(*y, ws_slots(h) * 4 / 16);
    h.n_mb = mb8.empty() ? 0 : (int)mb8.size() - 1;
    h.n_mb4 = mb4.empty() ? 0 : (int)mb4.size() - 1;
    if (h.n_mb) img.put(&DevPlan::mb_ptr, mb8);
    if (h.n_mb4) img.put(&DevPlan::mb_ptr4, mb4);
    img.put(&DevPlan::ppos, jp.ppos); img.put(&DevPlan::qpos, jp.qpos);
    // the same unknowns in pivot-step order (where the wavefront schedule's dx keeps them)
    std::vector<int> qinv(jp.m, -1), ppos_k(jp.ppos.size(), -1), qpos_k(jp.qpos.size(), -1), j_ci_k(jp.ci.size());
    for (int k = 0; k < jp.m; k++) qinv[d.q[k]] = k;
    for (size_t i = 0; i < jp.ppos.size(); i++) {
      if (jp.ppos[i] >= 0) ppos_k[i] = qinv[jp.ppos[i]];
      if (jp.qpos[i] >= 0) qpos_k[i] = qinv[jp.qpos[i]];
    }
    for (size_t e = 0; e < jp.ci.size(); e++) j_ci_k[e] = qinv[jp.ci[e]];
    img.put(&DevPlan::ppos_k, ppos_k); img.put(&DevPlan::qpos_k, qpos_k); img.put(&DevPlan::j_ci_k, j_ci_k);
  }
  if (spec) { img.put(&DevPlan::vm0, spec->vm); img.put(&DevPlan::va0, spec->va); }
  img.put(&DevPlan::j_rp, jp.rp); img.put(&DevPlan::j_ci, jp.ci);
  img.put(&DevPlan::pinv, d.pinv); img.put(&DevPlan::q, d.q); img.put(&DevPlan::prow, d.prow);
  img.put(&DevPlan::piv_pos, d.piv_pos);
  // the refactor's row stream, rounded up to whole rings of no-op rows plus one ring of look-ahead padding
  auto ring_up = [](int v) { return (v + kStreamPad - 1) / kStreamPad * kStreamPad; };
  h.wstride = d.wstride;
  h.n_a = d.n_a;
  h.n_rs = ring_up(d.n_rs);
  // (three rings of padding: a level-parallel run of one chunk starts at the even row at or before the chunk's
  // first, rounds its row range up to whole rings and loads one ring ahead of that). On the device the rows go in
  // pairs: pair q holds rows 2q and 2q + 1 of each lane side by side, so one 16-byte load brings a lane two rows.
  const int rs_rows = h.n_rs + kRsPadRings * kStreamPad;
  std::vector<int> rs_rec((size_t)rs_rows * 64 * 2, 0);
  for (int r = 0; r < d.n_rs; r++)
    for (int l = 0; l < 64; l++)
      for (int c = 0; c < 2; c++)
        rs_rec[(((size_t)(r / 2) * 64 + l) * 2 + (r % 2)) * 2 + c] = d.rs_rec[((size_t)r * 64 + l) * 2 + c];
  img.put(&DevPlan::rs_rec, rs_rec); img.put(&DevPlan::ch_a_src, d.ch_a_src);
  std::vector<int> wr0(d.rs_wr0);
  wr0.resize(rs_rows, d.n_wrows);
  img.put(&DevPlan::rs_wr0, wr0);
  // the solves' step streams (padded with steps that touch nothing) and the factor slots (64 per WRITE row)
  h.fs_steps = ring_up(d.fs_steps);
  h.bs_steps = ring_up(d.bs_steps);
  h.n_fac = d.n_fac;
  auto pad_steps = [](const std::vector<int>& rec, int steps) {
    std::vector<int> prec(rec);
    for (int t = (int)(rec.size() / 128); t < steps + 2 * kStreamPad; t++)
      for (int l = 0; l < 64; l++) { prec.push_back(-1); prec.push_back(0); }
    return prec;
  };
  img.put(&DevPlan::fs_rec, pad_steps(d.fs_rec, h.fs_steps));
  img.put(&DevPlan::bs_rec, pad_steps(d.bs_rec, h.bs_steps));
  // chunk and pass boundaries for level-parallel schedules
  img.put(&DevPlan::rs_chunk_row, d.rs_chunk_row);
  img.put(&DevPlan::fs_pass_step, d.fs_pass_step);
  img.put(&DevPlan::bs_pass_step, d.bs_pass_step);
  // the group schedule's segments: levels of more than `narrow` chunks (passes) split between the waves, runs of the
  // narrower levels walked by one wave
  std::vector<int> gseg;
  auto segs = [&gseg](const std::vector<int>& level_first, int narrow) {
    int run0 = -1, run1 = -1, n = 0;
    auto put = [&](int a, int b, int split) { gseg.insert(gseg.end(), {a, b, split, 0}); n++; };
    for (size_t l = 0; l + 1 < level_first.size(); l++) {
      const int a = level_first[l], b = level_first[l + 1];
      if (b <= a) continue;
      if (b - a <= narrow) {
        if (run0 < 0) run0 = a;
        run1 = b;
        continue;
      }
      if (run0 >= 0) { put(run0, run1, 0); run0 = -1; }
      put(a, b, 1);
    }
    if (run0 >= 0) put(run0, run1, 0);
    return n;
  };
  h.n_gseg_rs = segs(d.rs_level_chunk, kGroupNarrow);
  h.n_gseg_fs = segs(d.fs_level_pass, kGroupNarrowSolve);
  h.n_gseg_bs = segs(d.bs_level_pass, kGroupNarrowSolve);
  img.put(&DevPlan::gseg, gseg);
}

void upload_plan(DevicePlan& dp, const YBus* y, const JacobianPattern& jp, const BusSpec* spec, const Plan& plan,
                 double base_mva, hipStream_t s) {
  PlanImage img;
  DevPlan h;
  build_plan_image(img, h, dp.dph, y, jp, spec, plan, base_mva);
  dp.block.ensure(img.bytes.size());
  check(hipMemcpyAsync(dp.block.p, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice, s), "upload plan");
  // the image is a local: finish the copy before it goes
  check(hipStreamSynchronize(s), "upload plan");
  // (the branch fields, set by upload_branches, survive a plan upload)
  h.nb = dp.h.nb; h.br_ends = dp.h.br_ends; h.br_stamp = dp.h.br_stamp; h.br_rate = dp.h.br_rate;
  h.bus_lo = dp.h.bus_lo; h.bus_hi = dp.h.bus_hi;
  img.bind(h, dp.block.as<char>());
  dp.h = h;
}

void upload_branches(DevicePlan& dp, const Network& net, hipStream_t s) {
  DevPlan& h = dp.h;
  h.nb = net.nb;
  std::vector<int> ends(2 * (size_t)std::max(1, net.nb), -1);
  std::vector<double> stamp(8 * (size_t)std::max(1, net.nb), 0.0);
  for (int k = 0; k < net.nb; k++) {
    if (!net.br_in[k]) continue;
    ends[2 * k] = net.br_from[k];
    ends[2 * k + 1] = net.br_to[k];
    const Stamp st = branch_stamp(net.br_r[k], net.br_x[k], net.br_b[k], net.br_tap[k], net.br_shift[k]);
    const Cx v[4] = {st.yff, st.yft, st.ytf, st.ytt};
    for (int c = 0; c < 4; c++) { stamp[8 * (size_t)k + 2 * c] = v[c].re; stamp[8 * (size_t)k + 2 * c + 1] = v[c].im; }
  }
  PlanImage img;
  img.put(&DevPlan::br_ends, ends);
  img.put(&DevPlan::br_stamp, stamp);
  img.put(&DevPlan::br_rate, net.br_rate.empty() ? std::vector<double>(1, 0.0) : net.br_rate);
  img.put(&DevPlan::bus_lo, net.bus_lo);
  img.put(&DevPlan::bus_hi, net.bus_hi);
  dp.branch_block.ensure(img.bytes.size());
  check(hipMemcpyAsync(dp.branch_block.p, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice, s), "branches");
  check(hipStreamSynchronize(s), "branches");   // the image is a local
  img.bind(h, dp.branch_block.as<char>());
}

// The shape of a compiled plan as upload_plan lays it out in HBM (record streams padded the same way).
gs_plan_stats plan_stats(const DevicePlanHost& d, const YBus& y, const JacobianPattern& jp) {
  auto ring_up = [](int v) { return (v + kStreamPad - 1) / kStreamPad * kStreamPad; };
  gs_plan_stats st{};
  st.m = jp.m;
  st.nnz_j = jp.rp[jp.m];
  st.nnz_l = d.nnz_l;
  st.nnz_u = d.nnz_u;
  st.levels = d.n_levels;
  st.refactor_rows = ring_up(d.n_rs);
  st.forward_steps = ring_up(d.fs_steps);
  st.back_steps = ring_up(d.bs_steps);
  st.workspace_stride = d.wstride;
  st.factor_slots = d.n_fac;
  st.jacobian_transposed = d.y_tr.empty() ? 0 : 1;
  const int64_t n = y.n, nnz_y = y.rp.empty() ? 0 : y.rp[y.n], m = jp.m, rec_row = 64 * 8;
  int64_t b = 4 * (n + 1) + 4 * nnz_y * 2 + 4 * n + 3 * 16 * (nnz_y + n) + 2 * 4 * n;  // Y-bus, slots, pos
  b += 4 * (m + 1) + 4 * (int64_t)st.nnz_j + 2 * 8 * n + 4 * 4 * m;                            // J pattern, start, pivots
  b += rec_row * (int64_t)(st.refactor_rows + 2 * kStreamPad) + 4 * (int64_t)d.ch_a_src.size();
  b += rec_row * (int64_t)(st.forward_steps + 2 * kStreamPad + st.back_steps + 2 * kStreamPad);
  b += 4 * (int64_t)(d.rs_chunk_row.size() + d.fs_pass_step.size() + d.bs_pass_step.size());
  st.stream_bytes = b;
  for (int r = 0; r < d.n_rs; r++) {
    const int kind = d.rs_meta[r] & DevicePlanHost::kRsKindMask;
    if (kind < 5) st.rows_by_kind[kind]++;
    if (kind != DevicePlanHost::kRsUpdate) continue;
    for (int l = 0; l < 64; l++)
      st.update_lanes += (d.rs_rec[((size_t)r * 64 + l) * 2] & 0xFFFF) != d.spare(l);
  }
  auto product = [](int y) {
    return (((unsigned)y >> DevicePlanHost::kRsMetaShift) & DevicePlanHost::kStepKindMask) == DevicePlanHost::kStepProduct;
  };
  for (size_t k = 0; k < d.fs_rec.size(); k += 2) st.forward_lanes += product(d.fs_rec[k + 1]) && d.fs_rec[k] >= 0;
  for (size_t k = 0; k < d.bs_rec.size(); k += 2) st.back_lanes += product(d.bs_rec[k + 1]) && d.bs_rec[k] >= 0;
  return st;
}

// The symbolic pass: minimum-degree order of the Jacobian pattern and the Gilbert-Peierls pivot sequence of the
// Jacobian at (vm, va), with every structural cell kept and the fixed-pattern solver's diagonal preference
// (FixedPatternSparseLuSolver.cs:52,75-99).
constexpr double kDiagonalPreference = 0.1;

// growth_tol: the growth check the plan's pivots will face; the diagonal is preferred only where it also passes that
// check, so a plan always passes it at the state it was made at.
bool symbolic_pass(const YBus& y, const JacobianPattern& jp, const double* vm, const double* va, Plan& plan,
                   std::string& why, double growth_tol, const std::vector<int>* order) {
  std::vector<double> p(y.n), q(y.n), jv(jp.rp[jp.m]);
  host_injections(y, vm, va, p.data(), q.data());
  host_jacobian(y, jp, vm, va, p.data(), q.data(), jv.data());
  const std::vector<int> own = order ? std::vector<int>() : minimum_degree(jp.m, jp.rp.data(), jp.ci.data());
  return gp_plan(jp.m, jp.rp, jp.ci, jv.data(), order ? *order : own, std::max(kDiagonalPreference, growth_tol), true,
                 plan, why);
}

gs_nr_options normalise(const gs_nr_options* o) {
  gs_nr_options r{};
  r.tolerance = 1e-8;
  r.max_iterations = 30;
  r.flat_start = 0;
  r.diverge_limit = 10.0;
  r.lu_precision = GS_LU_FP32;
  r.refine_steps = 1;
  r.growth_tol = 1e-3;
  r.schedule = 0;
  r.flags = 0;
  if (o) r = *o;
  if (r.flags & ~(GS_FLAG_X_GLOBAL | GS_FLAG_Q_LIMITS | GS_FLAG_SINGLE_WAVE | GS_FLAG_DC_RETRY | GS_GROUP_WAVES(15)))
    fail(GS_INVALID_ARGUMENT, "unknown flags");
  if (GS_GROUP_WAVES_OF(r.flags) > kGroupWavesMax)
    fail(GS_INVALID_ARGUMENT, "a work-group per instance holds at most " + std::to_string(kGroupWavesMax) + " waves");
  if (r.schedule < GS_SCHEDULE_GROUP)
    fail(GS_INVALID_ARGUMENT, "schedule must be >= 0, GS_SCHEDULE_LEVELS or GS_SCHEDULE_GROUP");
  if (!(r.tolerance > 0)) fail(GS_INVALID_ARGUMENT, "tolerance must be positive");
  if (r.max_iterations < 0) fail(GS_INVALID_ARGUMENT, "max_iterations must be >= 0");
  if (r.refine_steps < 0 || r.refine_steps > 4) fail(GS_INVALID_ARGUMENT, "refine_steps must lie in 0..4");
  if (!(r.growth_tol >= 0.0 && r.growth_tol <= 1.0)) fail(GS_INVALID_ARGUMENT, "growth_tol must lie in [0, 1]");
  return r;
}

Layout instance_layout(const DevPlan& h, bool fp32, bool x_lds) {
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t t = fp32 ? 4 : 8;
  Layout L{};
  size_t o = 0;
  L.vm = o; o += al(8 * (size_t)h.n);
  L.va = o; o += al(8 * (size_t)h.n);
  L.p = o; o += al(8 * (size_t)h.n);
  L.q = o; o += al(8 * (size_t)h.n);
  L.f = o; o += al(8 * (size_t)h.m);
  L.dx = o; o += al(8 * (size_t)h.m);
  L.r = o; o += al(8 * (size_t)h.m);
  L.jv = o; o += al(8 * (size_t)h.nnz_j);
  L.tpq = o; o += al(16 * (size_t)h.nnz_y);
  L.vals = o; o += al(t * (size_t)vals_slots(h));
  L.dval = o; o += al(t * (size_t)dval_slots(h));
  L.xg = o; o += x_lds ? 0 : al(t * (size_t)h.m);
  L.spec = o; o += al(16 * (size_t)h.n);
  L.start = o; o += al(16 * (size_t)h.n);
  L.stride = o;
  return L;
}

// The outaged branches of an instance (outage_branch, then the list), checked: in range, each once, at most
// kMaxOutages.
OutageSet outage_set(const gs_instance& in, int n_branch) {
  OutageSet out;
  if (in.n_outages < 0 || in.n_outages > kMaxOutages || (in.n_outages > 0 && !in.outages))
    fail(GS_INVALID_ARGUMENT, "n_outages must lie in 0..GS_MAX_OUTAGES, with the list given");
  if (in.outage_branch < -1 || in.outage_branch >= n_branch) fail(GS_INVALID_ARGUMENT, "outage_branch out of range");
  if (in.outage_branch >= 0) out.b[out.n++] = in.outage_branch;
  for (int k = 0; k < in.n_outages; k++) {
    const int b = in.outages[k];
    if (b < 0 || b >= n_branch) fail(GS_INVALID_ARGUMENT, "outaged branch out of range");
    if (std::find(out.begin(), out.end(), b) != out.end()) fail(GS_INVALID_ARGUMENT, "a branch is outaged twice");
    out.b[out.n++] = b;
  }
  if (out.n > kMaxOutages) fail(GS_INVALID_ARGUMENT, "more than GS_MAX_OUTAGES outaged branches");
  return out;
}

// The instances of a call, checked (outage sets, load scale, start point): GS_INVALID_ARGUMENT naming the first bad one.
void check_instances(int n_branch, int count, const gs_instance* inst, const std::string& where) {
  for (int i = 0; i < count; i++) {
    try {
      (void)outage_set(inst[i], n_branch);
    } catch (const HipError& e) {
      fail(GS_INVALID_ARGUMENT, where + "instance " + std::to_string(i) + ": " + e.what);
    }
    if (!(inst[i].load_scale > 0) || !std::isfinite(inst[i].load_scale))
      fail(GS_INVALID_ARGUMENT, where + "instance " + std::to_string(i) + ": load_scale must be positive");
    if ((inst[i].vm0 == nullptr) != (inst[i].va0 == nullptr))
      fail(GS_INVALID_ARGUMENT, where + "instance " + std::to_string(i) + ": give both vm0 and va0, or neither");
  }
}

// Every Y cell the outaged branches touch, re-summed in insertion order without them: bit-identical to the Y-bus
// rebuilt without those branches (WithoutMany keeps model order and the builder sums from +0.0).
int outage_patches(const YBus& y, const Network& net, const OutageSet& outaged, Patch* out) {
  int n = 0;
  for (int br : outaged) {
    const int f = net.br_from[br], t = net.br_to[br];
    const int cells[4] = {y.find(f, f), y.find(t, t), y.find(f, t), y.find(t, f)};
    for (int k : cells) {
      if (k < 0) continue;
      bool dup = false;
      for (int q = 0; q < n; q++) dup |= out[q].idx == k;
      if (dup) continue;
      double g = 0.0, b = 0.0;
      bool any = false;
      for (int e = y.contrib_ptr[k]; e < y.contrib_ptr[k + 1]; e++) {
        if (std::find(outaged.begin(), outaged.end(), y.contrib_branch[e]) != outaged.end()) continue;
        if (!any) { g = 0.0 + y.contrib_g[e]; b = 0.0 + y.contrib_b[e]; any = true; }
        else { g += y.contrib_g[e]; b += y.contrib_b[e]; }
      }
      Patch p{};
      p.idx = k;
      p.g = any ? g : 0.0;
      p.b = any ? b : 0.0;
      out[n++] = p;
    }
  }
  return n;
}

// Instances of a batch whose outage touches more Y cells than the descriptor patches in place: they get Y value arrays
// of their own (16 bytes per cell each).
int own_y_count(const YBus* y, const Network* netw, int count, const gs_instance* inst) {
  if (!y || !netw) return 0;
  int c = 0;
  for (int i = 0; i < count; i++) {
    const OutageSet out = outage_set(inst[i], netw->nb);
    Patch patches[kMaxOutagePatches];
    if (out.n > 2 && outage_patches(*y, *netw, out, patches) > kMaxPatch) c++;
  }
  return c;
}

BatchRun plan_run(const DevPlan& h, const gs_nr_options& o, bool force_fp64) {
  BatchRun r;
  r.fp32 = !force_fp64 && o.lu_precision == GS_LU_FP32;
  size_t t = r.fp32 ? 4 : 8;
  size_t ws = t * (size_t)ws_slots(h);
  size_t x = t * (size_t)h.m;
  // (the level-parallel schedule shares the solve vector between work-groups: it lives in global memory)
  const bool levels = !force_fp64 && o.schedule == GS_SCHEDULE_LEVELS;
  if (!force_fp64 && o.schedule == GS_SCHEDULE_GROUP) {
    // a workspace per wave, as many waves as fit (at least one), then the solve vector when it fits beside them
    int w = GS_GROUP_WAVES_OF(o.flags) ? GS_GROUP_WAVES_OF(o.flags) : kGroupWavesDefault;
    while (w > 1 && w * ws > kLdsLimit) w--;
    if (ws > kLdsLimit) fail(GS_INVALID_ARGUMENT, "the refactor workspace exceeds the LDS a work-group may take");
    r.group_waves = w;
    r.x_lds = !(o.flags & GS_FLAG_X_GLOBAL) && w * ws + x <= kLdsLimit;
    r.lds_bytes = w * ws + (r.x_lds ? x : 0);
    return r;
  }
  r.x_lds = !levels && !(o.flags & GS_FLAG_X_GLOBAL) && ws + x <= kLdsLimit;
  r.lds_bytes = ws + (r.x_lds ? x : 0);
  // Two wavefronts per work-group sharing the solve vector, where that keeps more instances resident per compute
  // unit (at most kWavesPerCu, the register limit).
  if (r.x_lds && !force_fp64 && !(o.flags & GS_FLAG_SINGLE_WAVE)) {
    const size_t pair_bytes = 2 * ws + x + 16;
    const size_t single_res = std::min<size_t>(kWavesPerCu, kLdsPerCu / r.lds_bytes);
    const size_t pair_res = std::min<size_t>(kWavesPerCu, 2 * (kLdsPerCu / pair_bytes));
    if (pair_bytes <= kLdsLimit && pair_res > single_res) {
      r.pair = true;
      r.lds_bytes = pair_bytes;
    }
  }
  return r;
}

// Runs `count` instances on one plan. Fills results, optional vm/va. Handles paused instances (pivot growth) by a
// fresh host symbolic pass on the instance's own reduced network and a resumed solve on that plan.
struct HostFinal;
// (arena: the device buffers the instances are laid out in, or null: the network's own)
void run_batch(gs_net* net, int count, const gs_instance* inst, const gs_nr_options& o, gs_nr_result* results,
               double* vm_out, double* va_out, bool plain, std::vector<HostFinal>* finals = nullptr,
               std::vector<InstState>* final_states = nullptr, RunArena* arena = nullptr);

void fill_result(const InstState& s, gs_nr_result& r, int refreshes, double us_per_tick) {
  r.status = s.nonfinite ? GS_PF_NONFINITE : s.converged ? GS_PF_CONVERGED : s.diverged ? GS_PF_DIVERGED
                                                                                       : GS_PF_MAX_ITERATIONS;
  r.iterations = s.iter;
  r.max_mismatch = s.max_mis;
  r.symbolic_refreshes = refreshes;
  r.step_failures = s.step_failures;
  r.min_pivot_ratio = s.min_ratio;
  r.phase_us[0] = s.t_mismatch * us_per_tick;
  r.phase_us[1] = s.t_refactor * us_per_tick;
  r.phase_us[2] = s.t_solve * us_per_tick;
  r.phase_us[3] = s.t_update * us_per_tick;
  r.engine_stop = s.engine_stop;
  r.dc_retry = GS_RETRY_NOT_TRIED;
}

// `resume` (or null: start afresh) are paused states to continue from.
PreparedRun prepare_into(const DevPlan& h, const RunBuffers& rb, Staging& stg, int count, const gs_instance* inst,
                         const YBus* y, const Network* netw, const gs_nr_options& o, bool plain,
                         const std::vector<InstState>* resume,
                         const std::vector<std::vector<double>>* spec_override,
                         const std::vector<std::vector<double>>* start_override, hipStream_t s) {
  PreparedRun pr;
  pr.count = count;
  pr.desc = rb.desc;
  pr.run = plan_run(h, o, plain);
  const BatchRun& run = pr.run;
  pr.L = instance_layout(h, run.fp32, run.x_lds);
  const Layout& L = pr.L;
  char* base = rb.inst;
  InstState* dstate = rb.state;
  std::vector<InstDesc>& descs = stg.descs;
  descs.resize(count);   // (each record is cleared below)
  stg.stage.clear();
  stg.own_y.clear();
  int own_used = 0;
  for (int i = 0; i < count; i++) {
    InstDesc& d = descs[i];
    std::memset(&d, 0, sizeof(d));
    char* b = base + L.stride * i;
    d.state = dstate + i;
    d.vm = (double*)(b + L.vm); d.va = (double*)(b + L.va); d.p = (double*)(b + L.p); d.q = (double*)(b + L.q);
    d.f = (double*)(b + L.f); d.dx = (double*)(b + L.dx); d.r = (double*)(b + L.r); d.jv = (double*)(b + L.jv);
    d.tpq = (double*)(b + L.tpq);
    d.vals = b + L.vals; d.dval = b + L.dval;
    d.xg = run.x_lds ? nullptr : b + L.xg;
    const gs_instance& in = inst[i];
    d.n_patch = 0;
    d.n_out = 0;
    d.y_gb_own = nullptr;
    for (int t = 0; t < kMaxOutages; t++) d.out[t] = -1;
    if (y && netw) {
      const OutageSet outaged = outage_set(in, netw->nb);
      d.n_out = outaged.n;
      std::copy(outaged.begin(), outaged.end(), d.out);
      Patch patches[kMaxOutagePatches];
      const int np = outage_patches(*y, *netw, outaged, patches);
      if (np <= kMaxPatch) {
        d.n_patch = np;
        std::copy(patches, patches + np, d.patch);
      } else {
        // more cells than the descriptor holds: the instance's own Y values, the network's with the patches applied
        if (own_used >= rb.own_y_count) fail(GS_DEVICE_ERROR, "own Y-value arrays under-counted");
        // (laid out as pack_cells: the cells, then the buses' diagonal cells; the kernel patches neither)
        const size_t nnz = y->rp[y->n];
        std::vector<double> gb(2 * (nnz + y->n));
        for (size_t k = 0; k < nnz; k++) { gb[2 * k] = y->g[k]; gb[2 * k + 1] = y->b[k]; }
        for (int q = 0; q < np; q++) {
          gb[2 * (size_t)patches[q].idx] = patches[q].g;
          gb[2 * (size_t)patches[q].idx + 1] = patches[q].b;
        }
        for (int b = 0; b < y->n; b++) {
          gb[2 * (nnz + b)] = gb[2 * (size_t)y->diag[b]];
          gb[2 * (nnz + b) + 1] = gb[2 * (size_t)y->diag[b] + 1];
        }
        double* dst = rb.own_y + (size_t)own_used * 2 * (nnz + y->n);
        d.y_gb_own = dst;
        stg.own_y.emplace_back(dst, std::move(gb));
        own_used++;
      }
    }
    // specified injections: the network's, or a load-scaled copy
    const std::vector<double>* sp = spec_override ? &(*spec_override)[i] : nullptr;
    std::vector<double> scaled;
    if (!sp && netw && in.load_scale != 1.0) {
      BusSpec bs = build_bus_spec(*netw, o.flat_start != 0, in.load_scale);
      scaled.resize(2 * (size_t)h.n);
      std::copy(bs.pspec.begin(), bs.pspec.end(), scaled.begin());
      std::copy(bs.qspec.begin(), bs.qspec.end(), scaled.begin() + h.n);
      sp = &scaled;
    }
    const std::vector<double>* st = start_override ? &(*start_override)[i] : nullptr;
    std::vector<double> own;
    if (!st && in.vm0 && in.va0) {
      own.assign(in.vm0, in.vm0 + h.n);
      own.insert(own.end(), in.va0, in.va0 + h.n);
      st = &own;
    }
    if (sp || st) {
      stg.stage.emplace_back(i, std::vector<double>(4 * (size_t)h.n, 0.0));
      std::vector<double>& buf = stg.stage.back().second;
      if (sp) std::copy(sp->begin(), sp->end(), buf.begin());
      if (st) std::copy(st->begin(), st->end(), buf.begin() + 2 * (size_t)h.n);
    }
    d.pspec = sp ? (const double*)(b + L.spec) : nullptr;  // null: the network's spec, set below
    d.qspec = sp ? d.pspec + h.n : nullptr;
    d.vm_start = st ? (const double*)(b + L.start) : h.vm0;
    d.va_start = st ? d.vm_start + h.n : h.va0;
  }
  for (auto& d : descs)
    if (!d.pspec) { d.pspec = rb.base_spec; d.qspec = rb.base_spec + h.n; }
  for (auto& [i, buf] : stg.stage) {
    char* b = base + L.stride * i;
    if (descs[i].pspec == (const double*)(b + L.spec))
      check(hipMemcpyAsync(b + L.spec, buf.data(), 16 * (size_t)h.n, hipMemcpyHostToDevice, s), "stage spec");
    if (descs[i].vm_start == (const double*)(b + L.start))
      check(hipMemcpyAsync(b + L.start, buf.data() + 2 * h.n, 16 * (size_t)h.n, hipMemcpyHostToDevice, s),
            "stage start");
  }
  for (auto& [dst, gb] : stg.own_y)
    check(hipMemcpyAsync(dst, gb.data(), sizeof(double) * gb.size(), hipMemcpyHostToDevice, s), "own Y values");
  check(hipMemcpyAsync(rb.desc, descs.data(), sizeof(InstDesc) * count, hipMemcpyHostToDevice, s), "descs");
  // a fresh start needs no state upload: the kernel's init resets every counter it reads
  if (resume) {
    stg.states = *resume;
    check(hipMemcpyAsync(dstate, stg.states.data(), sizeof(InstState) * count, hipMemcpyHostToDevice, s), "states");
  }
  NrParams& prm = pr.prm;
  prm = NrParams{};
  prm.tolerance = o.tolerance;
  prm.diverge_limit = o.diverge_limit;
  prm.growth_tol = o.growth_tol;
  prm.max_iterations = o.max_iterations;
  prm.refine_steps = run.fp32 ? o.refine_steps : 0;
  prm.load_start = start_override ? 1 : 0;
  prm.steps_this_launch = 1 << 30;
  prm.init = resume ? 0 : 1;
  return pr;
}

// Lays `count` instances sharing one device plan out in HBM (the handle's own arenas, grown as needed) and uploads
// their descriptors and own data on stream s.
PreparedRun prepare_instances(gs_ctx* ctx, DevicePlan& dp, DevMem& inst_buf, DevMem& desc_buf, DevMem& state_buf,
                              DevMem& aux_buf, Staging& stg, int count, const gs_instance* inst, const YBus* y,
                              const Network* netw, const BusSpec& base_spec, const gs_nr_options& o, bool plain,
                              const std::vector<InstState>* resume,
                              const std::vector<std::vector<double>>* spec_override,
                              const std::vector<std::vector<double>>* start_override, hipStream_t s) {
  const DevPlan& h = dp.h;
  const BatchRun run = plan_run(h, o, plain);
  const Layout L = instance_layout(h, run.fp32, run.x_lds);
  size_t need = L.stride * (size_t)count;
  if (ctx->budget && need > ctx->budget) fail(GS_OUT_OF_MEMORY, "batch exceeds the HBM budget");
  inst_buf.ensure(need);
  desc_buf.ensure(sizeof(InstDesc) * (size_t)count);
  state_buf.ensure(sizeof(InstState) * (size_t)count);
  // the base specification, then the own Y-value arrays of the instances that need them
  const int own = own_y_count(y, netw, count, inst);
  const size_t spec_bytes = (16 * (size_t)h.n + 255) & ~size_t(255);
  aux_buf.ensure(spec_bytes + 16 * ((size_t)h.nnz_y + h.n) * own);
  stg.aux = base_spec.pspec;
  stg.aux.insert(stg.aux.end(), base_spec.qspec.begin(), base_spec.qspec.end());
  check(hipMemcpyAsync(aux_buf.p, stg.aux.data(), sizeof(double) * stg.aux.size(), hipMemcpyHostToDevice, s), "spec");
  RunBuffers rb;
  rb.inst = inst_buf.as<char>();
  rb.desc = desc_buf.as<InstDesc>();
  rb.state = state_buf.as<InstState>();
  rb.base_spec = aux_buf.as<double>();
  rb.own_y = reinterpret_cast<double*>(aux_buf.as<char>() + spec_bytes);
  rb.own_y_count = own;
  return prepare_into(h, rb, stg, count, inst, y, netw, o, plain, resume, spec_override, start_override, s);
}

void launch_prepared(const DevPlan& h, const PreparedRun& pr, int schedule, hipStream_t s) {
  if (pr.run.group_waves > 0) {
    check(launch_group(pr.run.fp32, pr.run.x_lds, pr.run.group_waves, pr.desc, h, pr.count, pr.prm, pr.run.lds_bytes,
                       s),
          "nr kernel (work-group per instance)");
    return;
  }
  if (schedule < 0) fail(GS_INVALID_ARGUMENT, "the level-parallel schedule runs one call at a time (gs_nr_solve_batch)");
  const int slice = schedule > 0 ? schedule : pr.count;
  for (int first = 0; first < pr.count; first += slice)
    check(launch_nr_batch(pr.run.fp32, pr.run.x_lds, pr.run.pair, pr.desc + first, h,
                          std::min(slice, pr.count - first), pr.prm, pr.run.lds_bytes, s),
          "nr kernel");
}

// The plan's split point, with the fork's side stream and events (created on first use) when the plan has a trunk.
WideSplit forked_split(const DevicePlanHost& d, WideGraph& wg) {
  WideSplit split = wide_split_of(d);
  if (split.trunk_level < d.n_levels && split.fs_level > 0) {
    if (!wg.side) check(hipStreamCreateWithFlags(&wg.side, hipStreamNonBlocking), "side stream");
    if (!wg.fork) check(hipEventCreateWithFlags(&wg.fork, hipEventDisableTiming), "fork event");
    if (!wg.join) check(hipEventCreateWithFlags(&wg.join, hipEventDisableTiming), "join event");
    split.side = wg.side;
    split.fork = wg.fork;
    split.join = wg.join;
  }
  return split;
}

// The level-parallel schedule: the start kernel, then replays of the captured Newton step until every instance has
// finished (or paused), polling the finished count after every second replay (a replay after the end still
// dispatches every kernel of the step); at most max_iterations + 1 replays (the last evaluates the mismatch of the
// final step). Returns the number of replays.
int launch_wide(DevicePlan& dp, const PreparedRun& pr, hipStream_t s) {
  if (!dp.wide) dp.wide = std::make_unique<WideGraph>();
  WideGraph& wg = *dp.wide;
  if (!wg.host) {
    wg.ctl.ensure(sizeof(WideCtl));
    check(hipHostMalloc((void**)&wg.host, sizeof(WideCtl), hipHostMallocDefault), "pinned control block");
  }
  const NrParams& p = pr.prm;
  const bool same = wg.exec && wg.desc == pr.desc && wg.count == pr.count && wg.fp32 == (int)pr.run.fp32 &&
                    wg.prm.tolerance == p.tolerance && wg.prm.diverge_limit == p.diverge_limit &&
                    wg.prm.growth_tol == p.growth_tol && wg.prm.max_iterations == p.max_iterations &&
                    wg.prm.refine_steps == p.refine_steps;
  WideCtl* ctl = wg.ctl.as<WideCtl>();
  if (!same) {
    if (wg.exec) { (void)hipGraphExecDestroy(wg.exec); wg.exec = nullptr; }
    if (wg.graph) { (void)hipGraphDestroy(wg.graph); wg.graph = nullptr; }
    const DevicePlanHost& d = dp.dph;
    const size_t ws_bytes = (pr.run.fp32 ? 4 : 8) * (size_t)ws_slots(dp.h);
    const WideSplit split = forked_split(d, wg);
    check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
    hipError_t e = launch_wide_step(pr.run.fp32, pr.desc, dp.h, pr.count, p, ctl, d.rs_level_chunk, d.fs_level_pass,
                                    d.bs_level_pass, ws_bytes, split, s);
    hipError_t e2 = hipStreamEndCapture(s, &wg.graph);
    check(e, "level-parallel step");
    check(e2, "end capture");
    check(hipGraphInstantiate(&wg.exec, wg.graph, nullptr, nullptr, 0), "instantiate");
    wg.desc = pr.desc;
    wg.count = pr.count;
    wg.fp32 = (int)pr.run.fp32;
    wg.prm = p;
  }
  check(launch_wide_start(pr.desc, dp.h, pr.count, ctl, s), "level-parallel start");
  const int total = p.max_iterations + 1;
  int replays = 0;
  while (replays < total) {
    const int k = std::min(2, total - replays);
    for (int i = 0; i < k; i++) check(hipGraphLaunch(wg.exec, s), "graph launch");
    replays += k;
    check(hipMemcpyAsync(wg.host, ctl, sizeof(int32_t), hipMemcpyDeviceToHost, s), "poll");
    check(hipStreamSynchronize(s), "poll sync");
    if (wg.host->finished >= pr.count) break;
  }
  return replays;
}

// Core launcher for instances sharing one device plan. `states` may be pre-filled (resume) when init == 0.
// trust_first: the plan was made at the instances' start state (a fresh symbolic pass), so their first step skips the
// growth and floor checks.
double launch_instances(gs_ctx* ctx, DevicePlan& dp, DevMem& inst_buf, DevMem& desc_buf, DevMem& state_buf,
                        DevMem& aux_buf, int count, const gs_instance* inst, const YBus* y, const Network* netw,
                        const BusSpec& base_spec, const gs_nr_options& o, bool plain,
                        std::vector<InstState>& states, bool init,
                        const std::vector<std::vector<double>>* spec_override,
                        const std::vector<std::vector<double>>* start_override, int* launches,
                        bool trust_first = false) {
  hipStream_t s = ctx->stream;
  const DevPlan& h = dp.h;
  Staging stg;
  PreparedRun pr = prepare_instances(ctx, dp, inst_buf, desc_buf, state_buf, aux_buf, stg, count, inst, y, netw,
                                     base_spec, o, plain, init ? nullptr : &states, spec_override, start_override, s);
  pr.prm.trust_first = trust_first ? 1 : 0;
  const BatchRun& run = pr.run;
  NrParams prm = pr.prm;
  InstState* dstate = state_buf.as<InstState>();
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  check(hipEventRecord(e0, s), "event");
  int nl = 0;
  if (!plain && o.schedule == GS_SCHEDULE_LEVELS) {
    if (!init) fail(GS_INVALID_ARGUMENT, "the level-parallel schedule starts afresh");
    nl = launch_wide(dp, pr, s);
  } else if (!plain) {
    launch_prepared(dp.h, pr, o.schedule, s);
    nl = o.schedule > 0 ? (count + o.schedule - 1) / o.schedule : 1;
  } else {
    // one instance per launch, one Newton step per launch, the host reading the verdict after each
    prm.steps_this_launch = 1;
    const int load_start = prm.load_start;
    for (int i = 0; i < count; i++) {
      bool first = true;
      for (;;) {
        prm.init = (first && init) ? 1 : 0;
        prm.load_start = first ? load_start : 0;
        prm.trust_first = first ? pr.prm.trust_first : 0;
        first = false;
        check(launch_nr_batch(0, run.x_lds, run.pair, desc_buf.as<InstDesc>() + i, h, 1, prm, run.lds_bytes, s),
              "plain kernel");
        nl++;
        InstState hs;
        check(hipMemcpyAsync(&hs, dstate + i, sizeof(InstState), hipMemcpyDeviceToHost, s), "plain state");
        check(hipStreamSynchronize(s), "plain sync");
        if (hs.done) break;
      }
    }
  }
  check(hipEventRecord(e1, s), "event");
  states.resize(count);
  check(hipMemcpyAsync(states.data(), dstate, sizeof(InstState) * count, hipMemcpyDeviceToHost, s), "states back");
  check(hipStreamSynchronize(s), "sync");
  float ms = 0;
  check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (launches) *launches += nl;
  return ms;
}

void copy_state_out(gs_ctx* ctx, const DevMem& inst_buf, const Layout& L, int i, int n, double* vm, double* va) {
  const char* b = inst_buf.as<char>() + L.stride * i;
  if (vm) check(hipMemcpyAsync(vm, b + L.vm, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "vm out");
  if (va) check(hipMemcpyAsync(va, b + L.va, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "va out");
}

std::vector<int> OrderCache::order(const std::vector<int>& rp, const std::vector<int>& ci, int m) {
  // FNV-1a over the four bytes of every integer, least significant first
  uint64_t h = 14695981039346656037ull;
  auto mix = [&h](int v) {
    const uint32_t u = (uint32_t)v;
    for (int b = 0; b < 4; b++) { h ^= (u >> (8 * b)) & 0xFF; h *= 1099511628211ull; }
  };
  mix(m);
  for (int v : rp) mix(v);
  for (int v : ci) mix(v);
  {
    std::lock_guard<std::mutex> lock(mu_);
    if (std::shared_ptr<Entry>* e = map_.find(h))
      if ((*e)->rp == rp && (*e)->ci == ci) {
        hits_++;
        return (*e)->order;
      }
    misses_++;
  }
  auto e = std::make_shared<Entry>();
  e->rp = rp;
  e->ci = ci;
  e->order = minimum_degree(m, rp.data(), ci.data());
  std::vector<int> order = e->order;
  std::lock_guard<std::mutex> lock(mu_);
  map_.put(h, e);
  return order;
}

void fresh_symbolic(const Network& base, bool flat_start, const gs_nr_options& o, int count, const gs_instance* inst,
                    const std::vector<const double*>& vm, const std::vector<const double*>& va,
                    const std::vector<int>* types, const std::vector<double>* qspec, int threads, OrderCache& cache,
                    std::vector<FreshSymbolic>& out) {
  out.assign(count, FreshSymbolic{});
  auto one = [&](int i) {
    FreshSymbolic& f = out[i];
    f.own = instance_network(base, inst[i]);
    f.y = build_ybus(f.own, -1);
    f.spec = build_bus_spec(f.own, flat_start, 1.0);
    if (types) f.spec.type = *types;
    if (qspec) f.spec.qspec = *qspec;
    f.jp = build_jacobian_pattern(f.y, f.spec.type);
    const std::vector<int> order = cache.order(f.jp.rp, f.jp.ci, f.jp.m);
    f.ok = symbolic_pass(f.y, f.jp, vm[i], va[i], f.plan, f.why, o.growth_tol, &order);
  };
  const int workers = std::max(1, std::min(threads, count));
  if (workers == 1) {
    for (int i = 0; i < count; i++) one(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  std::vector<std::string> errors(workers);
  for (int w = 0; w < workers; w++)
    pool.emplace_back([&, w]() {
      try {
        for (int i; (i = next.fetch_add(1)) < count;) one(i);
      } catch (const std::exception& e) {
        errors[w] = e.what();
      } catch (const HipError& e) {
        errors[w] = e.what;
      }
    });
  for (auto& t : pool) t.join();
  for (const std::string& e : errors)
    if (!e.empty()) fail(GS_INVALID_ARGUMENT, "symbolic pass: " + e);
}

// Host threads for the fresh symbolic passes of one call: the cores this process may use, at most 16.
int symbolic_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc ? hc : 1u, 16u));
}

// The network an instance solves: the uploaded one with its outaged branches out of service and every load and
// generator P scaled by its load multiplier.
Network instance_network(const Network& base, const gs_instance& in) {
  Network own = base;
  for (int br : outage_set(in, own.nb)) own.br_in[br] = 0;
  if (in.load_scale != 1.0) {
    for (int b = 0; b < own.n; b++) { own.pd[b] *= in.load_scale; own.qd[b] *= in.load_scale; }
    for (int g = 0; g < own.ng; g++) own.gen_pg[g] *= in.load_scale;
  }
  return own;
}

// A paused instance (a static pivot failed its growth check): fresh CPU symbolic passes on the instance's own network
// at its current state (vm, va: in and out), each followed by a resumed launch on that plan, until it finishes. The
// resumed run keeps its iteration count and start mismatch.
// The network pattern alone on the device (no LU plan): what the iterative step reads.
void upload_pattern(DevicePlan& dp, const YBus& y, const JacobianPattern& jp, double base, hipStream_t s) {
  DevPlan h{};
  h.n = y.n;
  h.m = jp.m;
  h.nnz_y = y.rp[y.n];
  h.nnz_j = jp.rp[jp.m];
  h.base_mva = base;
  std::vector<int> y_ij, aslots(jp.slots.size(), -1), jdiag(std::max(1, jp.m), -1);
  std::vector<double> y_gb;
  pack_cells(y, y_ij, y_gb);
  for (int r = 0; r < jp.m; r++)
    for (int k = jp.rp[r]; k < jp.rp[r + 1]; k++)
      if (jp.ci[k] == r) jdiag[r] = k;
  PlanImage img;
  img.put(&DevPlan::y_rp, y.rp); img.put(&DevPlan::y_ij, y_ij); img.put(&DevPlan::y_diag, y.diag);
  img.put(&DevPlan::y_gb, y_gb); img.put(&DevPlan::slots, with_diag_slots(y, jp.slots));
  img.put(&DevPlan::aslots, with_diag_slots(y, aslots));
  img.put(&DevPlan::ppos, jp.ppos); img.put(&DevPlan::qpos, jp.qpos); img.put(&DevPlan::j_rp, jp.rp);
  img.put(&DevPlan::j_ci, jp.ci); img.put(&DevPlan::j_diag, jdiag);
  dp.block.ensure(img.bytes.size());
  check(hipMemcpyAsync(dp.block.p, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice, s), "pattern upload");
  check(hipStreamSynchronize(s), "pattern upload");   // the image is a local
  img.bind(h, dp.block.as<char>());
  dp.h = h;
}

// One Newton iteration of one instance with SolveStep's ILU(0) + BiCGSTAB fallback on the device (the LU refused the
// Jacobian): state st and vm / va in and out; p / q (when non-null) the injections at the final state.
void iterative_step(gs_ctx* ctx, const YBus& y, const JacobianPattern& jp, const BusSpec& spec, double base,
                    const gs_nr_options& o, InstState& st, std::vector<double>& vm, std::vector<double>& va,
                    std::vector<double>* p_out, std::vector<double>* q_out) {
  hipStream_t s = ctx->stream;
  for (int r = 0; r < jp.m; r++) {
    bool diag = false;
    for (int k = jp.rp[r]; k < jp.rp[r + 1]; k++) diag |= jp.ci[k] == r;
    if (!diag) {  // ILU(0) needs every diagonal entry: the step fails and the run stops, as SolveStep's false
      st.done = 1;
      st.step_failures++;
      return;
    }
  }
  DevicePlan dp;
  upload_pattern(dp, y, jp, base, s);
  const size_t n = (size_t)y.n, m = (size_t)jp.m, nj = (size_t)jp.rp[jp.m], ny = (size_t)y.rp[y.n];
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t off = 0;
  const size_t o_vm = off; off += al(8 * n);
  const size_t o_va = off; off += al(8 * n);
  const size_t o_p = off; off += al(8 * n);
  const size_t o_q = off; off += al(8 * n);
  const size_t o_f = off; off += al(8 * m);
  const size_t o_dx = off; off += al(8 * m);
  const size_t o_jv = off; off += al(8 * nj);
  const size_t o_tpq = off; off += al(16 * ny);
  const size_t o_spec = off; off += al(16 * n);
  const size_t o_start = off; off += al(16 * n);
  const size_t o_work = off; off += al(8 * (nj + 8 * m));
  const size_t o_iw = off; off += al(4 * std::max<size_t>(1, m));
  const size_t o_state = off; off += al(sizeof(InstState));
  const size_t o_desc = off; off += al(sizeof(InstDesc));
  DevMem blk;
  blk.ensure(off);
  char* b = blk.as<char>();
  std::vector<double> specv(spec.pspec), start(vm);
  specv.insert(specv.end(), spec.qspec.begin(), spec.qspec.end());
  start.insert(start.end(), va.begin(), va.end());
  InstDesc d{};
  d.state = reinterpret_cast<InstState*>(b + o_state);
  d.pspec = reinterpret_cast<const double*>(b + o_spec);
  d.qspec = d.pspec + n;
  d.vm_start = reinterpret_cast<const double*>(b + o_start);
  d.va_start = d.vm_start + n;
  d.n_patch = 0;
  d.n_out = 0;
  d.y_gb_own = nullptr;
  d.vm = (double*)(b + o_vm); d.va = (double*)(b + o_va); d.p = (double*)(b + o_p); d.q = (double*)(b + o_q);
  d.f = (double*)(b + o_f); d.dx = (double*)(b + o_dx); d.r = nullptr; d.jv = (double*)(b + o_jv);
  d.tpq = (double*)(b + o_tpq);
  check(hipMemcpyAsync(b + o_spec, specv.data(), 16 * n, hipMemcpyHostToDevice, s), "iterative spec");
  check(hipMemcpyAsync(b + o_start, start.data(), 16 * n, hipMemcpyHostToDevice, s), "iterative start");
  check(hipMemcpyAsync(b + o_state, &st, sizeof(InstState), hipMemcpyHostToDevice, s), "iterative state");
  check(hipMemcpyAsync(b + o_desc, &d, sizeof(InstDesc), hipMemcpyHostToDevice, s), "iterative desc");
  NrParams prm{};
  prm.tolerance = o.tolerance;
  prm.diverge_limit = o.diverge_limit;
  prm.max_iterations = o.max_iterations;
  prm.load_start = 1;
  check(launch_iterative_step(reinterpret_cast<const InstDesc*>(b + o_desc), dp.h, prm,
                              reinterpret_cast<double*>(b + o_work), reinterpret_cast<int*>(b + o_iw), s),
        "iterative step");
  check(hipMemcpyAsync(&st, b + o_state, sizeof(InstState), hipMemcpyDeviceToHost, s), "iterative state back");
  check(hipMemcpyAsync(vm.data(), b + o_vm, 8 * n, hipMemcpyDeviceToHost, s), "iterative vm");
  check(hipMemcpyAsync(va.data(), b + o_va, 8 * n, hipMemcpyDeviceToHost, s), "iterative va");
  if (p_out) { p_out->resize(n); check(hipMemcpyAsync(p_out->data(), b + o_p, 8 * n, hipMemcpyDeviceToHost, s), "p"); }
  if (q_out) { q_out->resize(n); check(hipMemcpyAsync(q_out->data(), b + o_q, 8 * n, hipMemcpyDeviceToHost, s), "q"); }
  check(hipStreamSynchronize(s), "iterative step");
}

void finish_paused(gs_net* net, const gs_instance& in, const gs_nr_options& o, bool plain, InstState st,
                   std::vector<double>& vm, std::vector<double>& va, gs_nr_result& result, double& kms, int& launches,
                   std::vector<double>* p_out, std::vector<double>* q_out, const std::vector<int>* types,
                   const std::vector<double>* qspec, InstState* st_out, FreshSymbolic* first) {
  gs_ctx* ctx = net->ctx;
  const int n = net->net.n;
  int refreshes = 0;
  FreshSymbolic fs;
  if (first) {
    fs = std::move(*first);
  } else {
    std::vector<FreshSymbolic> one;
    fresh_symbolic(net->net, o.flat_start != 0, o, 1, &in, {vm.data()}, {va.data()}, types, qspec, 1, net->orders,
                   one);
    fs = std::move(one[0]);
  }
  const Network& own = fs.own;
  const YBus& y = fs.y;
  const BusSpec& spec = fs.spec;
  const JacobianPattern& jp = fs.jp;
  DevicePlan dp;
  DevMem ib, db, sb, ab;
  bool finished = false;
  int iterative = 0;
  bool have_first = true;   // the first pass is done (fs.plan, fs.ok)
  while (!finished) {
    Plan plan;
    std::string why;
    bool ok;
    if (have_first) {
      plan = std::move(fs.plan);
      ok = fs.ok;
      have_first = false;
    } else {
      const std::vector<int> order = net->orders.order(jp.rp, jp.ci, jp.m);
      ok = symbolic_pass(y, jp, vm.data(), va.data(), plan, why, o.growth_tol, &order);
    }
    if (!ok) {
      // singular at this state: the LU refuses the step, so SolveStep falls back to ILU(0) + BiCGSTAB for it (on the
      // device); the LU is tried again at the next step
      if (++iterative > o.max_iterations + 1) {
        st.done = 1;
        st.step_failures++;
        st.engine_stop = GS_STOP_FALLBACK_CAP;
        break;
      }
      iterative_step(ctx, y, jp, spec, own.base, o, st, vm, va, p_out, q_out);
      finished = st.done == 1;
      continue;
    }
    // (a fresh plan takes at least the step at the state it was made at, so this bound is a safety net only)
    if (++refreshes > o.max_iterations + 2) {
      // the fresh plans keep failing their growth checks: the run stops
      st.done = 1;
      st.step_failures++;
      st.engine_stop = GS_STOP_REFRESH_CAP;
      break;
    }
    upload_plan(dp, &y, jp, &spec, plan, own.base, ctx->stream);
    gs_instance one = in;
    one.outage_branch = -1;
    one.n_outages = 0;
    one.outages = nullptr;
    one.load_scale = 1.0;
    std::vector<std::vector<double>> start(1);
    start[0] = vm;
    start[0].insert(start[0].end(), va.begin(), va.end());
    std::vector<InstState> rs(1, st);
    rs[0].done = 0;
    gs_nr_options ow = o;
    ow.schedule = 0;  // a resumed run is one instance on one wavefront
    kms += launch_instances(ctx, dp, ib, db, sb, ab, 1, &one, nullptr, nullptr, spec, ow, plain, rs, false, nullptr,
                            &start, &launches, true);
    st = rs[0];
    BatchRun r2 = plan_run(dp.h, ow, plain);
    Layout L2 = instance_layout(dp.h, r2.fp32, r2.x_lds);
    copy_state_out(ctx, ib, L2, 0, n, vm.data(), va.data());
    if (p_out) {
      p_out->resize(n);
      check(hipMemcpyAsync(p_out->data(), ib.as<char>() + L2.p, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream),
            "p out");
    }
    if (q_out) {
      q_out->resize(n);
      check(hipMemcpyAsync(q_out->data(), ib.as<char>() + L2.q, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream),
            "q out");
    }
    check(hipStreamSynchronize(ctx->stream), "sync resume");
    finished = st.done == 1;
  }
  fill_result(st, result, refreshes, ctx->us_per_tick);
  if (st_out) *st_out = st;
}

// Results of a finished device run of `count` instances (states copied back): finished instances straight from the
// device, paused ones through finish_paused.
// Final state of an instance the host finished (a fresh symbolic pass): its index and vm, va, p, q.
struct HostFinal {
  int i;
  std::vector<double> vm, va, p, q;
};

void collect_results(gs_net* net, int count, const gs_instance* inst, const gs_nr_options& o, bool plain,
                     const std::vector<InstState>& states, const DevMem& inst_buf, const Layout& L,
                     gs_nr_result* results, double* vm_out, double* va_out, double& kms, int& launches,
                     std::vector<HostFinal>* finals = nullptr, std::vector<InstState>* final_states = nullptr) {
  if (final_states) *final_states = states;
  gs_ctx* ctx = net->ctx;
  const int n = net->net.n;
  for (int i = 0; i < count; i++)
    if (states[i].done == 1) fill_result(states[i], results[i], 0, ctx->us_per_tick);
  // every instance's voltages in one strided copy per array (the rows of paused instances are replaced below)
  const char* b = inst_buf.as<char>();
  if (vm_out && count > 0)
    check(hipMemcpy2DAsync(vm_out, 8 * (size_t)n, b + L.vm, L.stride, 8 * (size_t)n, (size_t)count,
                           hipMemcpyDeviceToHost, ctx->stream), "vm out");
  if (va_out && count > 0)
    check(hipMemcpy2DAsync(va_out, 8 * (size_t)n, b + L.va, L.stride, 8 * (size_t)n, (size_t)count,
                           hipMemcpyDeviceToHost, ctx->stream), "va out");
  check(hipStreamSynchronize(ctx->stream), "sync out");
  // paused instances: their states, their first fresh symbolic passes on host threads (independent of each other),
  // then each one's resumed launches in turn
  std::vector<int> paused;
  for (int i = 0; i < count; i++)
    if (states[i].done != 1) paused.push_back(i);
  if (paused.empty()) return;
  const int k = (int)paused.size();
  std::vector<std::vector<double>> vms(k, std::vector<double>(n)), vas(k, std::vector<double>(n));
  for (int j = 0; j < k; j++) copy_state_out(ctx, inst_buf, L, paused[j], n, vms[j].data(), vas[j].data());
  check(hipStreamSynchronize(ctx->stream), "sync pause");
  std::vector<gs_instance> pin(k);
  std::vector<const double*> pvm(k), pva(k);
  for (int j = 0; j < k; j++) { pin[j] = inst[paused[j]]; pvm[j] = vms[j].data(); pva[j] = vas[j].data(); }
  std::vector<FreshSymbolic> firsts;
  fresh_symbolic(net->net, o.flat_start != 0, o, k, pin.data(), pvm, pva, nullptr, nullptr, symbolic_threads(),
                 net->orders, firsts);
  for (int j = 0; j < k; j++) {
    const int i = paused[j];
    std::vector<double>& vm = vms[j];
    std::vector<double>& va = vas[j];
    std::vector<double> p, q;
    finish_paused(net, inst[i], o, plain, states[i], vm, va, results[i], kms, launches, finals ? &p : nullptr,
                  finals ? &q : nullptr, nullptr, nullptr, final_states ? &(*final_states)[i] : nullptr, &firsts[j]);
    if (finals) finals->push_back(HostFinal{i, vm, va, p, q});
    if (vm_out) std::copy(vm.begin(), vm.end(), vm_out + (size_t)n * i);
    if (va_out) std::copy(va.begin(), va.end(), va_out + (size_t)n * i);
  }
}

void run_batch(gs_net* net, int count, const gs_instance* inst, const gs_nr_options& o, gs_nr_result* results,
               double* vm_out, double* va_out, bool plain, std::vector<HostFinal>* finals,
               std::vector<InstState>* final_states, RunArena* arena) {
  gs_ctx* ctx = net->ctx;
  RunArena& a = arena ? *arena : net->arena;
  std::vector<InstState> states;
  auto t0 = std::chrono::steady_clock::now();
  int launches = 0;
  double kms = launch_instances(ctx, net->dev, a.inst_buf, a.desc_buf, a.state_buf, a.aux_buf, count, inst, &net->y,
                                &net->net, net->spec, o, plain, states, true, nullptr, nullptr, &launches);
  BatchRun run = plan_run(net->dev.h, o, plain);
  Layout L = instance_layout(net->dev.h, run.fp32, run.x_lds);
  collect_results(net, count, inst, o, plain, states, a.inst_buf, L, results, vm_out, va_out, kms, launches, finals,
                  final_states);
  net->last_kernel_ms = kms;
  net->last_launches = launches;
  net->last_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ================================================================================================ reactive limits
// The outer loop of SparseNewtonRaphsonPowerFlow.SolveAcAttempt (:118-160) over a batch: after a converged run each
// instance's PV buses are checked against their generators' limits (NewtonShell.ApplyQLimits :519-558, RevertQLimits
// off, then RevertBrokenPins :583-595); the instances that switched a bus are grouped by their bus types, each group
// gets the symbolic plan of its Jacobian pattern, and the group runs again on the device from its current state with
// its pinned Q specification. At most kMaxQLimitPasses runs; a switch after the last one gets one more run and the
// instance is reported not converged (as the reference does at its pass cap).
namespace {

enum { kPinNone = 0, kPinMax = 1, kPinMin = 2, kPinFixed = 3 };
constexpr double kQLimitTolMva = 1e-3, kRevertVoltTolPu = 1e-4;
constexpr int kMaxQLimitPasses = 100;  // PowerFlowOptions.MaxQLimitPasses

struct QState {
  BusSpec spec;                          // the instance's own (load-scaled) specification
  std::vector<int> eff, pinned, reverts;
  std::vector<double> vm, va, p, q;      // current state (pu, rad; injections pu)
};

bool apply_q_limits(QState& st, double base) {
  const BusSpec& s = st.spec;
  const int n = (int)st.eff.size();
  bool changed = false;
  for (int i = 0; i < n; i++) {
    if (s.type[i] != GS_BUS_PV || st.eff[i] != GS_BUS_PV) continue;
    const double qgen = st.q[i] * base + s.qd[i];
    if (qgen > s.qmax[i] + kQLimitTolMva) {
      st.eff[i] = GS_BUS_PQ;
      st.spec.qspec[i] = (s.qmax[i] - s.qd[i]) / base;
      st.pinned[i] = s.qmin[i] == s.qmax[i] ? kPinFixed : kPinMax;
      changed = true;
    } else if (qgen < s.qmin[i] - kQLimitTolMva) {
      st.eff[i] = GS_BUS_PQ;
      st.spec.qspec[i] = (s.qmin[i] - s.qd[i]) / base;
      st.pinned[i] = s.qmin[i] == s.qmax[i] ? kPinFixed : kPinMin;
      changed = true;
    }
  }
  if (changed) return true;
  for (int i = 0; i < n; i++) {
    if (s.type[i] != GS_BUS_PV || st.eff[i] != GS_BUS_PQ) continue;
    if (st.pinned[i] == kPinNone || st.pinned[i] == kPinFixed || st.reverts[i] >= 1) continue;
    const bool rmax = st.pinned[i] == kPinMax && st.vm[i] > s.vg[i] + kRevertVoltTolPu;
    const bool rmin = st.pinned[i] == kPinMin && st.vm[i] < s.vg[i] - kRevertVoltTolPu;
    if (!(rmax || rmin)) continue;
    st.eff[i] = GS_BUS_PV;
    st.vm[i] = s.vg[i];
    st.pinned[i] = kPinNone;
    st.reverts[i]++;
    changed = true;
  }
  return changed;
}

// A run's result added to the instance's earlier runs (iterations, refreshes, step failures and device time add up).
void add_run(gs_nr_result& total, const gs_nr_result& run) {
  const gs_nr_result before = total;
  total = run;
  total.iterations += before.iterations;
  total.symbolic_refreshes += before.symbolic_refreshes;
  total.step_failures += before.step_failures;
  total.min_pivot_ratio = std::min(before.min_pivot_ratio, run.min_pivot_ratio);
  for (int k = 0; k < 4; k++) total.phase_us[k] += before.phase_us[k];
}

}  // namespace

// After run_batch (the first run of every instance, its blocks still in inst_buf with layout L and paused
// instances' finals in `paused`): the reactive-limit passes. Returns the final state of every instance that ran
// again.
std::vector<HostFinal> q_limit_passes(gs_net* net, int count, const gs_instance* inst, const gs_nr_options& o,
                                      gs_nr_result* results, const Layout& L, const std::vector<HostFinal>& paused,
                                      const DevMem& inst_buf, std::vector<InstState>* final_states) {
  gs_ctx* ctx = net->ctx;
  hipStream_t s = ctx->stream;
  const int n = net->net.n;
  const double base = net->net.base;
  // every instance's state after its first run
  std::vector<double> vm((size_t)count * n), va((size_t)count * n), pp((size_t)count * n), qq((size_t)count * n);
  const size_t offs[4] = {L.vm, L.va, L.p, L.q};
  double* dsts[4] = {vm.data(), va.data(), pp.data(), qq.data()};
  for (int w = 0; w < 4; w++)
    check(hipMemcpy2DAsync(dsts[w], 8 * (size_t)n, inst_buf.as<char>() + offs[w], L.stride, 8 * (size_t)n,
                           (size_t)count, hipMemcpyDeviceToHost, s), "state for the limits");
  check(hipStreamSynchronize(s), "state for the limits");
  std::map<int, const HostFinal*> host_done;
  for (const HostFinal& f : paused) host_done[f.i] = &f;
  std::map<int, QState> st;   // the instances still switching
  for (int i = 0; i < count; i++) {
    if (results[i].status != GS_PF_CONVERGED) continue;
    QState q;
    q.spec = build_bus_spec(net->net, net->flat_start, inst[i].load_scale);
    q.eff = q.spec.type;
    q.pinned.assign(n, kPinNone);
    q.reverts.assign(n, 0);
    auto hf = host_done.find(i);
    if (hf != host_done.end() && (int)hf->second->q.size() == n) {
      q.vm = hf->second->vm; q.va = hf->second->va; q.p = hf->second->p; q.q = hf->second->q;
    } else {
      const size_t b = (size_t)i * n;
      q.vm.assign(vm.begin() + b, vm.begin() + b + n); q.va.assign(va.begin() + b, va.begin() + b + n);
      q.p.assign(pp.begin() + b, pp.begin() + b + n); q.q.assign(qq.begin() + b, qq.begin() + b + n);
    }
    if (apply_q_limits(q, base)) st.emplace(i, std::move(q));
  }
  std::map<int, HostFinal> out;
  gs_nr_options og = o;
  og.flags &= ~GS_FLAG_Q_LIMITS;
  for (int run = 1; run <= kMaxQLimitPasses && !st.empty(); run++) {
    const bool last = run == kMaxQLimitPasses;   // the run after a switch at the pass cap
    std::map<std::vector<int>, std::vector<int>> groups;
    for (auto& kv : st) groups[kv.second.eff].push_back(kv.first);
    std::vector<int> still;
    for (auto& g : groups) {
      const std::vector<int>& types = g.first;
      const std::vector<int>& ids = g.second;
      const int cnt = (int)ids.size();
      JacobianPattern jp = build_jacobian_pattern(net->y, types);
      Plan plan;
      std::string why;
      if (!symbolic_pass(net->y, jp, st[ids[0]].vm.data(), st[ids[0]].va.data(), plan, why, og.growth_tol)) {
        // singular at this state: the run cannot take a step (NewtonShell.SolveStep false) and stops there
        for (int i : ids) {
          results[i].status = GS_PF_MAX_ITERATIONS;
          results[i].step_failures++;
          QState& q = st[i];
          out[i] = HostFinal{i, q.vm, q.va, q.p, q.q};
          if (final_states) {
            // the pass could not take its first step (NewtonShell.SolveStep false): a run of no steps that ended at
            // the mismatch it started from
            InstState& e = (*final_states)[i];
            e.iter = 0; e.converged = 0; e.diverged = 0;
            e.start_mis = results[i].max_mismatch;
            e.last_from = __builtin_nan("");
          }
        }
        continue;
      }
      BusSpec gspec = net->spec;
      gspec.type = types;
      DevicePlan dp;
      upload_plan(dp, &net->y, jp, &gspec, plan, base, s);
      std::vector<gs_instance> gi(cnt);
      std::vector<std::vector<double>> spec_ov(cnt), start_ov(cnt);
      for (int k = 0; k < cnt; k++) {
        const QState& q = st[ids[k]];
        gi[k] = inst[ids[k]];
        gi[k].vm0 = gi[k].va0 = nullptr;
        spec_ov[k] = q.spec.pspec;
        spec_ov[k].insert(spec_ov[k].end(), q.spec.qspec.begin(), q.spec.qspec.end());
        start_ov[k] = q.vm;
        start_ov[k].insert(start_ov[k].end(), q.va.begin(), q.va.end());
      }
      DevMem ib, db, sb, ab;
      std::vector<InstState> states;
      int launches = 0;
      // (the plan was made at the first instance's state: only that instance's first step is the plan's own, so
      // the group's steps are checked as usual)
      double kms = launch_instances(ctx, dp, ib, db, sb, ab, cnt, gi.data(), &net->y, &net->net, gspec, og, false,
                                    states, true, &spec_ov, &start_ov, &launches);
      const BatchRun rg = plan_run(dp.h, og, false);
      const Layout Lg = instance_layout(dp.h, rg.fp32, rg.x_lds);
      for (int k = 0; k < cnt; k++) {
        const int i = ids[k];
        QState& q = st[i];
        gs_nr_result r{};
        const char* b = ib.as<char>() + Lg.stride * (size_t)k;
        double* dst[4] = {q.vm.data(), q.va.data(), q.p.data(), q.q.data()};
        const size_t src[4] = {Lg.vm, Lg.va, Lg.p, Lg.q};
        for (int w = 0; w < 4; w++)
          check(hipMemcpyAsync(dst[w], b + src[w], 8 * (size_t)n, hipMemcpyDeviceToHost, s), "state back");
        check(hipStreamSynchronize(s), "state back");
        InstState run_end = states[k];
        if (states[k].done == 1) {
          fill_result(states[k], r, 0, ctx->us_per_tick);
        } else {
          std::vector<double> p2, q2;
          finish_paused(net, gi[k], og, false, states[k], q.vm, q.va, r, kms, launches, &p2, &q2, &types,
                        &q.spec.qspec, &run_end);
          if ((int)p2.size() == n) q.p = p2;
          if ((int)q2.size() == n) q.q = q2;
        }
        add_run(results[i], r);
        out[i] = HostFinal{i, q.vm, q.va, q.p, q.q};
        if (final_states) (*final_states)[i] = run_end;
        if (last) {
          // the cap came right after a switch: this run's state is reported, not converged
          if (results[i].status == GS_PF_CONVERGED) results[i].status = GS_PF_MAX_ITERATIONS;
        } else if (r.status == GS_PF_CONVERGED && apply_q_limits(q, base)) {
          still.push_back(i);
        }
      }
    }
    std::map<int, QState> next;
    for (int i : still) next.emplace(i, std::move(st[i]));
    st.swap(next);
  }
  std::vector<HostFinal> finals;
  for (auto& kv : out) finals.push_back(std::move(kv.second));
  return finals;
}

// ================================================================================================ DC-start retry
// NewtonShell.SolveWithRetry (NewtonShell.cs:227-268): an instance whose last Newton run did not converge, and stalled
// or diverged rather than being stopped by the iteration cap while still converging (Stalled :162,200-206), is solved
// again from its ordinary start magnitudes and the DC angles of its own network (TrySeedDcAngles :282-299), with the
// same options; the retry's state is reported when it converges with every branch angle within MaxRetryBranchAngleDeg
// (:142), otherwise the first run's, with the outcome in gs_nr_result.dc_retry.
constexpr double kRetryStallFactor = 1e-3, kRetryLastStepFactor = 0.5, kMaxRetryBranchAngleDeg = 90.0;

// The retry's stall test on how the last Newton run ended (its final state e) and its final mismatch.
bool stalled(const InstState& e, double final_mis, int max_iterations) {
  const bool cap = !e.converged && !e.diverged && e.iter >= max_iterations;
  if (cap && e.iter == 0) return false;
  if (final_mis <= kRetryStallFactor * e.start_mis) return false;
  if (cap && e.iter >= 2 && final_mis <= kRetryLastStepFactor * e.last_from) return false;
  return true;
}

std::vector<HostFinal> solve_all(gs_net* net, int count, const gs_instance* inst, const gs_nr_options& o,
                                 gs_nr_result* results, double* vm_out, double* va_out, RunArena* arena);

// The retry of every instance that qualifies (ends: the final state of each instance's last Newton run). Returns the
// final states of the instances whose retry is reported.
std::vector<HostFinal> dc_retry_pass(gs_net* net, int count, const gs_instance* inst, const gs_nr_options& o,
                                     gs_nr_result* results, const std::vector<InstState>& ends) {
  const int n = net->net.n;
  std::vector<int> ids;
  std::vector<Network> nets;
  std::vector<std::vector<double>> starts;   // vm | va of each retried instance
  for (int i = 0; i < count; i++) {
    if (results[i].status == GS_PF_CONVERGED || ends[i].converged || o.max_iterations <= 1) continue;
    if (!stalled(ends[i], results[i].max_mismatch, o.max_iterations)) continue;
    Network own = instance_network(net->net, inst[i]);
    const BusSpec spec = build_bus_spec(own, net->flat_start, 1.0);
    std::vector<double> vm = inst[i].vm0 ? std::vector<double>(inst[i].vm0, inst[i].vm0 + n) : spec.vm;
    std::vector<double> va = inst[i].va0 ? std::vector<double>(inst[i].va0, inst[i].va0 + n) : spec.va;
    if (!dc_seed_angles(own, spec.type, va)) {
      results[i].dc_retry = GS_RETRY_UNAVAILABLE;
      continue;
    }
    vm.insert(vm.end(), va.begin(), va.end());
    ids.push_back(i);
    nets.push_back(std::move(own));
    starts.push_back(std::move(vm));
  }
  if (ids.empty()) return {};
  const int k = (int)ids.size();
  std::vector<gs_instance> ri(k);
  for (int j = 0; j < k; j++) {
    ri[j] = inst[ids[j]];
    ri[j].vm0 = starts[j].data();
    ri[j].va0 = starts[j].data() + n;
  }
  gs_nr_options ro = o;
  ro.flags &= ~GS_FLAG_DC_RETRY;
  if (ro.schedule == GS_SCHEDULE_LEVELS) ro.schedule = 0;   // a handful of instances: a wavefront each
  std::vector<gs_nr_result> rr(k);
  std::vector<double> rvm((size_t)k * n), rva((size_t)k * n);
  RunArena arena;
  const double kms = net->last_kernel_ms;
  const int nl = net->last_launches;
  std::map<int, HostFinal> fin;
  for (HostFinal& f : solve_all(net, k, ri.data(), ro, rr.data(), rvm.data(), rva.data(), &arena))
    fin[f.i] = std::move(f);
  net->last_kernel_ms += kms;
  net->last_launches += nl;
  // the injections of the retried instances the device finished, from their blocks
  const BatchRun run = plan_run(net->dev.h, ro, false);
  const Layout L = instance_layout(net->dev.h, run.fp32, run.x_lds);
  std::vector<double> pp((size_t)k * n), qq((size_t)k * n);
  hipStream_t s = net->ctx->stream;
  check(hipMemcpy2DAsync(pp.data(), 8 * (size_t)n, arena.inst_buf.as<char>() + L.p, L.stride, 8 * (size_t)n,
                         (size_t)k, hipMemcpyDeviceToHost, s), "retry p");
  check(hipMemcpy2DAsync(qq.data(), 8 * (size_t)n, arena.inst_buf.as<char>() + L.q, L.stride, 8 * (size_t)n,
                         (size_t)k, hipMemcpyDeviceToHost, s), "retry q");
  check(hipStreamSynchronize(s), "retry injections");
  std::vector<HostFinal> out;
  for (int j = 0; j < k; j++) {
    const int i = ids[j];
    if (rr[j].status != GS_PF_CONVERGED) {
      results[i].dc_retry = GS_RETRY_DID_NOT_CONVERGE;
      continue;
    }
    const double* va = rva.data() + (size_t)j * n;
    if (max_branch_angle_deg(nets[j], va) > kMaxRetryBranchAngleDeg) {
      results[i].dc_retry = GS_RETRY_REJECTED;
      continue;
    }
    results[i] = rr[j];
    results[i].dc_retry = GS_RETRY_CONVERGED;
    HostFinal f{i, std::vector<double>(rvm.begin() + (size_t)j * n, rvm.begin() + (size_t)(j + 1) * n),
                std::vector<double>(va, va + n), {}, {}};
    auto it = fin.find(j);
    if (it != fin.end() && (int)it->second.p.size() == n) {
      f.p = it->second.p;
      f.q = it->second.q;
    } else {
      f.p.assign(pp.begin() + (size_t)j * n, pp.begin() + (size_t)(j + 1) * n);
      f.q.assign(qq.begin() + (size_t)j * n, qq.begin() + (size_t)(j + 1) * n);
    }
    out.push_back(std::move(f));
  }
  return out;
}

// The whole solve of a call: the first run of every instance (laid out in `arena`, or the network's own when null),
// then, as the flags ask, the reactive-limit passes and the DC-start retry. vm_out / va_out (when non-null) get every
// instance's final state; the returned list holds the final states (vm, va, and p, q when known) of the instances the
// host finished or ran again, whose device blocks still hold their first run.
std::vector<HostFinal> solve_all(gs_net* net, int count, const gs_instance* inst, const gs_nr_options& o,
                                 gs_nr_result* results, double* vm_out, double* va_out, RunArena* arena) {
  const bool qlim = o.flags & GS_FLAG_Q_LIMITS, retry = o.flags & GS_FLAG_DC_RETRY;
  RunArena& a = arena ? *arena : net->arena;
  std::vector<HostFinal> first;
  std::vector<InstState> ends;
  run_batch(net, count, inst, o, results, vm_out, va_out, false, &first, retry ? &ends : nullptr, &a);
  std::map<int, HostFinal> fin;
  for (HostFinal& f : first) fin[f.i] = std::move(f);
  if (qlim) {
    const BatchRun run = plan_run(net->dev.h, o, false);
    const Layout L = instance_layout(net->dev.h, run.fp32, run.x_lds);
    std::vector<HostFinal> paused;
    for (auto& kv : fin) paused.push_back(kv.second);
    for (HostFinal& f : q_limit_passes(net, count, inst, o, results, L, paused, a.inst_buf, retry ? &ends : nullptr))
      fin[f.i] = std::move(f);
  }
  if (retry)
    for (HostFinal& f : dc_retry_pass(net, count, inst, o, results, ends)) fin[f.i] = std::move(f);
  const int n = net->net.n;
  std::vector<HostFinal> out;
  for (auto& kv : fin) {
    const HostFinal& f = kv.second;
    if (vm_out) std::copy(f.vm.begin(), f.vm.end(), vm_out + (size_t)n * f.i);
    if (va_out) std::copy(f.va.begin(), f.va.end(), va_out + (size_t)n * f.i);
    out.push_back(std::move(kv.second));
  }
  return out;
}

// ================================================================================================ seam solves
struct SeamPlan {
  DevicePlan dev;
};

// Byte offsets of the seam's per-call device buffers for a system of h.m rows and nnz stored entries on plan h: the
// matrix values and right-hand side as given, the solution, the factor arrays, the global solve vector, the row
// floors, the status word and the kernel's descriptor.
struct SeamLayout {
  size_t a, b, x, vals, dval, xs, rf, status, desc, total;
};

SeamLayout seam_layout(const DevPlan& h, int nnz) {
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  const size_t m = (size_t)h.m;
  SeamLayout L{};
  size_t o = 0;
  L.a = o; o += al(8 * (size_t)std::max(nnz, 1));
  L.b = o; o += al(8 * m);
  L.x = o; o += al(8 * m);
  L.vals = o; o += al(8 * (size_t)vals_slots(h));
  L.dval = o; o += al(8 * (size_t)dval_slots(h));
  L.xs = o; o += al(8 * m);
  L.rf = o; o += al(8 * m);
  L.status = o; o += al(sizeof(int32_t));
  L.desc = o; o += al(sizeof(SeamDesc));
  L.total = o;
  return L;
}

// One seam system on plan h with its per-call buffers at `base` (device memory laid out by seam_layout): the values
// and right-hand side go up, the kernel replays the factorisation exactly, x comes back. Returns the kernel's status
// (0 solved, 1 refused).
int seam_run(const DevPlan& h, char* base, const SeamLayout& L, int nnz, const double* val, const double* b,
             std::vector<double>& xt, hipStream_t s) {
  const size_t m = (size_t)h.m;
  SeamDesc d{reinterpret_cast<double*>(base + L.a),    reinterpret_cast<double*>(base + L.b),
             reinterpret_cast<double*>(base + L.x),    reinterpret_cast<double*>(base + L.vals),
             reinterpret_cast<double*>(base + L.dval), reinterpret_cast<double*>(base + L.xs),
             reinterpret_cast<double*>(base + L.rf),   reinterpret_cast<int32_t*>(base + L.status)};
  const int32_t refused = 1;
  check(hipMemcpyAsync(base + L.a, val, 8 * (size_t)nnz, hipMemcpyHostToDevice, s), "seam a");
  check(hipMemcpyAsync(base + L.b, b, 8 * m, hipMemcpyHostToDevice, s), "seam b");
  check(hipMemcpyAsync(base + L.status, &refused, sizeof(refused), hipMemcpyHostToDevice, s), "seam status");
  check(hipMemcpyAsync(base + L.desc, &d, sizeof(SeamDesc), hipMemcpyHostToDevice, s), "seam desc");
  check(launch_seam(reinterpret_cast<const SeamDesc*>(base + L.desc), h, 1, 8 * (size_t)ws_slots(h), s), "seam kernel");
  int32_t status = 1;
  xt.resize(m);
  check(hipMemcpyAsync(&status, base + L.status, sizeof(status), hipMemcpyDeviceToHost, s), "seam status");
  check(hipMemcpyAsync(xt.data(), base + L.x, 8 * m, hipMemcpyDeviceToHost, s), "seam x");
  // the host arguments and the descriptor are read by the stream: finish before they go
  check(hipStreamSynchronize(s), "seam sync");
  return status;
}

// `count` systems of one structure on plan h in one launch, a wave each, with their buffers one after another at
// `base` (count x L.total bytes, then the descriptors and the status words): the values and right-hand sides go up,
// the kernel replays the factorisation of each, the solutions come back into xt. status[i]: 0 solved, 1 refused.
void seam_run_many(const DevPlan& h, char* base, const SeamLayout& L, int nnz, int count, const double* const* val,
                   const double* const* b, std::vector<std::vector<double>>& xt, std::vector<int32_t>& status,
                   hipStream_t s) {
  const size_t m = (size_t)h.m;
  char* desc_base = base + L.total * (size_t)count;
  char* status_base = desc_base + ((sizeof(SeamDesc) * (size_t)count + 255) & ~size_t(255));
  std::vector<SeamDesc> descs(count);
  for (int i = 0; i < count; i++) {
    char* p = base + L.total * (size_t)i;
    descs[i] = SeamDesc{reinterpret_cast<double*>(p + L.a),    reinterpret_cast<double*>(p + L.b),
                        reinterpret_cast<double*>(p + L.x),    reinterpret_cast<double*>(p + L.vals),
                        reinterpret_cast<double*>(p + L.dval), reinterpret_cast<double*>(p + L.xs),
                        reinterpret_cast<double*>(p + L.rf),   reinterpret_cast<int32_t*>(status_base) + i};
    check(hipMemcpyAsync(p + L.a, val[i], 8 * (size_t)nnz, hipMemcpyHostToDevice, s), "seam a");
    check(hipMemcpyAsync(p + L.b, b[i], 8 * m, hipMemcpyHostToDevice, s), "seam b");
  }
  status.assign(count, 1);
  check(hipMemcpyAsync(status_base, status.data(), 4 * (size_t)count, hipMemcpyHostToDevice, s), "seam status");
  check(hipMemcpyAsync(desc_base, descs.data(), sizeof(SeamDesc) * (size_t)count, hipMemcpyHostToDevice, s),
        "seam desc");
  check(launch_seam(reinterpret_cast<const SeamDesc*>(desc_base), h, count, 8 * (size_t)ws_slots(h), s),
        "seam kernel");
  check(hipMemcpyAsync(status.data(), status_base, 4 * (size_t)count, hipMemcpyDeviceToHost, s), "seam status");
  xt.resize(count);
  for (int i = 0; i < count; i++) {
    xt[i].resize(m);
    check(hipMemcpyAsync(xt[i].data(), base + L.total * (size_t)i + L.x, 8 * m, hipMemcpyDeviceToHost, s), "seam x");
  }
  // the host arguments and the descriptors are read by the stream: finish before they go
  check(hipStreamSynchronize(s), "seam sync");
}

// Bytes seam_run_many needs for `count` systems of layout L.
size_t seam_many_bytes(const SeamLayout& L, int count) {
  return L.total * (size_t)count + ((sizeof(SeamDesc) * (size_t)count + 255) & ~size_t(255)) + 4 * (size_t)count;
}

// The seam's checks of its arguments (CpuSparseLuSolver.Solve and SparseLuFactorization.Factor): GS_OK, or the
// status to return with ctx->err set; a malformed call throws GS_INVALID_ARGUMENT.
gs_status seam_check(gs_ctx* ctx, int n, const int* rowptr, const int* colidx, const double* val, const double* b,
                     const double* x) {
  if (n < 0 || !rowptr || (n > 0 && (!colidx || !val || !b || !x))) fail(GS_INVALID_ARGUMENT, "null argument");
  if (n == 0) return GS_OK;
  if (rowptr[0] != 0) fail(GS_INVALID_ARGUMENT, "rowptr[0] must be 0");
  for (int i = 0; i < n; i++) {
    if (rowptr[i + 1] < rowptr[i]) fail(GS_INVALID_ARGUMENT, "rowptr must be non-decreasing");
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      if (colidx[k] < 0 || colidx[k] >= n) fail(GS_INVALID_ARGUMENT, "column index out of range");
      if (k > rowptr[i] && colidx[k] <= colidx[k - 1]) fail(GS_INVALID_ARGUMENT, "columns must be sorted per row");
    }
  }
  for (int k = 0; k < rowptr[n]; k++)
    if (!std::isfinite(val[k])) {
      ctx->err = "Sparse LU: the matrix contains a non-finite entry, so it cannot be factorised.";
      return GS_NONFINITE_INPUT;
    }
  return GS_OK;
}

gs_status seam_rhs_check(gs_ctx* ctx, int n, const double* b) {
  for (int k = 0; k < n; k++)
    if (!std::isfinite(b[k])) {
      ctx->err = "Sparse LU: the right-hand side has a non-finite entry.";
      return GS_NONFINITE_INPUT;
    }
  return GS_OK;
}

// The minimum-degree order of a pattern, cached per pattern as CpuSparseLuSolver caches it (CpuSparseLuSolver.cs:34-39).
const std::vector<int>& seam_order(gs_ctx* ctx, int n, const int* rowptr, const int* colidx) {
  std::vector<int> key;
  key.reserve(1 + (size_t)n + 1 + (size_t)rowptr[n]);
  key.push_back(n);
  key.insert(key.end(), rowptr, rowptr + n + 1);
  key.insert(key.end(), colidx, colidx + rowptr[n]);
  if (std::vector<int>* hit = ctx->seam_orders.find(key)) return *hit;
  return ctx->seam_orders.put(key, minimum_degree(n, rowptr, colidx));
}

}  // namespace gs

namespace {

// A device plan is keyed by the whole structure it was made for: the pattern and the positions of the nonzero
// values (the factorisation drops exact zeros, as SparseLu.ToCsc does).
std::vector<int> seam_key(int n, const int* rowptr, const int* colidx, const double* val) {
  const int nnz = rowptr[n];
  std::vector<int> key;
  key.reserve(2 + (size_t)n + 2 * (size_t)nnz);
  key.push_back(n);
  key.insert(key.end(), rowptr, rowptr + n + 1);
  key.insert(key.end(), colidx, colidx + nnz);
  key.push_back(-1);
  for (int k = 0; k < nnz; k++) if (val[k] != 0.0) key.push_back(k);
  return key;
}

gs_status seam_solve_one(gs_ctx* ctx, int n, const int* rowptr, const int* colidx, const double* val,
                         const double* b, double* x) {
  if (gs_status st = seam_check(ctx, n, rowptr, colidx, val, b, x)) return st;
  if (n == 0) return GS_OK;
  const int nnz = rowptr[n];
  const std::vector<int>& order = seam_order(ctx, n, rowptr, colidx);
  const std::vector<int> key = seam_key(n, rowptr, colidx, val);
  hipStream_t s = ctx->stream;
  std::vector<double> xt;
  if (std::shared_ptr<SeamPlan>* hit = ctx->seam_plans.find(key)) {
    std::shared_ptr<SeamPlan> sp = *hit;
    const SeamLayout L = seam_layout(sp->dev.h, nnz);
    ctx->seam_buf.ensure(L.total);
    if (seam_run(sp->dev.h, ctx->seam_buf.as<char>(), L, nnz, val, b, xt, s) == 0) {
      if (gs_status st = seam_rhs_check(ctx, n, b)) return st;
      ctx->seam_plan_hits++;
      std::copy(xt.begin(), xt.end(), x);
      return GS_OK;
    }
  }
  // no plan for this structure, or a pivot of the cached one would move: the factorisation afresh on the host
  Plan plan;
  std::string why;
  std::vector<int> rp(rowptr, rowptr + n + 1), ci(colidx, colidx + nnz);
  if (!gp_plan(n, rp, ci, val, order, 0.0, false, plan, why)) {
    ctx->err = why;
    return GS_SINGULAR;
  }
  if (gs_status st = seam_rhs_check(ctx, n, b)) return st;
  JacobianPattern jp;
  jp.m = n;
  jp.rp = std::move(rp);
  jp.ci = std::move(ci);
  auto sp = std::make_shared<SeamPlan>();
  upload_plan(sp->dev, nullptr, jp, nullptr, plan, 100.0, s);
  const SeamLayout L = seam_layout(sp->dev.h, nnz);
  ctx->seam_buf.ensure(L.total);
  if (seam_run(sp->dev.h, ctx->seam_buf.as<char>(), L, nnz, val, b, xt, s) != 0)
    fail(GS_DEVICE_ERROR, "the device refused the replay of its own factorisation");
  ctx->seam_plans.put(key, sp);
  ctx->seam_plan_builds++;
  std::copy(xt.begin(), xt.end(), x);
  return GS_OK;
}

// gs_sparse_solve with no device allocation: the factorisation's plan is made on the host as gs_sparse_solve makes a
// new one, and its image and the call's buffers are written into the context's device's scratch array (address
// ctx->scratch, from hipGetSymbolAddress at gs_init), then the same kernel replays the factorisation and solves. The
// same arithmetic as gs_sparse_solve, so the same bits.
gs_status seam_solve_noalloc(gs_ctx* ctx, int n, const int* rowptr, const int* colidx, const double* val,
                             const double* b, double* x) {
  if (gs_status st = seam_check(ctx, n, rowptr, colidx, val, b, x)) return st;
  if (n == 0) return GS_OK;
  if (!ctx->scratch) fail(GS_DEVICE_ERROR, "the scratch array has no device address on this device");
  const int nnz = rowptr[n];
  hipStream_t s = ctx->stream;
  std::vector<int> key = seam_key(n, rowptr, colidx, val);
  {
    // the plan the array already holds, when it was made for this structure: replay it
    ScratchPlan resident;
    {
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      auto it = g_scratch_plan.find(ctx->device);
      if (it != g_scratch_plan.end() && it->second.key == key) resident = it->second;
    }
    if (!resident.key.empty()) {
      const SeamLayout L = seam_layout(resident.h, nnz);
      std::vector<double> xt;
      if (seam_run(resident.h, ctx->scratch + resident.plan_bytes, L, nnz, val, b, xt, s) == 0) {
        if (gs_status st = seam_rhs_check(ctx, n, b)) return st;
        std::copy(xt.begin(), xt.end(), x);
        return GS_OK;
      }
    }
  }
  const std::vector<int>& order = seam_order(ctx, n, rowptr, colidx);
  Plan plan;
  std::string why;
  std::vector<int> rp(rowptr, rowptr + n + 1), ci(colidx, colidx + nnz);
  if (!gp_plan(n, rp, ci, val, order, 0.0, false, plan, why)) {
    ctx->err = why;
    return GS_SINGULAR;
  }
  if (gs_status st = seam_rhs_check(ctx, n, b)) return st;
  JacobianPattern jp;
  jp.m = n;
  jp.rp = std::move(rp);
  jp.ci = std::move(ci);
  PlanImage img;
  DevPlan h;
  DevicePlanHost dph;
  build_plan_image(img, h, dph, nullptr, jp, nullptr, plan, 100.0);
  const size_t plan_bytes = (img.bytes.size() + 255) & ~size_t(255);
  const SeamLayout L = seam_layout(h, nnz);
  if (plan_bytes + L.total > kNoAllocScratchBytes)
    fail(GS_INVALID_ARGUMENT, "the system needs " + std::to_string(plan_bytes + L.total) +
                                  " bytes of device scratch; gs_sparse_solve_noalloc holds " +
                                  std::to_string(kNoAllocScratchBytes));
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    g_scratch_plan[ctx->device].key.clear();   // the array's image is about to change
  }
  check(hipMemcpyAsync(ctx->scratch, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice, s), "scratch plan");
  img.bind(h, ctx->scratch);
  std::vector<double> xt;
  // (seam_run finishes the stream before it returns, so the image above is consumed before it goes)
  if (seam_run(h, ctx->scratch + plan_bytes, L, nnz, val, b, xt, s) != 0)
    fail(GS_DEVICE_ERROR, "the device refused the replay of its own factorisation");
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ScratchPlan& r = g_scratch_plan[ctx->device];
    r.key = std::move(key);
    r.h = h;
    r.plan_bytes = plan_bytes;
  }
  std::copy(xt.begin(), xt.end(), x);
  return GS_OK;
}

}  // namespace

// ================================================================================================ multi-network batches
struct gs_batch {
  gs_ctx* ctx = nullptr;
  struct Group {
    gs_net* net = nullptr;
    int count = 0;
    std::vector<gs_instance> inst;
    std::vector<std::vector<int32_t>> outage_lists;   // the instances' outage lists (inst[i].outages points here)
    gs_nr_options opts{};
    DevMem inst_buf, desc_buf, state_buf, aux_buf;
    Staging stg;
    PreparedRun pr;
    hipStream_t stream = nullptr;
    hipEvent_t joined = nullptr;
    std::unique_ptr<WideGraph> wide;     // the level-parallel schedule's control block, side stream and events
    std::vector<InstState> states;
    std::vector<gs_nr_result> results;
    std::vector<HostFinal> finals;      // instances the host finished (fresh symbolic passes): their final state
  };
  std::vector<std::unique_ptr<Group>> groups;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t fork = nullptr, e0 = nullptr, e1 = nullptr;
  bool ran = false;
  ~gs_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& g : groups) {
      if (g->joined) (void)hipEventDestroy(g->joined);
      if (g->stream) (void)hipStreamDestroy(g->stream);
    }
    for (hipEvent_t e : {fork, e0, e1})
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

// A group's launches on its stream. The level-parallel schedule cannot poll the host from inside a graph, so its
// Newton steps are unrolled: the start kernel, then max_iterations + 1 steps (as many as launch_wide replays at most);
// every kernel of a step returns at once for an instance that has finished or paused, so the steps after the last
// instance ends cost only their dispatches.
void enqueue_group(gs_batch::Group& g) {
  const DevicePlan& dp = g.net->dev;
  if (g.opts.schedule != GS_SCHEDULE_LEVELS) {
    launch_prepared(dp.h, g.pr, g.opts.schedule, g.stream);
    return;
  }
  const DevicePlanHost& d = dp.dph;
  WideGraph& wg = *g.wide;
  WideCtl* ctl = wg.ctl.as<WideCtl>();
  const WideSplit split = forked_split(d, wg);
  const size_t ws_bytes = (g.pr.run.fp32 ? 4 : 8) * (size_t)ws_slots(dp.h);
  check(launch_wide_start(g.pr.desc, dp.h, g.count, ctl, g.stream), "level-parallel start");
  for (int k = 0; k <= g.pr.prm.max_iterations; k++)
    check(launch_wide_step(g.pr.run.fp32, g.pr.desc, dp.h, g.count, g.pr.prm, ctl, d.rs_level_chunk, d.fs_level_pass,
                           d.bs_level_pass, ws_bytes, split, g.stream),
          "level-parallel step");
}

// Every group's launches on its own stream, forked from and joined back into the context's stream.
void enqueue_batch(gs_batch* b) {
  hipStream_t s0 = b->ctx->stream;
  check(hipEventRecord(b->fork, s0), "fork");
  for (auto& g : b->groups) {
    check(hipStreamWaitEvent(g->stream, b->fork, 0), "fork wait");
    enqueue_group(*g);
    check(hipEventRecord(g->joined, g->stream), "join");
    check(hipStreamWaitEvent(s0, g->joined, 0), "join wait");
  }
}

gs_status batch_run(gs_batch* b, double* device_ms) {
  hipStream_t s0 = b->ctx->stream;
  if (b->use_graph && !b->exec) {
    check(hipStreamBeginCapture(s0, hipStreamCaptureModeGlobal), "capture");
    enqueue_batch(b);
    check(hipStreamEndCapture(s0, &b->graph), "end capture");
    check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate");
  }
  check(hipEventRecord(b->e0, s0), "event");
  if (b->use_graph) check(hipGraphLaunch(b->exec, s0), "graph launch");
  else enqueue_batch(b);
  check(hipEventRecord(b->e1, s0), "event");
  for (auto& g : b->groups) {
    g->states.resize(g->count);
    check(hipMemcpyAsync(g->states.data(), g->state_buf.p, sizeof(InstState) * g->count, hipMemcpyDeviceToHost, s0),
          "states back");
  }
  check(hipStreamSynchronize(s0), "batch sync");
  float ms = 0;
  check(hipEventElapsedTime(&ms, b->e0, b->e1), "elapsed");
  if (device_ms) *device_ms = ms;
  for (auto& g : b->groups) {
    g->results.assign(g->count, gs_nr_result{});
    g->finals.clear();
    double kms = 0;
    int launches = 0;
    // the voltages stay in HBM until gs_batch_results asks for them
    collect_results(g->net, g->count, g->inst.data(), g->opts, false, g->states, g->inst_buf, g->pr.L,
                    g->results.data(), nullptr, nullptr, kms, launches, &g->finals);
  }
  b->ran = true;
  return GS_OK;
}

// The autotuner's candidates: arithmetic (LU precision, refinement steps) and launch slicing. A candidate must keep
// every converged voltage within kAutotuneDv (pu, rad) of the FP64 LU's.
struct Candidate { int32_t lu, refine, schedule, flags; };
constexpr double kAutotuneDv = 1e-9;

}  // namespace

// ================================================================================================ C ABI
extern "C" {

GS_API const char* gs_version(void) { return "gridsim-hip 0.1 (gfx950)"; }

GS_API gs_status gs_init(int32_t device, size_t hbm_budget_bytes, gs_ctx** out) {
  if (!out) return GS_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return GS_NO_DEVICE;
  if (device < 0 || device >= count) return GS_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return GS_DEVICE_ERROR;
  auto* ctx = new gs_ctx();
  ctx->device = device;
  ctx->budget = hbm_budget_bytes;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
    ctx->us_per_tick = 1000.0 / khz;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return GS_DEVICE_ERROR;
  }
  // set-up of the no-allocation entry point: this device's address of its scratch array
  if (scratch_address(device, &ctx->scratch) != GS_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GS_DEVICE_ERROR;
  }
  *out = ctx;
  return GS_OK;
}

GS_API void gs_shutdown(gs_ctx* ctx) {
  if (!ctx) return;
  ctx->seam_plans.clear();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

GS_API const char* gs_last_error(const gs_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

GS_API gs_status gs_device_info_get(gs_ctx* ctx, gs_device_info* out) {
  return guarded(ctx, [&]() -> gs_status {
    if (!out) return GS_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    hipDeviceProp_t p;
    check(hipGetDeviceProperties(&p, ctx->device), "properties");
    std::snprintf(out->name, sizeof(out->name), "%s", p.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
    out->compute_units = p.multiProcessorCount;
    out->hbm_bytes = p.totalGlobalMem;
    DevMem d;
    d.ensure(3 * sizeof(double));
    check(launch_fp64_probe(d.as<double>(), ctx->stream), "probe");
    double h[3];
    check(hipMemcpyAsync(h, d.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream), "probe back");
    check(hipStreamSynchronize(ctx->stream), "probe sync");
    out->fp64_ok = h[0] == 9.094947017729282e-13;
    double s, c;
    gs_fdlibm::sincos(3.141592653589793 / 7.0, &s, &c);
    out->fdlibm_ok = (h[1] == s && h[2] == c);
    return GS_OK;
  });
}

GS_API gs_status gs_sparse_solve(gs_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                                 const double* val, const double* b, double* x) {
  return guarded(ctx, [&]() { return seam_solve_one(ctx, n, rowptr, colidx, val, b, x); });
}

GS_API gs_status gs_sparse_solve_noalloc(gs_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* colidx,
                                         const double* val, const double* b, double* x) {
  return guarded(ctx, [&]() { return seam_solve_noalloc(ctx, n, rowptr, colidx, val, b, x); });
}

GS_API gs_status gs_sparse_solve_batch(gs_ctx* ctx, int32_t count, const gs_csr* systems, const double* const* b,
                                       double* const* x, gs_status* per_system) {
  return guarded(ctx, [&]() -> gs_status {
    if (count < 0 || (count > 0 && (!systems || !b || !x || !per_system))) fail(GS_INVALID_ARGUMENT, "null argument");
    // The systems whose structure has a cached device plan go to the device together, one launch per structure
    // (a wave per system); the rest, and any whose pivots the cached plan cannot keep, one at a time as
    // gs_sparse_solve solves them (a fresh host factorisation, then a plan cached for the systems after them).
    std::map<std::vector<int>, std::vector<int>> by_plan;
    std::vector<int> alone;
    for (int i = 0; i < count; i++) {
      const gs_csr& sy = systems[i];
      per_system[i] = GS_OK;
      try {
        const gs_status st = seam_check(ctx, sy.n, sy.rowptr, sy.colidx, sy.val, b[i], x[i]);
        if (st != GS_OK) { per_system[i] = st; continue; }
        if (sy.n == 0) continue;
      } catch (const HipError& e) {
        per_system[i] = e.code;
        continue;
      }
      bool finite_rhs = true;
      for (int k = 0; k < sy.n && finite_rhs; k++) finite_rhs = std::isfinite(b[i][k]);
      std::vector<int> key = seam_key(sy.n, sy.rowptr, sy.colidx, sy.val);
      // (a non-finite right-hand side takes the single-system path, which orders its checks as the reference does)
      if (finite_rhs && ctx->seam_plans.find(key)) by_plan[std::move(key)].push_back(i);
      else alone.push_back(i);
    }
    hipStream_t s = ctx->stream;
    for (auto& [key, members] : by_plan) {
      std::shared_ptr<SeamPlan> sp = *ctx->seam_plans.find(key);
      const DevPlan& h = sp->dev.h;
      const int nnz = systems[members[0]].rowptr[systems[members[0]].n];
      const SeamLayout L = seam_layout(h, nnz);
      const int chunk = (int)std::max<size_t>(1, std::min<size_t>(members.size(), (size_t)(1u << 30) / L.total));
      for (size_t c0 = 0; c0 < members.size(); c0 += chunk) {
        const int cnt = (int)std::min<size_t>(chunk, members.size() - c0);
        std::vector<const double*> vals(cnt), rhs(cnt);
        for (int j = 0; j < cnt; j++) {
          vals[j] = systems[members[c0 + j]].val;
          rhs[j] = b[members[c0 + j]];
        }
        ctx->seam_buf.ensure(seam_many_bytes(L, cnt));
        std::vector<std::vector<double>> xt;
        std::vector<int32_t> status;
        seam_run_many(h, ctx->seam_buf.as<char>(), L, nnz, cnt, vals.data(), rhs.data(), xt, status, s);
        for (int j = 0; j < cnt; j++) {
          const int i = members[c0 + j];
          if (status[j] == 0) {
            std::copy(xt[j].begin(), xt[j].end(), x[i]);
            ctx->seam_plan_hits++;
          } else {
            alone.push_back(i);   // a pivot of the cached plan would move: afresh on the host
          }
        }
      }
    }
    std::sort(alone.begin(), alone.end());
    for (int i : alone) {
      try {
        per_system[i] = seam_solve_one(ctx, systems[i].n, systems[i].rowptr, systems[i].colidx, systems[i].val,
                                       b[i], x[i]);
      } catch (const HipError& e) {
        if (e.code == GS_DEVICE_ERROR) throw;
        per_system[i] = e.code;
      }
    }
    return GS_OK;
  });
}

GS_API gs_status gs_network_upload(gs_ctx* ctx, const gs_network* net, const gs_nr_options* opts, gs_net** out) {
  return guarded(ctx, [&]() -> gs_status {
    if (!net || !out || net->n_bus <= 0 || net->n_branch < 0 || net->n_gen < 0) fail(GS_INVALID_ARGUMENT, "network");
    *out = nullptr;
    gs_nr_options o = normalise(opts);
    auto h = std::make_unique<gs_net>();
    h->ctx = ctx;
    h->net = copy_network(*net);
    for (int k = 0; k < h->net.nb; k++)
      if (h->net.br_from[k] < 0 || h->net.br_from[k] >= h->net.n || h->net.br_to[k] < 0 || h->net.br_to[k] >= h->net.n)
        fail(GS_INVALID_ARGUMENT, "branch end outside the bus range");
    for (int g = 0; g < h->net.ng; g++)
      if (h->net.gen_bus[g] < 0 || h->net.gen_bus[g] >= h->net.n) fail(GS_INVALID_ARGUMENT, "generator bus");
    h->flat_start = o.flat_start != 0;
    h->y = build_ybus(h->net, -1);
    h->spec = build_bus_spec(h->net, h->flat_start, 1.0);
    h->jp = build_jacobian_pattern(h->y, h->spec.type);
    std::string why;
    if (!symbolic_pass(h->y, h->jp, h->spec.vm.data(), h->spec.va.data(), h->plan, why, o.growth_tol)) {
      ctx->err = why;
      return GS_SINGULAR;
    }
    upload_plan(h->dev, &h->y, h->jp, &h->spec, h->plan, h->net.base, ctx->stream);
    upload_branches(h->dev, h->net, ctx->stream);
    check(hipStreamSynchronize(ctx->stream), "upload sync");
    *out = h.release();
    return GS_OK;
  });
}

GS_API void gs_network_free(gs_net* net) { delete net; }

GS_API gs_status gs_nr_solve_batch(gs_net* net, int32_t count, const gs_instance* inst, const gs_nr_options* opts,
                                   gs_nr_result* results, double* vm_out, double* va_out) {
  if (!net) return GS_INVALID_ARGUMENT;
  return guarded(net->ctx, [&]() -> gs_status {
    if (count < 0 || (count > 0 && (!inst || !results))) fail(GS_INVALID_ARGUMENT, "null argument");
    gs_nr_options o = normalise(opts);
    if ((o.flat_start != 0) != net->flat_start) fail(GS_INVALID_ARGUMENT, "flat_start differs from the upload");
    check_instances(net->net.nb, count, inst, "");
    if (count == 0) return GS_OK;
    solve_all(net, count, inst, o, results, vm_out, va_out, nullptr);
    return GS_OK;
  });
}

GS_API gs_status gs_nr_screen_batch(gs_net* net, int32_t count, const gs_instance* inst, const gs_nr_options* opts,
                                   gs_nr_result* results, gs_screen_summary* summaries, double* loading,
                                   double* flows, double* vm_out, double* va_out, double* p_inj, double* q_inj) {
  if (!net) return GS_INVALID_ARGUMENT;
  return guarded(net->ctx, [&]() -> gs_status {
    if (count < 0 || (count > 0 && (!inst || !results || !summaries))) fail(GS_INVALID_ARGUMENT, "null argument");
    gs_nr_options o = normalise(opts);
    if ((o.flat_start != 0) != net->flat_start) fail(GS_INVALID_ARGUMENT, "flat_start differs from the upload");
    check_instances(net->net.nb, count, inst, "");
    if (count == 0) return GS_OK;
    const std::vector<HostFinal> finals = solve_all(net, count, inst, o, results, vm_out, va_out, nullptr);
    gs_ctx* ctx = net->ctx;
    hipStream_t s = ctx->stream;
    const BatchRun run = plan_run(net->dev.h, o, false);
    const Layout L = instance_layout(net->dev.h, run.fp32, run.x_lds);
    const int n = net->net.n, nb = net->net.nb;
    // instances the host finished or ran again: their final state back into their blocks, where the screen reads it
    for (const HostFinal& f : finals) {
      char* b = net->arena.inst_buf.as<char>() + L.stride * (size_t)f.i;
      check(hipMemcpyAsync(b + L.vm, f.vm.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s), "final vm");
      check(hipMemcpyAsync(b + L.va, f.va.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s), "final va");
      // (a run the host stopped before any resumed launch has no new injections: the block keeps its last ones)
      if ((int)f.p.size() == n)
        check(hipMemcpyAsync(b + L.p, f.p.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s), "final p");
      if ((int)f.q.size() == n)
        check(hipMemcpyAsync(b + L.q, f.q.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s), "final q");
    }
    net->screen_buf.ensure(sizeof(gs_screen_summary) * (size_t)count);
    if (loading) net->loading_buf.ensure(8 * (size_t)count * std::max(1, nb));
    if (flows) net->flow_buf.ensure(32 * (size_t)count * std::max(1, nb));
    check(launch_screen(net->arena.desc_buf.as<InstDesc>(), net->dev.h, count, net->screen_buf.as<gs_screen_summary>(),
                        loading ? net->loading_buf.as<double>() : nullptr, flows ? net->flow_buf.as<double>() : nullptr,
                        s),
          "screen kernel");
    check(hipMemcpyAsync(summaries, net->screen_buf.p, sizeof(gs_screen_summary) * (size_t)count,
                         hipMemcpyDeviceToHost, s), "summaries");
    if (loading)
      check(hipMemcpyAsync(loading, net->loading_buf.p, 8 * (size_t)count * nb, hipMemcpyDeviceToHost, s), "loading");
    if (flows)
      check(hipMemcpyAsync(flows, net->flow_buf.p, 32 * (size_t)count * nb, hipMemcpyDeviceToHost, s), "flows");
    // the injections at the final states (per unit on the device), scaled to MW / MVAr as AssembleResult does
    for (int w = 0; w < 2; w++) {
      double* dst = w == 0 ? p_inj : q_inj;
      if (!dst) continue;
      check(hipMemcpy2DAsync(dst, 8 * (size_t)n, net->arena.inst_buf.as<char>() + (w == 0 ? L.p : L.q), L.stride,
                             8 * (size_t)n, (size_t)count, hipMemcpyDeviceToHost, s), "injections");
    }
    check(hipStreamSynchronize(s), "screen sync");
    for (int w = 0; w < 2; w++) {
      double* dst = w == 0 ? p_inj : q_inj;
      if (dst)
        for (size_t e = 0; e < (size_t)count * n; e++) dst[e] *= net->net.base;
    }
    return GS_OK;
  });
}

GS_API gs_status gs_nr_solve_plain(gs_net* net, int32_t count, const gs_instance* inst, const gs_nr_options* opts,
                                   gs_nr_result* results, double* vm_out, double* va_out) {
  if (!net) return GS_INVALID_ARGUMENT;
  return guarded(net->ctx, [&]() -> gs_status {
    if (count < 0 || (count > 0 && (!inst || !results))) fail(GS_INVALID_ARGUMENT, "null argument");
    gs_nr_options o = normalise(opts);
    o.lu_precision = GS_LU_FP64;
    if ((o.flat_start != 0) != net->flat_start) fail(GS_INVALID_ARGUMENT, "flat_start differs from the upload");
    if (o.flags & (GS_FLAG_Q_LIMITS | GS_FLAG_DC_RETRY))
      fail(GS_INVALID_ARGUMENT, "the plain path runs neither the reactive-limit passes nor the DC-start retry");
    if (count == 0) return GS_OK;
    run_batch(net, count, inst, o, results, vm_out, va_out, true);
    return GS_OK;
  });
}

GS_API gs_status gs_network_plan_sizes(const gs_net* net, int32_t* m, int32_t* nnz_j, int32_t* nnz_l,
                                       int32_t* nnz_u) {
  if (!net) return GS_INVALID_ARGUMENT;
  if (m) *m = net->jp.m;
  if (nnz_j) *nnz_j = net->jp.rp[net->jp.m];
  if (nnz_l) *nnz_l = net->plan.lp[net->plan.m];
  if (nnz_u) *nnz_u = net->plan.up[net->plan.m];
  return GS_OK;
}

GS_API gs_status gs_network_plan_get(const gs_net* net, int32_t* q, int32_t* pivot_row, int32_t* lp, int32_t* li,
                                     int32_t* up, int32_t* ui, int32_t* j_rowptr, int32_t* j_colidx) {
  if (!net) return GS_INVALID_ARGUMENT;
  auto cp = [](int32_t* dst, const std::vector<int>& v) { if (dst) std::copy(v.begin(), v.end(), dst); };
  cp(q, net->plan.q); cp(pivot_row, net->plan.pivot_row); cp(lp, net->plan.lp); cp(li, net->plan.li);
  cp(up, net->plan.up); cp(ui, net->plan.ui); cp(j_rowptr, net->jp.rp); cp(j_colidx, net->jp.ci);
  return GS_OK;
}

GS_API gs_status gs_network_plan_stats(const gs_net* net, gs_plan_stats* out) {
  if (!net || !out) return GS_INVALID_ARGUMENT;
  *out = plan_stats(net->dev.dph, net->y, net->jp);
  return GS_OK;
}

}  // extern "C"

struct gs_plan {
  Network net;
  YBus y;
  BusSpec spec;
  JacobianPattern jp;
  Plan plan;
  DevicePlanHost dph;
};

extern "C" {

GS_API gs_status gs_plan_shape(const gs_plan* plan, gs_plan_stats* out) {
  if (!plan || !out) return GS_INVALID_ARGUMENT;
  *out = plan_stats(plan->dph, plan->y, plan->jp);
  return GS_OK;
}

GS_API gs_status gs_plan_build(const gs_network* net, const gs_nr_options* opts, gs_plan** out) {
  if (!net || !out || net->n_bus <= 0) return GS_INVALID_ARGUMENT;
  *out = nullptr;
  try {
    gs_nr_options o = normalise(opts);
    auto p = std::make_unique<gs_plan>();
    p->net = copy_network(*net);
    p->y = build_ybus(p->net, -1);
    p->spec = build_bus_spec(p->net, o.flat_start != 0, 1.0);
    p->jp = build_jacobian_pattern(p->y, p->spec.type);
    std::string why;
    if (!symbolic_pass(p->y, p->jp, p->spec.vm.data(), p->spec.va.data(), p->plan, why, o.growth_tol)) return GS_SINGULAR;
    compile_newton_plan(p->plan, p->dph, p->y, p->jp);
    *out = p.release();
    return GS_OK;
  } catch (const HipError& e) {
    return e.code;
  } catch (...) {
    return GS_INVALID_ARGUMENT;
  }
}

GS_API gs_status gs_plan_fresh_batch(const gs_network* net, const gs_nr_options* opts, int32_t count,
                                     const gs_instance* inst, const double* vm, const double* va, int32_t threads,
                                     gs_plan** plans, int64_t* order_cache_hits) {
  if (!net || net->n_bus <= 0 || count < 0 || (count > 0 && (!inst || !plans))) return GS_INVALID_ARGUMENT;
  try {
    const gs_nr_options o = normalise(opts);
    const Network base = copy_network(*net);
    check_instances(base.nb, count, inst, "");
    for (int i = 0; i < count; i++) plans[i] = nullptr;
    const int n = base.n;
    // each instance's state: the given one, or its own start point
    std::vector<std::vector<double>> own_vm(count), own_va(count);
    std::vector<const double*> pvm(count), pva(count);
    for (int i = 0; i < count; i++) {
      if (vm && va) {
        pvm[i] = vm + (size_t)n * i;
        pva[i] = va + (size_t)n * i;
      } else {
        const BusSpec sp = build_bus_spec(instance_network(base, inst[i]), o.flat_start != 0, 1.0);
        own_vm[i] = inst[i].vm0 ? std::vector<double>(inst[i].vm0, inst[i].vm0 + n) : sp.vm;
        own_va[i] = inst[i].va0 ? std::vector<double>(inst[i].va0, inst[i].va0 + n) : sp.va;
        pvm[i] = own_vm[i].data();
        pva[i] = own_va[i].data();
      }
    }
    OrderCache cache;
    std::vector<FreshSymbolic> fs;
    fresh_symbolic(base, o.flat_start != 0, o, count, inst, pvm, pva, nullptr, nullptr, std::max(1, threads), cache,
                   fs);
    if (order_cache_hits) *order_cache_hits = cache.hits();
    gs_status worst = GS_OK;
    for (int i = 0; i < count; i++) {
      if (!fs[i].ok) { worst = GS_SINGULAR; continue; }
      auto p = std::make_unique<gs_plan>();
      p->net = std::move(fs[i].own);
      p->y = std::move(fs[i].y);
      p->spec = std::move(fs[i].spec);
      p->jp = std::move(fs[i].jp);
      p->plan = std::move(fs[i].plan);
      compile_newton_plan(p->plan, p->dph, p->y, p->jp);
      plans[i] = p.release();
    }
    return worst;
  } catch (const HipError& e) {
    return e.code;
  } catch (...) {
    return GS_INVALID_ARGUMENT;
  }
}

GS_API gs_status gs_plan_sizes(const gs_plan* p, int32_t* m, int32_t* nnz_j, int32_t* nnz_l, int32_t* nnz_u,
                               int32_t* n_levels) {
  if (!p) return GS_INVALID_ARGUMENT;
  if (m) *m = p->jp.m;
  if (nnz_j) *nnz_j = p->jp.rp[p->jp.m];
  if (nnz_l) *nnz_l = p->plan.lp[p->plan.m];
  if (nnz_u) *nnz_u = p->plan.up[p->plan.m];
  if (n_levels) *n_levels = p->dph.n_levels;
  return GS_OK;
}

GS_API gs_status gs_plan_get(const gs_plan* p, int32_t* q, int32_t* pivot_row, int32_t* lp, int32_t* li,
                             int32_t* up, int32_t* ui, int32_t* j_rowptr, int32_t* j_colidx) {
  if (!p) return GS_INVALID_ARGUMENT;
  auto cp = [](int32_t* dst, const std::vector<int>& v) { if (dst) std::copy(v.begin(), v.end(), dst); };
  cp(q, p->plan.q); cp(pivot_row, p->plan.pivot_row); cp(lp, p->plan.lp); cp(li, p->plan.li);
  cp(up, p->plan.up); cp(ui, p->plan.ui); cp(j_rowptr, p->jp.rp); cp(j_colidx, p->jp.ci);
  return GS_OK;
}

GS_API void gs_plan_free(gs_plan* p) { delete p; }

GS_API gs_status gs_plan_selfcheck(const gs_plan* p, double* rel_residual, char* why) {
  if (!p || !rel_residual) return GS_INVALID_ARGUMENT;
  const int n = p->net.n, m = p->jp.m;
  std::vector<double> pp(n), qq(n), jv(p->jp.rp[m]), f(m), x(m);
  host_injections(p->y, p->spec.vm.data(), p->spec.va.data(), pp.data(), qq.data());
  host_jacobian(p->y, p->jp, p->spec.vm.data(), p->spec.va.data(), pp.data(), qq.data(), jv.data());
  for (int i = 0; i < n; i++) {
    if (p->jp.ppos[i] >= 0) f[p->jp.ppos[i]] = p->spec.pspec[i] - pp[i];
    if (p->jp.qpos[i] >= 0) f[p->jp.qpos[i]] = p->spec.qspec[i] - qq[i];
  }
  std::string w;
  // With a symmetric Y-bus, Y cell k's lane forms the Jacobian values of cell y_tr[k] (its transpose) from its own Y
  // value, the two magnitudes swapped and the sine of the opposite angle difference, exactly as the mismatch kernel
  // does: they equal that cell's entries bit for bit (compared with host_jacobian, which forms every cell's own).
  const DevicePlanHost& dt = p->dph;
  const double *vm = p->spec.vm.data(), *va = p->spec.va.data();
  for (size_t k = 0; k < dt.y_tr.size(); k++) {
    const int row = (int)(std::upper_bound(p->y.rp.begin(), p->y.rp.end(), (int)k) - p->y.rp.begin()) - 1;
    const int col = p->y.ci[k];
    if (row == col) continue;
    double sn, cs;
    gs_fdlibm::sincos(va[row] - va[col], &sn, &cs);
    const double g = p->y.g[k], b = p->y.b[k];
    const double vr = vm[col], vc = vm[row], st = sn == 0.0 ? va[col] - va[row] : 0.0 - sn;
    const double v[4] = {vr * vc * (g * st - b * cs), vr * (g * cs + b * st), -vr * vc * (g * cs + b * st),
                         vr * (g * st - b * cs)};
    for (int e = 0; e < 4; e++) {
      const int slot = p->jp.slots[4 * (size_t)dt.y_tr[k] + e];
      if (slot >= 0 && std::memcmp(&v[e], &jv[slot], sizeof(double)) != 0) {
        if (why) std::snprintf(why, 128, "cell %d's lane forms a transposed Jacobian value unlike the cell's own", (int)k);
        return GS_INVALID_ARGUMENT;
      }
    }
  }
  if (!emulate_device_lu(p->dph, p->jp.rp[m], jv.data(), f.data(), x.data(), w)) {
    if (why) std::snprintf(why, 128, "%s", w.c_str());
    return GS_INVALID_ARGUMENT;
  }
  // the level-parallel decomposition of the same streams, against the padding upload_plan gives them
  auto ring_up = [](int v) { return (v + kStreamPad - 1) / kStreamPad * kStreamPad; };
  const DevicePlanHost& d = p->dph;
  if (!check_level_schedule(d, kStreamPad, ring_up(d.n_rs) + kRsPadRings * kStreamPad, ring_up(d.fs_steps) + 2 * kStreamPad,
                            ring_up(d.bs_steps) + 2 * kStreamPad, w)) {
    if (why) std::snprintf(why, 128, "level-parallel schedule: %s", w.c_str());
    return GS_INVALID_ARGUMENT;
  }
  double rmax = 0, fmax = 0;
  for (int i = 0; i < m; i++) {
    double s = 0;
    for (int k = p->jp.rp[i]; k < p->jp.rp[i + 1]; k++) s += jv[k] * x[p->jp.ci[k]];
    rmax = std::max(rmax, std::fabs(s - f[i]));
    fmax = std::max(fmax, std::fabs(f[i]));
  }
  *rel_residual = fmax > 0 ? rmax / fmax : rmax;
  if (why) why[0] = 0;
  return GS_OK;
}

GS_API gs_status gs_network_launch_shape(const gs_net* net, const gs_nr_options* opts, int32_t* lu_fp32,
                                         int32_t* x_in_lds, int32_t* waves_per_group, int64_t* lds_bytes) {
  if (!net) return GS_INVALID_ARGUMENT;
  return guarded(net->ctx, [&]() -> gs_status {
    const gs_nr_options o = normalise(opts);
    const BatchRun r = plan_run(net->dev.h, o, false);
    if (lu_fp32) *lu_fp32 = r.fp32;
    if (x_in_lds) *x_in_lds = r.x_lds;
    if (waves_per_group) *waves_per_group = r.group_waves > 0 ? r.group_waves : r.pair ? 2 : 1;
    if (lds_bytes) *lds_bytes = (int64_t)r.lds_bytes;
    return GS_OK;
  });
}

GS_API gs_status gs_last_timing(const gs_net* net, double* total_ms, double* kernel_ms, int32_t* launches) {
  if (!net) return GS_INVALID_ARGUMENT;
  if (total_ms) *total_ms = net->last_total_ms;
  if (kernel_ms) *kernel_ms = net->last_kernel_ms;
  if (launches) *launches = net->last_launches;
  return GS_OK;
}

GS_API gs_status gs_device_sincos(gs_ctx* ctx, int64_t n, const double* x, double* s, double* c) {
  return guarded(ctx, [&]() -> gs_status {
    if (n < 0 || (n > 0 && (!x || !s || !c))) fail(GS_INVALID_ARGUMENT, "null argument");
    if (n == 0) return GS_OK;
    DevMem dx, ds, dc;
    dx.ensure(8 * (size_t)n); ds.ensure(8 * (size_t)n); dc.ensure(8 * (size_t)n);
    check(hipMemcpyAsync(dx.p, x, 8 * (size_t)n, hipMemcpyHostToDevice, ctx->stream), "sincos in");
    check(launch_sincos(dx.as<double>(), ds.as<double>(), dc.as<double>(), n, ctx->stream), "sincos kernel");
    check(hipMemcpyAsync(s, ds.p, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "sincos s");
    check(hipMemcpyAsync(c, dc.p, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "sincos c");
    check(hipStreamSynchronize(ctx->stream), "sincos sync");
    return GS_OK;
  });
}

GS_API void gs_host_sincos(int64_t n, const double* x, double* s, double* c) {
  for (int64_t i = 0; i < n; i++) gs_fdlibm::sincos(x[i], &s[i], &c[i]);
}

GS_API gs_status gs_batch_create(gs_ctx* ctx, int32_t n_groups, gs_net* const* nets, const int32_t* counts,
                                 const gs_instance* const* inst, const gs_nr_options* opts, int32_t use_graph,
                                 gs_batch** out) {
  return guarded(ctx, [&]() -> gs_status {
    if (!out || n_groups <= 0 || !nets || !counts || !inst) fail(GS_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto b = std::make_unique<gs_batch>();
    b->ctx = ctx;
    b->use_graph = use_graph != 0;
    check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "event");
    check(hipEventCreate(&b->e0), "event");
    check(hipEventCreate(&b->e1), "event");
    for (int gi = 0; gi < n_groups; gi++) {
      gs_net* net = nets[gi];
      if (!net || net->ctx != ctx) fail(GS_INVALID_ARGUMENT, "group " + std::to_string(gi) + ": network");
      if (counts[gi] <= 0 || !inst[gi]) fail(GS_INVALID_ARGUMENT, "group " + std::to_string(gi) + ": instances");
      gs_nr_options o = normalise(opts ? &opts[gi] : nullptr);
      if ((o.flat_start != 0) != net->flat_start) fail(GS_INVALID_ARGUMENT, "flat_start differs from the upload");
      if (o.flags & (GS_FLAG_Q_LIMITS | GS_FLAG_DC_RETRY))
        fail(GS_INVALID_ARGUMENT, "group " + std::to_string(gi) +
                                      ": reactive limits and the DC-start retry run through gs_nr_solve_batch");
      check_instances(net->net.nb, counts[gi], inst[gi], "group " + std::to_string(gi) + ", ");
      auto g = std::make_unique<gs_batch::Group>();
      g->net = net;
      g->count = counts[gi];
      g->inst.assign(inst[gi], inst[gi] + counts[gi]);
      // the outage lists are the caller's: the group keeps its own copies (the host reads them after the call)
      g->outage_lists.resize(g->count);
      for (int i = 0; i < g->count; i++) {
        gs_instance& in = g->inst[i];
        if (in.n_outages > 0) g->outage_lists[i].assign(in.outages, in.outages + in.n_outages);
        in.outages = in.n_outages > 0 ? g->outage_lists[i].data() : nullptr;
      }
      g->opts = o;
      check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "stream");
      check(hipEventCreateWithFlags(&g->joined, hipEventDisableTiming), "event");
      if (o.schedule == GS_SCHEDULE_LEVELS) {
        g->wide = std::make_unique<WideGraph>();
        g->wide->ctl.ensure(sizeof(WideCtl));
      }
      g->pr = prepare_instances(ctx, net->dev, g->inst_buf, g->desc_buf, g->state_buf, g->aux_buf, g->stg, g->count,
                                g->inst.data(), &net->y, &net->net, net->spec, o, false, nullptr, nullptr, nullptr,
                                ctx->stream);
      for (auto& in : g->inst) in.vm0 = in.va0 = nullptr;  // start points now live in HBM
      b->groups.push_back(std::move(g));
    }
    check(hipStreamSynchronize(ctx->stream), "batch upload");
    *out = b.release();
    return GS_OK;
  });
}

GS_API gs_status gs_batch_run(gs_batch* batch, double* device_ms) {
  if (!batch) return GS_INVALID_ARGUMENT;
  return guarded(batch->ctx, [&]() { return batch_run(batch, device_ms); });
}

GS_API gs_status gs_batch_results(gs_batch* batch, int32_t group, gs_nr_result* results, double* vm_out,
                                  double* va_out) {
  if (!batch) return GS_INVALID_ARGUMENT;
  return guarded(batch->ctx, [&]() -> gs_status {
    if (group < 0 || group >= (int)batch->groups.size() || !results) fail(GS_INVALID_ARGUMENT, "group");
    if (!batch->ran) fail(GS_INVALID_ARGUMENT, "the batch has not run");
    const auto& g = *batch->groups[group];
    std::copy(g.results.begin(), g.results.end(), results);
    // voltages of the last run: the device's rows (the instance blocks are untouched until the next run), then the
    // rows of the instances the host finished
    const int n = g.net->net.n;
    const size_t row = 8 * (size_t)n;
    const char* b = g.inst_buf.as<char>();
    hipStream_t s = batch->ctx->stream;
    if (vm_out) check(hipMemcpy2DAsync(vm_out, row, b + g.pr.L.vm, g.pr.L.stride, row, (size_t)g.count,
                                       hipMemcpyDeviceToHost, s), "batch vm");
    if (va_out) check(hipMemcpy2DAsync(va_out, row, b + g.pr.L.va, g.pr.L.stride, row, (size_t)g.count,
                                       hipMemcpyDeviceToHost, s), "batch va");
    check(hipStreamSynchronize(s), "batch results");
    for (const HostFinal& f : g.finals) {
      if (vm_out) std::copy(f.vm.begin(), f.vm.end(), vm_out + (size_t)n * f.i);
      if (va_out) std::copy(f.va.begin(), f.va.end(), va_out + (size_t)n * f.i);
    }
    return GS_OK;
  });
}

GS_API void gs_batch_free(gs_batch* batch) {
  if (!batch) return;
  std::lock_guard<std::mutex> lock(batch->ctx->mu);
  delete batch;
}

GS_API gs_status gs_nr_autotune(gs_net* net, int32_t count, const gs_instance* sample, gs_nr_options* opts,
                                double* best_ms) {
  if (!net) return GS_INVALID_ARGUMENT;
  return guarded(net->ctx, [&]() -> gs_status {
    if (count <= 0 || !sample || !opts) fail(GS_INVALID_ARGUMENT, "null argument");
    gs_nr_options base = normalise(opts);
    if ((base.flat_start != 0) != net->flat_start) fail(GS_INVALID_ARGUMENT, "flat_start differs from the upload");
    if (base.flags & GS_FLAG_Q_LIMITS) fail(GS_INVALID_ARGUMENT, "the autotuner times single runs (no limit passes)");
    // verdicts of the FP64 LU: the reference every candidate must reproduce
    const size_t nv = (size_t)count * net->net.n;
    std::vector<gs_nr_result> ref(count), res(count);
    std::vector<double> ref_vm(nv), ref_va(nv), vm(nv), va(nv);
    gs_nr_options o = base;
    o.lu_precision = GS_LU_FP64;
    o.refine_steps = 0;
    o.schedule = 0;
    run_batch(net, count, sample, o, ref.data(), ref_vm.data(), ref_va.data(), false);
    std::vector<Candidate> cands;
    for (int32_t flags : {0, (int32_t)GS_FLAG_SINGLE_WAVE, (int32_t)GS_FLAG_X_GLOBAL})
      for (int32_t sched : {0, 512, 256})
        // FP32 LU always with one or two FP64 refinement steps: the refined FP64 residual is what sets the answer
        for (Candidate c : {Candidate{GS_LU_FP32, 1, 0, 0}, Candidate{GS_LU_FP32, 2, 0, 0},
                            Candidate{GS_LU_FP64, 0, 0, 0}}) {
          if (sched != 0 && sched >= count) continue;
          if (flags == GS_FLAG_SINGLE_WAVE) {  // only where the default would pair the instances
            gs_nr_options t = base;
            t.lu_precision = c.lu;
            t.flags = 0;
            if (!plan_run(net->dev.h, t, false).pair) continue;
          }
          cands.push_back({c.lu, c.refine, sched, flags});
        }
    // asked for: a work-group of 2, 4 or 8 waves per instance
    if (base.schedule == GS_SCHEDULE_GROUP || base.schedule == GS_SCHEDULE_LEVELS)
      for (int32_t w : {2, 4, 8}) {
        cands.push_back({GS_LU_FP64, 0, GS_SCHEDULE_GROUP, GS_GROUP_WAVES(w)});
        cands.push_back({GS_LU_FP32, 1, GS_SCHEDULE_GROUP, GS_GROUP_WAVES(w)});
      }
    if (base.schedule == GS_SCHEDULE_LEVELS) {   // asked for: the level-parallel schedule too
      cands.push_back({GS_LU_FP64, 0, GS_SCHEDULE_LEVELS, 0});
      cands.push_back({GS_LU_FP32, 1, GS_SCHEDULE_LEVELS, 0});
    }
    double best = 1e300;
    gs_nr_options chosen = base;
    for (const Candidate& c : cands) {
      o = base;
      o.lu_precision = c.lu;
      o.refine_steps = c.refine;
      o.schedule = c.schedule;
      o.flags = c.flags;
      run_batch(net, count, sample, o, res.data(), vm.data(), va.data(), false);  // warm-up, and the states
      const double kernel_ms = (run_batch(net, count, sample, o, res.data(), nullptr, nullptr, false),
                                net->last_kernel_ms);
      // the parity bar: same verdicts, converged instances within the tolerance and within kAutotuneDv of the FP64
      // LU's voltages
      bool ok = true;
      for (int i = 0; i < count && ok; i++) {
        ok = res[i].status == ref[i].status &&
             (res[i].status != GS_PF_CONVERGED || res[i].max_mismatch < base.tolerance);
        if (ok && res[i].status == GS_PF_CONVERGED)
          for (size_t b = (size_t)i * net->net.n; b < (size_t)(i + 1) * net->net.n && ok; b++)
            ok = std::fabs(vm[b] - ref_vm[b]) <= kAutotuneDv && std::fabs(va[b] - ref_va[b]) <= kAutotuneDv;
      }
      if (ok && kernel_ms < best) {
        best = kernel_ms;
        chosen = o;
      }
    }
    chosen.flags |= base.flags & GS_FLAG_DC_RETRY;   // the caller's retry choice is not the autotuner's to make
    *opts = chosen;
    if (best_ms) *best_ms = best;
    return GS_OK;
  });
}

}  // extern "C"
